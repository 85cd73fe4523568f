// C ABI of libtga: error handling, device checks, level-1/2 entry points, replica helpers.
//
// Every entry point catches tga::Error and returns its status; nothing aborts.  The messages
// of the reference's panics are reproduced where the reference panics (correlate.rs:103,
// acquisition.rs:131-134, observation.rs:141-146, correlate.rs:138,153).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include "internal.h"
#include "plan.h"
#include "search.h"

namespace tga {

namespace {
thread_local std::string g_last_error;

template <class F>
int guard(F&& fn) {
  try {
    g_last_error.clear();
    fn();
    return TGA_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return TGA_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return TGA_ERR_INVALID_ARGUMENT;
  } catch (...) {
    g_last_error = "unknown error";
    return TGA_ERR_INVALID_ARGUMENT;
  }
}

void need(const void* p, const char* name) {
  if (p == nullptr) fail(TGA_ERR_INVALID_ARGUMENT, std::string(name) + " must not be null");
}

// A per-thread stream so concurrent level-1 callers (rayon workers) do not serialise on one.
struct ThreadStream {
  hipStream_t s = nullptr;
  int dev = -1;
  ~ThreadStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};
thread_local ThreadStream t_stream;

hipStream_t thread_stream() {
  const int dev = current_device_checked();
  if (t_stream.s == nullptr || t_stream.dev != dev) {
    if (t_stream.s) (void)hipStreamDestroy(t_stream.s);
    t_stream.s = nullptr;
    TGA_HIP(hipStreamCreateWithFlags(&t_stream.s, hipStreamNonBlocking));
    t_stream.dev = dev;
  }
  return t_stream.s;
}

template <class T>
void upload(DevBuf& b, const T* h, size_t count, hipStream_t st) {
  b.alloc(count * sizeof(T));
  TGA_HIP(hipMemcpyAsync(b.p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
}

}  // namespace

[[noreturn]] void fail(int code, const std::string& msg) { throw Error(code, msg); }

int plan_guard_impl(void (*fn)(void*), void* ctx) {
  return guard([&] { fn(ctx); });
}

void check_hip(hipError_t e, const char* what, const char* file, int line) {
  if (e == hipSuccess) return;
  // a failed call also leaves its code as the thread's last error: take it, so the next
  // launch check of a later, unrelated call does not report it again
  (void)hipGetLastError();
  const char* fname = std::strrchr(file, '/') ? std::strrchr(file, '/') + 1 : file;
  std::string msg = std::string(what) + ": " + hipGetErrorString(e) + " (" + fname + ":" +
                    std::to_string(line) + ")";
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    fail(TGA_ERR_OUT_OF_MEMORY, msg);
  if (e == hipErrorNoDevice) fail(TGA_ERR_NO_DEVICE, msg);
  fail(TGA_ERR_DEVICE, msg);
}

int current_device_checked() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    fail(TGA_ERR_NO_DEVICE, "no HIP device is visible");
  int dev = 0;
  TGA_HIP(hipGetDevice(&dev));
  static std::mutex mu;
  static std::vector<int> checked;  // devices known to be gfx950
  std::lock_guard<std::mutex> lk(mu);
  if (std::find(checked.begin(), checked.end(), dev) == checked.end()) {
    hipDeviceProp_t prop;
    TGA_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      fail(TGA_ERR_UNSUPPORTED, std::string("libtga is built for gfx950; device ") +
                                    std::to_string(dev) + " is " + prop.gcnArchName);
    checked.push_back(dev);
  }
  return dev;
}

void DevBuf::alloc(size_t b) {
  release();
  if (b == 0) return;
  TGA_HIP(hipMalloc(&p, b));
  bytes = b;
}

void DevBuf::release() {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
}

// N-point correlation of `n_bin` wiped rows against the two replica spectra rep[2][N]
// (FFT of the time-domain replicas, i.e. the conjugate of what correlate_code receives),
// returning the selected peak.  Implements correlate.rs:107-183 on the generic engine.
static tga_peak correlate_generic(const float* d_sig, const double* d_pp, const double* d_fc,
                                  int n_bin, size_t N, size_t P, double fs, const float2* d_rep,
                                  hipStream_t st) {
  const FftPlan* plan = fft_plan(N);
  DevBuf X(static_cast<size_t>(n_bin) * N * sizeof(float2));
  DevBuf S(static_cast<size_t>(n_bin) * N * sizeof(float2));
  launch_wipe_signal(d_sig, d_pp, d_fc, n_bin, static_cast<int>(N), X.as<float2>(), st);
  float2* Xf = fft_forward(plan, X.as<float2>(), S.as<float2>(), n_bin, st);
  const long long n_rows = 2LL * n_bin;
  DevBuf Z(static_cast<size_t>(n_rows) * N * sizeof(float2));
  DevBuf Zs(static_cast<size_t>(n_rows) * N * sizeof(float2));
  launch_mul_rows(Xf, d_rep, static_cast<int>(N), 1, n_bin, false, 0, n_rows, Z.as<float2>(), st);
  float2* Zf = fft_forward(plan, Z.as<float2>(), Zs.as<float2>(), n_rows, st);
  DevBuf rows(static_cast<size_t>(n_rows) * sizeof(RowOut));
  const int cw = static_cast<int>(std::ceil(fs / 1.023e6));
  launch_row_reduce(Zf, static_cast<int>(N), static_cast<int>(P), cw, n_rows, rows.as<RowOut>(),
                    st);
  PeakSlot slot{0, 0, 0, 0};
  DevBuf dslot(sizeof(PeakSlot));
  TGA_HIP(hipMemcpyAsync(dslot.p, &slot, sizeof slot, hipMemcpyHostToDevice, st));
  DevBuf dpk(sizeof(tga_peak));
  DevBuf dnan(sizeof(int));
  TGA_HIP(hipMemsetAsync(dnan.p, 0, sizeof(int), st));
  const float scale = 1.0f / std::sqrt(static_cast<float>(N));
  launch_select(rows.as<RowOut>(), dslot.as<PeakSlot>(), 1, 1, n_bin, 1.0f, scale,
                dpk.as<tga_peak>(), dnan.as<int>(), st);
  tga_peak pk;
  int nan = 0;
  TGA_HIP(hipMemcpyAsync(&pk, dpk.p, sizeof pk, hipMemcpyDeviceToHost, st));
  TGA_HIP(hipMemcpyAsync(&nan, dnan.p, sizeof nan, hipMemcpyDeviceToHost, st));
  TGA_HIP(hipStreamSynchronize(st));
  if (nan)
    fail(TGA_ERR_NAN, "called `Option::unwrap()` on a `None` value (NaN in the correlation)");
  return pk;
}

// Time-domain replica variants -> spectra on the GPU: out[2][N] = FFT(code), FFT(alt).
static void replica_spectra(const double* code, size_t N, size_t P, DevBuf& out,
                            hipStream_t st) {
  std::vector<float2> h(2 * N);
  const size_t split = std::min(P, N);
  for (size_t i = 0; i < N; ++i) {
    const float c = static_cast<float>(code[i]);
    h[i] = make_float2(c, 0.f);
    h[N + i] = make_float2(i < split ? c : -c, 0.f);
  }
  DevBuf a;
  upload(a, h.data(), h.size(), st);
  DevBuf s(2 * N * sizeof(float2));
  float2* r = fft_forward(fft_plan(N), a.as<float2>(), s.as<float2>(), 2, st);
  TGA_HIP(hipStreamSynchronize(st));
  out = std::move(r == a.as<float2>() ? a : s);
}

}  // namespace tga

using namespace tga;

extern "C" {

const char* tga_last_error(void) { return g_last_error.c_str(); }

const char* tga_version(void) { return "tga 0.1.0 (gfx950)"; }

int tga_device_count(int* count) {
  return guard([&] {
    need(count, "count");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
  });
}

int tga_code_length(int32_t system, size_t* chips) {
  return guard([&] {
    need(chips, "chips");
    *chips = system_info(system).chips;
  });
}

int tga_code_chips(int32_t system, int32_t prn, int8_t* out, size_t cap) {
  return guard([&] {
    need(out, "out");
    std::vector<int8_t> c = code_chips(system, prn);
    if (cap < c.size()) fail(TGA_ERR_INVALID_ARGUMENT, "output capacity below code length");
    std::memcpy(out, c.data(), c.size());
  });
}

int tga_code_resampled(int32_t system, int32_t prn, double samples_per_code, size_t num_samples,
                       double* out) {
  return guard([&] {
    need(out, "out");
    if (!(samples_per_code > 0)) fail(TGA_ERR_INVALID_ARGUMENT, "samples_per_code must be > 0");
    resample(code_chips(system, prn), samples_per_code, num_samples, out);
  });
}

int tga_code_spectra(const double* code, size_t n, size_t period, float* cc, float* cc_alt) {
  return guard([&] {
    need(code, "code");
    need(cc, "code_complex");
    need(cc_alt, "code_complex_alt");
    if (n == 0) fail(TGA_ERR_INVALID_ARGUMENT, "code length must be positive");
    hipStream_t st = thread_stream();
    DevBuf spec;
    replica_spectra(code, n, period, spec, st);
    std::vector<float2> h(2 * n);
    TGA_HIP(hipMemcpyAsync(h.data(), spec.p, 2 * n * sizeof(float2), hipMemcpyDeviceToHost, st));
    TGA_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < n; ++i) {  // conjugate (correlate.rs:65-67)
      cc[2 * i] = h[i].x;
      cc[2 * i + 1] = -h[i].y;
      cc_alt[2 * i] = h[n + i].x;
      cc_alt[2 * i + 1] = -h[n + i].y;
    }
  });
}

int tga_correlate_code(const float* signal, const double* phasepoints, const float* code_complex,
                       const float* code_complex_alt, const double* fc, size_t n_fc,
                       size_t num_samples, size_t code_period_samples, double sampling_freq,
                       tga_peak* out) {
  return guard([&] {
    need(out, "out");
    if (n_fc == 0 || fc == nullptr)
      fail(TGA_ERR_EMPTY_BINS, "frequency bin list must be non-empty");
    need(signal, "signal_f32");
    need(phasepoints, "phasepoints");
    need(code_complex, "code_complex");
    need(code_complex_alt, "code_complex_alt");
    if (num_samples == 0) fail(TGA_ERR_INVALID_ARGUMENT, "num_samples must be positive");
    if (code_period_samples == 0)
      fail(TGA_ERR_INVALID_ARGUMENT, "code_period_samples must be positive");
    if (n_fc > (1u << 20)) fail(TGA_ERR_INVALID_ARGUMENT, "too many frequency bins");
    const size_t N = num_samples;
    hipStream_t st = thread_stream();
    DevBuf dsig, dpp, dfc, drep;
    upload(dsig, signal, N, st);
    upload(dpp, phasepoints, N, st);
    upload(dfc, fc, n_fc, st);
    // rep = conj(given): the kernels correlate conj(X) . rep, whose FFT has |.| equal to the
    // reference's IFFT(X . code_complex).
    std::vector<float2> h(2 * N);
    for (size_t i = 0; i < N; ++i) {
      h[i] = make_float2(code_complex[2 * i], -code_complex[2 * i + 1]);
      h[N + i] = make_float2(code_complex_alt[2 * i], -code_complex_alt[2 * i + 1]);
    }
    upload(drep, h.data(), h.size(), st);
    *out = correlate_generic(dsig.as<float>(), dpp.as<double>(), dfc.as<double>(),
                             static_cast<int>(n_fc), N, code_period_samples, sampling_freq,
                             drep.as<float2>(), st);
  });
}

int tga_acquire_single(int32_t system, const float* signal, const double* phasepoints,
                       size_t n_samples, double sampling_freq, double center_freq,
                       double search_band, int32_t prn, size_t samples_per_period,
                       tga_acq_result* out) {
  return guard([&] {
    need(out, "out");
    need(signal, "signal_f32");
    need(phasepoints, "phasepoints");
    if (n_samples == 0) fail(TGA_ERR_INVALID_ARGUMENT, "signal must not be empty");
    if (samples_per_period == 0)
      fail(TGA_ERR_INVALID_ARGUMENT, "samples_per_period must be positive");
    check_prn(system, prn);
    // 300 Hz grid (acquisition.rs:189-198)
    const double nb_f = 2.0 * search_band / 300.0;
    const long long nb = static_cast<long long>(std::round(nb_f)) + 1;
    if (nb < 1 || nb > (1 << 20)) fail(TGA_ERR_INVALID_ARGUMENT, "invalid search band");
    std::vector<double> fc;
    if (nb == 1) {
      fc.push_back(center_freq);
    } else {
      for (long long i = 0; i < nb; ++i)
        fc.push_back(center_freq - search_band +
                     static_cast<double>(i) * (2.0 * search_band) / (static_cast<double>(nb) - 1.0));
    }
    const size_t N = n_samples;
    std::vector<double> code(N);
    resample(code_chips(system, prn), static_cast<double>(samples_per_period), N, code.data());
    hipStream_t st = thread_stream();
    DevBuf rep;
    replica_spectra(code.data(), N, samples_per_period, rep, st);
    DevBuf dsig, dpp, dfc;
    upload(dsig, signal, N, st);
    upload(dpp, phasepoints, N, st);
    upload(dfc, fc.data(), fc.size(), st);
    const tga_peak pk = correlate_generic(dsig.as<float>(), dpp.as<double>(), dfc.as<double>(),
                                          static_cast<int>(fc.size()), N, samples_per_period,
                                          sampling_freq, rep.as<float2>(), st);
    out->prn = prn;
    out->pad = 0;
    out->signal_strength = static_cast<double>(pk.best_peak);
    out->second_peak = static_cast<double>(pk.second_peak);
    out->codephase_frac = static_cast<double>(pk.best_codephase % samples_per_period) /
                          static_cast<double>(samples_per_period);
    out->frequency = fc[pk.best_freq_idx] - center_freq;
  });
}

int tga_estimate_cn0(double r_a, int32_t table, double* cn0, int32_t* valid) {
  return guard([&] {
    need(cn0, "cn0");
    need(valid, "valid");
    double c = 0;
    const bool ok = estimate_cn0(r_a, table, &c);
    *valid = ok ? 1 : 0;
    *cn0 = ok ? c : 0.0;
  });
}

int tga_acr_cn0(double v_m, double v_s, int32_t table, double* cn0, int32_t* valid) {
  return guard([&] {
    need(cn0, "cn0");
    need(valid, "valid");
    *valid = 0;
    *cn0 = 0.0;
    if (v_s > 0.0 && v_m > v_s) {  // acquisition.rs:346-347
      const double r = (v_m / v_s) * (v_m / v_s);
      double c = 0;
      if (estimate_cn0(r, table, &c)) {
        *valid = 1;
        *cn0 = c;
      }
    }
  });
}

}  // extern "C"
