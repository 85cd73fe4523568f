// Level-1 correlation for callers that cannot allocate device memory (code under hipGraph
// capture, tight serving loops, harnesses that forbid hipMalloc): tga_correlate_code_noalloc.
//
// It runs the same kernels as tga_correlate_code (wipe-off, generic FFT passes, conj(X) . R
// rows, row reduction, selection: correlate.rs:107-183) on device-resident inputs, and keeps its
// working set in `scratch`, a file-scope device array of kScratchBytes that every GPU holds until
// the process ends.  The Doppler bins are processed in chunks sized to the array, so every
// reference shape up to N = 2 x 163680 (B1C / L1C, 20 ms) fits.  Layout of the array:
//   twiddles   fft_twiddle_count(N) float2 of the N-point transform (<= kTwCap)
//   rows       RowOut[2 n_fc]                 (n_fc <= kMaxBins)
//   slot       PeakSlot                       (the one output slot of the selection)
//   nan        int                            (when the caller passes no NaN flag)
//   rep        float2[2][N]                   conj(code_complex), conj(code_complex_alt)
//   X, S       float2[B][N] each              forward spectra of a chunk of B bins + FFT scratch
//   Z, Zs      float2[2B][N] each             correlation rows of the chunk + FFT scratch
// Calls that overlap on one GPU share the array and race; one call at a time per GPU.
#include <mutex>

#include "internal.h"
#include "plan.h"
#include "search.h"

namespace {

constexpr size_t kScratchBytes = 22u << 20;  // 22 MiB per GPU
constexpr size_t kMaxN = 327360;             // 2 x 163680
constexpr int kMaxBins = 4096;
constexpr int kTwCap = 4096;                 // float2 twiddles (2 sqrt(kMaxN) + 2 < 1200)
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace tga {
namespace {

std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on device `dev` (current), looked up once per device: each GPU
// holds its own copy of the array at its own address.
unsigned char* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) fail(TGA_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    TGA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch[dev];
}

constexpr size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// rep[0][i] = conj(cc[i]), rep[1][i] = conj(ca[i]) (the kernels correlate conj(X) . rep,
// whose transform has |.| equal to the reference's IFFT(X . code_complex)); clears the NaN
// flag and writes the selection's output slot.
__global__ __launch_bounds__(256) void k_noalloc_prep(const float2* __restrict__ cc,
                                                      const float2* __restrict__ ca, long long n,
                                                      float2* __restrict__ rep,
                                                      PeakSlot* __restrict__ slot,
                                                      int* __restrict__ nan) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i == 0) {
    *slot = PeakSlot{0, 0, 0, 0};
    *nan = 0;
  }
  if (i < n) {
    const float2 a = cc[i], b = ca[i];
    rep[i] = make_float2(a.x, -a.y);
    rep[n + i] = make_float2(b.x, -b.y);
  }
}

void correlate_noalloc(const float* d_sig, const double* d_pp, const float2* d_cc,
                       const float2* d_ca, const double* d_fc, size_t n_fc, size_t N, size_t P,
                       double fs, tga_peak* d_out, int* d_nan, hipStream_t st) {
  int radix[64];
  const int n_radix = fft_codelet_passes(N, radix, 64);
  if (n_radix < 0)
    fail(TGA_ERR_UNSUPPORTED, "num_samples " + std::to_string(N) +
                                  " has a prime factor without a register codelet (only 2, 3, "
                                  "5, 7, 11, 13, 31 are served without allocation)");
  if (fft_twiddle_count(N) > kTwCap) fail(TGA_ERR_UNSUPPORTED, "twiddle table exceeds its slot");
  const int dev = current_device_checked();
  unsigned char* base = scratch_on(dev);
  // carve the array (every offset 256-byte aligned)
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += align256(bytes);
    return p;
  };
  float2* tw = reinterpret_cast<float2*>(take(kTwCap * sizeof(float2)));
  RowOut* rows = reinterpret_cast<RowOut*>(take(2 * kMaxBins * sizeof(RowOut)));
  PeakSlot* slot = reinterpret_cast<PeakSlot*>(take(sizeof(PeakSlot)));
  int* own_nan = reinterpret_cast<int*>(take(sizeof(int)));
  float2* rep = reinterpret_cast<float2*>(take(2 * N * sizeof(float2)));
  const size_t left = kScratchBytes - off;
  const size_t per_bin = 6 * align256(N * sizeof(float2));  // X, S, and two rows each of Z, Zs
  const size_t B = std::min<size_t>(n_fc, left / per_bin);
  if (B == 0) fail(TGA_ERR_UNSUPPORTED, "num_samples too large for the scratch array");
  float2* X = reinterpret_cast<float2*>(take(B * N * sizeof(float2)));
  float2* S = reinterpret_cast<float2*>(take(B * N * sizeof(float2)));
  float2* Z = reinterpret_cast<float2*>(take(2 * B * N * sizeof(float2)));
  float2* Zs = reinterpret_cast<float2*>(take(2 * B * N * sizeof(float2)));
  if (off > kScratchBytes) fail(TGA_ERR_UNSUPPORTED, "scratch layout exceeds the array");
  int* nan = d_nan ? d_nan : own_nan;

  const long long n = static_cast<long long>(N);
  hipLaunchKernelGGL(k_noalloc_prep, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256),
                     0, st, d_cc, d_ca, n, rep, slot, nan);
  TGA_HIP(hipGetLastError());
  launch_fft_twiddles(N, tw, st);
  const int cw = static_cast<int>(std::ceil(fs / 1.023e6));
  for (size_t b0 = 0; b0 < n_fc; b0 += B) {
    const int nb = static_cast<int>(std::min(B, n_fc - b0));
    launch_wipe_signal(d_sig, d_pp, d_fc + b0, nb, static_cast<int>(N), X, st);
    float2* Xf = fft_forward_tables(N, radix, n_radix, tw, X, S, nb, st);
    const long long n_rows = 2LL * nb;
    launch_mul_rows(Xf, rep, static_cast<int>(N), 1, nb, false, 0, n_rows, Z, st);
    float2* Zf = fft_forward_tables(N, radix, n_radix, tw, Z, Zs, n_rows, st);
    launch_row_reduce(Zf, static_cast<int>(N), static_cast<int>(P), cw, n_rows, rows + 2 * b0,
                      st);
  }
  const float scale = 1.0f / std::sqrt(static_cast<float>(N));
  launch_select(rows, slot, 1, 1, static_cast<int>(n_fc), 1.0f, scale, d_out, nan, st);
}

}  // namespace
}  // namespace tga

extern "C" int tga_correlate_code_noalloc(const float* d_signal, const double* d_phasepoints,
                                          const float* d_code_complex,
                                          const float* d_code_complex_alt, const double* d_fc,
                                          size_t n_fc, size_t num_samples,
                                          size_t code_period_samples, double sampling_freq,
                                          tga_peak* d_out, int32_t* d_nan, void* stream) {
  using namespace tga;
  return plan_guard([&] {
    if (!d_out) fail(TGA_ERR_INVALID_ARGUMENT, "out must not be null");
    if (n_fc == 0 || d_fc == nullptr)
      fail(TGA_ERR_EMPTY_BINS, "frequency bin list must be non-empty");
    if (!d_signal || !d_phasepoints || !d_code_complex || !d_code_complex_alt)
      fail(TGA_ERR_INVALID_ARGUMENT, "null input pointer");
    if (num_samples < 2 || num_samples > kMaxN)
      fail(TGA_ERR_UNSUPPORTED, "num_samples must be in [2, " + std::to_string(kMaxN) +
                                    "] without allocation (got " + std::to_string(num_samples) +
                                    ")");
    if (n_fc > static_cast<size_t>(kMaxBins))
      fail(TGA_ERR_UNSUPPORTED, "at most " + std::to_string(kMaxBins) +
                                    " frequency bins without allocation (got " +
                                    std::to_string(n_fc) + ")");
    if (code_period_samples == 0)
      fail(TGA_ERR_INVALID_ARGUMENT, "code_period_samples must be positive");
    if (!(sampling_freq > 0.0)) fail(TGA_ERR_INVALID_ARGUMENT, "sampling_freq must be positive");
    correlate_noalloc(d_signal, d_phasepoints, reinterpret_cast<const float2*>(d_code_complex),
                      reinterpret_cast<const float2*>(d_code_complex_alt), d_fc, n_fc,
                      num_samples, code_period_samples, sampling_freq, d_out, d_nan,
                      static_cast<hipStream_t>(stream));
  });
}
