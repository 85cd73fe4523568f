// Narrowband RFI report (rfi.rs:83-184) on the GPU, every antenna of an observation at once.
//
// Per antenna: decode + de-mean in f64, variance (dead channel below 1e-9), the f64 FFT of the
// whole record, P[k] = |X[k]|^2 for k < n/2, the median floor sorted[half/2] by an exact
// 8-bit-digit radix select on the f64 bit patterns (non-negative doubles order like their
// bits), local maxima P[k] > P[k-1] && P[k] >= P[k+1] with P[k] > 10^1.5 * floor gathered as
// (k, P[k]) candidates, and the ACF at lags 1, 3, 16.  The host then computes
// 10 log10(P/floor) in f64, sorts candidates (stable, strongest first), clusters at 10 kHz
// and keeps three lines, exactly as detect_lines (rfi.rs:110-137).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"
#include "plan.h"

namespace tga {

namespace {

constexpr int kCandCap = 1 << 16;  // line candidates kept per antenna
constexpr double kMinExcessRatio = 31.622776601683793;  // 10^(15/10): 15 dB (rfi.rs:36)

__device__ __forceinline__ int level_at(const uint8_t* row, int fmt, long long n) {
  if (fmt == TGA_INPUT_1BIT) return ((row[n >> 3] >> (7 - static_cast<int>(n & 7))) & 1) ? 1 : -1;
  const int pair = (row[n >> 2] >> (6 - 2 * static_cast<int>(n & 3))) & 3;
  const int mag = (pair & 1) ? 3 : 1;
  return (pair & 2) ? mag : -mag;
}

// Integer moments per antenna: sum b, and sum b_i b_{i+lag} for lag 1, 3, 16, plus the head
// and tail sums needed to de-mean the products exactly.
struct Moments {
  long long s;        // sum_{i<n} b_i
  long long s2;       // sum b_i^2
  long long prod[3];  // sum_{i<n-lag} b_i b_{i+lag}
  long long head[3];  // sum_{i<lag} b_i
  long long tail[3];  // sum_{i>=n-lag} b_i
};

__global__ __launch_bounds__(256) void k_moments(const uint8_t* __restrict__ packed,
                                                 size_t row_bytes, long long n, int fmt,
                                                 Moments* __restrict__ out) {
  const uint8_t* row = packed + static_cast<size_t>(blockIdx.x) * row_bytes;
  const int lags[3] = {1, 3, 16};
  long long v[11] = {0};
  for (long long i = threadIdx.x; i < n; i += 256) {
    const int b = level_at(row, fmt, i);
    v[0] += b;
    v[1] += b * b;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (i + lags[q] < n) v[2 + q] += b * level_at(row, fmt, i + lags[q]);
      if (i < lags[q]) v[5 + q] += b;
      if (i >= n - lags[q]) v[8 + q] += b;
    }
  }
  __shared__ long long red[4][11];
#pragma unroll
  for (int q = 0; q < 11; ++q)
    for (int off = 32; off >= 1; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 11; ++q) red[threadIdx.x >> 6][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t[11];
    for (int q = 0; q < 11; ++q) t[q] = red[0][q] + red[1][q] + red[2][q] + red[3][q];
    Moments m;
    m.s = t[0];
    m.s2 = t[1];
    for (int q = 0; q < 3; ++q) {
      m.prod[q] = t[2 + q];
      m.head[q] = t[5 + q];
      m.tail[q] = t[8 + q];
    }
    out[blockIdx.x] = m;
  }
}

// z[i] = b_i - mean as complex f64 rows.
__global__ void k_demean(const uint8_t* __restrict__ packed, size_t row_bytes, long long n,
                         int fmt, const double* __restrict__ means, long long total,
                         double2* __restrict__ z) {
  const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const long long a = gid / n;
  const long long i = gid - a * n;
  const double v = static_cast<double>(level_at(packed + a * row_bytes, fmt, i)) - means[a];
  z[gid] = make_double2(v, 0.0);
}

__global__ void k_power(const double2* __restrict__ X, long long n, long long half,
                        long long total, double* __restrict__ p) {
  const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const long long a = gid / half;
  const long long k = gid - a * half;
  const double2 x = X[a * n + k];
  p[gid] = x.x * x.x + x.y * x.y;  // norm_sqr (rfi.rs:103)
}

// Radix select state per antenna: prefix of the selected key, its mask, remaining rank.
struct SelState {
  unsigned long long prefix;
  unsigned long long mask;
  long long rank;
};

__global__ __launch_bounds__(256) void k_hist(const double* __restrict__ p, long long half,
                                              int shift, const SelState* __restrict__ st,
                                              unsigned int* __restrict__ hist) {
  __shared__ unsigned int h[256];
  const int a = blockIdx.y;
  h[threadIdx.x] = 0;
  __syncthreads();
  const SelState s = st[a];
  const unsigned long long* key = reinterpret_cast<const unsigned long long*>(p) + a * half;
  for (long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; i < half;
       i += static_cast<long long>(gridDim.x) * 256) {
    const unsigned long long k = key[i];
    if ((k & s.mask) == s.prefix) atomicAdd(&h[(k >> shift) & 255], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[a * 256 + threadIdx.x], h[threadIdx.x]);
}

__global__ void k_pick(unsigned int* __restrict__ hist, int shift, SelState* __restrict__ st,
                       int n_ant) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_ant) return;
  SelState s = st[a];
  unsigned int* h = hist + a * 256;
  long long r = s.rank;
  int d = 0;
  for (; d < 255; ++d) {
    if (r < h[d]) break;
    r -= h[d];
  }
  s.prefix |= static_cast<unsigned long long>(d) << shift;
  s.mask |= 255ULL << shift;
  s.rank = r;
  st[a] = s;
  for (int q = 0; q < 256; ++q) h[q] = 0;
}

struct Cand {
  long long k;
  double p;
};

__global__ void k_lines(const double* __restrict__ p, long long half, long long total,
                        const SelState* __restrict__ st, Cand* __restrict__ cand,
                        int* __restrict__ count) {
  const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const long long a = gid / half;
  const long long k = gid - a * half;
  if (k < 1 || k >= half - 1) return;  // rfi.rs:111
  const double* pa = p + a * half;
  const double v = pa[k];
  if (!(v > pa[k - 1] && v >= pa[k + 1])) return;
  const double floor = __longlong_as_double(static_cast<long long>(st[a].prefix));
  // prefilter at 15 dB (the host recomputes the exact excess in f64)
  if (!(v > floor * (kMinExcessRatio * 0.999))) return;
  const int slot = atomicAdd(&count[a], 1);
  if (slot < kCandCap) cand[a * kCandCap + slot] = Cand{k, v};
}

inline unsigned blocks_for(long long total, int b) {
  return static_cast<unsigned>((total + b - 1) / b);
}

}  // namespace

// packed: host memory (copied in on the call's stream) or, with on_device, device memory
void rfi_run(const uint8_t* packed, bool on_device, size_t n_ant, size_t row_bytes,
             long long n, int fmt, double fs, tga_rfi_channel* out) {
  const int dev = current_device_checked();
  (void)dev;
  hipStream_t st;
  TGA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{st};
  DevBuf din(on_device ? 0 : n_ant * row_bytes);
  if (!on_device)
    TGA_HIP(hipMemcpyAsync(din.p, packed, n_ant * row_bytes, hipMemcpyHostToDevice, st));
  const uint8_t* d_in = on_device ? packed : din.as<uint8_t>();
  DevBuf dmom(n_ant * sizeof(Moments));
  hipLaunchKernelGGL(k_moments, dim3(static_cast<unsigned>(n_ant)), dim3(256), 0, st,
                     d_in, row_bytes, n, fmt, dmom.as<Moments>());
  TGA_HIP(hipGetLastError());
  std::vector<Moments> mom(n_ant);
  TGA_HIP(hipMemcpyAsync(mom.data(), dmom.p, n_ant * sizeof(Moments), hipMemcpyDeviceToHost, st));
  TGA_HIP(hipStreamSynchronize(st));

  // f64 statistics on the host from exact integer moments
  std::vector<double> means(n_ant), var(n_ant);
  for (size_t a = 0; a < n_ant; ++a) {
    const double nn = static_cast<double>(n);
    means[a] = static_cast<double>(mom[a].s) / nn;  // rfi.rs:143
    // sum (b - m)^2 = sum b^2 - 2 m sum b + n m^2
    const double m = means[a];
    var[a] = (static_cast<double>(mom[a].s2) - 2.0 * m * static_cast<double>(mom[a].s) +
              nn * m * m) / nn;
  }
  const long long half = n / 2;
  const bool spectrum = n >= 8;  // detect_lines returns nothing below 8 samples (rfi.rs:93)
  std::vector<std::vector<Cand>> cands(n_ant);
  std::vector<double> floors(n_ant, 0.0);
  if (spectrum) {
    DevBuf dmeans;
    dmeans.alloc(n_ant * sizeof(double));
    TGA_HIP(hipMemcpyAsync(dmeans.p, means.data(), n_ant * sizeof(double), hipMemcpyHostToDevice,
                           st));
    const long long total = static_cast<long long>(n_ant) * n;
    DevBuf z(static_cast<size_t>(total) * sizeof(double2));
    DevBuf zs(static_cast<size_t>(total) * sizeof(double2));
    hipLaunchKernelGGL(k_demean, dim3(blocks_for(total, 256)), dim3(256), 0, st,
                       d_in, row_bytes, n, fmt, dmeans.as<double>(), total,
                       z.as<double2>());
    TGA_HIP(hipGetLastError());
    double2* X = fft_forward_d(fft_plan_d(static_cast<size_t>(n)), z.as<double2>(),
                               zs.as<double2>(), n_ant, st);
    const long long ptot = static_cast<long long>(n_ant) * half;
    DevBuf pw(static_cast<size_t>(ptot) * sizeof(double));
    hipLaunchKernelGGL(k_power, dim3(blocks_for(ptot, 256)), dim3(256), 0, st, X, n, half,
                       ptot, pw.as<double>());
    TGA_HIP(hipGetLastError());
    // radix select of rank half/2 (rfi.rs:106-108)
    std::vector<SelState> init(n_ant);
    for (auto& s : init) s = SelState{0ULL, 0ULL, half / 2};
    DevBuf dst(n_ant * sizeof(SelState));
    TGA_HIP(hipMemcpyAsync(dst.p, init.data(), n_ant * sizeof(SelState), hipMemcpyHostToDevice,
                           st));
    DevBuf hist(n_ant * 256 * sizeof(unsigned int));
    TGA_HIP(hipMemsetAsync(hist.p, 0, hist.bytes, st));
    const unsigned gx = static_cast<unsigned>(std::min<long long>(1024, (half + 255) / 256));
    for (int shift = 56; shift >= 0; shift -= 8) {
      hipLaunchKernelGGL(k_hist, dim3(gx, static_cast<unsigned>(n_ant)), dim3(256), 0, st,
                         pw.as<double>(), half, shift, dst.as<SelState>(),
                         hist.as<unsigned int>());
      TGA_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_pick, dim3(blocks_for(static_cast<long long>(n_ant), 64)), dim3(64),
                         0, st, hist.as<unsigned int>(), shift, dst.as<SelState>(),
                         static_cast<int>(n_ant));
      TGA_HIP(hipGetLastError());
    }
    DevBuf dcand(n_ant * kCandCap * sizeof(Cand));
    DevBuf dcount(n_ant * sizeof(int));
    TGA_HIP(hipMemsetAsync(dcount.p, 0, dcount.bytes, st));
    hipLaunchKernelGGL(k_lines, dim3(blocks_for(ptot, 256)), dim3(256), 0, st, pw.as<double>(),
                       half, ptot, dst.as<SelState>(), dcand.as<Cand>(), dcount.as<int>());
    TGA_HIP(hipGetLastError());
    std::vector<SelState> sel(n_ant);
    std::vector<int> cnt(n_ant);
    TGA_HIP(hipMemcpyAsync(sel.data(), dst.p, n_ant * sizeof(SelState), hipMemcpyDeviceToHost,
                           st));
    TGA_HIP(hipMemcpyAsync(cnt.data(), dcount.p, n_ant * sizeof(int), hipMemcpyDeviceToHost, st));
    TGA_HIP(hipStreamSynchronize(st));
    for (size_t a = 0; a < n_ant; ++a) {
      if (cnt[a] > kCandCap)
        fail(TGA_ERR_UNSUPPORTED, "more than 65536 spectral-line candidates on antenna " +
                                      std::to_string(a));
      unsigned long long bits = sel[a].prefix;
      std::memcpy(&floors[a], &bits, sizeof bits);
      cands[a].resize(cnt[a]);
      if (cnt[a])
        TGA_HIP(hipMemcpy(cands[a].data(), dcand.as<Cand>() + a * kCandCap,
                          cnt[a] * sizeof(Cand), hipMemcpyDeviceToHost));
    }
  }
  const double bin_hz = fs / static_cast<double>(n);
  for (size_t a = 0; a < n_ant; ++a) {
    tga_rfi_channel& c = out[a];
    std::memset(&c, 0, sizeof c);
    c.antenna = static_cast<int32_t>(a);
    const double nan = std::nan("");
    c.acf_lag1 = c.acf_lag3 = c.acf_lag16 = c.period_samples = nan;
    if (var[a] < 1e-9) {  // rfi.rs:146-156
      c.dead = 1;
      continue;
    }
    const int lags[3] = {1, 3, 16};
    double acf[3];
    for (int q = 0; q < 3; ++q) {
      // sum_{i<n-lag} (b_i - m)(b_{i+lag} - m), expanded over exact integer sums
      const double m = means[a];
      const double cnt = static_cast<double>(n - lags[q]);
      const double s_lo = static_cast<double>(mom[a].s - mom[a].tail[q]);  // i < n-lag
      const double s_hi = static_cast<double>(mom[a].s - mom[a].head[q]);  // i >= lag
      acf[q] = (static_cast<double>(mom[a].prod[q]) - m * (s_lo + s_hi) + cnt * m * m) / cnt;
    }
    c.acf_lag1 = acf[0];
    c.acf_lag3 = acf[1];
    c.acf_lag16 = acf[2];
    if (!spectrum) continue;
    // detect_lines tail (rfi.rs:111-137) on the candidates, in ascending k like the reference
    std::vector<Cand>& cv = cands[a];
    std::sort(cv.begin(), cv.end(), [](const Cand& x, const Cand& y) { return x.k < y.k; });
    struct Line {
      double f, ex;
    };
    std::vector<Line> lines;
    for (const Cand& q : cv) {
      const double ex = 10.0 * std::log10(q.p / floors[a]);
      if (ex > 15.0) lines.push_back(Line{static_cast<double>(q.k) * bin_hz, ex});
    }
    std::stable_sort(lines.begin(), lines.end(), [](const Line& x, const Line& y) {
      return x.ex > y.ex;  // descending excess; total_cmp order for finite values
    });
    int nl = 0;
    for (const Line& l : lines) {
      bool far = true;
      for (int q = 0; q < nl; ++q)
        if (!(std::fabs(c.lines[q].frequency - l.f) > 10000.0)) far = false;
      if (far) {
        c.lines[nl].frequency = l.f;
        c.lines[nl].excess_db = l.ex;
        if (++nl == 3) break;
      }
    }
    c.n_lines = nl;
    if (nl > 0) c.period_samples = fs / c.lines[0].frequency;
  }
}

}  // namespace tga

namespace tga {
namespace {
void rfi_checked(const uint8_t* packed, bool on_device, size_t n_ant, size_t row_bytes,
                 int64_t record_samples, int32_t input_format, double sampling_freq,
                 tga_rfi_channel* out) {
  if (!packed || !out) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
  if (n_ant == 0 || record_samples <= 16)
    fail(TGA_ERR_INVALID_ARGUMENT, "need at least one antenna and more than 16 samples");
  if (input_format != TGA_INPUT_1BIT && input_format != TGA_INPUT_2BIT)
    fail(TGA_ERR_INVALID_ARGUMENT, "input_format must be TGA_INPUT_1BIT or TGA_INPUT_2BIT");
  const int bits = input_format == TGA_INPUT_1BIT ? 1 : 2;
  if (row_bytes * 8 < static_cast<size_t>(record_samples) * bits)
    fail(TGA_ERR_INVALID_ARGUMENT, "row_bytes too small for record_samples");
  if (!(sampling_freq > 0)) fail(TGA_ERR_INVALID_ARGUMENT, "sampling_freq must be positive");
  rfi_run(packed, on_device, n_ant, row_bytes, record_samples, input_format, sampling_freq, out);
}
}  // namespace
}  // namespace tga

extern "C" int tga_rfi(const uint8_t* packed, size_t n_ant, size_t row_bytes,
                       int64_t record_samples, int32_t input_format, double sampling_freq,
                       tga_rfi_channel* out) {
  return tga::plan_guard([&] {
    tga::rfi_checked(packed, false, n_ant, row_bytes, record_samples, input_format,
                     sampling_freq, out);
  });
}

extern "C" int tga_rfi_device(const void* d_packed, size_t n_ant, size_t row_bytes,
                              int64_t record_samples, int32_t input_format,
                              double sampling_freq, tga_rfi_channel* out) {
  return tga::plan_guard([&] {
    tga::rfi_checked(static_cast<const uint8_t*>(d_packed), true, n_ant, row_bytes,
                     record_samples, input_format, sampling_freq, out);
  });
}
