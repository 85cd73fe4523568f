// Generic batched mixed-radix Stockham FFT through global memory (gfx950).
//
// One kernel launch per radix pass; each thread computes one radix-R butterfly of one row:
//   j in [0, n/R), k = j mod ns (ns = product of the radices already applied)
//   v[r] = src[j + r n/R] * W_{ns R}^{r k}           (twiddle skipped in the first pass)
//   v    = DFT_R(v)
//   dst[(j / ns) ns R + k + r ns] = v[r]
// Reads of consecutive j are consecutive addresses; so are the writes once ns > 1.  Radices
// 16, 8, 4, 2, 3, 5, 7, 11, 13, 31 have register codelets; any other prime factor p up to
// kMaxDirectPrime runs through a direct O(p^2) butterfly.  A length with a larger prime factor
// (a ragged record, an odd sampling rate) is computed with Bluestein's chirp-z identity
//   X_k = w_k sum_j (x_j w_j) conj(w_{k-j}),   w_k = exp(-i pi k^2 / n),
// as a cyclic convolution of power-of-two length m >= 2n - 1 on this same engine, so every
// length costs O(n log n).  This engine serves every length that has no fused LDS kernel:
// level-1 calls of any N, replica spectra, RFI records, long code periods.
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>

#include "fft_device.h"
#include "internal.h"

namespace tga {

struct FftPlan {
  size_t n = 0;
  std::vector<int> radix;  // pass order
  int split = 1;
  DevBuf tw;  // tw1[split] then tw2[ceil(n/split)]
  std::vector<DevBuf> prime_roots;  // per pass with a generic radix: W_R^m, m < R
  // Bluestein form (m > 0): m-point sub-plan, chirp w_k (k < n), FFT_m(b) / m
  size_t m = 0;
  const FftPlan* sub = nullptr;
  DevBuf chirp, bspec;
};

struct FftPlanD {
  size_t n = 0;
  std::vector<int> radix;
  int split = 1;
  DevBuf tw;
  std::vector<DevBuf> prime_roots;
  size_t m = 0;
  const FftPlanD* sub = nullptr;
  DevBuf chirp, bspec;
};

namespace {

constexpr int kBlock = 256;
constexpr size_t kMaxDirectPrime = 127;  // larger prime factors switch to Bluestein

bool has_codelet(int r) {
  switch (r) {
    case 2: case 3: case 4: case 5: case 7: case 8: case 11: case 13: case 16: case 31:
      return true;
  }
  return false;
}

void push_odd_primes(size_t m, std::vector<int>& out) {
  for (size_t p = 3; p * p <= m; p += 2)
    while (m % p == 0) { out.push_back(static_cast<int>(p)); m /= p; }
  if (m > 1) out.push_back(static_cast<int>(m));
}

size_t strip_pow2(size_t m) {
  while (m % 2 == 0) m /= 2;
  return m;
}

size_t largest_prime_factor(size_t m) {
  size_t best = 1;
  for (size_t p = 2; p * p <= m; ++p)
    while (m % p == 0) { best = p; m /= p; }
  return m > 1 ? m : best;
}

// ---- Bluestein (chirp-z) stages; T is float2 or double2 --------------------------------------
template <class T>
__device__ __forceinline__ T bmul(T a, T b) {
  T r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// a[row][k] = x[row][k] w_k for k < n, 0 up to m
template <class T>
__global__ __launch_bounds__(kBlock) void k_blu_pre(const T* __restrict__ x,
                                                    const T* __restrict__ chirp,
                                                    T* __restrict__ a, long long n, long long m,
                                                    long long total) {
  const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (gid >= total) return;
  const long long row = gid / m, k = gid - row * m;
  T v;
  v.x = 0;
  v.y = 0;
  if (k < n) v = bmul(x[row * n + k], chirp[k]);
  a[gid] = v;
}

// A <- conj(A * B / m) (B already carries 1/m): the inverse transform as a forward one
template <class T>
__global__ __launch_bounds__(kBlock) void k_blu_mid(T* __restrict__ A, const T* __restrict__ B,
                                                    long long m, long long total) {
  const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (gid >= total) return;
  T v = bmul(A[gid], B[gid % m]);
  v.y = -v.y;
  A[gid] = v;
}

// y[row][k] = w_k conj(R[row][k]) for k < n
template <class T>
__global__ __launch_bounds__(kBlock) void k_blu_post(const T* __restrict__ R,
                                                     const T* __restrict__ chirp,
                                                     T* __restrict__ y, long long n, long long m,
                                                     long long total) {
  const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (gid >= total) return;
  const long long row = gid / n, k = gid - row * n;
  T r = R[row * m + k];
  r.y = -r.y;
  y[gid] = bmul(chirp[k], r);
}

unsigned grid_for(long long total) {
  return static_cast<unsigned>((total + kBlock - 1) / kBlock);
}

// Single precision pass order: odd radices first (ascending), then 16s and one 8/4/2
// remainder; written to radix[0..cap), returns the pass count (-1 if more than cap).
int factorize_into(size_t n, int* radix, int cap) {
  int k = 0;
  auto push = [&](size_t r) {
    if (k < cap) radix[k] = static_cast<int>(r);
    ++k;
  };
  size_t m = strip_pow2(n);
  for (size_t p = 3; p * p <= m; p += 2)
    while (m % p == 0) { push(p); m /= p; }
  if (m > 1) push(m);
  size_t q = n / strip_pow2(n);
  while (q % 16 == 0) { push(16); q /= 16; }
  if (q > 1) push(q);  // 8, 4 or 2
  return k <= cap ? k : -1;
}

std::vector<int> factorize(size_t n) {
  int buf[64];
  const int k = factorize_into(n, buf, 64);  // at most log2(n) passes
  return std::vector<int>(buf, buf + k);
}

// Double precision pass order: radix 4 (butterfly path), one 2, then odd primes (direct).
std::vector<int> factorize_d(size_t n) {
  std::vector<int> out;
  size_t m = n / strip_pow2(n);
  while (m % 4 == 0) { out.push_back(4); m /= 4; }
  if (m == 2) out.push_back(2);
  push_odd_primes(strip_pow2(n), out);
  return out;
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_pass(const float2* __restrict__ src,
                                                 float2* __restrict__ dst, int n, int ns,
                                                 long long nb_total, const float2* tw1,
                                                 const float2* tw2, int split) {
  const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (gid >= nb_total) return;
  const int nb = n / R;
  const long long row = gid / nb;
  const int j = static_cast<int>(gid - row * nb);
  const float2* s = src + row * n;
  float2* d = dst + row * n;
  float2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = s[j + r * nb];
  const int k = j % ns;
  if (ns > 1) {
    const int e1 = k * (n / (ns * R));
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle2(tw1, tw2, split, r * e1));
  }
  Dft<R>::run(v);
  const int base = (j / ns) * ns * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) d[base + r * ns] = v[r];
}

// Direct butterfly for a prime radix without a codelet.
__global__ __launch_bounds__(kBlock) void k_pass_generic(const float2* __restrict__ src,
                                                         float2* __restrict__ dst, int n,
                                                         int ns, int R, long long nb_total,
                                                         const float2* tw1, const float2* tw2,
                                                         int split, const float2* roots) {
  const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (gid >= nb_total) return;
  const int nb = n / R;
  const long long row = gid / nb;
  const int j = static_cast<int>(gid - row * nb);
  const float2* s = src + row * n;
  float2* d = dst + row * n;
  const int k = j % ns;
  const int e1 = k * (n / (ns * R));
  const int base = (j / ns) * ns * R + k;
  for (int q = 0; q < R; ++q) {
    float2 acc = make_float2(0.f, 0.f);
    for (int r = 0; r < R; ++r) {
      float2 x = s[j + r * nb];
      if (ns > 1 && r > 0) x = cmul(x, twiddle2(tw1, tw2, split, r * e1));
      acc = cadd(acc, cmul(x, roots[(static_cast<long long>(r) * q) % R]));
    }
    d[base + q * ns] = acc;
  }
}

// ---- double precision twins (RFI spectrum) ------------------------------------------------
__device__ __forceinline__ double2 dcmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(kBlock) void k_pass_d(const double2* __restrict__ src,
                                                   double2* __restrict__ dst, int n, int ns,
                                                   int R, long long nb_total,
                                                   const double2* tw1, const double2* tw2,
                                                   int split, const double2* roots) {
  const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (gid >= nb_total) return;
  const int nb = n / R;
  const long long row = gid / nb;
  const int j = static_cast<int>(gid - row * nb);
  const double2* s = src + row * n;
  double2* d = dst + row * n;
  const int k = j % ns;
  const int e1 = k * (n / (ns * R));
  const int base = (j / ns) * ns * R + k;
  if (R == 4) {
    double2 v[4];
    for (int r = 0; r < 4; ++r) {
      v[r] = s[j + r * nb];
      if (ns > 1 && r > 0) {
        const int e = r * e1, hi = e / split, lo = e - hi * split;
        v[r] = dcmul(v[r], dcmul(tw1[lo], tw2[hi]));
      }
    }
    double2 a = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
    double2 b = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
    double2 c = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
    double2 e = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
    d[base] = make_double2(a.x + c.x, a.y + c.y);
    d[base + ns] = make_double2(b.x + e.y, b.y - e.x);
    d[base + 2 * ns] = make_double2(a.x - c.x, a.y - c.y);
    d[base + 3 * ns] = make_double2(b.x - e.y, b.y + e.x);
    return;
  }
  for (int q = 0; q < R; ++q) {
    double2 acc = make_double2(0.0, 0.0);
    for (int r = 0; r < R; ++r) {
      double2 x = s[j + r * nb];
      if (ns > 1 && r > 0) {
        const int e = r * e1, hi = e / split, lo = e - hi * split;
        x = dcmul(x, dcmul(tw1[lo], tw2[hi]));
      }
      const double2 w = roots[(static_cast<long long>(r) * q) % R];
      acc = make_double2(acc.x + x.x * w.x - x.y * w.y, acc.y + x.x * w.y + x.y * w.x);
    }
    d[base + q * ns] = acc;
  }
}

template <class T>
std::vector<T> host_roots(size_t n, size_t count, size_t step) {
  // exp(-2 pi i (m * step) / n) for m < count, computed in long double-free f64 with exact
  // integer reduction of the angle.
  std::vector<T> out(count);
  for (size_t m = 0; m < count; ++m) {
    const unsigned long long e = (static_cast<unsigned long long>(m) * step) % n;
    const double a = -2.0 * kPiD * static_cast<double>(e) / static_cast<double>(n);
    out[m].x = static_cast<decltype(out[m].x)>(std::cos(a));
    out[m].y = static_cast<decltype(out[m].y)>(std::sin(a));
  }
  return out;
}

inline float2* forward_any(const FftPlan* p, float2* d, float2* s, size_t b, hipStream_t st) {
  return fft_forward(p, d, s, b, st);
}
inline double2* forward_any(const FftPlanD* p, double2* d, double2* s, size_t b,
                            hipStream_t st) {
  return fft_forward_d(p, d, s, b, st);
}
inline const FftPlan* plan_any(size_t n, const FftPlan*) { return fft_plan(n); }
inline const FftPlanD* plan_any(size_t n, const FftPlanD*) { return fft_plan_d(n); }

// Bluestein tables: w_k = exp(-i pi k^2 / n) with k^2 reduced mod 2n exactly, and the
// spectrum of b (b_l = conj(w_l) for |l| < n, cyclic over m, scaled by 1/m, a power of two).
template <class Plan, class T>
void init_bluestein(Plan& p, size_t n) {
  size_t m = 1;
  while (m < 2 * n - 1) m <<= 1;
  p.m = m;
  p.sub = plan_any(m, static_cast<const Plan*>(nullptr));
  std::vector<T> w(n), b(m);
  for (size_t k = 0; k < n; ++k) {
    const unsigned long long e =
        (static_cast<unsigned long long>(k) * k) % (2ull * static_cast<unsigned long long>(n));
    const double a = -kPiD * static_cast<double>(e) / static_cast<double>(n);
    w[k].x = static_cast<decltype(w[k].x)>(std::cos(a));
    w[k].y = static_cast<decltype(w[k].y)>(std::sin(a));
  }
  const double inv_m = 1.0 / static_cast<double>(m);
  for (size_t l = 0; l < m; ++l) {
    b[l].x = 0;
    b[l].y = 0;
  }
  for (size_t l = 0; l < n; ++l) {
    const unsigned long long e =
        (static_cast<unsigned long long>(l) * l) % (2ull * static_cast<unsigned long long>(n));
    const double a = kPiD * static_cast<double>(e) / static_cast<double>(n);
    T v;
    v.x = static_cast<decltype(v.x)>(std::cos(a) * inv_m);
    v.y = static_cast<decltype(v.y)>(std::sin(a) * inv_m);
    b[l] = v;
    if (l > 0) b[m - l] = v;
  }
  p.chirp.alloc(n * sizeof(T));
  TGA_HIP(hipMemcpy(p.chirp.p, w.data(), n * sizeof(T), hipMemcpyHostToDevice));
  DevBuf d(m * sizeof(T)), sc(m * sizeof(T));
  TGA_HIP(hipMemcpy(d.p, b.data(), m * sizeof(T), hipMemcpyHostToDevice));
  T* r = forward_any(p.sub, d.as<T>(), sc.as<T>(), 1, nullptr);
  p.bspec.alloc(m * sizeof(T));
  TGA_HIP(hipMemcpyAsync(p.bspec.p, r, m * sizeof(T), hipMemcpyDeviceToDevice, nullptr));
  TGA_HIP(hipStreamSynchronize(nullptr));
}

template <class Plan, class T>
void init_plan(Plan& p, size_t n, bool codelets) {
  p.n = n;
  if (n > 1 && largest_prime_factor(n) > kMaxDirectPrime) {
    init_bluestein<Plan, T>(p, n);
    return;
  }
  p.radix = codelets ? factorize(n) : factorize_d(n);
  p.split = static_cast<int>(std::ceil(std::sqrt(static_cast<double>(n))));
  if (p.split < 1) p.split = 1;
  const size_t n_hi = (n + p.split - 1) / p.split + 1;
  std::vector<T> t1 = host_roots<T>(n, p.split, 1);
  std::vector<T> t2 = host_roots<T>(n, n_hi, p.split);
  p.tw.alloc((t1.size() + t2.size()) * sizeof(T));
  TGA_HIP(hipMemcpy(p.tw.p, t1.data(), t1.size() * sizeof(T), hipMemcpyHostToDevice));
  TGA_HIP(hipMemcpy(static_cast<T*>(p.tw.p) + t1.size(), t2.data(), t2.size() * sizeof(T),
                    hipMemcpyHostToDevice));
  for (int r : p.radix) {
    if (codelets && has_codelet(r)) {
      p.prime_roots.emplace_back();
      continue;
    }
    std::vector<T> roots = host_roots<T>(r, r, 1);
    DevBuf b(roots.size() * sizeof(T));
    TGA_HIP(hipMemcpy(b.p, roots.data(), roots.size() * sizeof(T), hipMemcpyHostToDevice));
    p.prime_roots.push_back(std::move(b));
  }
}

std::recursive_mutex g_plan_mu;  // Bluestein plans build their sub-plan inside
std::map<std::pair<int, size_t>, std::unique_ptr<FftPlan>> g_plans;
std::map<std::pair<int, size_t>, std::unique_ptr<FftPlanD>> g_plans_d;

template <int R>
void launch_pass(const float2* src, float2* dst, int n, int ns, long long nbt,
                 const float2* tw1, const float2* tw2, int split, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>((nbt + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_pass<R>, dim3(grid), dim3(kBlock), 0, st, src, dst, n, ns, nbt, tw1,
                     tw2, split);
}

}  // namespace

const FftPlan* fft_plan(size_t n) {
  if (n == 0) fail(TGA_ERR_INVALID_ARGUMENT, "FFT length must be positive");
  if (n > (1u << 30)) fail(TGA_ERR_UNSUPPORTED, "FFT length too large");
  const int dev = current_device_checked();
  std::lock_guard<std::recursive_mutex> lk(g_plan_mu);
  auto key = std::make_pair(dev, n);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) return it->second.get();
  auto p = std::make_unique<FftPlan>();
  init_plan<FftPlan, float2>(*p, n, true);
  const FftPlan* raw = p.get();
  g_plans.emplace(key, std::move(p));
  return raw;
}

size_t fft_plan_size(const FftPlan* plan) { return plan->n; }

bool fft_plan_bluestein(const FftPlan* plan) { return plan->m > 0; }

// ---- heap-free form: codelet radices only, twiddle table supplied by the caller ------------
namespace {

// The table of init_plan on the GPU: tw[m] = W_n^m for m < split, then W_n^{split m} for
// m < n_hi, each from its exactly reduced exponent in f64 and rounded to f32 once.
__global__ __launch_bounds__(kBlock) void k_fft_twiddles(long long n, int split, int n_hi,
                                                         float2* __restrict__ tw) {
  const int i = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
  if (i >= split + n_hi) return;
  const unsigned long long m = i < split ? i : i - split;
  const unsigned long long step = i < split ? 1ull : static_cast<unsigned long long>(split);
  const unsigned long long e = (m * step) % static_cast<unsigned long long>(n);
  const double a = -2.0 * kPiD * static_cast<double>(e) / static_cast<double>(n);
  tw[i] = make_float2(static_cast<float>(cos(a)), static_cast<float>(sin(a)));
}

}  // namespace

int fft_codelet_passes(size_t n, int* radix, int cap) {
  if (n < 2 || n > (1u << 30)) return -1;
  const int k = factorize_into(n, radix, cap);
  if (k < 0) return -1;
  for (int i = 0; i < k; ++i)
    if (!has_codelet(radix[i])) return -1;
  return k;
}

int fft_twiddle_count(size_t n) {
  const int split = std::max(1, static_cast<int>(std::ceil(std::sqrt(static_cast<double>(n)))));
  return split + static_cast<int>((n + split - 1) / split + 1);
}

void launch_fft_twiddles(size_t n, float2* tw, hipStream_t st) {
  const int split = std::max(1, static_cast<int>(std::ceil(std::sqrt(static_cast<double>(n)))));
  const int n_hi = static_cast<int>((n + split - 1) / split + 1);
  hipLaunchKernelGGL(k_fft_twiddles, dim3(grid_for(split + n_hi)), dim3(kBlock), 0, st,
                     static_cast<long long>(n), split, n_hi, tw);
  TGA_HIP(hipGetLastError());
}

float2* fft_forward_tables(size_t n_, const int* radix, int n_radix, const float2* tw,
                           float2* data, float2* scratch, size_t batch, hipStream_t st) {
  const int n = static_cast<int>(n_);
  if (n == 1 || batch == 0) return data;
  const int split = std::max(1, static_cast<int>(std::ceil(std::sqrt(static_cast<double>(n_)))));
  const float2* tw1 = tw;
  const float2* tw2 = tw + split;
  float2* src = data;
  float2* dst = scratch;
  int ns = 1;
  for (int pi = 0; pi < n_radix; ++pi) {
    const int R = radix[pi];
    const long long nbt = static_cast<long long>(n / R) * static_cast<long long>(batch);
    switch (R) {
      case 2: launch_pass<2>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      case 3: launch_pass<3>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      case 4: launch_pass<4>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      case 5: launch_pass<5>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      case 7: launch_pass<7>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      case 8: launch_pass<8>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      case 11: launch_pass<11>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      case 13: launch_pass<13>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      case 16: launch_pass<16>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      case 31: launch_pass<31>(src, dst, n, ns, nbt, tw1, tw2, split, st); break;
      default: fail(TGA_ERR_UNSUPPORTED, "radix without a register codelet");
    }
    TGA_HIP(hipGetLastError());
    ns *= R;
    std::swap(src, dst);
  }
  return src;
}

TwiddleView fft_twiddles(const FftPlan* plan) {
  TwiddleView v;
  v.tw1 = plan->tw.as<float2>();
  v.tw2 = plan->tw.as<float2>() + plan->split;
  v.split = plan->split;
  return v;
}

namespace {

// The m-point work buffers are allocated per call and released after the stream drains: the
// Bluestein form only serves lengths with a large prime factor, never the fused search.
template <class Plan, class T>
T* bluestein_forward(const Plan* plan, T* data, size_t batch, hipStream_t st) {
  const long long n = static_cast<long long>(plan->n), m = static_cast<long long>(plan->m);
  const long long tm = m * static_cast<long long>(batch), tn = n * static_cast<long long>(batch);
  DevBuf a(static_cast<size_t>(tm) * sizeof(T)), s(static_cast<size_t>(tm) * sizeof(T));
  hipLaunchKernelGGL(k_blu_pre<T>, dim3(grid_for(tm)), dim3(kBlock), 0, st, data,
                     plan->chirp.template as<T>(), a.as<T>(), n, m, tm);
  TGA_HIP(hipGetLastError());
  T* A = forward_any(plan->sub, a.as<T>(), s.as<T>(), batch, st);
  hipLaunchKernelGGL(k_blu_mid<T>, dim3(grid_for(tm)), dim3(kBlock), 0, st, A,
                     plan->bspec.template as<T>(), m, tm);
  TGA_HIP(hipGetLastError());
  T* R = forward_any(plan->sub, A, A == a.as<T>() ? s.as<T>() : a.as<T>(), batch, st);
  hipLaunchKernelGGL(k_blu_post<T>, dim3(grid_for(tn)), dim3(kBlock), 0, st, R,
                     plan->chirp.template as<T>(), data, n, m, tn);
  TGA_HIP(hipGetLastError());
  TGA_HIP(hipStreamSynchronize(st));
  return data;
}

}  // namespace

float2* fft_forward(const FftPlan* plan, float2* data, float2* scratch, size_t batch,
                    hipStream_t st) {
  const int n = static_cast<int>(plan->n);
  if (n == 1 || batch == 0) return data;
  if (plan->m) return bluestein_forward<FftPlan, float2>(plan, data, batch, st);
  const float2* tw1 = plan->tw.as<float2>();
  const float2* tw2 = tw1 + plan->split;
  float2* src = data;
  float2* dst = scratch;
  int ns = 1;
  for (size_t pi = 0; pi < plan->radix.size(); ++pi) {
    const int R = plan->radix[pi];
    const long long nbt = static_cast<long long>(n / R) * static_cast<long long>(batch);
    switch (R) {
      case 2: launch_pass<2>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      case 3: launch_pass<3>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      case 4: launch_pass<4>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      case 5: launch_pass<5>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      case 7: launch_pass<7>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      case 8: launch_pass<8>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      case 11: launch_pass<11>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      case 13: launch_pass<13>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      case 16: launch_pass<16>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      case 31: launch_pass<31>(src, dst, n, ns, nbt, tw1, tw2, plan->split, st); break;
      default: {
        const unsigned grid = static_cast<unsigned>((nbt + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_pass_generic, dim3(grid), dim3(kBlock), 0, st, src, dst, n, ns, R,
                           nbt, tw1, tw2, plan->split, plan->prime_roots[pi].as<float2>());
      }
    }
    TGA_HIP(hipGetLastError());
    ns *= R;
    std::swap(src, dst);
  }
  return src;
}

const FftPlanD* fft_plan_d(size_t n) {
  if (n == 0) fail(TGA_ERR_INVALID_ARGUMENT, "FFT length must be positive");
  if (n > (1u << 30)) fail(TGA_ERR_UNSUPPORTED, "FFT length too large");
  const int dev = current_device_checked();
  std::lock_guard<std::recursive_mutex> lk(g_plan_mu);
  auto key = std::make_pair(dev, n);
  auto it = g_plans_d.find(key);
  if (it != g_plans_d.end()) return it->second.get();
  auto p = std::make_unique<FftPlanD>();
  init_plan<FftPlanD, double2>(*p, n, false);
  const FftPlanD* raw = p.get();
  g_plans_d.emplace(key, std::move(p));
  return raw;
}

double2* fft_forward_d(const FftPlanD* plan, double2* data, double2* scratch, size_t batch,
                       hipStream_t st) {
  const int n = static_cast<int>(plan->n);
  if (n == 1 || batch == 0) return data;
  if (plan->m) return bluestein_forward<FftPlanD, double2>(plan, data, batch, st);
  const double2* tw1 = plan->tw.as<double2>();
  const double2* tw2 = tw1 + plan->split;
  double2* src = data;
  double2* dst = scratch;
  int ns = 1;
  for (size_t pi = 0; pi < plan->radix.size(); ++pi) {
    const int R = plan->radix[pi];
    const long long nbt = static_cast<long long>(n / R) * static_cast<long long>(batch);
    const unsigned grid = static_cast<unsigned>((nbt + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_pass_d, dim3(grid), dim3(kBlock), 0, st, src, dst, n, ns, R, nbt, tw1,
                       tw2, plan->split, plan->prime_roots[pi].as<double2>());
    TGA_HIP(hipGetLastError());
    ns *= R;
    std::swap(src, dst);
  }
  return src;
}

}  // namespace tga
