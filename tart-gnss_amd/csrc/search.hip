// Acquisition-search kernels for gfx950.
//
// The reference correlates, per (antenna, PRN, bin, hypothesis h), an N = 2P point wiped
// signal against the replica [c,c] (h=0) or [c,-c] (h=1) (correlate.rs:107-146).  In exact
// arithmetic both reduce to P-point cyclic problems:
//   h=0:  |y[m]| = |IFFT_P(FFT_P(s) . conj FFT_P(c))[m]|,      s[n] = x[n] + x[n+P]
//   h=1:  |y[m]| = |IFFT_P(FFT_P(e) . conj FFT_P(c tau))[m]|,  e[n] = (x[n] - x[n+P]) tau[n]
// with tau[n] = exp(-i pi n / P), |y[m+P]| = |y[m]|, and an extra factor 2 on the P-point
// results (unnormalised 2P-point transforms carry twice the gain).  For P = 16368 = 31*11*3*16 a whole
// P-point transform fits the 160 KB LDS of a gfx950 CU (128 KB of complex f32), so one
// workgroup runs the four Stockham passes on chip:
//   forward kernel  (antenna, bin, h): decode + de-mean + wipe-off + fold + FFT -> X in HBM
//   rows kernel     (antenna, bin, h) x PRN group: conj(X) . R -> FFT -> |.|^2 -> argmax and
//                   masked second maximum, X kept in registers across the group's PRNs.
// |IFFT(Z)| = |FFT(conj Z)| so every transform here is a forward one.
#include <type_traits>

#include "fft_device.h"
#include "search.h"

namespace tga {

namespace {

// ---- sample decoding ------------------------------------------------------------------------
// 1-bit: MSB-first, bit 1 -> +1, 0 -> -1 (observation.rs:140-151, 232-240).
// 2-bit: MSB-first pairs (sign, magnitude) -> (sign ? +1 : -1) * (magnitude ? 3 : 1).
__device__ __forceinline__ int decode_level(const uint8_t* row, int fmt, long long n) {
  if (fmt == TGA_INPUT_1BIT) {
    const uint8_t b = row[n >> 3];
    return ((b >> (7 - static_cast<int>(n & 7))) & 1) ? 1 : -1;
  }
  const uint8_t b = row[n >> 2];
  const int pair = (b >> (6 - 2 * static_cast<int>(n & 3))) & 3;
  const int mag = (pair & 1) ? 3 : 1;
  return (pair & 2) ? mag : -mag;
}

// Wiped sample x[n] = s[n] e^{+j phi}, phi = f32(f64(pc * n) * f), s = f32(level - mean)
// (correlate.rs:110-115, acquisition.rs:284-291); cos/sin of the f32 phase evaluated in f64
// and rounded once, i.e. correctly rounded like a <1 ULP libm cosf/sinf.
__device__ __forceinline__ float2 wipe_value(float s, double pp, double f) {
  const float ph = __double2float_rn(__dmul_rn(pp, f));
  double sd, cd;
  sincos(static_cast<double>(ph), &sd, &cd);
  const float c = __double2float_rn(cd);
  const float sn = __double2float_rn(sd);
  return make_float2(c * s, sn * s);
}

__device__ __forceinline__ float2 wipe_obs(const uint8_t* row, int fmt, long long n, double mean,
                                           double pc, double f) {
  const float s = __double2float_rn(static_cast<double>(decode_level(row, fmt, n)) - mean);
  return wipe_value(s, __dmul_rn(pc, static_cast<double>(n)), f);
}

// The same sample with the carrier phase kept in f64 (no f32 rounding of phi): the precision
// form's NCO, whose phase error stays ~1e-16 rad over any record length.
__device__ __forceinline__ float2 wipe_obs_f64(const uint8_t* row, int fmt, long long n,
                                               double mean, double pc, double f) {
  const float s = __double2float_rn(static_cast<double>(decode_level(row, fmt, n)) - mean);
  double sd, cd;
  sincos(__dmul_rn(__dmul_rn(pc, static_cast<double>(n)), f), &sd, &cd);
  return make_float2(__double2float_rn(cd) * s, __double2float_rn(sd) * s);
}
// (cos, sin) of an f64 phase rounded to f32 (the FP16 precision form's carrier)
__device__ __forceinline__ float2 cossin_f64(double ph) {
  double sd, cd;
  sincos(ph, &sd, &cd);
  return make_float2(__double2float_rn(cd), __double2float_rn(sd));
}

// ---- block reductions (wave64 shuffles, then LDS across waves) --------------------------------
__device__ __forceinline__ void arg_combine(float& p, int& i, float q, int j) {
  if (q > p || (q == p && j > i)) {
    p = q;
    i = j;
  }
}

// Workgroup barrier that orders LDS traffic only (__syncthreads() also waits for outstanding
// global loads); the block reductions and the fused kernels exchange data through LDS alone.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Raw buffer view of `bytes` bytes at p: loads take a 32-bit lane offset plus a scalar offset,
// so a strided gather needs no 64-bit address arithmetic, and reads past the end return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_view(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, 0x00020000);
}
__device__ __forceinline__ float2 buffer_load_f2(__amdgpu_buffer_rsrc_t rs, int voff, int soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
}

template <int NT>
__device__ __forceinline__ void block_argmax(float& p, int& i, float* sp, int* si) {
  constexpr int NW = (NT + 63) / 64;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float q = __shfl_xor(p, off, 64);
    const int j = __shfl_xor(i, off, 64);
    arg_combine(p, i, q, j);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sp[wave] = p;
    si[wave] = i;
  }
  lds_barrier();
  p = sp[0];
  i = si[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) arg_combine(p, i, sp[w], si[w]);
  lds_barrier();
}

template <int NT>
__device__ __forceinline__ float block_max(float v, float* sp) {
  constexpr int NW = (NT + 63) / 64;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sp[wave] = v;
  lds_barrier();
  v = sp[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) v = fmaxf(v, sp[w]);
  lds_barrier();
  return v;
}

// Long rows are enumerated in tiles of TX xrows x TP PRNs (tile = TX | TP << 16, picked by the
// autotuner), xrow fastest inside a tile, with xrow = (ant * n_bin + bin) * 2 + h taken h = 0
// first, then h = 1; hs = 2 enumerates h = 0 only (unfolded rows, whose two hypotheses
// coincide).  The rows of a tile share TX forward spectra and TP replica spectra, which stay in
// the caches while the tile passes, so the pre-pass does not fetch one replica spectrum from
// memory per row.  Tiles at the end of the xrow or PRN range are narrower.
struct BigItem {
  long long xrow;
  int prn;
};
__device__ __forceinline__ BigItem big_item(long long I, int n_prn, int n_bin, int n_ant,
                                            int hs, int tile) {
  const int tx = tile & 0xffff, tp = tile >> 16;
  const long long n_xe = static_cast<long long>(n_ant) * n_bin * 2 / hs;  // enumerated xrows
  const long long blk = static_cast<long long>(tx) * n_prn;               // rows per xrow block
  const long long xb = I / blk;
  long long rem = I - xb * blk;
  const long long left = n_xe - xb * tx;
  const int a = static_cast<int>(left < tx ? left : tx);  // xrows in this block
  const long long tl = static_cast<long long>(a) * tp;
  const int pt = static_cast<int>(rem / tl);
  rem -= pt * tl;
  const int q = static_cast<int>(rem / a);
  // enumerated xrow xe -> xrow with h outermost (hs = 1), so a tile's xrows share h and with it
  // the replica spectra (prn, h) of its rows
  const long long xe = xb * tx + (rem - static_cast<long long>(q) * a);
  const long long n_ab = static_cast<long long>(n_ant) * n_bin;
  BigItem it;
  it.xrow = hs == 1 ? (xe >= n_ab ? (xe - n_ab) * 2 + 1 : xe * 2) : xe * 2;
  it.prn = pt * tp + q;
  return it;
}
// the row index ((ant * n_prn + prn) * n_bin + bin) * 2 + h of enumerated row I
__device__ __forceinline__ long long big_row(long long I, int n_prn, int n_bin, int n_ant,
                                             int hs, int tile) {
  const BigItem it = big_item(I, n_prn, n_bin, n_ant, hs, tile);
  const long long ant = it.xrow / (2 * n_bin);
  const int bh = static_cast<int>(it.xrow - ant * 2 * n_bin);
  return ((ant * n_prn + it.prn) * n_bin) * 2 + bh;
}

__device__ __forceinline__ int ring_dist(int a, int b, int P) {
  int d = a > b ? a - b : b - a;
  return d < P - d ? d : P - d;
}

// ---- antenna sums ---------------------------------------------------------------------------
// Exact integer sums of each antenna's decoded levels, in kAntSplit slices per antenna so the
// grid fills the GPU even for one antenna: partial[ant * kAntSplit + slice].
__global__ __launch_bounds__(256) void k_ant_partials(const uint8_t* __restrict__ packed,
                                                      size_t row_bytes, long long L, int fmt,
                                                      long long* __restrict__ partial) {
  __shared__ long long red[4];
  const int ant = blockIdx.x / kAntSplit;
  const int slice = blockIdx.x - ant * kAntSplit;
  const long long chunk = (L + kAntSplit - 1) / kAntSplit;
  const long long n0 = slice * chunk;
  const long long n1 = n0 + chunk < L ? n0 + chunk : L;
  const uint8_t* row = packed + static_cast<size_t>(ant) * row_bytes;
  long long acc = 0;
  for (long long n = n0 + threadIdx.x; n < n1; n += 256) acc += decode_level(row, fmt, n);
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Per-antenna mean: the slices' exact sum over one wave, then one f64 division.  Also clears
// the search's NaN flag (block 0), so a search has no memset between kernels and a captured
// graph of it holds kernel nodes only.
__global__ __launch_bounds__(kAntSplit) void k_ant_means(const long long* __restrict__ partial,
                                                         long long L, double* __restrict__ means,
                                                         int* __restrict__ nan_flag) {
  if (nan_flag != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *nan_flag = 0;
  long long v = partial[static_cast<size_t>(blockIdx.x) * kAntSplit + threadIdx.x];
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if (threadIdx.x == 0) means[blockIdx.x] = static_cast<double>(v) / static_cast<double>(L);
}

// ================================ fused P = 16368 path ===================================
namespace fused {

constexpr int P = kFusedP;  // 31 * 11 * 3 * 16
// 512 threads = 8 waves = 2 per SIMD: every SIMD gets the same number of waves, and the
// register budget is 256 VGPRs (no spills).  Per pass the butterflies per thread are
// R31: 1 (+ one direct output of the 16 left-over butterflies), R11: 3, R3: 11, R16: 2.
constexpr int T = 512;
constexpr int NB1 = P / 31;  // 528
constexpr int NB4 = P / 16;  // 1023
constexpr int NW = T / 64;
constexpr int LEFT = NB1 - T;  // 16 radix-31 butterflies beyond one per thread

__device__ __forceinline__ float2 tw_lds(const float2* stw, int e) {
  return cmul(stw[e & 127], stw[128 + (e >> 7)]);
}

// Prime-factor (Good-Thomas) order of the FP16 form's rows.  16368 = 33 x 16 x 31 with pairwise
// coprime factors: with the input map n = (496 n1 + 1023 n2 + 528 n3) mod P and the output map
// k = (496 k1 + 15345 k2 + 528 k3) mod P (the CRT number with k mod 33 = k1, k mod 16 = k2,
// k mod 31 = k3), the P-point DFT is a plain 33 x 16 x 31 DFT with no twiddle factors between
// the passes.  Storage position s = 496 n1 + 31 n2 + n3 holds the value of natural index
// pfa_nat(s).
__device__ __forceinline__ int pfa_nat(int s) {
  const int n1 = s / 496, b = s - 496 * n1, n2 = b / 31, n3 = b - 31 * n2;
  const int n = 496 * n1 + 1023 * n2 + 528 * n3;  // < 3 P
  return n >= 2 * P ? n - 2 * P : n >= P ? n - P : n;
}
// natural index of output k3 of the 31-point pass for column m = k1 + 33 k2: base(m) + 528 k3
// reduced mod P, base(m) = (496 k1 + 15345 k2) mod P
__device__ __forceinline__ int pfa_base(int m) {
  const int k2 = m / 33, k1 = m - 33 * k2;
  return (496 * k1 + 15345 * k2) % P;
}
__device__ __forceinline__ int pfa_wrap(int k) { return k >= P ? k - P : k; }

// In-LDS Stockham pass of radix R after NS points of sub-transforms (in place: every thread
// holds its butterflies in registers across the barrier).
template <int R, int NS>
__device__ __forceinline__ void lds_pass(float2* buf, const float2* stw, int tid) {
  // Opaque copy of the thread index: inside the PRN loop of the rows kernel this keeps the
  // LDS addresses and twiddle loads of every pass from being hoisted out of the loop as
  // loop-invariant registers.
  asm volatile("" : "+v"(tid));
  constexpr int NB = P / R;
  constexpr int BPT = (NB + T - 1) / T;
  constexpr int STRIDE = P / (NS * R);
  float2 v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = tid + b * T;
    if ((b + 1) * T <= NB || j < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = buf[j + r * NB];
    }
  }
  lds_barrier();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = tid + b * T;
    if ((b + 1) * T <= NB || j < NB) {
      const int k = j % NS;
      if constexpr (NS > 1) {
        const int e1 = k * STRIDE;
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tw_lds(stw, r * e1));
      }
      Dft<R>::run(v[b]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[base + r * NS] = v[b][r];
    }
  }
  lds_barrier();
}

// First pass, radix 31 (NS = 1, no twiddles): butterfly j = tid by the symmetric codelet, and
// the 16 left-over butterflies 512..527 as 496 single outputs (one per thread) by direct sums,
// so that no wave runs a second full radix-31 round.
__device__ __forceinline__ void lds_pass31(float2* buf, const float2* stw, int tid) {
  asm volatile("" : "+v"(tid));
  float2 v[31];
#pragma unroll
  for (int r = 0; r < 31; ++r) v[r] = buf[tid + r * NB1];
  const bool left = tid < LEFT * 31;
  const int q = T + tid / 31;
  const int kk = tid - (tid / 31) * 31;
  float2 w[31];
  if (left) {
#pragma unroll
    for (int r = 0; r < 31; ++r) w[r] = buf[q + r * NB1];
  }
  __syncthreads();
  Dft<31>::run(v);
#pragma unroll
  for (int r = 0; r < 31; ++r) buf[tid * 31 + r] = v[r];
  if (left) {
    float2 acc = w[0];
    int e = 0;  // r * kk mod 31, as a multiple of P/31 in the W_P table
#pragma unroll
    for (int r = 1; r < 31; ++r) {
      e += kk;
      if (e >= 31) e -= 31;
      const float2 t = cmul(w[r], tw_lds(stw, e * NB1));
      acc = cadd(acc, t);
    }
    buf[q * 31 + kk] = acc;
  }
  __syncthreads();
}

// Passes 31, 11, 3 in LDS; the caller runs the last radix-16 pass.
__device__ __forceinline__ void passes_to_last(float2* buf, const float2* stw, int tid) {
  lds_pass31(buf, stw, tid);
  lds_pass<11, 31>(buf, stw, tid);
  lds_pass<3, 341>(buf, stw, tid);
}

// The 16368-point forward DFT of buf in place (Stockham passes 31, 11, 3 and a last radix 16
// after 1023 points), output in natural order; the caller synchronises before and after.
__device__ __forceinline__ void lds_fft_natural(float2* buf, const float2* stw, int tid) {
  passes_to_last(buf, stw, tid);
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int j = tid + b * T;
    if (j < NB4) {
      float2 v[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = buf[j + r * NB4];
#pragma unroll
      for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], tw_lds(stw, r * j));
      Dft<16>::run(v);
      // back into LDS in natural order: slots j + r NB4 are read and written by this thread
      // only, so no barrier is needed before the writes
#pragma unroll
      for (int r = 0; r < 16; ++r) buf[j + r * NB4] = v[r];
    }
  }
}

struct FwdArgs {
  const uint8_t* packed;
  size_t row_bytes;
  int fmt;
  const double* means;
  const double* fc;
  int n_bin;
  double pc;
  const float2* tau;
  const float2* tw;
  float2* X;     // FP32 form
  uint32_t* X16;  // FP16 form: f16 (re, im) pairs of X * kF16Scale
  int epochs;     // non-coherent form: one-period epochs per (antenna, bin)
  // the carrier (cos, sin) of every (bin, sample), [n_bin][clen]: the same for all antennas
  const float2* carrier = nullptr;
  long long clen = 0;
};

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// f16 (re, im) pair of v * s, rounded to nearest
__device__ __forceinline__ uint32_t pack_f16(float2 v, float s) {
  return __builtin_bit_cast(uint32_t, half2_t{static_cast<_Float16>(v.x * s),
                                              static_cast<_Float16>(v.y * s)});
}

// Forward spectra.  Reference form: xrow = (ant * n_bin + bin) * 2 + h, the folded sum (h = 0)
// or twisted difference (h = 1) of the two periods.  Non-coherent form (NC): xrow =
// (ant * n_bin + bin) * epochs + e, the e-th one-period epoch.
template <bool F16, bool NC = false>
__global__ __launch_bounds__(T) void k_forward(FwdArgs a) {
  __shared__ float2 buf[P];
  __shared__ float2 stw[256];
  const int tid = threadIdx.x;
  for (int i = tid; i < 256; i += T) stw[i] = a.tw[i];
  const int xrow = blockIdx.x;
  const int h = NC ? 0 : (xrow & 1);
  const int e = NC ? xrow % a.epochs : 0;
  const int ab = NC ? xrow / a.epochs : (xrow >> 1);
  const int bin = ab % a.n_bin;
  const int ant = ab / a.n_bin;
  const uint8_t* row = a.packed + static_cast<size_t>(ant) * a.row_bytes;
  const double mean = a.means[ant];
  // The packed bytes of the block's samples (2P, or one epoch of P) are staged in LDS by
  // independent coalesced loads; decoding straight from global memory made every sample of the
  // wipe loop wait for its own load.  The first sample e P is byte aligned (P = 8 x 2046).
  __shared__ uint8_t raw[2 * P / 4];
  const long long s0 = NC ? static_cast<long long>(e) * P : 0;
  const int bits = a.fmt == TGA_INPUT_1BIT ? 1 : 2;
  const int nbytes = ((NC ? P : 2 * P) * bits) >> 3;
  const uint8_t* src = row + ((s0 * bits) >> 3);
#pragma unroll 4
  for (int i = tid; i < nbytes; i += T) raw[i] = src[i];
  __syncthreads();
  auto level = [&](int n) {  // f32(level - mean) of block-local sample n
    return __double2float_rn(static_cast<double>(decode_level(raw, a.fmt, n)) - mean);
  };
  // the carrier comes from the plan's table (computed once with the same rounding as
  // wipe_value / the f64 phase of the FP16 form), so no antenna repeats its cos / sin
  const float2* car = a.carrier + static_cast<size_t>(bin) * a.clen + s0;
  auto wipe = [&](int n) {
    const float l = level(n);
    const float2 c = car[n];
    return make_float2(c.x * l, c.y * l);
  };
  // eight samples per thread at a time, so that their carrier loads are in flight together
#pragma unroll 8
  for (int n = tid; n < P; n += T) {
    if constexpr (NC) {
      buf[n] = wipe(n);
    } else {
      const float2 x0 = wipe(n), x1 = wipe(n + P);
      buf[n] = h == 0 ? cadd(x0, x1) : cmul(csub(x0, x1), a.tau[n]);
    }
  }
  __syncthreads();
  lds_fft_natural(buf, stw, tid);
  // X is stored in the prime-factor order of the rows kernels, written coalesced
  __syncthreads();
  if constexpr (F16) {
    uint32_t* out = a.X16 + static_cast<size_t>(xrow) * P;
    for (int s = tid; s < P; s += T) out[s] = pack_f16(buf[pfa_nat(s)], kF16Scale);
  } else {
    float2* out = a.X + static_cast<size_t>(xrow) * P;
    for (int s = tid; s < P; s += T) out[s] = buf[pfa_nat(s)];
  }
}


// Pre-pass of positions t, t + 1 of one long row: conj(X) . R at the R0 positions
// t + 16368 n0 (16-byte loads), an R0-point DFT, the twiddle W_P^{n1 k0}, and the R0 sub-rows
// at Yrow + k0 * 16368 + t (non-temporal stores; f16 pairs scaled by kSplitF16Scale with F16).
template <int R0, bool F16>
__device__ __forceinline__ void big_pre_core(const float2* __restrict__ xrow_p,
                                             const float2* __restrict__ crow_p,
                                             const float2* __restrict__ twb, int t,
                                             void* __restrict__ Yrow) {
  constexpr int P1 = kFusedP;
  const float4* x = reinterpret_cast<const float4*>(xrow_p);
  const float4* c = reinterpret_cast<const float4*>(crow_p);
  float2 v[R0], u[R0];
#pragma unroll
  for (int n0 = 0; n0 < R0; ++n0) {
    const float4 xa = x[(t + P1 * n0) >> 1], ca = c[(t + P1 * n0) >> 1];
    v[n0] = cconjmul(make_float2(xa.x, xa.y), make_float2(ca.x, ca.y));
    u[n0] = cconjmul(make_float2(xa.z, xa.w), make_float2(ca.z, ca.w));
  }
  Dft<R0>::run(v);
  Dft<R0>::run(u);
  // W_P^{n1 k0} = (W_P^{n1})^k0 by repeated products from one lookup of W_P^{n1} per position
  // in a table kept in the prime-factor order of the positions (at most R0 - 1 = 9 products:
  // relative error below 1e-6)
  if constexpr (R0 > 1) {  // one segment (R0 = 1): the products alone, no table
    const float4 wq = reinterpret_cast<const float4*>(twb)[t >> 1];
    const float2 w1 = make_float2(wq.x, wq.y), w2 = make_float2(wq.z, wq.w);
    float2 wk = w1, wl = w2;
#pragma unroll
    for (int k0 = 1; k0 < R0; ++k0) {
      v[k0] = cmul(v[k0], wk);
      u[k0] = cmul(u[k0], wl);
      if (k0 + 1 < R0) {
        wk = cmul(wk, w1);
        wl = cmul(wl, w2);
      }
    }
  }
  // Non-temporal stores: the sub-rows are read once, by the transforms, and should not displace
  // the tile's spectra from L2 (the float stores merge into one non-temporal dwordx4 per
  // sub-row; a 64-bit integer store through the same builtin would lose the non-temporal bit).
  if constexpr (F16) {
    uint32_t* y = static_cast<uint32_t*>(Yrow);
#pragma unroll
    for (int k0 = 0; k0 < R0; ++k0) {
      uint32_t* d = y + static_cast<long long>(k0) * P1 + t;
      __builtin_nontemporal_store(pack_f16(v[k0], kSplitF16Scale), d);
      __builtin_nontemporal_store(pack_f16(u[k0], kSplitF16Scale), d + 1);
    }
  } else {
    float2* y = static_cast<float2*>(Yrow);
#pragma unroll
    for (int k0 = 0; k0 < R0; ++k0) {
      float* d = reinterpret_cast<float*>(y + static_cast<long long>(k0) * P1 + t);
      __builtin_nontemporal_store(v[k0].x, d);
      __builtin_nontemporal_store(v[k0].y, d + 1);
      __builtin_nontemporal_store(u[k0].x, d + 2);
      __builtin_nontemporal_store(u[k0].y, d + 3);
    }
  }
}

// Pre-pass of positions [p_begin, p_end) of row rr of a chunk (enumerated row r0 + rr, whose
// forward-spectrum row, replica row and RowOut row come from the plan's row table): thread lt
// takes the two consecutive positions p_begin + 2 lt (+1).
template <int R0, bool F16>
__device__ __forceinline__ void big_pre_positions(const float2* __restrict__ X,
                                                  const float2* __restrict__ rep,
                                                  const float2* __restrict__ twb,
                                                  const int4* __restrict__ rowtab,
                                                  long long r0, long long rr, int p_begin,
                                                  int p_end, int lt,
                                                  void* __restrict__ Yv,
                                                  unsigned long long* __restrict__ rowkey,
                                                  int* __restrict__ rownan,
                                                  int* __restrict__ rowmap) {
  constexpr int P = R0 * kFusedP;
  const int t = p_begin + 2 * lt;
  if (t >= p_end) return;
  const int4 ref = rowtab[r0 + rr];  // (X row, replica row, RowOut row), block-uniform
  void* yrow = F16 ? static_cast<void*>(static_cast<uint32_t*>(Yv) + rr * P)
                   : static_cast<void*>(static_cast<float2*>(Yv) + rr * P);
  big_pre_core<R0, F16>(X + static_cast<long long>(ref.x) * P,
                        rep + static_cast<long long>(ref.y) * P, twb, t, yrow);
  if (t == 0) {
    rowkey[ref.z] = 0ull;
    rownan[ref.z] = 0;
    rowmap[rr] = ref.z;  // the chunk's row index for its transforms
  }
}

struct RowArgs {
  const float2* X;
  const float2* rep;
  const float2* tw;
  int n_bin;
  int n_prn;
  int groups;
  int per_group;
  int cw;
  RowOut* rows;
  // sub-row mode (code periods R0 x 16368): sub-rows Y[s][P] of rows r0 + s / R0
  const float2* Y;
  long long n_sub;
  long long r0;
  unsigned long long* rowkey;
  int* rownan;
  int epochs;  // non-coherent form: one-period epochs per (antenna, bin)
  int hs;      // sub-row mode: 2 when only h = 0 rows are enumerated (unfolded rows)
  int n_ant;   // sub-row mode: antennas (for the tiled row order)
  int tile;    // sub-row mode: tile shape of the row order (big_item)
  // surface form (DUMP): every output power also goes to dump[row * period + sample], row the
  // RowOut index of the correlation row, period 16368 (rows) or SUB x 16368 (sub-rows)
  float* dump;
  // sub-row mode: the row index of each of the chunk's rows (written by the pre-pass)
  const int* rowmap;
  // masked mode of k_subrows (second peaks of the winning rows): sub-rows of winner w =
  // ant * n_prn + prn, whose peak (slot of prn) names the window centre; the largest power
  // outside it goes to m2key[w]
  const tga_peak* peaks;
  const PeakSlot* slots;
  unsigned int* m2key;
};

typedef float floatx4 __attribute__((ext_vector_type(4)));

// (power, index) as one 64-bit key: for p >= 0 the float bits order like the value, and the
// index in the low word breaks ties towards the larger index (last maximum).
__device__ __forceinline__ unsigned long long pkey(float p, int idx) {
  return (static_cast<unsigned long long>(__float_as_uint(p)) << 32) |
         static_cast<unsigned int>(idx);
}

// Running top two of a set of powers whose indices are pairwise more than 2 cw apart: `key`
// holds the best (power, index), `second` the best power other than it.
struct Top2 {
  unsigned long long key = 0;  // power 0 at index 0 is below every real key but equal-safe
  float second = -1.0f;
  bool any = false;
  __device__ __forceinline__ void add(float p, int idx) {
    const unsigned long long k = pkey(p, idx);
    const float best = any ? __uint_as_float(static_cast<unsigned int>(key >> 32)) : -1.0f;
    const bool up = !any || k > key;
    second = up ? best : fmaxf(second, p);
    key = up ? k : key;
    any = true;
  }
  __device__ __forceinline__ float best() const {
    return any ? __uint_as_float(static_cast<unsigned int>(key >> 32)) : -1.0f;
  }
  // a set holding one (power, index) pair
  __device__ __forceinline__ static Top2 leaf(float p, int idx) {
    Top2 t;
    t.key = pkey(p, idx);
    t.any = true;
    return t;
  }
  // union of two sets whose members are pairwise more than 2 cw apart
  __device__ __forceinline__ static Top2 merge(const Top2& a, const Top2& b) {
    const bool up = b.key > a.key;
    Top2 t;
    t.key = up ? b.key : a.key;
    const float lose = __uint_as_float(static_cast<unsigned int>((up ? a.key : b.key) >> 32));
    t.second = fmaxf(fmaxf(a.second, b.second), lose);
    t.any = true;
    return t;
  }
  __device__ __forceinline__ int index() const {
    return any ? static_cast<int>(key & 0xffffffffu) : -1;
  }
  // max over the set without the element inside the exclusion window around m (at most one)
  __device__ __forceinline__ float masked(int m, int cw) const {
    if (!any) return -1.0f;
    return ring_dist(index(), m, P) > cw ? best() : second;
  }
};

// Order-preserving map of NaN-free floats onto unsigned ints (larger float, larger image), and
// its inverse: maxima of images are maxima of the floats, with no float compares.
__device__ __forceinline__ unsigned int ford(float p) {
  const unsigned int b = __float_as_uint(p);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float funord(unsigned int u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Unsigned maximum over each row of 16 lanes, delivered to every lane of the row, by four DPP
// moves: swap neighbours (quad_perm 1,0,3,2), swap pairs (quad_perm 2,3,0,1), mirror each half
// row of 8 (row_half_mirror), mirror the row (row_mirror).  Every lane of the wave must be
// active: DPP reads the source lane's register whether or not that lane executes.
__device__ __forceinline__ unsigned int row16_umax(unsigned int v) {
  const int x = static_cast<int>(v);
  v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(x, x, 0xB1, 0xF, 0xF, false)));
  const int y = static_cast<int>(v);
  v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(y, y, 0x4E, 0xF, 0xF, false)));
  const int z = static_cast<int>(v);
  v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(z, z, 0x141, 0xF, 0xF, false)));
  const int w = static_cast<int>(v);
  v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(w, w, 0x140, 0xF, 0xF, false)));
  return v;
}

// Unsigned maximum over the wave (wave-uniform result): rows of 16 by DPP, then the four row
// maxima read out of lanes 0, 16, 32 and 48.
__device__ __forceinline__ unsigned int wave_umax(unsigned int v) {
  v = row16_umax(v);
  const unsigned int a = static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(v), 0));
  const unsigned int b = static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(v), 16));
  const unsigned int c = static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(v), 32));
  const unsigned int d = static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(v), 48));
  return max(max(a, b), max(c, d));
}

// Argmax with the last-maximum rule (largest power, then largest index) over the wave, as
// (power image, 1 + index): two unsigned maxima instead of a 64-bit key.
__device__ __forceinline__ void wave_argmax(float p, int i, unsigned int& pw, unsigned int& iw) {
  const unsigned int pk = ford(p);
  pw = wave_umax(pk);
  iw = wave_umax(pk == pw ? static_cast<unsigned int>(i + 1) : 0u);
}

// The same over at most 16 per-wave results held by lanes 0..15 (other lanes hold 0 / 0).
__device__ __forceinline__ void lanes_argmax(unsigned int pk, unsigned int ik, unsigned int& pw,
                                             unsigned int& iw) {
  pw = static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(row16_umax(pk)), 0));
  iw = static_cast<unsigned int>(
      __builtin_amdgcn_readlane(static_cast<int>(row16_umax(pk == pw ? ik : 0u)), 0));
}

// LDS scratch of one row's reduction: per-wave argmax (power image, 1 + index), masked maximum
// (image) and bad-value flag.
template <int NW_>
struct RowRed {
  static_assert(NW_ <= 16, "cross-wave reductions read one wave per lane of a 16-lane row");
  unsigned int pk[NW_];
  unsigned int ik[NW_];
  unsigned int m[NW_];
  int bad[NW_];
};

// A row whose result waits for its masked maximum: thread 0 writes it in row_flush.
struct RowPending {
  RowOut* out = nullptr;
  float p1 = 0.0f;
  int i1 = 0;
  int nan = 0;
};

// End of one correlation row, first half: block argmax (last maximum), the NaN / overflow flag
// and each wave's masked maximum around the argmax, with one workgroup barrier.  Waves reduce
// with DPP moves; across waves, lane w reads wave w's slot and the same 16-lane reduction
// combines them, so no thread walks the slots one by one.  The result waits in `pend` until
// row_flush, which runs after the kernel's next workgroup barrier and before the second one
// (so each LDS slot is read before the barrier that precedes its next write).  Called by every
// thread of the block in uniform control flow.
template <int NT>
__device__ __forceinline__ void row_finish_begin(const Top2& main_set, const Top2& extra_set,
                                                 bool bad, int cw, RowRed<NT / 64>& red,
                                                 RowOut* out, RowPending& pend) {
  constexpr int NWR = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float bp = main_set.best();
  int bi = main_set.index();
  arg_combine(bp, bi, extra_set.best(), extra_set.index());
  unsigned int pw, iw;
  wave_argmax(bp, bi, pw, iw);
  const bool wave_bad = __ballot(bad) != 0;
  if (lane == 0) {
    red.pk[wave] = pw;
    red.ik[wave] = iw;
    red.bad[wave] = wave_bad ? 1 : 0;
  }
  lds_barrier();
  const bool mine = lane < NWR;
  lanes_argmax(mine ? red.pk[lane] : 0u, mine ? red.ik[lane] : 0u, pw, iw);
  const bool any_bad = __ballot(mine && red.bad[mine ? lane : 0] != 0) != 0;
  bp = funord(pw);
  bi = static_cast<int>(iw) - 1;
  const unsigned int mw =
      wave_umax(ford(fmaxf(main_set.masked(bi, cw), extra_set.masked(bi, cw))));
  if (lane == 0) red.m[wave] = mw;
  pend.out = out;
  pend.p1 = bp;
  pend.i1 = bi;
  pend.nan = any_bad ? 1 : 0;
}

// Second half: the block's masked maximum and the row's RowOut (wave 0), once a workgroup
// barrier separates it from every wave's row_finish_begin.
template <int NT>
__device__ __forceinline__ void row_flush(const RowRed<NT / 64>& red, RowPending& pend) {
  constexpr int NWR = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave == 0 && pend.out != nullptr) {
    const unsigned int m2 = static_cast<unsigned int>(__builtin_amdgcn_readlane(
        static_cast<int>(row16_umax(lane < NWR ? red.m[lane] : 0u)), 0));
    if (lane == 0) {
      RowOut o;
      o.p1 = pend.p1;
      o.p2 = funord(m2);
      o.i1 = pend.i1;
      o.nan = pend.nan;
      *pend.out = o;
    }
  }
  pend.out = nullptr;
}

// Both halves with their own second barrier.
template <int NT>
__device__ __forceinline__ void row_finish(const Top2& main_set, const Top2& extra_set,
                                           bool bad, int cw, RowRed<NT / 64>& red, RowOut* out) {
  RowPending pend;
  row_finish_begin<NT>(main_set, extra_set, bad, cw, red, out, pend);
  lds_barrier();
  row_flush<NT>(red, pend);
}

// End of one sub-row of a long row: block argmax of the threads' best keys (power bits, sample
// index: kmax), folded into the row's 64-bit key by an atomic maximum, so the last maximum over
// all sub-rows wins.  Only the argmax is kept: the masked second peak of long rows comes from a
// recomputation of the winning row.
template <int NT>
__device__ __forceinline__ void row_finish_key(unsigned long long kmax, bool bad,
                                               RowRed<NT / 64>& red, int k0, int r0f,
                                               unsigned long long* key, int* nan) {
  constexpr int NWR = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float bp = __uint_as_float(static_cast<unsigned int>(kmax >> 32));
  const int bi = static_cast<int>(kmax & 0xffffffffu);
  unsigned int pw, iw;
  wave_argmax(bp, bi, pw, iw);
  const bool wave_bad = __ballot(bad) != 0;
  if (lane == 0) {
    red.pk[wave] = pw;
    red.ik[wave] = iw;
    red.bad[wave] = wave_bad ? 1 : 0;
  }
  lds_barrier();
  if (wave == 0) {
    const bool mine = lane < NWR;
    lanes_argmax(mine ? red.pk[lane] : 0u, mine ? red.ik[lane] : 0u, pw, iw);
    const bool any_bad = __ballot(mine && red.bad[mine ? lane : 0] != 0) != 0;
    if (lane == 0) {
      atomicMax(key, pkey(funord(pw), k0 + r0f * (static_cast<int>(iw) - 1)));
      if (any_bad) atomicOr(nan, 1);
    }
  }
}

// Correlation rows.  Per (antenna, bin, hypothesis) block and per PRN of its group:
//   pass A  radix 33 (NS = 1), 496 butterflies, operands conj(X) . R gathered straight from
//           global memory (buffer loads) into registers
//   pass B  radix 16 (NS = 33) through LDS
//   pass C  radix 31 (NS = 528); outputs stay in registers as |z|^2.
// Rows (SUB == 0: conj(X) . R from the spectra) and sub-rows of long rows (SUB > 0: the
// pre-pass output) are all stored in the prime-factor order of 16368 = 33 x 16 x 31, for which
// the passes need no twiddle factors and output r of pass-C butterfly j is sample out_idx(j, r),
// the CRT number of (j mod 33, j / 33, r).  Pass C has two implementations:
//     VARIANT 0 (VALU): one butterfly per thread by the symmetric codelet (its twiddles in
//           registers); the 16 butterflies beyond one per thread are split into
//           256 (butterfly, output pair) units, one per thread of waves 0-3.
//     VARIANT 1 (matrix cores): with a_j = v_j + v_{31-j}, b_j = v_j - v_{31-j} (a_0 = v_0),
//           Y_k = A_k - i B_k and Y_{31-k} = A_k + i B_k where A = C a, B = S b for the 16x16
//           real matrices C[k][j] = cos(2 pi jk/31) (C[k][0] = 1), S[k][j] = sin(2 pi jk/31)
//           (S[k][0] = 0): a tile of 16 butterflies is 16 v_mfma_f32_16x16x4f32, and the 33
//           tiles are dealt to the 8 waves (wave 0 takes tile 32 too).
// Each thread keeps the top two powers of sets whose indices lie more than 32 samples apart
// pairwise, so at most one member of a set falls inside the +-cw exclusion window around the
// block's argmax and the masked second maximum needs no second pass over the row.
// DUMP: the surface form, which also stores every output power (a diagnostic, not the search).
template <int VARIANT, int SUB = 0, bool NC = false, bool DUMP = false>
__global__ __launch_bounds__(T) void k_rows(RowArgs a) {
  constexpr int NBA = P / 33;  // 496
  constexpr int NBB = P / 16;  // 1023
  constexpr int NBC = P / 31;  // 528
  constexpr int TILES = NBC / 16;  // 33
  __shared__ float2 buf[P];
  __shared__ float2 stw[256];
  __shared__ float2 r31[31];  // variant 0: (cos, sin)(2 pi m / 31)
  __shared__ RowRed<NW> red;
  const int tid = threadIdx.x;
  for (int i = tid; i < 256; i += T) stw[i] = a.tw[i];
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63, ln16 = lane & 15, grp = lane >> 4;
  // variant 1: GEMM constants cA/sA and the lane's radix-31 inputs rr
  float cA[4], sA[4];
  int rr[8];
  if constexpr (VARIANT == 0) {
    if (tid < 31) {
      const float2 w = tw_lds(stw, tid * NBC);  // exp(-2 pi i m / 31)
      r31[tid] = make_float2(w.x, -w.y);
    }
  } else {
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const int j = 4 * st + grp;
      const float2 w = tw_lds(stw, ((j * ln16) % 31) * NBC);  // exp(-2 pi i jk/31)
      cA[st] = j == 0 ? 1.0f : w.x;
      sA[st] = j == 0 ? 0.0f : -w.y;
      rr[2 * st] = j;
      rr[2 * st + 1] = j == 0 ? 0 : 31 - j;
    }
  }
  // Sample index of output r of pass-C butterfly j: the CRT number of (j mod 33, j / 33, r).
  auto out_idx = [](int j, int r) { return pfa_wrap(pfa_base(j) + NBC * r); };
  // Lane offset of pass A's operand loads: threads beyond the 496 butterflies load at an offset
  // past the end of the view, which the buffer load returns as 0 (no branch per load).
  const int voa = tid < NBA ? tid * 8 : P * 8;
  // Rows and sub-rows with the matrix-core radix 31 keep only each mirrored output pair's
  // larger power, max(|Y_kk|^2, |Y_{31-kk}|^2) with Y = A -/+ iB, A = C a and B = S b of output
  // row kk, and the sign of the difference, which names its sample.
  constexpr bool kPairs = SUB > 0 && VARIANT == 1;
  // Rows of the matrix-core variant keep only the lane's running maximum of the 64-bit keys
  // (power bits, sample) of their outputs and end with a one-barrier argmax; their second peak
  // comes from a masked recomputation of the winning rows.
  constexpr bool kRowKey = SUB == 0 && VARIANT == 1 && !NC;
  // One transform of a correlation row (or sub-row): `operand(ta, r)` gives pass A's input r of
  // butterfly ta; every output power goes to `sink(slot, p, idx)` with slot < 31 for the
  // thread's main set (outputs more than 32 samples apart) and slot 31 / 32 for its extra set,
  // or, with kPairs, every output pair to `pair(ts, e, q, kk, p1, p2)` (tile slot ts, GEMM row
  // e, powers p1 = |Y_kk|^2, p2 = |Y_{31-kk}|^2).
  // `after_a()` runs once pass A has consumed its operands (prefetches for the next row).
  auto transform = [&](auto operand, auto sink, auto after_a, auto pair) {
    // ---- pass A
    int ta = tid;
    asm volatile("" : "+v"(ta));
    if (ta < NBA) {
      float2 v[33];
#pragma unroll
      for (int r = 0; r < 33; ++r) v[r] = operand(ta, r);
      Dft<33>::run(v);
#pragma unroll
      for (int r = 0; r < 33; ++r) buf[ta * 33 + r] = v[r];
    }
    after_a();
    lds_barrier();
    // ---- pass B
    {
      int tb2 = tid;
      asm volatile("" : "+v"(tb2));
      float2 v[2][16];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int j = tb2 + b * T;
        if (b == 0 || j < NBB) {
#pragma unroll
          for (int r = 0; r < 16; ++r) v[b][r] = buf[j + r * NBB];
        }
      }
      lds_barrier();
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int j = tb2 + b * T;
        if (b == 0 || j < NBB) {
          const int q33 = j / 33;
          const int k = j - q33 * 33;
          Dft<16>::run(v[b]);
          const int base = q33 * NBC + k;
#pragma unroll
          for (int r = 0; r < 16; ++r) buf[base + r * 33] = v[b][r];
        }
      }
      lds_barrier();
    }
    // ---- pass C
    if constexpr (VARIANT == 0) {
      int tc = tid;
      asm volatile("" : "+v"(tc));
      float2 v[31];
#pragma unroll
      for (int r = 0; r < 31; ++r) v[r] = buf[tc + r * NBC];
      if (tc < 16 * LEFT) {  // left-over unit: butterfly T + b, output pair kp
        const int b = tc & (LEFT - 1);
        const int kp = tc >> 4;
        const int q = T + b;
        float2 w[31];
#pragma unroll
        for (int r = 0; r < 31; ++r) w[r] = buf[q + r * NBC];
        float ar = w[0].x, ai = w[0].y, br = 0.f, bi = 0.f;
        int m = 0;
#pragma unroll
        for (int jj = 1; jj <= 15; ++jj) {
          const float2 s = cadd(w[jj], w[31 - jj]);
          const float2 d = csub(w[jj], w[31 - jj]);
          m += kp;
          if (m >= 31) m -= 31;
          const float2 cs = r31[m];
          ar = fmaf(s.x, cs.x, ar);
          ai = fmaf(s.y, cs.x, ai);
          br = fmaf(d.x, cs.y, br);
          bi = fmaf(d.y, cs.y, bi);
        }
        // kp = 0: every root is 1 and the sine sums vanish, so Y_0 = ar + i ai
        const float y1r = ar + bi, y1i = ai - br;  // Y_kp
        const float p1 = y1r * y1r + y1i * y1i;
        sink(31, p1, out_idx(q, kp));
        if (kp > 0) {
          const float y2r = ar - bi, y2i = ai + br;  // Y_{31-kp}
          const float p2 = y2r * y2r + y2i * y2i;
          // the two outputs of a unit lie a nonzero multiple of 528 samples apart
          sink(32, p2, out_idx(q, 31 - kp));
        }
      }
      Dft<31>::run(v);
#pragma unroll
      for (int r = 0; r < 31; ++r) {
        const float pv = v[r].x * v[r].x + v[r].y * v[r].y;
        sink(r, pv, out_idx(tc, r));  // outputs of one butterfly lie multiples of 528 apart
      }
    } else {
      int gc = grp;
      asm volatile("" : "+v"(gc));
      const int ntile = wave == 0 ? 5 : 4;  // tiles wave, wave+8, ..., 32
      // two batches of tile slots (0-2, 3-4) bound the live GEMM operands and accumulators
#pragma unroll
      for (int b0 = 0; b0 < 5; b0 += 3) {
        constexpr int NB = 3;
        float fr[NB][4][4];  // [tile][step][a.re, a.im, b.re, b.im]
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const int ts = b0 + u;
          if (ts < ntile) {
            const int t = wave + 8 * ts;
            const int q = 16 * t + ln16;
#pragma unroll
            for (int st = 0; st < 4; ++st) {
              const int r1 = rr[2 * st], r2 = rr[2 * st + 1];
              const float2 v1 = buf[q + r1 * NBC];
              float2 v2 = buf[q + r2 * NBC];
              if (r1 == 0) v2 = make_float2(0.f, 0.f);  // a_0 = v_0 (no partner input)
              fr[u][st][0] = v1.x + v2.x;
              fr[u][st][1] = v1.y + v2.y;
              fr[u][st][2] = v1.x - v2.x;
              fr[u][st][3] = v1.y - v2.y;
            }
          }
        }
        floatx4 acc[NB][4];
#pragma unroll
        for (int u = 0; u < NB; ++u)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[u][c] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < 4; ++st) {
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            if (b0 + u < ntile) {
              acc[u][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(cA[st], fr[u][st][0], acc[u][0], 0, 0, 0);
              acc[u][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(cA[st], fr[u][st][1], acc[u][1], 0, 0, 0);
              acc[u][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(sA[st], fr[u][st][2], acc[u][2], 0, 0, 0);
              acc[u][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(sA[st], fr[u][st][3], acc[u][3], 0, 0, 0);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const int ts = b0 + u;
          if (ts < ntile) {
            const int q = 16 * (wave + 8 * ts) + ln16;
            // butterflies q0 + 128 ts: in the prime-factor order the outputs of butterflies 256
            // apart can lie 8 samples apart, so tiles 0, 1 (and 4) and tiles 2, 3 keep separate
            // top-2 sets (within each, outputs lie at least 49 samples apart)
            const int slot = (ts == 2 || ts == 3) ? 31 : 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int kk = 4 * gc + e;  // output row of the GEMM tile
              const float ar = acc[u][0][e], ai = acc[u][1][e];
              const float br = acc[u][2][e], bi = acc[u][3][e];
              if constexpr (kPairs) {
                // kk = 0: B = 0 (row 0 of S is zero), so p2 = p1 = |Y_0|^2
                const float y1r = ar + bi, y1i = ai - br, y2r = ar - bi, y2i = ai + br;
                pair(ts, e, q, kk, fmaf(y1r, y1r, y1i * y1i), fmaf(y2r, y2r, y2i * y2i));
              } else {
                const float y1r = ar + bi, y1i = ai - br;  // Y_kk
                const float p1 = y1r * y1r + y1i * y1i;
                sink(slot, p1, out_idx(q, kk));
                if (e > 0 || kk > 0) {
                  const float y2r = ar - bi, y2i = ai + br;  // Y_{31-kk}
                  const float p2 = y2r * y2r + y2i * y2i;
                  sink(slot, p2, out_idx(q, 31 - kk));
                }
              }
            }
          }
        }
      }
    }
  };
  __syncthreads();
  if constexpr (NC) {
    // Non-coherent form: block (antenna, bin) x PRN group; per PRN, a.epochs one-period
    // transforms whose powers are summed per output before the top-2 selection.
    const int g = blockIdx.x % a.groups;
    const int ab = blockIdx.x / a.groups;
    const int bin = ab % a.n_bin;
    const int ant = ab / a.n_bin;
    const int p0 = g * a.per_group;
    const int p1 = min(p0 + a.per_group, a.n_prn);
    for (int prn = p0; prn < p1; ++prn) {
      const __amdgpu_buffer_rsrc_t rv =
          buffer_view(a.rep + static_cast<size_t>(prn) * 2 * P, P * 8);
      float acc[33];
#pragma unroll
      for (int i = 0; i < 33; ++i) acc[i] = 0.0f;
      for (int e = 0; e < a.epochs; ++e) {
        const __amdgpu_buffer_rsrc_t xv = buffer_view(
            a.X + (static_cast<size_t>(ab) * a.epochs + e) * P, P * 8);
        transform(
            [&](int ta, int r) {
              return cconjmul(buffer_load_f2(xv, ta * 8, r * NBA * 8),
                              buffer_load_f2(rv, ta * 8, r * NBA * 8));
            },
            [&](int slot, float p, int) { acc[slot] += p; }, [] {},
            [](int, int, int, int, float, float) {});
        lds_barrier();  // the next epoch's pass A overwrites the LDS row
      }
      // the outputs of each slot (same indices in every epoch, see pass C)
      Top2 ms, es;
      float accn = 0.0f;
#pragma unroll
      for (int r = 0; r < 31; ++r) {
        accn += acc[r];
        ms.add(acc[r], out_idx(tid, r));
      }
      if (tid < 16 * LEFT) {
        const int b = tid & (LEFT - 1), kp = tid >> 4, q = T + b;
        accn += acc[31];
        es.add(acc[31], out_idx(q, kp));
        if (kp > 0) {
          accn += acc[32];
          es.add(acc[32], out_idx(q, 31 - kp));
        }
      }
      RowOut* out = a.rows + ((static_cast<size_t>(ant) * a.n_prn + prn) * a.n_bin + bin) * 2;
      row_finish<T>(ms, es, accn != accn, a.cw, red, out);
      if (tid == 0) out[1] = RowOut{0.0f, -1.0f, 0, 0};  // single hypothesis: h = 1 never wins
    }
  } else if constexpr (SUB == 0) {
    // Blocks are dealt round-robin to the 8 XCDs: the PRN group follows blockIdx % groups, so
    // with 8 or more groups the replica spectra a group touches stay in one XCD's L2 (with
    // fewer, a group spans 8 / groups XCDs and more PRNs share each block's forward spectrum;
    // the autotuner picks the count).  One block per (xrow, group):
    // an outer loop over several items per block (the persistent form of the FP16 kernel)
    // measured 45% slower here, even when it ran once per block.
    const int g = blockIdx.x % a.groups;
    const int xrow = blockIdx.x / a.groups;
    const int h = xrow & 1;
    const int bin = (xrow >> 1) % a.n_bin;
    const int ant = (xrow >> 1) / a.n_bin;
    const int p0 = g * a.per_group;
    const int p1 = min(p0 + a.per_group, a.n_prn);
    const __amdgpu_buffer_rsrc_t xv = buffer_view(a.X + static_cast<size_t>(xrow) * P, P * 8);
    // The block's forward spectrum is the same for every PRN.  With the matrix-core radix-31
    // (few VGPRs elsewhere) each thread keeps its 33 pass-A inputs in registers (66 VGPRs)
    // instead of re-reading them per PRN through L2, where the 32 X rows in flight on an XCD
    // do not fit beside the replicas; the VALU codelet needs those registers itself (255 VGPRs
    // and 20% slower with them) and re-reads.
    constexpr bool kKeepX = VARIANT == 1;
    float2 xr[kKeepX ? 33 : 1];
    if constexpr (kKeepX) {
#pragma unroll
      for (int r = 0; r < 33; ++r)
        xr[r] = buffer_load_f2(xv, voa, r * NBA * 8);
    }
    // With X in registers the next PRN's replica inputs are loaded right after pass A, so
    // their latency hides behind passes B and C.
    auto rep_view = [&](int prn) {
      return buffer_view(a.rep + (static_cast<size_t>(prn) * 2 + h) * P, P * 8);
    };
    float2 rn[kKeepX ? 33 : 1];
    auto load_rep = [&](int prn) {
      const __amdgpu_buffer_rsrc_t rv = rep_view(prn);
#pragma unroll
      for (int r = 0; r < 33; ++r)
        rn[r] = buffer_load_f2(rv, voa, r * NBA * 8);
    };
    if constexpr (kKeepX) {
      if (p0 < p1) load_rep(p0);
    }
    for (int prn = p0; prn < p1; ++prn) {
      const __amdgpu_buffer_rsrc_t rv = rep_view(prn);
      const size_t orow = ((static_cast<size_t>(ant) * a.n_prn + prn) * a.n_bin + bin) * 2 + h;
      Top2 ms, es;
      float accn = 0.0f;  // sum of powers: NaN-propagating detector
      // kRowKey: the lane's running maximum of the keys (power bits, sample); powers are >= 0,
      // so their f32 bits order like the values, a NaN above everything
      unsigned long long km = 0ull;
      float2 rc[kKeepX ? 33 : 1];
      if constexpr (kKeepX) {
#pragma unroll
        for (int r = 0; r < 33; ++r) rc[r] = rn[r];
      }
      transform(
          [&](int ta, int r) {
            if constexpr (kKeepX) {
              return cconjmul(xr[r], rc[r]);
            } else {
              const float2 rr = buffer_load_f2(rv, ta * 8, r * NBA * 8);
              return cconjmul(buffer_load_f2(xv, ta * 8, r * NBA * 8), rr);
            }
          },
          [&](int slot, float p, int idx) {
            if constexpr (kRowKey) {
              const unsigned long long k = pkey(p, idx);
              km = k > km ? k : km;
            } else {
              accn += p;
              (slot < 31 ? ms : es).add(p, idx);
            }
            if constexpr (DUMP) a.dump[orow * P + idx] = p;
          },
          [&] {
            if constexpr (kKeepX) {
              if (prn + 1 < p1) load_rep(prn + 1);
            }
          },
          [](int, int, int, int, float, float) {});
      if constexpr (kRowKey) {
        // block argmax with the last-maximum rule: per wave (power bits, 1 + sample) by two
        // unsigned maxima, across waves in wave 0 after one barrier (which also orders this
        // row's pass-C reads before the next row's pass-A writes; the slots are next written
        // three barriers later)
        const unsigned int kb = static_cast<unsigned int>(km >> 32);
        const unsigned int pw = wave_umax(kb);
        const unsigned int iw =
            wave_umax(kb == pw ? static_cast<unsigned int>(km & 0xffffffffu) + 1u : 0u);
        if (lane == 0) {
          red.pk[wave] = pw;
          red.ik[wave] = iw;
        }
        lds_barrier();
        if (wave == 0) {
          const bool mine = lane < NW;
          unsigned int M, I;
          lanes_argmax(mine ? red.pk[lane] : 0u, mine ? red.ik[lane] : 0u, M, I);
          if (lane == 0) {
            RowOut o;
            o.p1 = __uint_as_float(M);
            o.p2 = -1.0f;
            o.i1 = static_cast<int32_t>(I) - 1;
            o.nan = M > 0x7f800000u ? 1 : 0;
            a.rows[orow] = o;
          }
        }
      } else {
        // the two-barrier finish: deferring its second half to the next row's first barrier
        // measured 6% slower here (wave 0 then carries the flush into the next pass)
        row_finish<T>(ms, es, accn != accn, a.cw, red, a.rows + orow);
      }
    }
  } else {
    // sub-row s of the chunk: row r0 + s / SUB, output samples k0 + SUB m (k0 = s mod SUB).
    // The matrix-core variant keeps the sub-row's 33 pass-A inputs per thread in registers and
    // loads the next sub-row's right after pass A, so the HBM latency hides behind passes B, C.
    constexpr bool kPrefetch = VARIANT == 1;
    float2 yn[kPrefetch ? 33 : 1];
    auto load_sub = [&](long long sidx) {
      const __amdgpu_buffer_rsrc_t yv = buffer_view(a.Y + sidx * P, P * 8);
#pragma unroll
      for (int r = 0; r < 33; ++r)
        yn[r] = buffer_load_f2(yv, voa, r * NBA * 8);
    };
    if constexpr (kPrefetch) {
      if (static_cast<long long>(blockIdx.x) < a.n_sub) load_sub(blockIdx.x);
    }
    for (long long sidx = blockIdx.x; sidx < a.n_sub; sidx += gridDim.x) {
      const long long rq = sidx / SUB;
      const int k0 = static_cast<int>(sidx - rq * SUB);
      const long long r = big_row(a.r0 + rq, a.n_prn, a.n_bin, a.n_ant, a.hs, a.tile);
      const __amdgpu_buffer_rsrc_t yv = buffer_view(a.Y + sidx * P, P * 8);
      float2 yc[kPrefetch ? 33 : 1];
      if constexpr (kPrefetch) {
#pragma unroll
        for (int r = 0; r < 33; ++r) yc[r] = yn[r];
      }
      // the sub-row's argmax only: a running 64-bit key maximum per thread (a NaN power has
      // the largest key, so it reaches the row's key and is flagged there)
      unsigned long long km = 0ull;
      // kPairs: each output pair's larger power (its f32 bits: powers are >= 0, so the bits
      // order like the values and a NaN has the largest) with the difference of the pair's two
      // powers, and the thread's maximum over them
      unsigned int pmx = 0u;
      unsigned int pmb[kPairs ? 5 : 1][4];
      float pdd[kPairs ? 5 : 1][4];
      if constexpr (kPairs) {
#pragma unroll
        for (int ts = 0; ts < 5; ++ts)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            pmb[ts][e] = 0u;
            pdd[ts][e] = 0.0f;
          }
      }
      auto dump_at = [&](int idx, float p) {
        if constexpr (DUMP) a.dump[static_cast<size_t>(r) * (SUB * P) + k0 + SUB * idx] = p;
      };
      transform(
          [&](int ta, int r) {
            if constexpr (kPrefetch)
              return yc[r];
            else
              return buffer_load_f2(yv, ta * 8, r * NBA * 8);
          },
          [&](int, float p, int idx) {
            const unsigned long long k = pkey(p, idx);
            km = k > km ? k : km;
            dump_at(idx, p);
          },
          [&] {
            if constexpr (kPrefetch) {
              if (sidx + gridDim.x < a.n_sub) load_sub(sidx + gridDim.x);
            }
          },
          [&](int ts, int e, int q, int kk, float p1, float p2) {
            if constexpr (kPairs) {
              if (ts >= 5) return;  // tile slots 0..4 (wave 0 takes 5 tiles, the others 4)
              const unsigned int b = max(__float_as_uint(p1), __float_as_uint(p2));
              pmb[ts][e] = b;
              pdd[ts][e] = p1 - p2;
              pmx = max(pmx, b);
              if constexpr (DUMP) {
                dump_at(out_idx(q, kk), p1);
                if (kk > 0) dump_at(out_idx(q, 31 - kk), p2);
              }
            }
          });
      if constexpr (kPairs) {
        // block maximum M of the pair maxima (one barrier, which also orders this sub-row's
        // pass-C reads before the next sub-row's pass-A writes), then each wave holding M finds
        // the largest sample with power M among its pairs (p1 > p2: Y_kk, p1 < p2: Y_{31-kk},
        // equal: both) and folds (M, sample) into the row's key: several waves holding M fold
        // their own, and the atomic maximum keeps the last maximum
        const unsigned int wmax = wave_umax(pmx);
        if (lane == 0) red.pk[wave] = wmax;
        lds_barrier();
        unsigned int M;
        {
          const unsigned int wv = lane < NW ? red.pk[lane] : 0u;
          M = static_cast<unsigned int>(__builtin_amdgcn_readlane(
              static_cast<int>(row16_umax(wv)), 0));
        }
        const bool mine = pmx == M;
        if (__ballot(mine) != 0) {
          unsigned int cand = 0u;  // 1 + sample, 0 for none
          if (mine) {
            const int ntile = wave == 0 ? 5 : 4;
#pragma unroll
            for (int ts = 0; ts < 5; ++ts) {
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                if (ts < ntile && pmb[ts][e] == M) {
                  const int q = 16 * (wave + 8 * ts) + ln16;
                  const int kk = 4 * grp + e;
                  const float d = pdd[ts][e];
                  const int i1 = out_idx(q, kk);
                  const int i2 = kk > 0 ? out_idx(q, 31 - kk) : i1;
                  const int id = d > 0.0f ? i1 : d < 0.0f ? i2 : max(i1, i2);
                  cand = max(cand, static_cast<unsigned int>(id + 1));
                }
              }
            }
          }
          const unsigned int iw = wave_umax(cand);
          if (lane == 0) {
            atomicMax(a.rowkey + r, (static_cast<unsigned long long>(M) << 32) |
                                        static_cast<unsigned int>(k0 + SUB * (static_cast<int>(iw) - 1)));
            if (M > 0x7f800000u) atomicOr(a.rownan + r, 1);
          }
        }
      } else {
        row_finish_key<T>(km, (km >> 32) > 0x7f800000ull, red, k0, SUB, a.rowkey + r,
                          a.rownan + r);
      }
    }
  }
}


// ---- Sub-row transform of the long rows (FP32 split variant 2) ----------------------------------
// The sub-row transforms of k_rows<1, SUB> restructured for fewer barriers and less per-output
// work, with the same prime-factor passes (radix 33 from the prefetched registers, radix 16,
// radix 31 on the matrix cores):
//  * pass B runs in place (a butterfly's 16 outputs go back to its inputs' slots), so a
//    sub-row needs three workgroup barriers, not four;
//  * each lane keeps only its best output pair (the larger of |Y_kk|^2 and |Y_{31-kk}|^2, its
//    slot and which of the two), not a 64-bit key per output; an exact rescan runs only when a
//    lane's maximum occurs twice;
//  * the row index of each sub-row comes from the pre-pass's row map (no 64-bit divisions).
// LDS positions (sample-space indices n1, n2, n3 -> k1, k2, k3 of 16368 = 33 x 16 x 31):
//   after pass A   1023 n2 + 33 n3 + k1   (pass-A butterfly ta = 31 n2 + n3, output k1)
//   after pass B   1023 k2 + 33 n3 + k1   (pass-B butterfly j = 33 n3 + k1, output k2)
//   pass C         butterfly q = 33 k2 + k1 reads n3 = 0..30 at c31_pos(q) + 33 n3
// In pass C the four 16-lane groups of a wave take the symmetric input pairs (j, 31 - j) as
// j = 2s, 15 - 2s, 2s + 1, 14 - 2s at GEMM step s, the odd groups loading v_{31-j} first (their
// sine column negated), so the two groups of each half-wave load addresses 16 x 33 x 8 B apart,
// on disjoint LDS banks.
__device__ __forceinline__ int c31_pos(int q) { return q + 990 * (q / 33); }

// Pass C (radix 31) of k_subrows and k_rowsx, per lane: the GEMM constants of A row k = ln16
// and the LDS offsets of the lane's two loads at each of the 4 GEMM steps (input pairs in the
// bank-disjoint order above; step 0 of group 0 is j = 0, a_0 = v_0 alone).
struct C31Lane {
  float cA[4], sA[4];
  int off1[4], off2[4];
  bool z0;
  __device__ __forceinline__ void init(const float2* stw, int ln16, int grp) {
    const bool odd = (grp & 1) != 0;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const int j = grp == 0 ? 2 * st : grp == 1 ? 15 - 2 * st : grp == 2 ? 2 * st + 1 : 14 - 2 * st;
      const float2 w = tw_lds(stw, ((j * ln16) % 31) * (P / 31));  // exp(-2 pi i jk/31)
      const int m = j == 0 ? 0 : 31 - j;
      cA[st] = j == 0 ? 1.0f : w.x;
      sA[st] = j == 0 ? 0.0f : (odd ? w.y : -w.y);
      off1[st] = 33 * (odd ? m : j);
      off2[st] = 33 * (odd ? j : m);
    }
    z0 = grp == 0;
  }
  // the 16 MFMAs of the tile of butterflies q0..q0+15 (this lane: q = q0 + ln16): acc[0], acc[1]
  // = C a (re, im), acc[2], acc[3] = S b (re, im); row kk = 4 grp + e of the tile in element e
  __device__ __forceinline__ void tile(const float2* buf, int q, floatx4 (&acc)[4]) const {
    const int base = c31_pos(q);
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const float2 x1 = buf[base + off1[st]];
      float2 x2 = buf[base + off2[st]];
      if (st == 0 && z0) x2 = make_float2(0.f, 0.f);
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(cA[st], x1.x + x2.x, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(cA[st], x1.y + x2.y, acc[1], 0, 0, 0);
      acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(sA[st], x1.x - x2.x, acc[2], 0, 0, 0);
      acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(sA[st], x1.y - x2.y, acc[3], 0, 0, 0);
    }
  }
};

// powers p1 = |Y_kk|^2 = |A - iB|^2 and p2 = |Y_{31-kk}|^2 = |A + iB|^2 of row e of a tile's
// accumulators (each from its own output, so a small power carries no cancellation error)
__device__ __forceinline__ void c31_pair(const floatx4 (&acc)[4], int e, float& p1, float& p2) {
  const float ar = acc[0][e], ai = acc[1][e];
  const float br = acc[2][e], bi = acc[3][e];
  const float y1r = ar + bi, y1i = ai - br, y2r = ar - bi, y2i = ai + br;
  p1 = fmaf(y1r, y1r, y1i * y1i);
  p2 = fmaf(y2r, y2r, y2i * y2i);
}

// sample of the larger output of pair (butterfly q, row kk) with d = p1 - p2: Y_kk for d > 0,
// Y_{31-kk} for d < 0, the larger of the two samples for d = 0 (equal powers: last maximum)
__device__ __forceinline__ int c31_pair_sample(int q, int kk, float d) {
  const int i1 = pfa_wrap(pfa_base(q) + (P / 31) * kk);
  const int i2 = kk > 0 ? pfa_wrap(pfa_base(q) + (P / 31) * (31 - kk)) : i1;
  return d > 0.0f ? i1 : d < 0.0f ? i2 : max(i1, i2);
}

template <int SUB, bool DUMP = false, bool MASK = false>
__global__ __launch_bounds__(T) void k_subrows(RowArgs a) {
  constexpr int NBA = P / 33;  // 496
  constexpr int NBB = P / 16;  // 1023
  constexpr int NBC = P / 31;  // 528
  __shared__ float2 buf[P];
  __shared__ float2 stw[256];
  __shared__ RowRed<NW> red;
  const int tid = threadIdx.x;
  for (int i = tid; i < 256; i += T) stw[i] = a.tw[i];
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63, ln16 = lane & 15, grp = lane >> 4;
  C31Lane c31;
  c31.init(stw, ln16, grp);
  auto out_idx = [](int q, int r) { return pfa_wrap(pfa_base(q) + NBC * r); };
  // sample of pair (tile slot ts, GEMM row e) of this lane with d = p1 - p2
  auto pair_sample = [&](int ts, int e, float d) {
    return c31_pair_sample(16 * (wave + 8 * ts) + ln16, 4 * grp + e, d);
  };
  // lanes beyond the 496 pass-A butterflies load past the end of the view: 0, no branch
  const int voa = tid < NBA ? tid * 8 : P * 8;
  auto load_sub = [&](long long s, float2 (&y)[33]) {
    const __amdgpu_buffer_rsrc_t yv = buffer_view(a.Y + s * P, P * 8);
#pragma unroll
    for (int r = 0; r < 33; ++r) y[r] = buffer_load_f2(yv, voa, r * NBA * 8);
  };
  const long long G = gridDim.x;
  long long sidx = blockIdx.x;
  if (sidx >= a.n_sub) return;
  // `cur` holds the sub-row's pass-A inputs (loaded during the previous sub-row); they are
  // transformed and stored at the top, and the registers reloaded with the next sub-row's
  // inputs, which then have a whole sub-row to arrive.
  float2 cur[33];
  // the row index of the sub-row (big_row(a.r0 + rq, ...), from the pre-pass's map) is loaded
  // one sub-row ahead like its inputs, so its latency is not exposed where the atomic needs it
  int rmap = MASK ? 0 : a.rowmap[sidx / SUB];
  load_sub(sidx, cur);
  for (;;) {
    const long long s = sidx;
    const long long rq = s / SUB;
    const int k0 = static_cast<int>(s - rq * SUB);
    // the row index, or in the masked mode the winner and the centre of its exclusion window
    const long long r = MASK ? rq : __builtin_amdgcn_readfirstlane(rmap);
    int centre = 0;
    if constexpr (MASK) {
      const int ant = static_cast<int>(rq / a.n_prn);
      const PeakSlot sl = a.slots[rq - static_cast<long long>(ant) * a.n_prn];
      centre = static_cast<int>(
          a.peaks[sl.base + static_cast<long long>(ant) * sl.stride + sl.local].best_codephase %
          static_cast<uint64_t>(SUB * P));
    }
    // ---- pass A (radix 33) -> LDS; the next sub-row's inputs in flight
    Dft<33>::run(cur);
    {
      int ta = tid;
      asm volatile("" : "+v"(ta));
      if (ta < NBA) {
#pragma unroll
        for (int q = 0; q < 33; ++q) buf[ta * 33 + q] = cur[q];
      }
    }
    if (s + G < a.n_sub) {
      if constexpr (!MASK) rmap = a.rowmap[(s + G) / SUB];
      load_sub(s + G, cur);
    }
    lds_barrier();
    // ---- pass B (radix 16) in place
    {
      int tb = tid;
      asm volatile("" : "+v"(tb));
      // one butterfly after the other (the memory clobber keeps the second one's loads after
      // the first one's stores): 32 VGPRs of values instead of 64
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int j = tb + b * T;
        if (b == 0 || j < NBB) {
          float2 v[16];
#pragma unroll
          for (int q = 0; q < 16; ++q) v[q] = buf[j + q * NBB];
          Dft<16>::run(v);
#pragma unroll
          for (int q = 0; q < 16; ++q) buf[j + q * NBB] = v[q];
        }
        asm volatile("" ::: "memory");
      }
    }
    lds_barrier();
    // ---- pass C (radix 31 on the matrix cores).  Each lane keeps its largest pair maximum max(p1, p2) (f32 bits: powers are >= 0, so the
    // bits order like the values and a NaN has the largest), the pair's slot 4 ts + e and
    // p1 - p2; `tmask` marks the lanes whose maximum another of their pairs equals (no branch here:
    // such a tie is resolved at the end of the sub-row by an exact rescan).
    unsigned int pmx = 0u;
    int psl = 0;
    float pdv = 0.0f;
    unsigned long long tmask = 0ull;
    int gc = grp;
    asm volatile("" : "+v"(gc));
    auto dump_at = [&](int idx, float p) {
      if constexpr (DUMP) a.dump[static_cast<size_t>(r) * (SUB * P) + k0 + SUB * idx] = p;
    };
    // the 16 MFMAs of tile slot TS (tile wave + 8 TS) into acc
    auto tile_mfma = [&](auto ts_c, floatx4 (&acc)[4]) {
      c31.tile(buf, 16 * (wave + 8 * decltype(ts_c)::value) + ln16, acc);
    };
    // one tile: its MFMAs, then its outputs' pair maxima
    auto tile = [&](auto ts_c) {
      constexpr int TS = decltype(ts_c)::value;
      floatx4 acc[4];
      tile_mfma(ts_c, acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float p1, p2;
        c31_pair(acc, e, p1, p2);
        // the pair's maximum (bits: a NaN has the largest); d > 0: Y_kk, d < 0: Y_{31-kk}
        if constexpr (MASK) {
          // the largest power outside the window of `centre` (ring distance > cw)
          const int q = 16 * (wave + 8 * TS) + ln16;
          const int kk = 4 * grp + e;
          const int sa = k0 + SUB * out_idx(q, kk);
          const int sb = k0 + SUB * out_idx(q, 31 - kk);
          const unsigned int ua = ring_dist(sa, centre, SUB * P) > a.cw ? __float_as_uint(p1) : 0u;
          const unsigned int ub =
              kk > 0 && ring_dist(sb, centre, SUB * P) > a.cw ? __float_as_uint(p2) : 0u;
          pmx = max(pmx, max(ua, ub));
          continue;
        }
        const unsigned int bits = max(__float_as_uint(p1), __float_as_uint(p2));
        const float d = p1 - p2;
        const bool up = bits > pmx;
        tmask = (tmask | __ballot(bits == pmx)) & ~__ballot(up);
        pmx = up ? bits : pmx;
        psl = up ? 4 * TS + e : psl;
        pdv = up ? d : pdv;
        if constexpr (DUMP) {
          const int q = 16 * (wave + 8 * TS) + ln16;
          const int kk = 4 * gc + e;
          dump_at(out_idx(q, kk), p1);
          if (kk > 0) dump_at(out_idx(q, 31 - kk), p2);
        }
      }
    };
    tile(std::integral_constant<int, 0>{});
    tile(std::integral_constant<int, 1>{});
    tile(std::integral_constant<int, 2>{});
    tile(std::integral_constant<int, 3>{});
    if (wave == 0) tile(std::integral_constant<int, 4>{});  // tile 32
    // ---- end of the sub-row: block maximum M of the lanes' maxima (the barrier also orders
    // this sub-row's pass-C loads before the next sub-row's pass-A store); the waves holding M
    // fold (M, largest sample with power M) into the row's key (several waves: several atomic
    // maxima, the largest sample wins)
    const unsigned int wmax = wave_umax(pmx);
    const bool wtie = (__ballot(pmx == wmax) & tmask) != 0ull;
    if (lane == 0) {
      red.pk[wave] = wmax;
      red.bad[wave] = wtie ? 1 : 0;
    }
    lds_barrier();
    const bool mine = lane < NW;
    const unsigned int M = static_cast<unsigned int>(__builtin_amdgcn_readlane(
        static_cast<int>(row16_umax(mine ? red.pk[lane] : 0u)), 0));
    if constexpr (MASK) {
      if (tid == 0) atomicMax(a.m2key + r, M);
      sidx += G;
      if (sidx >= a.n_sub) break;
      continue;
    }
    const bool tie = __ballot(mine && red.pk[mine ? lane : 0] == M &&
                              red.bad[mine ? lane : 0] != 0) != 0ull;
    unsigned int cand = 0u;  // 1 + sample, 0 for none
    if (tie) {
      // Rare (block-uniform): a lane holding M reached it in more than one pair.  Recompute the
      // wave's tiles (the pass-B values are still in LDS: no wave stores the next sub-row's
      // pass A before the barrier below) and take every pair of power M.
      if (wmax == M) {
        auto rescan = [&](auto ts_c) {
          constexpr int TS = decltype(ts_c)::value;
          floatx4 acc[4];
          tile_mfma(ts_c, acc);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            float p1, p2;
            c31_pair(acc, e, p1, p2);
            if (max(__float_as_uint(p1), __float_as_uint(p2)) == M)
              cand = max(cand, static_cast<unsigned int>(pair_sample(TS, e, p1 - p2) + 1));
          }
        };
        rescan(std::integral_constant<int, 0>{});
        rescan(std::integral_constant<int, 1>{});
        rescan(std::integral_constant<int, 2>{});
        rescan(std::integral_constant<int, 3>{});
        if (wave == 0) rescan(std::integral_constant<int, 4>{});
      }
      lds_barrier();
    } else if (pmx == M) {
      cand = static_cast<unsigned int>(pair_sample(psl >> 2, psl & 3, pdv) + 1);
    }
    if (wmax == M) {
      const unsigned int iw = wave_umax(cand);
      if (lane == 0) {
        atomicMax(a.rowkey + r,
                  (static_cast<unsigned long long>(M) << 32) |
                      static_cast<unsigned int>(k0 + SUB * (static_cast<int>(iw) - 1)));
        if (M > 0x7f800000u) atomicOr(a.rownan + r, 1);
      }
    }
    sidx += G;
    if (sidx >= a.n_sub) break;
  }
}

// ---- Half sub-rows of the long rows (FP32 split variant 5) ------------------------------------
// P = R0 x 16368 = R x 8184 with R = 2 R0.  With n = n1 + 8184 n0 and k = k0 + R k1,
//   Z[k0 + R k1] = FFT_8184( W_P^{n1 k0} sum_{n0 < R} z[n1 + 8184 n0] W_R^{n0 k0} )[k1],
// so a long row becomes R sub-rows of 8184 points: 65 KB of f32 complex values, so that two
// transforming workgroups share one CU's LDS and overlap their passes (a 16368-point sub-row
// takes the whole of it, and its workgroup's barriers leave the SIMDs idle).
// 8184 = 33 x 8 x 31 with pairwise coprime factors.  A half sub-row is stored in the CRT order:
// position s = 248 c1 + 31 c2 + c3 holds natural index m with m mod 33 = c1, m mod 8 = c2,
// m mod 31 = c3, i.e. m = (496 c1 + 7161 c2 + 528 c3) mod 8184, and output (k1, k2, k3) of the
// three passes (radix 33 over c1, 8 over c2, 31 over c3) is sample (248 k1 + 1023 k2 + 264 k3)
// mod 8184 (Good-Thomas: no twiddles between the passes).  The spectra keep the 16368-segment
// prime-factor order of the other split forms, where natural n1 + 8184 b of a segment sits at
// 496 c1 + 31 ((-n1 - 8 b) mod 16) + c3: the two values of half position s in every segment are
// the pair lo = 496 c1 + 31 ((-c2) mod 8) + c3 and lo + 248, b = 0 being lo exactly when
// (-n1) mod 16 < 8, so the pre-pass reads runs of 31 contiguous values.
constexpr int kHalfS = kFusedP / 2;  // 8184
constexpr int TH = 256;
constexpr int NWH = TH / 64;

__device__ __forceinline__ int half_nat(int s) {
  const int c1 = s / 248, b = s - 248 * c1, c2 = b / 31, c3 = b - 31 * c2;
  return (496 * c1 + 7161 * c2 + 528 * c3) % kHalfS;
}
// sample of output k3 = 0 of pass-C butterfly q = 33 k2 + k1 (add 264 k3, wrap once)
__device__ __forceinline__ int half_base(int q) {
  const int k2 = q / 33, k1 = q - 33 * k2;
  return (248 * k1 + 1023 * k2) % kHalfS;
}
__device__ __forceinline__ int half_wrap(int k) { return k >= kHalfS ? k - kHalfS : k; }
// sample of the larger output of pair (butterfly q, row kk) with d = p1 - p2 (c31_pair_sample)
__device__ __forceinline__ int half_pair_sample(int q, int kk, float d) {
  const int i1 = half_wrap(half_base(q) + (kHalfS / 31) * kk);
  const int i2 = kk > 0 ? half_wrap(half_base(q) + (kHalfS / 31) * (31 - kk)) : i1;
  return d > 0.0f ? i1 : d < 0.0f ? i2 : max(i1, i2);
}

// Pre-pass of a chunk of rows into half sub-rows: 32 blocks of 256 positions per row (the XCD
// dealing of k_big_pre); per position s, conj(X) R at the R = 2 R0 natural points m + 8184 n0
// (the lo / lo + 248 pair of every segment), an R-point DFT, the twiddles W_P^{m k0} from two
// table values W_P^m, W_P^{R0 m} (at most R0 products each: relative error below 1e-6), and R
// non-temporal stores into the row's R sub-rows; position 0 clears the row's key and records it
// in the chunk's row map.
template <int R0>
__global__ __launch_bounds__(256) void k_half_pre(const float2* __restrict__ X,
                                                  const float2* __restrict__ rep,
                                                  const float2* __restrict__ twh,
                                                  const int4* __restrict__ rowtab,
                                                  long long r0, long long n_rows,
                                                  float2* __restrict__ Y,
                                                  unsigned long long* __restrict__ rowkey,
                                                  int* __restrict__ rownan,
                                                  int* __restrict__ rowmap) {
  constexpr int R = 2 * R0;
  constexpr long long PL = static_cast<long long>(R0) * kFusedP;
  const long long rr = blockIdx.x >> 5;
  if (rr >= n_rows) return;
  const int chunk = 4 * static_cast<int>(blockIdx.x & 7) + static_cast<int>((blockIdx.x >> 3) & 3);
  const int s = 256 * chunk + static_cast<int>(threadIdx.x);
  if (s >= kHalfS) return;
  const int4 ref = rowtab[r0 + rr];  // (X row, replica row, RowOut row), block-uniform
  const float2* xr = X + static_cast<long long>(ref.x) * PL;
  const float2* cr = rep + static_cast<long long>(ref.y) * PL;
  const int c1 = s / 248, b = s - 248 * c1, c2 = b / 31, c3 = b - 31 * c2;
  const int m = (496 * c1 + 7161 * c2 + 528 * c3) % kHalfS;
  const int lo = 496 * c1 + 31 * ((8 - c2) & 7) + c3;
  const bool swap = ((-m) & 15) >= 8;  // natural m of each segment sits at lo + 248
  float2 v[R];
#pragma unroll
  for (int g = 0; g < R0; ++g) {
    const float2 pl = cconjmul(xr[kFusedP * g + lo], cr[kFusedP * g + lo]);
    const float2 ph = cconjmul(xr[kFusedP * g + lo + 248], cr[kFusedP * g + lo + 248]);
    v[2 * g] = swap ? ph : pl;
    v[2 * g + 1] = swap ? pl : ph;
  }
  Dft<R>::run(v);
  const float2 w1 = twh[s], wh = twh[kHalfS + s];  // W_P^m, W_P^{R0 m}
  v[R0] = cmul(v[R0], wh);
  float2 wk = w1;
#pragma unroll
  for (int k0 = 1; k0 < R0; ++k0) {
    v[k0] = cmul(v[k0], wk);
    v[k0 + R0] = cmul(v[k0 + R0], cmul(wk, wh));
    if (k0 + 1 < R0) wk = cmul(wk, w1);
  }
  float* y = reinterpret_cast<float*>(Y + rr * PL + s);
#pragma unroll
  for (int k0 = 0; k0 < R; ++k0) {
    __builtin_nontemporal_store(v[k0].x, y + 2 * kHalfS * k0);
    __builtin_nontemporal_store(v[k0].y, y + 2 * kHalfS * k0 + 1);
  }
  if (s == 0) {
    rowkey[ref.z] = 0ull;
    rownan[ref.z] = 0;
    rowmap[rr] = ref.z;
  }
}

// Transform of the half sub-rows (the passes of k_subrows on 8184 points): 256 threads, two
// workgroups per CU within 256 VGPRs and 68 KB of LDS each.
//   pass A  radix 33 (codelet), 248 butterflies ta = 31 c2 + c3 from the prefetched registers
//           (inputs c1 at ta + 248 c1) -> LDS 33 ta + k1 = 1023 c2 + 33 c3 + k1
//   pass B  radix 8 in place, butterflies j = 33 c3 + k1 < 1023 over c2 at j + 1023 c2
//   pass C  radix 31 on the matrix cores (C31Lane), butterflies q = 33 k2 + k1 < 264 at
//           c31_pos(q) + 33 c3: tiles w + 4 ts of wave w, and tile 16 (butterflies 256..263 in
//           lanes 0..7; its other 8 columns are ignored) on wave 0
// Each lane keeps its best output pair; an exact rescan runs only on ties (as k_subrows); one
// atomic maximum per sub-row folds (power, k0 + R sample) into the row's key.
template <int R, bool DUMP = false>
__global__ __launch_bounds__(TH, 2) void k_halfrows(RowArgs a) {
  constexpr int S = kHalfS;
  constexpr int NBA = S / 33;  // 248
  constexpr int NBB = S / 8;   // 1023
  constexpr int NBC = S / 31;  // 264
  __shared__ float2 buf[S];
  __shared__ float2 stw[256];
  __shared__ RowRed<NWH> red;
  const int tid = threadIdx.x;
  for (int i = tid; i < 256; i += TH) stw[i] = a.tw[i];
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63, ln16 = lane & 15, grp = lane >> 4;
  C31Lane c31;
  c31.init(stw, ln16, grp);
  // lanes beyond the 248 pass-A butterflies load past the end of the view: 0, no branch
  const int voa = tid < NBA ? tid * 8 : S * 8;
  auto load_sub = [&](long long s, float2 (&y)[33]) {
    const __amdgpu_buffer_rsrc_t yv = buffer_view(a.Y + s * S, S * 8);
#pragma unroll
    for (int r = 0; r < 33; ++r) y[r] = buffer_load_f2(yv, voa, r * NBA * 8);
  };
  auto c_tile = [&](int ts) { return ts < 4 ? wave + 4 * ts : 16; };
  const long long G = gridDim.x;
  long long sidx = blockIdx.x;
  if (sidx >= a.n_sub) return;
  // the sub-row's row index (from the pre-pass's map) and inputs, loaded one sub-row ahead
  int rmap = a.rowmap[sidx / R];
  float2 cur[33];
  load_sub(sidx, cur);
  for (;;) {
    const long long s = sidx;
    const long long rq = s / R;
    const int k0 = static_cast<int>(s - rq * R);
    const long long r = __builtin_amdgcn_readfirstlane(rmap);
    // ---- pass A (radix 33) -> LDS; the next sub-row's inputs in flight
    Dft<33>::run(cur);
    {
      int ta = tid;
      asm volatile("" : "+v"(ta));
      if (ta < NBA) {
#pragma unroll
        for (int q = 0; q < 33; ++q) buf[ta * 33 + q] = cur[q];
      }
    }
    if (s + G < a.n_sub) {
      rmap = a.rowmap[(s + G) / R];
      load_sub(s + G, cur);
    }
    lds_barrier();
    // ---- pass B (radix 8) in place, butterflies tid + 256 i
    {
      int tb = tid;
      asm volatile("" : "+v"(tb));
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int j = tb + i * TH;
        if (i < 3 || j < NBB) {
          float2 v[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = buf[j + q * NBB];
          Dft<8>::run(v);
#pragma unroll
          for (int q = 0; q < 8; ++q) buf[j + q * NBB] = v[q];
        }
      }
    }
    lds_barrier();
    // ---- pass C (radix 31, matrix cores): each lane's best pair as in k_subrows; the columns
    // of tile 16 beyond butterfly 263 take no part
    unsigned int pmx = 0u;
    int psl = 0;
    float pdv = 0.0f;
    unsigned long long tmask = 0ull;
    auto dump_at = [&](int idx, float p) {
      if constexpr (DUMP) a.dump[static_cast<size_t>(r) * (R * S) + k0 + R * idx] = p;
    };
    auto tile = [&](int ts) {
      floatx4 acc[4];
      const int q = 16 * c_tile(ts) + ln16;
      const bool valid = q < NBC;
      c31.tile(buf, valid ? q : NBC - 1, acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float p1, p2;
        c31_pair(acc, e, p1, p2);
        const unsigned int bits = max(__float_as_uint(p1), __float_as_uint(p2));
        const float d = p1 - p2;
        const bool up = valid && bits > pmx;
        tmask = (tmask | __ballot(valid && bits == pmx)) & ~__ballot(up);
        pmx = up ? bits : pmx;
        psl = up ? 4 * ts + e : psl;
        pdv = up ? d : pdv;
        if constexpr (DUMP) {
          if (valid) {
            const int kk = 4 * grp + e;
            dump_at(half_wrap(half_base(q) + NBC * kk), p1);
            if (kk > 0) dump_at(half_wrap(half_base(q) + NBC * (31 - kk)), p2);
          }
        }
      }
    };
    tile(0);
    tile(1);
    tile(2);
    tile(3);
    if (wave == 0) tile(4);
    // ---- end of the sub-row: block maximum, the waves holding it fold (M, largest sample with
    // power M) into the row's key (the barrier also orders this sub-row's pass-C loads before
    // the next sub-row's pass-A stores)
    const unsigned int wmax = wave_umax(pmx);
    const bool wtie = (__ballot(pmx == wmax) & tmask) != 0ull;
    if (lane == 0) {
      red.pk[wave] = wmax;
      red.bad[wave] = wtie ? 1 : 0;
    }
    lds_barrier();
    const bool mine = lane < NWH;
    const unsigned int M = static_cast<unsigned int>(__builtin_amdgcn_readlane(
        static_cast<int>(row16_umax(mine ? red.pk[lane] : 0u)), 0));
    const bool tie = __ballot(mine && red.pk[mine ? lane : 0] == M &&
                              red.bad[mine ? lane : 0] != 0) != 0ull;
    unsigned int cand = 0u;  // 1 + sample, 0 for none
    if (tie) {
      // rare (block-uniform): a lane holding M reached it in more than one pair; the waves
      // holding M recompute their tiles from the pass-B values still in LDS
      if (wmax == M) {
        int l16 = ln16;
        asm volatile("" : "+v"(l16));
        auto rescan = [&](int ts) {
          floatx4 acc[4];
          const int q = 16 * c_tile(ts) + l16;
          const bool valid = q < NBC;
          c31.tile(buf, valid ? q : NBC - 1, acc);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            float p1, p2;
            c31_pair(acc, e, p1, p2);
            if (valid && max(__float_as_uint(p1), __float_as_uint(p2)) == M)
              cand = max(cand,
                         static_cast<unsigned int>(half_pair_sample(q, 4 * grp + e, p1 - p2) + 1));
          }
        };
        rescan(0);
        rescan(1);
        rescan(2);
        rescan(3);
        if (wave == 0) rescan(4);
      }
      lds_barrier();
    } else if (pmx == M) {
      int l16 = ln16;
      asm volatile("" : "+v"(l16));
      cand = static_cast<unsigned int>(
          half_pair_sample(16 * c_tile(psl >> 2) + l16, 4 * grp + (psl & 3), pdv) + 1);
    }
    if (wmax == M) {
      const unsigned int iw = wave_umax(cand);
      if (lane == 0) {
        atomicMax(a.rowkey + r,
                  (static_cast<unsigned long long>(M) << 32) |
                      static_cast<unsigned int>(k0 + R * (static_cast<int>(iw) - 1)));
        if (M > 0x7f800000u) atomicOr(a.rownan + r, 1);
      }
    }
    sidx += G;
    if (sidx >= a.n_sub) break;
  }
}

// ---- Forward spectra of long code periods (P = R0 x 16368), four-step through LDS ------------
// With n = R0 m + r (m < 16368, r < R0) and k = t + 16368 k0 (t < 16368, k0 < R0),
//   X[t + 16368 k0] = sum_r W_R0^{r k0} (W_P^{r t} Y_r[t]),   Y_r = FFT_16368(x[R0 m + r]),
// so step 1 (k_fwd_long) runs one LDS transform per (antenna, bin, fold, r) of the decimated,
// decoded, wiped and folded samples, and step 2 (k_fwd_twist) the twiddles and the R0-point
// DFTs, writing X segment by segment in the prime-factor order the split's pre-pass reads
// (position 16368 k0 + s holds natural index 16368 k0 + pfa_nat(s)).  The carrier comes from
// a plan-time table in the decimated order, [bin][half][r][m] for sample n = half P + R0 m + r
// (same rounding as k_carrier), and the twist tau_n = exp(-i pi n / P) from [r][m].
struct LongFwdArgs {
  const float* lev;       // [n_ant][nh][R0][16368] f32(level - mean) (k_levels_dec)
  const float2* carrier;  // [n_bin][nh][R0][16368]
  const float2* tau;      // [R0][16368] (folded rows)
  const float2* tw;       // fused twiddles of the LDS passes
  int n_ant;
  int n_bin;
  float2* Y;  // [xrow][R0][16368] in the prime-factor order of 16368
};

template <int R0, bool FOLD>
__global__ __launch_bounds__(T) void k_fwd_long(LongFwdArgs a) {
  constexpr int NH = FOLD ? 2 : 1;
  constexpr int NBA = P / 33;  // 496
  constexpr int NBB = P / 16;  // 1023
  // the natural-order row and the output scatter use a padded layout, one slot in 17
  // (pad(m) = m + m / 16), so the gathers at lane strides 528 and 496 and the scatter at 496
  // fall on distinct banks; the passes in between use the first P slots unpadded
  __shared__ float2 buf[P + P / 16 + 1];
  __shared__ float2 stw[256];
  auto pad = [](int m) { return m + (m >> 4); };
  const int tid = threadIdx.x;
  for (int i = tid; i < 256; i += T) stw[i] = a.tw[i];
  // blocks (bin, r, ant, h), ant and h fastest so that the blocks reading one slice of the
  // carrier table run together; each XCD takes one contiguous share of them
  const int nblk = a.n_bin * R0 * a.n_ant * NH;
  const int per_xcd = (nblk + 7) / 8;
  const int lb = static_cast<int>(blockIdx.x & 7) * per_xcd + static_cast<int>(blockIdx.x >> 3);
  if (lb >= nblk) return;
  const int h = FOLD ? lb % NH : 0;
  const int ant = (lb / NH) % a.n_ant;
  const int br = lb / (NH * a.n_ant);
  const int r = br % R0;
  const int bin = br / R0;
  const float* l0p = a.lev + (static_cast<size_t>(ant) * NH * R0 + r) * P;
  const float2* c0 = a.carrier + (static_cast<size_t>(bin) * NH * R0 + r) * P;
  const float2* tau = a.tau + static_cast<size_t>(r) * P;
  // ---- wipe, fold / twist: the decimated samples in natural order, eight per thread at a
  // time so that their loads are in flight together
  for (int m0 = tid; m0 < P; m0 += 8 * T) {
    float2 w0[8], w1[8], tw[8];
    float v0[8], v1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int m = m0 + u * T;
      const int mm = m < P ? m : 0;
      v0[u] = l0p[mm];
      w0[u] = c0[mm];
      if constexpr (FOLD) {
        v1[u] = l0p[static_cast<size_t>(R0) * P + mm];
        w1[u] = c0[static_cast<size_t>(R0) * P + mm];
        tw[u] = tau[mm];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int m = m0 + u * T;
      if (m < P) {
        const float l0 = v0[u];
        const float2 x0 = make_float2(w0[u].x * l0, w0[u].y * l0);
        if constexpr (FOLD) {
          const float l1 = v1[u];
          const float2 x1 = make_float2(w1[u].x * l1, w1[u].y * l1);
          buf[pad(m)] = h == 0 ? cadd(x0, x1) : cmul(csub(x0, x1), tw[u]);
        } else {
          buf[pad(m)] = x0;
        }
      }
    }
  }
  __syncthreads();
  // ---- the 16368-point DFT in the prime-factor order of the rows kernels (33 x 16 x 31, no
  // twiddles).  Pass A: butterfly ta = 31 n2 + n3 gathers its inputs at natural indices
  // (496 n1 + 1023 n2 + 528 n3) mod P, all before any thread stores its outputs at
  // 1023 n2 + 33 n3 + k1 (the layout of k_subrows after pass A).
  C31Lane c31;
  c31.init(stw, tid & 15, (tid & 63) >> 4);
  {
    int ta = tid;
    asm volatile("" : "+v"(ta));
    float2 v[33];
    if (ta < NBA) {
      const int n2 = ta / 31, n3 = ta - 31 * n2;
      int m = (1023 * n2 + 528 * n3) % P;
#pragma unroll
      for (int n1 = 0; n1 < 33; ++n1) {
        v[n1] = buf[pad(m)];
        m += 496;
        m = m >= P ? m - P : m;
      }
    }
    __syncthreads();
    if (ta < NBA) {
      Dft<33>::run(v);
#pragma unroll
      for (int q = 0; q < 33; ++q) buf[ta * 33 + q] = v[q];
    }
  }
  __syncthreads();
  // pass B (radix 16) in place
  {
    int tb = tid;
    asm volatile("" : "+v"(tb));
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = tb + b * T;
      if (b == 0 || j < NBB) {
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = buf[j + q * NBB];
        Dft<16>::run(v);
#pragma unroll
        for (int q = 0; q < 16; ++q) buf[j + q * NBB] = v[q];
      }
      asm volatile("" ::: "memory");
    }
  }
  __syncthreads();
  // pass C (radix 31 on the matrix cores): the lane's outputs Y_kk = A - iB and
  // Y_{31-kk} = A + iB of its tiles (tile wave + 8 ts), kept until every pass-C load is done
  const int wave = tid >> 6, lane = tid & 63, ln16 = lane & 15, grp = lane >> 4;
  const int ntile = wave == 0 ? 5 : 4;
  float2 yo[5][4][2];
#pragma unroll
  for (int ts = 0; ts < 5; ++ts) {
    if (ts < ntile) {
      floatx4 acc[4];
      c31.tile(buf, 16 * (wave + 8 * ts) + ln16, acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float ar = acc[0][e], ai = acc[1][e], br_ = acc[2][e], bi = acc[3][e];
        yo[ts][e][0] = make_float2(ar + bi, ai - br_);
        yo[ts][e][1] = make_float2(ar - bi, ai + br_);
      }
    }
  }
  __syncthreads();
  // output k (k mod 33 = q mod 33, k mod 16 = q / 33, k mod 31 = kk) goes to its prime-factor
  // position pfa_pos(k) = 496 (k mod 33) + 31 ((-k) mod 16) + (k mod 31)
#pragma unroll
  for (int ts = 0; ts < 5; ++ts) {
    if (ts < ntile) {
      const int q = 16 * (wave + 8 * ts) + ln16;
      const int k2 = q / 33, k1 = q - 33 * k2;
      const int base = 496 * k1 + 31 * ((16 - k2) & 15);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int kk = 4 * grp + e;
        buf[pad(base + kk)] = yo[ts][e][0];
        if (kk > 0) buf[pad(base + 31 - kk)] = yo[ts][e][1];
      }
    }
  }
  __syncthreads();
  const size_t xrow = (static_cast<size_t>(ant) * a.n_bin + bin) * NH + h;
  float2* out = a.Y + (xrow * R0 + r) * P;
  for (int s2 = tid; s2 < P; s2 += T) out[s2] = buf[pad(s2)];
}

// Step 2: per position s of a row (natural index t = pfa_nat(s)), the R0 values Y_r[s] times
// W_P^{r t} (from the two-level table twl: W_P^e = twl[e & 1023] twl[1024 + (e >> 10)]), an
// R0-point DFT, and X[xrow][k0][s]; two positions per thread with 16-byte accesses.
template <int R0>
__global__ __launch_bounds__(256) void k_fwd_twist(const float2* __restrict__ Y,
                                                   const float2* __restrict__ twl,
                                                   long long n_pairs, float2* __restrict__ X) {
  constexpr int PL = R0 * P;
  const long long gid = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (gid >= n_pairs) return;
  const long long xrow = gid / (P / 2);
  const int s = 2 * static_cast<int>(gid - xrow * (P / 2));
  const float4* y = reinterpret_cast<const float4*>(Y + xrow * PL + s);
  const int t0 = pfa_nat(s), t1 = pfa_nat(s + 1);
  float2 v[R0], u[R0];
  int e0 = 0, e1 = 0;  // r t mod PL
#pragma unroll
  for (int r = 0; r < R0; ++r) {
    const float4 q = y[static_cast<long long>(r) * (P / 2)];
    v[r] = make_float2(q.x, q.y);
    u[r] = make_float2(q.z, q.w);
    if (r > 0) {
      v[r] = cmul(v[r], cmul(twl[e0 & 1023], twl[1024 + (e0 >> 10)]));
      u[r] = cmul(u[r], cmul(twl[e1 & 1023], twl[1024 + (e1 >> 10)]));
    }
    e0 += t0;
    e0 = e0 >= PL ? e0 - PL : e0;
    e1 += t1;
    e1 = e1 >= PL ? e1 - PL : e1;
  }
  Dft<R0>::run(v);
  Dft<R0>::run(u);
  float4* x = reinterpret_cast<float4*>(X + xrow * PL + s);
#pragma unroll
  for (int k0 = 0; k0 < R0; ++k0)
    x[static_cast<long long>(k0) * (P / 2)] = make_float4(v[k0].x, v[k0].y, u[k0].x, u[k0].y);
}

// f32(level - mean) of sample n = half P + R0 m + r of every antenna in the decimated order
// [ant][half][r][m] of k_fwd_long, decoded once per search (every bin's blocks read it).
__global__ void k_levels_dec(const uint8_t* __restrict__ packed, size_t row_bytes, int fmt,
                             const double* __restrict__ means, int R0, int nh, long long total,
                             float* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= total) return;
  const long long per_ant = static_cast<long long>(nh) * R0 * P;
  const long long ant = i / per_ant;
  const long long j = i - ant * per_ant;
  const long long half = j / (static_cast<long long>(R0) * P);
  const long long k = j - half * R0 * P;
  const long long r = k / P, m = k - r * P;
  const long long n = half * R0 * P + R0 * m + r;
  out[i] = __double2float_rn(
      static_cast<double>(decode_level(packed + ant * row_bytes, fmt, n)) - means[ant]);
}

// Carrier (cos, sin) of sample n = half P + R0 m + r in the decimated order [bin][half][r][m]
// of k_fwd_long, rounded like k_carrier (f32 phase, or the f64 phase of the FP16 form).
template <bool PHASE64>
__global__ void k_carrier_dec(const double* __restrict__ fc, double pc, int R0, int nh,
                              long long total, float2* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= total) return;
  const long long per_bin = static_cast<long long>(nh) * R0 * P;
  const long long bin = i / per_bin;
  const long long j = i - bin * per_bin;
  const long long half = j / (static_cast<long long>(R0) * P);
  const long long k = j - half * R0 * P;
  const long long r = k / P, m = k - r * P;
  const long long n = half * R0 * P + R0 * m + r;
  const double f = fc[bin];
  if constexpr (PHASE64) {
    out[i] = cossin_f64(__dmul_rn(__dmul_rn(pc, static_cast<double>(n)), f));
  } else {
    out[i] = wipe_value(1.0f, __dmul_rn(pc, static_cast<double>(n)), f);
  }
}

// ---- FP16 form of the rows kernel ------------------------------------------------------------
// Operands X and R are stored as f16 (re, im) pairs scaled by kF16Scale, conj(X) . R is formed
// with packed f16 arithmetic, and every radix-R pass is a real GEMM on the matrix cores
// (v_mfma_f32_16x16x32_f16, f32 accumulation):
//   [Re Y_q; Im Y_q] = sum_r [[wr, -wi], [wi, wr]]_{qr} [Re v_r; Im v_r],  wr + i wi = W_R^{qr}
// rows 2q / 2q+1, columns 2r / 2r+1, zero-padded to 16-row tiles and 32-column k-blocks:
// radix 33 -> 5 x 3 fragments, radix 16 -> 2 x 1, radix 31 -> 4 x 2.  A B tile is 16
// butterflies: lane l supplies butterfly l%16 and k-rows 8(l/16)..+7 of a k-block (inputs
// r = 4(l/16)..+3), and receives D rows 4(l/16)..+3 of a tile (outputs q = 2(l/16), 2(l/16)+1
// of each 8-output tile).  Rows and sub-rows are stored in the prime-factor order of
// 16368 = 33 x 16 x 31, so the pass outputs are rounded to f16 in LDS with no twiddle product.
// Pass C (31) keeps its outputs in registers: output q of butterfly j is sample
// pfa_base(j) + 528 q (mod P), and |.|^2 feeds the same top-2 sets as the FP32 form.
struct Row16Args {
  const uint32_t* X;    // [xrow][P] f16 pairs
  const uint32_t* rep;  // [prn][2][P] f16 pairs
  const float2* tw;
  int n_bin;
  int n_prn;
  int groups;
  int per_group;
  int cw;
  RowOut* rows;
  // sub-row form (SUB > 0, long rows, see k_big_pre): the pass-A operands are the f16 words
  // of sub-row s of the chunk, the results go to the row keys like k_rows<0, SUB>
  const uint32_t* Y = nullptr;  // [sub-row][P] f16 pairs
  long long n_sub = 0, r0 = 0;
  unsigned long long* rowkey = nullptr;
  int* rownan = nullptr;
  int hs = 1;
  int n_ant = 0;  // sub-row mode: antennas (for the tiled row order)
  int tile = 0;   // sub-row mode: tile shape of the row order (big_item)
  long long n_items = 0;  // SUB == 0: (xrow, PRN group) items, n_xrow * groups
  float* dump = nullptr;  // surface form: as RowArgs::dump
  const int* rowmap = nullptr;  // sub-row mode: as RowArgs::rowmap
};

// conj(x) * r on f16 pairs (two packed instructions)
__device__ __forceinline__ uint32_t conjmul16(uint32_t xw, uint32_t rw) {
  const half2_t x = __builtin_bit_cast(half2_t, xw);
  const half2_t r = __builtin_bit_cast(half2_t, rw);
  const half2_t a = half2_t{x.x, x.x} * r;
  const half2_t b = half2_t{r.y, -r.x};
  return __builtin_bit_cast(uint32_t, a + half2_t{x.y, x.y} * b);
}

// Fragment (tile t, k-block kb) of the real DFT matrix of radix R held by this lane.
template <int R>
__device__ __forceinline__ halfx8 dft_frag(const float2* stw, int t, int kb, int lane) {
  const int rho = 16 * t + (lane & 15);
  const int q = rho >> 1, po = rho & 1;
  halfx8 f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int kap = 32 * kb + 8 * (lane >> 4) + i;
    const int r = kap >> 1, pi = kap & 1;
    float v = 0.0f;
    if (q < R && r < R) {
      const float2 w = tw_lds(stw, ((q * r) % R) * (P / R));  // W_R^{qr}
      v = po == 0 ? (pi == 0 ? w.x : -w.y) : (pi == 0 ? w.y : w.x);
    }
    f[i] = static_cast<_Float16>(v);
  }
  return f;
}

// Raw operand words of one pass-A tile for this lane: X and R at j + 496 r for the lane's
// inputs r = 16 kb + 4 grp + i2 (kb < 2) and r = 32 (kb = 2, i2 = 0); r >= 33 lies past the
// row, where buffer loads return 0.
struct TileA {
  uint32_t x[9], r[9];
};

__device__ __forceinline__ void load_tile_a(TileA& d, __amdgpu_buffer_rsrc_t xv,
                                            __amdgpu_buffer_rsrc_t rv, int vo) {
  constexpr int NBA = P / 33;
#pragma unroll
  for (int u = 0; u < 9; ++u) {
    const int so = NBA * (16 * (u >> 2) + (u & 3)) * 4;
    d.x[u] = __builtin_amdgcn_raw_buffer_load_b32(xv, vo, so, 0);
    d.r[u] = __builtin_amdgcn_raw_buffer_load_b32(rv, vo, so, 0);
  }
}

// The same lane's words of one sub-row (already conj(X) . R, twiddled, in f16).
__device__ __forceinline__ void load_tile_y(uint32_t (&y)[9], __amdgpu_buffer_rsrc_t yv, int vo) {
  constexpr int NBA = P / 33;
#pragma unroll
  for (int u = 0; u < 9; ++u)
    y[u] = __builtin_amdgcn_raw_buffer_load_b32(yv, vo, NBA * (16 * (u >> 2) + (u & 3)) * 4, 0);
}

// 1024 threads (16 waves, four per SIMD): the matrix-core form keeps few values per lane, so
// more waves can hide the LDS and matrix-core latencies.
constexpr int T16 = 1024;
constexpr int NW16 = T16 / 64;

template <int SUB, bool DUMP = false>
__global__ __launch_bounds__(T16) void k_rows16(Row16Args a) {
  constexpr int NBA = P / 33;  // 496
  constexpr int NBB = P / 16;  // 1023
  constexpr int NBC = P / 31;  // 528
  // LDS rows are padded so the four 16-lane groups of a wave, which read inputs r 4 apart,
  // start 16 banks apart: pass B input r of butterfly j at j + SB r, pass C at j + SC r.
  constexpr int SB = NBB + 5;  // 1028
  constexpr int SC = NBC + 4;  // 532
  __shared__ uint32_t buf0[16 * SB];
  __shared__ uint32_t buf1[31 * SC];
  __shared__ u32x4 fa_lds[15 * 64];  // radix-33 fragments [t*3 + kb][lane]
  __shared__ u32x4 fc_lds[8 * 64];   // radix-31 fragments [t*2 + kb][lane] (registers would spill)
  __shared__ float2 stw[256];
  __shared__ RowRed<NW16> red;
  const int tid = threadIdx.x;
  for (int i = tid; i < 256; i += T16) stw[i] = a.tw[i];
  __syncthreads();
  for (int i = tid; i < 15 * 64; i += T16)
    fa_lds[i] = __builtin_bit_cast(u32x4, dft_frag<33>(stw, (i >> 6) / 3, (i >> 6) % 3, i & 63));
  for (int i = tid; i < 8 * 64; i += T16)
    fc_lds[i] = __builtin_bit_cast(u32x4, dft_frag<31>(stw, (i >> 6) >> 1, (i >> 6) & 1, i & 63));
  halfx8 fb[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) fb[t] = dft_frag<16>(stw, t, 0, tid & 63);
  __syncthreads();
  // SUB == 0: persistent blocks over the items (xrow, PRN group) b, b + gridDim, ... with
  // item = xrow * groups + g; gridDim is a multiple of groups, so a block keeps g = b mod groups
  // and, blocks being dealt round-robin to the 8 XCDs, a group's replicas stay in one XCD's L2.
  // SUB > 0: one pass over the sub-rows, grid-stride.
  const long long n_items = SUB == 0 ? a.n_items : 1;
  for (long long item = SUB == 0 ? blockIdx.x : 0; item < n_items;
       item += SUB == 0 ? gridDim.x : 1) {
  const int g = SUB == 0 ? static_cast<int>(item % a.groups) : 0;
  const int xrow = SUB == 0 ? static_cast<int>(item / a.groups) : 0;
  const int h = xrow & 1;
  const int bin = SUB == 0 ? (xrow >> 1) % a.n_bin : 0;
  const int ant = SUB == 0 ? (xrow >> 1) / a.n_bin : 0;
  const long long it0 = SUB == 0 ? g * a.per_group : blockIdx.x;
  const long long it1 = SUB == 0 ? min(g * a.per_group + a.per_group, a.n_prn) : a.n_sub;
  const long long step = SUB == 0 ? 1 : gridDim.x;
  const __amdgpu_buffer_rsrc_t xv =
      buffer_view(SUB == 0 ? a.X + static_cast<size_t>(xrow) * P : a.Y, P * 4);
  // the X words of the lane's first pass-A tile are the same for every PRN of the item: kept
  // in 9 VGPRs (the second tile's would not fit the 128-VGPR budget of 16 waves)
  uint32_t xk[9];
  if constexpr (SUB == 0) {
    const int w0 = tid >> 6, l16 = tid & 15, gp = (tid & 63) >> 4;
    const int vo = (16 * w0 + l16 + (P / 33) * 4 * gp) * 4;
#pragma unroll
    for (int u = 0; u < 9; ++u)
      xk[u] = __builtin_amdgcn_raw_buffer_load_b32(xv, vo, (P / 33) * (16 * (u >> 2) + (u & 3)) * 4, 0);
  }
  // SUB > 0: the next sub-row's pass-A words (tiles wave and wave + 16) and row index are
  // loaded right after this sub-row's pass A, so their latency hides behind passes B and C
  uint32_t ypf[SUB > 0 ? 2 : 1][9];
  int rmap = 0;
  auto load_next = [&](long long itx) {
    if constexpr (SUB > 0) {
      const __amdgpu_buffer_rsrc_t yv = buffer_view(a.Y + itx * P, P * 4);
      const int w0 = tid >> 6, l16 = tid & 15, gp = (tid & 63) >> 4;
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2)
        if (w0 + 16 * s2 < NBA / 16)
          load_tile_y(ypf[s2], yv, (16 * (w0 + 16 * s2) + l16 + NBA * 4 * gp) * 4);
      rmap = a.rowmap[itx / SUB];
    }
  };
  if constexpr (SUB > 0) {
    if (it0 < it1) load_next(it0);
  }
  for (long long it = it0; it < it1; it += step) {
    const int rcur = rmap;  // SUB > 0: this sub-row's row index
    const int prn = static_cast<int>(it);
    // surface form: output sample idx of this row (sub-row) is stored at dbase + dstride idx
    size_t dbase = 0;
    int dstride = 1;
    if constexpr (DUMP) {
      if constexpr (SUB == 0) {
        dbase = ((((static_cast<size_t>(ant) * a.n_prn + prn) * a.n_bin + bin) * 2 + h)) * P;
      } else {
        const long long rq = it / SUB;
        const long long r = a.rowmap[rq];
        dbase = static_cast<size_t>(r) * (SUB * P) + static_cast<size_t>(it - rq * SUB);
        dstride = SUB;
      }
    }
    auto dump_at = [&](float p, int idx) {
      if constexpr (DUMP) a.dump[dbase + static_cast<size_t>(dstride) * idx] = p;
    };
    // SUB == 0: replica row; SUB > 0: the sub-row itself
    const __amdgpu_buffer_rsrc_t rv =
        SUB == 0 ? buffer_view(a.rep + (static_cast<size_t>(prn) * 2 + h) * P, P * 4)
                 : buffer_view(a.Y + it * P, P * 4);
    // Lane coordinates are re-derived per PRN through an opaque copy: otherwise the compiler
    // hoists every lane-dependent twiddle and address out of the PRN loop and spills them.
    int tid2 = tid;
    asm volatile("" : "+v"(tid2));
    const int wave = tid2 >> 6, lane = tid2 & 63, ln16 = lane & 15, grp = lane >> 4;
    // ---- pass A: radix 33 from HBM, 31 tiles of 16 butterflies: tiles wave and wave + 16
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int tile = wave + 16 * s;
      if (tile < NBA / 16) {
        const int j = 16 * tile + ln16;
        halfx8 b[3];
        if constexpr (SUB == 0) {
          TileA cur;
          if (s == 0) {
            const int vo = (j + NBA * 4 * grp) * 4;
#pragma unroll
            for (int u = 0; u < 9; ++u) {
              cur.x[u] = xk[u];
              cur.r[u] = __builtin_amdgcn_raw_buffer_load_b32(rv, vo, NBA * (16 * (u >> 2) + (u & 3)) * 4, 0);
            }
          } else {
            load_tile_a(cur, xv, rv, (j + NBA * 4 * grp) * 4);
          }
#pragma unroll
          for (int kb = 0; kb < 3; ++kb) {
            u32x4 w;
#pragma unroll
            for (int i2 = 0; i2 < 4; ++i2) {
              const int u = 4 * kb + i2;
              w[i2] = u < 9 ? conjmul16(cur.x[u], cur.r[u]) : 0u;
            }
            b[kb] = __builtin_bit_cast(halfx8, w);
          }
        } else {
          const uint32_t(&y)[9] = ypf[s];
#pragma unroll
          for (int kb = 0; kb < 3; ++kb) {
            u32x4 w;
#pragma unroll
            for (int i2 = 0; i2 < 4; ++i2) {
              const int u = 4 * kb + i2;
              w[i2] = u < 9 ? y[u] : 0u;
            }
            b[kb] = __builtin_bit_cast(halfx8, w);
          }
        }
        // output q = 8t + 2 grp (+1) of butterfly j = 31 n2 + n3 is pass-B input r = j / 31 of
        // butterfly 33 (j mod 31) + q (prime-factor order: no twiddle)
        const int jd = j / 31;
        uint32_t* dst = buf0 + 33 * (j - 31 * jd) + SB * jd + 2 * grp;
#pragma unroll
        for (int t = 0; t < 5; ++t) {
          floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kb = 0; kb < 3; ++kb)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                __builtin_bit_cast(halfx8, fa_lds[(t * 3 + kb) * 64 + lane]), b[kb], acc, 0, 0,
                0);
          if (t < 4 || grp == 0) dst[8 * t] = pack_f16(make_float2(acc[0], acc[1]), 1.0f);
          if (t < 4) dst[8 * t + 1] = pack_f16(make_float2(acc[2], acc[3]), 1.0f);
        }
      }
    }
    if constexpr (SUB > 0) {
      if (it + step < it1) load_next(it + step);
    }
    lds_barrier();
    // ---- pass B: radix 16, 64 tiles (the last holds 15 butterflies): wave + 16 s
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int j = 16 * (wave + 16 * s) + ln16;
      const bool ok = j < NBB;
      u32x4 w;
#pragma unroll
      for (int i2 = 0; i2 < 4; ++i2) w[i2] = ok ? buf0[j + SB * (4 * grp + i2)] : 0u;
      const halfx8 b = __builtin_bit_cast(halfx8, w);
      floatx4 acc0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(fb[0], b, floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      floatx4 acc1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(fb[1], b, floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      if (ok) {
        // output q of butterfly j (a = j / 33, k = j mod 33) is pass-C input r = a of
        // butterfly k + 33 q
        const int aa = j / 33;
        const int k = j - 33 * aa;
        uint32_t* dst = buf1 + aa * SC + k + 66 * grp;
        dst[0] = pack_f16(make_float2(acc0[0], acc0[1]), 1.0f);
        dst[33] = pack_f16(make_float2(acc0[2], acc0[3]), 1.0f);
        dst[264] = pack_f16(make_float2(acc1[0], acc1[1]), 1.0f);
        dst[297] = pack_f16(make_float2(acc1[2], acc1[3]), 1.0f);
      }
    }
    lds_barrier();
    // ---- pass C: radix 31, 33 tiles: tiles t0 and t0 + 8 (t0 = wave, or wave + 8 for waves
    // 8-15) on every wave, and the four output tiles t of tile 32 on waves 0-3 (wave t takes
    // output tile t).  All products are issued before the first power is used.  Output q of
    // butterfly j is sample pfa_base(j) + 528 q mod P.
    // Rows keep the lane's running maximum of the 64-bit keys (power bits, sample) of its
    // outputs (powers are >= 0: the bits order like the values, inf and NaN the largest); the
    // second peak of the winning rows comes from a masked recomputation.
    unsigned long long km = 0ull;
    auto keyed = [&](float p, int idx) {
      const unsigned long long k = pkey(p, idx);
      km = k > km ? k : km;
      dump_at(p, idx);
    };
    // sub-rows keep only their argmax: the lane keeps its powers (outputs 8 t and 8 t + 1 of
    // butterflies j0 and j1, and of tile 32 on waves 0-3) and the maximum of their f32 bits
    // (powers are >= 0: the bits order like the values, inf and NaN the largest); the samples
    // are found only where the block maximum is (end of the sub-row)
    float pj0[4][2], pj1[4][2], pex[2];
    unsigned int pmx = 0u;
    auto keep = [&](float& slot, float p, int idx) {
      slot = p;
      pmx = max(pmx, __float_as_uint(p));
      dump_at(p, idx);
    };
    {
      auto load_c = [&](int j, halfx8* b) {
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
          u32x4 w;
#pragma unroll
          for (int i2 = 0; i2 < 4; ++i2) {
            const int r = 16 * kb + 4 * grp + i2;
            w[i2] = r < 31 ? buf1[j + SC * r] : 0u;
          }
          b[kb] = __builtin_bit_cast(halfx8, w);
        }
      };
      const int j0 = 16 * (wave < 8 ? wave : wave + 8) + ln16, j1 = j0 + 128;
      // sample index of output 2 grp of butterflies j0, j1 (add 528 q' and wrap for q')
      const int base0 = pfa_base(j0) + NBC * 2 * grp;
      const int base1 = pfa_base(j1) + NBC * 2 * grp;
      auto sample = [&](int base, int dq) { return pfa_wrap(base + NBC * dq); };
      halfx8 b0[2], b1[2];
      load_c(j0, b0);
      load_c(j1, b1);
      auto fcf = [&](int t, int kb) {
        return __builtin_bit_cast(halfx8, fc_lds[(t * 2 + kb) * 64 + lane]);
      };
      floatx4 c0[4], c1[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const halfx8 f = fcf(t, 0);
        c0[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f, b0[0], floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        c1[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f, b1[0], floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const halfx8 f = fcf(t, 1);
        c0[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f, b0[1], c0[t], 0, 0, 0);
        c1[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f, b1[1], c1[t], 0, 0, 0);
      }
      if constexpr (SUB == 0) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          keyed(c0[t][0] * c0[t][0] + c0[t][1] * c0[t][1], sample(base0, 8 * t));
          keyed(c1[t][0] * c1[t][0] + c1[t][1] * c1[t][1], sample(base1, 8 * t));
          if (t < 3 || grp < 3) {  // output 8 t + 2 grp + 1 = 31 does not exist
            keyed(c0[t][2] * c0[t][2] + c0[t][3] * c0[t][3], sample(base0, 8 * t + 1));
            keyed(c1[t][2] * c1[t][2] + c1[t][3] * c1[t][3], sample(base1, 8 * t + 1));
          }
        }
      } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          keep(pj0[t][0], c0[t][0] * c0[t][0] + c0[t][1] * c0[t][1], sample(base0, 8 * t));
          keep(pj1[t][0], c1[t][0] * c1[t][0] + c1[t][1] * c1[t][1], sample(base1, 8 * t));
          pj0[t][1] = pj1[t][1] = 0.0f;
          if (t < 3 || grp < 3) {  // output 8 t + 2 grp + 1 = 31 does not exist
            keep(pj0[t][1], c0[t][2] * c0[t][2] + c0[t][3] * c0[t][3], sample(base0, 8 * t + 1));
            keep(pj1[t][1], c1[t][2] * c1[t][2] + c1[t][3] * c1[t][3], sample(base1, 8 * t + 1));
          }
        }
      }
      pex[0] = pex[1] = 0.0f;
      if (wave < 4) {
        const int j = 512 + ln16;
        halfx8 b[2];
        load_c(j, b);
        // output tile t = wave
        const halfx8 f0 = fcf(wave, 0), f1 = fcf(wave, 1);
        floatx4 acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(f0, b[0], floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(f1, b[1], acc, 0, 0, 0);
        const int q = 8 * wave + 2 * grp;
        const int be = pfa_base(j) + NBC * q;
        const float pa = acc[0] * acc[0] + acc[1] * acc[1];
        const float pb = acc[2] * acc[2] + acc[3] * acc[3];
        const bool okb = wave < 3 || grp < 3;
        if constexpr (SUB == 0) {
          keyed(pa, sample(be, 0));
          if (okb) keyed(pb, sample(be, 1));
        } else {
          keep(pex[0], pa, sample(be, 0));
          if (okb) keep(pex[1], pb, sample(be, 1));
        }
      }
    }
    if constexpr (SUB == 0) {
      // block argmax with the last-maximum rule: per wave (power bits, 1 + sample) by two
      // unsigned maxima, across waves in wave 0 after one barrier (which also orders this row's
      // pass-C reads before the next row's pass-A writes; the slots are next written two
      // barriers later); inf / NaN powers are flagged
      const unsigned int kb = static_cast<unsigned int>(km >> 32);
      const unsigned int pw = wave_umax(kb);
      const unsigned int iw =
          wave_umax(kb == pw ? static_cast<unsigned int>(km & 0xffffffffu) + 1u : 0u);
      if (lane == 0) {
        red.pk[wave] = pw;
        red.ik[wave] = iw;
      }
      lds_barrier();
      if (wave == 0) {
        unsigned int M, I;
        lanes_argmax(lane < NW16 ? red.pk[lane] : 0u, lane < NW16 ? red.ik[lane] : 0u, M, I);
        if (lane == 0) {
          RowOut o;
          o.p1 = __uint_as_float(M);
          o.p2 = -1.0f;
          o.i1 = static_cast<int32_t>(I) - 1;
          o.nan = M >= 0x7f800000u ? 1 : 0;
          a.rows[((static_cast<size_t>(ant) * a.n_prn + prn) * a.n_bin + bin) * 2 + h] = o;
        }
      }
    } else {
      // sub-row it of the chunk: row r0 + it / SUB (the pre-pass's row map), output samples
      // k0 + SUB m (k0 = it mod SUB).  Block maximum M of the lanes' maxima (one barrier); in
      // the waves holding M, the lanes holding it take the largest sample of power M among
      // their outputs, and the wave folds (M, sample) into the row's key (several waves holding
      // M: several atomic maxima, the largest sample wins).  inf / NaN powers are flagged.
      const long long rq = it / SUB;
      const int k0 = static_cast<int>(it - rq * SUB);
      const long long r = __builtin_amdgcn_readfirstlane(rcur);
      const unsigned int wmax = wave_umax(pmx);
      if (lane == 0) red.pk[wave] = wmax;
      lds_barrier();
      const unsigned int M = static_cast<unsigned int>(__builtin_amdgcn_readlane(
          static_cast<int>(row16_umax(lane < NW16 ? red.pk[lane] : 0u)), 0));
      if (wmax == M) {
        unsigned int cand = 0u;  // 1 + sample, 0 for none
        if (pmx == M) {
          const int j0 = 16 * (wave < 8 ? wave : wave + 8) + ln16, j1 = j0 + 128;
          const int base0 = pfa_base(j0) + NBC * 2 * grp, base1 = pfa_base(j1) + NBC * 2 * grp;
          auto take = [&](float p, int idx) {
            if (__float_as_uint(p) == M) cand = max(cand, static_cast<unsigned int>(idx + 1));
          };
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            take(pj0[t][0], pfa_wrap(base0 + NBC * 8 * t));
            take(pj1[t][0], pfa_wrap(base1 + NBC * 8 * t));
            if (t < 3 || grp < 3) {
              take(pj0[t][1], pfa_wrap(base0 + NBC * (8 * t + 1)));
              take(pj1[t][1], pfa_wrap(base1 + NBC * (8 * t + 1)));
            }
          }
          if (wave < 4) {
            const int be = pfa_base(512 + ln16) + NBC * (8 * wave + 2 * grp);
            take(pex[0], pfa_wrap(be));
            if (wave < 3 || grp < 3) take(pex[1], pfa_wrap(be + NBC));
          }
        }
        const unsigned int iw = wave_umax(cand);
        if (lane == 0) {
          atomicMax(a.rowkey + r, (static_cast<unsigned long long>(M) << 32) |
                                      static_cast<unsigned int>(k0 + SUB * (static_cast<int>(iw) - 1)));
          if (M >= 0x7f800000u) atomicOr(a.rownan + r, 1);
        }
      }
    }
  }
  }
}

// rows of P f32 pairs (natural order) -> the same in the prime-factor order
__global__ void k_to_pfa(const float2* __restrict__ in, float2* __restrict__ out, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const long long row = i / P;
  const int pos = static_cast<int>(i - row * P);
  out[i] = in[row * P + pfa_nat(pos)];
}

// carrier table [n_bin][clen]: (cos, sin) of phi = f32(f64(pc * n) * fc[bin]) in f64 rounded to
// f32 (wipe_value), or of the f64 phase itself (PHASE64, the FP16 form)
template <bool PHASE64>
__global__ void k_carrier(const double* __restrict__ fc, double pc, long long clen,
                          long long total, float2* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= total) return;
  const long long bin = i / clen;
  const long long n = i - bin * clen;
  const double f = fc[bin];
  if constexpr (PHASE64) {
    out[i] = cossin_f64(__dmul_rn(__dmul_rn(pc, static_cast<double>(n)), f));
  } else {
    out[i] = wipe_value(1.0f, __dmul_rn(pc, static_cast<double>(n)), f);
  }
}

// rows of P f32 pairs (natural order) -> f16 pairs scaled by s in the prime-factor order
__global__ void k_to_f16(const float2* __restrict__ in, uint32_t* __restrict__ out, long long n,
                         float s) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const long long row = i / P;
  const int pos = static_cast<int>(i - row * P);
  out[i] = pack_f16(in[row * P + pfa_nat(pos)], s);
}

// ---- long rows as N1 x 496 (N1 = R0 x 33): the four-step split with small transforms ----------
// A long row of P = R0 x 16368 = N1 x 496 points in natural order (folded conj(X) R) is
//   Z[k1 + N1 k2] = sum_{n2} W_496^{n2 k2} . W_P^{n2 k1} . sum_{n1} z[496 n1 + n2] W_N1^{n1 k1}:
//   step 1 (k_4s_pre)  per column n2 the N1-point DFT (prime-factor R0 x 33: n1 = (33 a + R0 b)
//                      mod N1, k1 = (F1 ka + F2 kb) mod N1) and the twiddle W_P^{n2 k1},
//                      written as rows Y[k1][n2] of 496 values;
//   step 2 (k_4s_rows) per k1 the 496-point DFT (prime-factor 16 x 31: n2 = (31 a + 16 b)
//                      mod 496, k2 = (465 ka + 32 kb) mod 496; radix 31 on the matrix cores as
//                      k_rows<1>), |.|^2 and the argmax of sample k1 + N1 k2 folded into the long
//                      row's key.
// Both kernels hold a few KB of LDS per wave, so several workgroups share a CU; the sub-row
// variants above need a whole CU's LDS per 16368-point transform.
constexpr int k4N2 = 496;
constexpr int k4Cols = 16;  // columns (consecutive n2) per step-1 block: 128-byte segments
constexpr int k4TA = 256;   // step-1 threads
constexpr int k4TB = 256;   // step-2 threads: four waves, each two k1 rows at a time

template <int R0>
struct FourStep {
  static constexpr int N1 = R0 * 33;
  static constexpr int PL = N1 * k4N2;
  static constexpr int F1 = 33 * inv_mod(33 % R0, R0);
  static constexpr int F2 = R0 * inv_mod(R0 % 33, 33);
  static_assert(F1 % R0 == 1 && F1 % 33 == 0 && F2 % 33 == 1 && F2 % R0 == 0, "coprime");
};

// Lanes of one wave exchanging values through its own LDS region: every LDS access of the wave
// issued so far has completed (and the compiler moves none across).
__device__ __forceinline__ void wave_lds_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

template <int R0>
__global__ __launch_bounds__(k4TA) void k_4s_pre(const float2* __restrict__ X,
                                                  const float2* __restrict__ rep,
                                                  const float2* __restrict__ tw4, int n_prn,
                                                  int n_bin, int n_ant, int hs, int tile,
                                                  long long r0, long long n_rows,
                                                  float2* __restrict__ Y,
                                                  unsigned long long* __restrict__ rowkey,
                                                  int* __restrict__ rownan) {
  using FS = FourStep<R0>;
  constexpr int N1 = FS::N1, PL = FS::PL;
  constexpr int NCB = k4N2 / k4Cols;  // 31 column blocks per row
  __shared__ float2 l1[N1 * k4Cols];  // pass-1 outputs [ka][b][column]
  const long long rr = blockIdx.x / NCB;
  if (rr >= n_rows) return;
  const int c0 = k4Cols * static_cast<int>(blockIdx.x - rr * NCB);
  const long long I = r0 + rr;
  const BigItem it = big_item(I, n_prn, n_bin, n_ant, hs, tile);
  const int h = static_cast<int>(it.xrow & 1);
  const float2* x = X + (hs == 1 ? it.xrow : it.xrow >> 1) * PL;
  const float2* c = rep + (static_cast<long long>(it.prn) * 2 + h) * PL;
  // pass 1: radix R0 from the spectra, units (column, b), consecutive threads on consecutive
  // columns
  constexpr int NU = (k4Cols * 33 + k4TA - 1) / k4TA;  // units per thread (the last partly)
  float2 v[NU][R0];
#pragma unroll
  for (int s = 0; s < NU; ++s) {  // every unit's operands in flight before the first DFT
    const int u = threadIdx.x + s * k4TA;
    const int cc = u % k4Cols, b = u / k4Cols;
    const int n2 = c0 + cc;
    if (u < k4Cols * 33) {
#pragma unroll
      for (int a = 0; a < R0; ++a) {
        const int n = k4N2 * ((33 * a + R0 * b) % N1) + n2;
        v[s][a] = cconjmul(x[n], c[n]);
      }
    }
  }
#pragma unroll
  for (int s = 0; s < NU; ++s) {
    const int u = threadIdx.x + s * k4TA;
    const int cc = u % k4Cols, b = u / k4Cols;
    if (u < k4Cols * 33) {
      Dft<R0>::run(v[s]);
#pragma unroll
      for (int ka = 0; ka < R0; ++ka) l1[(ka * 33 + b) * k4Cols + cc] = v[s][ka];
    }
  }
  __syncthreads();
  // pass 2: radix 33, units (column, ka); twiddle W_P^{n2 k1} from the [k1][n2] table, rows of
  // the output written with non-temporal stores (read once, by step 2)
  float2* y = Y + rr * PL;
  for (int u = threadIdx.x; u < k4Cols * R0; u += k4TA) {
    const int cc = u % k4Cols, ka = u / k4Cols;
    const int n2 = c0 + cc;
    float2 v[33];
#pragma unroll
    for (int b = 0; b < 33; ++b) v[b] = l1[(ka * 33 + b) * k4Cols + cc];
    Dft<33>::run(v);
#pragma unroll
    for (int kb = 0; kb < 33; ++kb) {
      const int k1 = (FS::F1 * ka + FS::F2 * kb) % N1;
      const float2 t = cmul(v[kb], tw4[k1 * k4N2 + n2]);
      float* d = reinterpret_cast<float*>(y + k1 * k4N2 + n2);
      __builtin_nontemporal_store(t.x, d);
      __builtin_nontemporal_store(t.y, d + 1);
    }
  }
  if (c0 == 0 && threadIdx.x == 0) {
    const long long r = big_row(I, n_prn, n_bin, n_ant, hs, tile);
    rowkey[r] = 0ull;
    rownan[r] = 0;
  }
}

// c31[m] = (cos, sin)(2 pi m / 31).  dump (DUMP): every power also goes to
// dump[row * P + sample], row the RowOut index of the long row.
template <int R0, bool DUMP>
__global__ __launch_bounds__(k4TB, 4) void k_4s_rows(const float2* __restrict__ Y,
                                                   const float2* __restrict__ c31, int n_prn,
                                                   int n_bin, int n_ant, int hs, int tile,
                                                   long long r0, long long n_pairs,
                                                   unsigned long long* __restrict__ rowkey,
                                                   int* __restrict__ rownan,
                                                   float* __restrict__ dump) {
  using FS = FourStep<R0>;
  constexpr int N1 = FS::N1, PL = FS::PL;
  constexpr int NP = 2 * k4N2;  // values of a wave's pair of k1 rows
  // LDS position of value i: one pad slot per 16 values, so the radix-16 lanes (16 values apart)
  // fall on different banks
  auto pad = [](int i) { return i + (i >> 4); };
  __shared__ float2 lds[k4TB / 64][NP + NP / 16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, ln16 = lane & 15, grp = lane >> 4;
  float2* L = lds[wave];
  // the lane's radix-31 GEMM constants (A = [l%16][l/16] of the cos / sin matrices, as k_rows<1>)
  float cA[4], sA[4];
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const int j = 4 * st + grp;
    const float2 w = c31[(j * ln16) % 31];
    cA[st] = j == 0 ? 1.0f : w.x;
    sA[st] = j == 0 ? 0.0f : w.y;
  }
  // the next pair's values are loaded into registers while the current pair is transformed
  constexpr int NV = (NP + 63) / 64;  // 16 values per lane (the last partly)
  const long long pstep = static_cast<long long>(gridDim.x) * (k4TB / 64);
  float2 nx[NV];
  auto load_pair = [&](long long pi) {
    const float2* src = Y + 2 * pi * k4N2;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = lane + 64 * i;
      nx[i] = idx < NP ? src[idx] : make_float2(0.f, 0.f);
    }
  };
  const long long p0 = static_cast<long long>(blockIdx.x) * (k4TB / 64) + wave;
  if (p0 < n_pairs) load_pair(p0);
  for (long long pi = p0; pi < n_pairs; pi += pstep) {
    // k1 rows 2 pi, 2 pi + 1 of the chunk (N1 is even: both belong to one long row)
    const long long g = 2 * pi;
    const long long rq = g / N1;
    const int k1 = static_cast<int>(g - rq * N1);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = lane + 64 * i;
      if (idx < NP) L[pad(idx)] = nx[i];
    }
    wave_lds_sync();
    if (pi + pstep < n_pairs) load_pair(pi + pstep);
    // radix 16 over a, in place: lane (q, b) owns the 16 positions (31 a + 16 b) mod 496 of
    // row q and writes output ka where input a was
    int ln = lane;
    asm volatile("" : "+v"(ln));
    if (ln < 62) {
      const int q = ln >= 31 ? 1 : 0;
      const int b = ln - 31 * q;
      float2 v[16];
#pragma unroll
      for (int a = 0; a < 16; ++a) v[a] = L[pad(q * k4N2 + (31 * a + 16 * b) % k4N2)];
      Dft<16>::run(v);
#pragma unroll
      for (int a = 0; a < 16; ++a) L[pad(q * k4N2 + (31 * a + 16 * b) % k4N2)] = v[a];
    }
    wave_lds_sync();
    // radix 31 over b on the matrix cores: one tile of 16 butterflies (ka = lane % 16) per row
    unsigned long long km = 0ull;
    long long rowoff = 0;
    if constexpr (DUMP) rowoff = big_row(r0 + rq, n_prn, n_bin, n_ant, hs, tile) * PL;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      floatx4 acc[4];
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) acc[cc] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const int j1 = 4 * st + grp, j2 = j1 == 0 ? 0 : 31 - j1;
        const float2 v1 = L[pad(q * k4N2 + (31 * ln16 + 16 * j1) % k4N2)];
        float2 v2 = L[pad(q * k4N2 + (31 * ln16 + 16 * j2) % k4N2)];
        if (j1 == 0) v2 = make_float2(0.f, 0.f);  // a_0 = v_0 (no partner input)
        acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(cA[st], v1.x + v2.x, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(cA[st], v1.y + v2.y, acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(sA[st], v1.x - v2.x, acc[2], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(sA[st], v1.y - v2.y, acc[3], 0, 0, 0);
      }
      const int kq = k1 + q;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int kb = 4 * grp + e;  // output row of the GEMM tile
        const float ar = acc[0][e], ai = acc[1][e];
        const float br = acc[2][e], bi = acc[3][e];
        const float y1r = ar + bi, y1i = ai - br;  // Y_kb
        const int s1 = kq + N1 * ((465 * ln16 + 32 * kb) % k4N2);
        const float p1 = y1r * y1r + y1i * y1i;
        const unsigned long long t1 = pkey(p1, s1);
        km = t1 > km ? t1 : km;
        if constexpr (DUMP) dump[rowoff + s1] = p1;
        if (kb > 0) {
          const float y2r = ar - bi, y2i = ai + br;  // Y_{31-kb}
          const int s2 = kq + N1 * ((465 * ln16 + 32 * (31 - kb)) % k4N2);
          const float p2 = y2r * y2r + y2i * y2i;
          const unsigned long long t2 = pkey(p2, s2);
          km = t2 > km ? t2 : km;
          if constexpr (DUMP) dump[rowoff + s2] = p2;
        }
      }
    }
    // the pair's argmax into the long row's key (a NaN power has the largest key)
    unsigned int pw, iw;
    wave_argmax(__uint_as_float(static_cast<unsigned int>(km >> 32)),
                static_cast<int>(km & 0xffffffffu), pw, iw);
    const bool bad = __ballot((km >> 32) > 0x7f800000ull) != 0;
    if (lane == 0) {
      const long long r = big_row(r0 + rq, n_prn, n_bin, n_ant, hs, tile);
      atomicMax(rowkey + r, pkey(funord(pw), static_cast<int>(iw) - 1));
      if (bad) atomicOr(rownan + r, 1);
    }
    wave_lds_sync();  // the next pair's staging overwrites L
  }
}

}  // namespace fused

// ================================ generic path ===========================================
// PHASE64: the FP16 precision form's carrier phase, f64 throughout (no f32 rounding of phi).
template <bool PHASE64>
__global__ void k_fold_obs(const uint8_t* __restrict__ packed, size_t row_bytes, int fmt,
                           const double* __restrict__ means, const double* __restrict__ fc,
                           int n_bin, int P, long long total, double pc,
                           const float2* __restrict__ tau, float2* __restrict__ out) {
  const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const long long xrow = gid / P;
  const int n = static_cast<int>(gid - xrow * P);
  const int h = static_cast<int>(xrow & 1);
  const int bin = static_cast<int>((xrow >> 1) % n_bin);
  const int ant = static_cast<int>((xrow >> 1) / n_bin);
  const uint8_t* row = packed + static_cast<size_t>(ant) * row_bytes;
  const long long n2 = static_cast<long long>(n) + P;
  const float2 x0 = PHASE64 ? wipe_obs_f64(row, fmt, n, means[ant], pc, fc[bin])
                            : wipe_obs(row, fmt, n, means[ant], pc, fc[bin]);
  const float2 x1 = PHASE64 ? wipe_obs_f64(row, fmt, n2, means[ant], pc, fc[bin])
                            : wipe_obs(row, fmt, n2, means[ant], pc, fc[bin]);
  out[gid] = h == 0 ? cadd(x0, x1) : cmul(csub(x0, x1), tau[n]);
}

template <bool PHASE64>
__global__ void k_wipe_obs(const uint8_t* __restrict__ packed, size_t row_bytes, int fmt,
                           const double* __restrict__ means, const double* __restrict__ fc,
                           int n_bin, int N, long long total, double pc,
                           float2* __restrict__ out) {
  const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const long long xrow = gid / N;
  const int n = static_cast<int>(gid - xrow * N);
  const int bin = static_cast<int>(xrow % n_bin);
  const int ant = static_cast<int>(xrow / n_bin);
  const uint8_t* row = packed + static_cast<size_t>(ant) * row_bytes;
  out[gid] = PHASE64 ? wipe_obs_f64(row, fmt, n, means[ant], pc, fc[bin])
                    : wipe_obs(row, fmt, n, means[ant], pc, fc[bin]);
}

__global__ void k_wipe_signal(const float* __restrict__ sig, const double* __restrict__ pp,
                              const double* __restrict__ fc, int N, long long total,
                              float2* __restrict__ out) {
  const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const long long bin = gid / N;
  const int n = static_cast<int>(gid - bin * N);
  out[gid] = wipe_value(sig[n], pp[n], fc[bin]);
}

__global__ void k_mul_rows(const float2* __restrict__ X, const float2* __restrict__ rep, int M,
                           int n_prn, int n_bin, int folded, long long r0, long long total,
                           float2* __restrict__ Z) {
  const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const long long rr = gid / M;
  const int i = static_cast<int>(gid - rr * M);
  const long long r = r0 + rr;
  const int h = static_cast<int>(r & 1);
  const long long q = r >> 1;  // (ant*n_prn + prn)*n_bin + bin
  const int bin = static_cast<int>(q % n_bin);
  const long long ap = q / n_bin;
  const int prn = static_cast<int>(ap % n_prn);
  const int ant = static_cast<int>(ap / n_prn);
  const long long xrow = folded ? (static_cast<long long>(ant) * n_bin + bin) * 2 + h
                                : static_cast<long long>(ant) * n_bin + bin;
  Z[gid] = cconjmul(X[xrow * M + i], rep[(static_cast<long long>(prn) * 2 + h) * M + i]);
}

constexpr int kRedT = 256;

__global__ __launch_bounds__(kRedT) void k_row_reduce(const float2* __restrict__ Z, int M, int P,
                                                      int cw, RowOut* __restrict__ out) {
  __shared__ float sp[kRedT / 64];
  __shared__ int si[kRedT / 64];
  const float2* z = Z + static_cast<size_t>(blockIdx.x) * M;
  float bp = -1.0f;
  int bi = -1;
  int nan = 0;
  for (int i = threadIdx.x; i < M; i += kRedT) {
    const float2 v = z[i];
    const float p = v.x * v.x + v.y * v.y;
    nan |= (p != p);
    arg_combine(bp, bi, p, i);
  }
  block_argmax<kRedT>(bp, bi, sp, si);
  const int main_cp = bi % P;
  float m2 = -1.0f;
  for (int i = threadIdx.x; i < M; i += kRedT) {
    if (ring_dist(i % P, main_cp, P) > cw) {
      const float2 v = z[i];
      m2 = fmaxf(m2, v.x * v.x + v.y * v.y);
    }
  }
  m2 = block_max<kRedT>(m2, sp);
  nan = __syncthreads_or(nan);
  if (threadIdx.x == 0) {
    RowOut o;
    o.p1 = bp;
    o.p2 = m2;
    o.i1 = bi;
    o.nan = nan;
    out[blockIdx.x] = o;
  }
}

// ---- long code periods P = R0 x 16368 (four-step split of the row transform) --------------
// With n = n1 + 16368 n0 and k = k0 + R0 k1, the P-point transform of z = conj(X) R is
//   Z[k0 + R0 k1] = FFT_16368( W_P^{n1 k0} sum_{n0} z[n1 + 16368 n0] W_R0^{n0 k0} )[k1]:
// k_big_pre forms the R0 twiddled sub-rows y_k0 of each row (one R0-point DFT per n1), and the
// fused rows kernel in sub-row mode transforms each sub-row in LDS.  X and R of long rows are
// stored segment by segment in the prime-factor order of 16368 (position 16368 n0 + t holds
// natural index 16368 n0 + pfa_nat(t)), so the pre-pass reads and writes position t of every
// segment with contiguous accesses and its sub-rows arrive in the order the rows kernels
// transform without twiddle factors.  F16: the sub-rows are stored as f16 (re, im) pairs
// scaled by kSplitF16Scale for the matrix-core rows kernel (the products and the R0-point DFT
// stay f32).
// Standalone pre-pass of a chunk: 32 blocks of 512 positions per row.  Blocks are dealt
// round-robin to the 8 XCDs, so block b takes chunk 4 (b mod 8) + (b / 8) mod 4 of row b / 32:
// each XCD works on one contiguous eighth of every row and its L2 holds contiguous slices of
// the tile's spectra.
template <int R0, bool F16>
__global__ __launch_bounds__(256) void k_big_pre(const float2* __restrict__ X,
                                                 const float2* __restrict__ rep,
                                                 const float2* __restrict__ twb,
                                                 const int4* __restrict__ rowtab,
                                                 long long r0, long long n_rows,
                                                 void* __restrict__ Yv,
                                                 unsigned long long* __restrict__ rowkey,
                                                 int* __restrict__ rownan,
                                                 int* __restrict__ rowmap) {
  const long long rr = blockIdx.x >> 5;
  if (rr >= n_rows) return;
  const int chunk = 4 * static_cast<int>(blockIdx.x & 7) + static_cast<int>((blockIdx.x >> 3) & 3);
  const int p0 = 512 * chunk;
  fused::big_pre_positions<R0, F16>(X, rep, twb, rowtab, r0, rr, p0,
                             p0 + 512 < kFusedP ? p0 + 512 : kFusedP,
                             static_cast<int>(threadIdx.x), Yv, rowkey, rownan, rowmap);
}

// The row table of a split group's tiled row order (big_item / big_row), enumerated rows [0, n):
// (forward-spectrum row: xrow, or xrow / 2 for unfolded rows; replica row 2 prn + h; RowOut row)
// per row, built once per plan, so a pre-pass block takes its row with one scalar load instead
// of 64-bit divisions in every wave.
__global__ void k_big_rowtab(int4* __restrict__ tab, long long n, int n_prn, int n_bin,
                             int n_ant, int hs, int tile) {
  const long long I = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (I >= n) return;
  const BigItem it = big_item(I, n_prn, n_bin, n_ant, hs, tile);
  const int h = static_cast<int>(it.xrow & 1);
  tab[I] = make_int4(static_cast<int>(hs == 1 ? it.xrow : it.xrow >> 1), it.prn * 2 + h,
                     static_cast<int>(big_row(I, n_prn, n_bin, n_ant, hs, tile)), 0);
}

// Pre-pass of the winning row of every (antenna, PRN) pair w (win[w] = 2 bin + h): its R0
// sub-rows into Zw[w], 32 blocks of 512 positions per row; clears the masked-maximum key.
template <int R0>
__global__ __launch_bounds__(256) void k_win_pre(const float2* __restrict__ X,
                                                 const float2* __restrict__ rep,
                                                 const float2* __restrict__ twb, int n_prn,
                                                 int n_bin, int folded,
                                                 const int* __restrict__ win, int nw,
                                                 float2* __restrict__ Zw,
                                                 unsigned int* __restrict__ m2key) {
  constexpr int PL = R0 * kFusedP;
  const int w = static_cast<int>(blockIdx.x >> 5);
  if (w >= nw) return;
  const int t = 512 * static_cast<int>(blockIdx.x & 31) + 2 * static_cast<int>(threadIdx.x);
  if (t >= kFusedP) return;
  const int ant = w / n_prn;
  const int prn = w - ant * n_prn;
  const int bh = win[w];
  const long long xrow = static_cast<long long>(ant) * n_bin * 2 + bh;
  fused::big_pre_core<R0, false>(X + (folded ? xrow : xrow >> 1) * PL,
                                 rep + (static_cast<long long>(prn) * 2 + (bh & 1)) * PL, twb, t,
                                 Zw + static_cast<long long>(w) * PL);
  if (t == 0) m2key[w] = 0u;
}

// The same for the FP16 fused rows (one segment): Zw[w] = conj(X) R of the winning row in f32
// from the f16 spectra (both scaled by kF16Scale, as the FP16 rows kernel multiplies them), in
// their prime-factor order.
__global__ __launch_bounds__(256) void k_win_pre16(const uint32_t* __restrict__ X16,
                                                   const uint32_t* __restrict__ rep16,
                                                   int n_prn, int n_bin,
                                                   const int* __restrict__ win, int nw,
                                                   float2* __restrict__ Zw,
                                                   unsigned int* __restrict__ m2key) {
  const int w = static_cast<int>(blockIdx.x >> 5);
  if (w >= nw) return;
  const int t = 512 * static_cast<int>(blockIdx.x & 31) + 2 * static_cast<int>(threadIdx.x);
  if (t >= kFusedP) return;
  const int ant = w / n_prn;
  const int prn = w - ant * n_prn;
  const int bh = win[w];
  const long long xrow = static_cast<long long>(ant) * n_bin * 2 + bh;
  const uint2 xw = *reinterpret_cast<const uint2*>(X16 + xrow * kFusedP + t);
  const uint2 rw = *reinterpret_cast<const uint2*>(
      rep16 + (static_cast<long long>(prn) * 2 + (bh & 1)) * kFusedP + t);
  auto f2 = [](uint32_t u) {
    const fused::half2_t h = __builtin_bit_cast(fused::half2_t, u);
    return make_float2(static_cast<float>(h.x), static_cast<float>(h.y));
  };
  float4 z;
  const float2 z0 = cconjmul(f2(xw.x), f2(rw.x)), z1 = cconjmul(f2(xw.y), f2(rw.y));
  z.x = z0.x;
  z.y = z0.y;
  z.z = z1.x;
  z.w = z1.y;
  *reinterpret_cast<float4*>(Zw + static_cast<long long>(w) * kFusedP + t) = z;
  if (t == 0) m2key[w] = 0u;
}

// Second peak of every (antenna, PRN) pair from its masked maximum (k_second_win's rounding).
__global__ void k_win_second(const unsigned int* __restrict__ m2key, int nw, int n_prn,
                             const PeakSlot* __restrict__ slots, float gain, float scale,
                             tga_peak* __restrict__ out) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  const int ant = w / n_prn;
  const int prn = w - ant * n_prn;
  const PeakSlot sl = slots[prn];
  const float m2 = __uint_as_float(m2key[w]);
  float second = (__fsqrt_rn(m2) * gain) * scale;
  if (!(second > 0.0f)) second = 1e-6f;  // correlate.rs:172-174
  out[sl.base + static_cast<long long>(ant) * sl.stride + sl.local].second_peak = second;
}

// Row keys of long rows -> RowOut (p2 is filled later from the winning row only).
__global__ void k_big_rowout(const unsigned long long* __restrict__ rowkey,
                             const int* __restrict__ rownan, long long r0, long long n_rows,
                             int n_prn, int n_bin, int n_ant, int hs, int tile,
                             RowOut* __restrict__ rows) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n_rows) return;
  const long long r = big_row(r0 + i, n_prn, n_bin, n_ant, hs, tile);
  const unsigned long long k = rowkey[r];
  RowOut o;
  o.p1 = __uint_as_float(static_cast<unsigned int>(k >> 32));
  o.p2 = -1.0f;
  o.i1 = static_cast<int32_t>(k & 0xffffffffu);
  o.nan = rownan[r];
  rows[r] = o;
  if (hs == 2) rows[r + 1] = RowOut{0.0f, -1.0f, 0, 0};  // the coinciding hypothesis never wins
}

// Position of natural index n (< 16368) in the prime-factor order of 16368: the inverse of
// fused::pfa_nat, 496 (n mod 33) + 31 ((-n) mod 16) + (n mod 31).
__device__ __forceinline__ int pfa_pos(int n) {
  return 496 * (n % 33) + 31 * ((16 - (n & 15)) & 15) + n % 31;
}

// Z[w][i] = conj(X) R of the winning row win[w] of (antenna, PRN) pair w, in natural order
// (X and R of long rows are stored segment by segment in the prime-factor order).
__global__ void k_mul_win(const float2* __restrict__ X, const float2* __restrict__ rep, int M,
                          int n_prn, int n_bin, int folded, const int* __restrict__ win,
                          long long total, float2* __restrict__ Z, int natural) {
  const long long gid = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (gid >= total) return;
  const long long w = gid / M;
  const int i = static_cast<int>(gid - w * M);
  const int seg = i / kFusedP;
  const int pos = natural ? i : seg * kFusedP + pfa_pos(i - seg * kFusedP);
  const int ant = static_cast<int>(w / n_prn);
  const int prn = static_cast<int>(w - static_cast<long long>(ant) * n_prn);
  const int bh = win[w];  // bin * 2 + h
  const long long xrow = static_cast<long long>(ant) * n_bin * 2 + bh;
  Z[gid] = cconjmul(X[(folded ? xrow : xrow >> 1) * M + pos],
                    rep[(static_cast<long long>(prn) * 2 + (bh & 1)) * M + pos]);
}

// Second peak of each winning row around the selected code phase: max |z|^2 more than cw
// samples (ring distance mod P) from it, scaled like the best peak, floor 1e-6
// (correlate.rs:156-175).
__global__ __launch_bounds__(kRedT) void k_second_win(const float2* __restrict__ Z, int M, int P,
                                                      int cw, int n_prn,
                                                      const PeakSlot* __restrict__ slots,
                                                      float gain, float scale,
                                                      tga_peak* __restrict__ out) {
  __shared__ float sp[kRedT / 64];
  const int w = blockIdx.x;
  const int ant = w / n_prn;
  const int prn = w - ant * n_prn;
  const PeakSlot sl = slots[prn];
  tga_peak* pk = out + sl.base + static_cast<long long>(ant) * sl.stride + sl.local;
  const int main_cp = static_cast<int>(pk->best_codephase % static_cast<uint64_t>(P));
  const float2* z = Z + static_cast<size_t>(w) * M;
  float m2 = -1.0f;
  for (int i = threadIdx.x; i < M; i += kRedT) {
    if (ring_dist(i % P, main_cp, P) > cw) {
      const float2 v = z[i];
      m2 = fmaxf(m2, v.x * v.x + v.y * v.y);
    }
  }
  m2 = block_max<kRedT>(m2, sp);
  if (threadIdx.x == 0) {
    float second = m2 >= 0.f ? (__fsqrt_rn(m2) * gain) * scale : -1.0f;
    if (!(second > 0.0f)) second = 1e-6f;
    pk->second_peak = second;
  }
}

// gain = 2 for folded rows: an unnormalised P-point transform pair carries half the gain of
// the reference's 2P-point pair (FFT_2P([c,c]) at bin 2k is 2 FFT_P(c) at bin k).  The factor
// is a power of two, so corr = (sqrt(p) * gain) * scale rounds exactly like sqrt(4p) * scale.
__global__ void k_select(const RowOut* __restrict__ rows, const PeakSlot* __restrict__ slots,
                         int n_ant, int n_prn, int n_bin, float gain, float scale,
                         tga_peak* __restrict__ out, int* nan_flag, int* __restrict__ win) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n_ant * n_prn) return;
  const int ant = gid / n_prn;
  const int prn = gid - ant * n_prn;
  const RowOut* r = rows + static_cast<size_t>(gid) * n_bin * 2;
  float best = 0.f;
  int bbin = 0, bh = 0;
  int nan = 0;
  for (int bin = 0; bin < n_bin; ++bin) {
    const RowOut r0 = r[2 * bin], r1 = r[2 * bin + 1];
    nan |= r0.nan | r1.nan;
    const float c0 = (__fsqrt_rn(r0.p1) * gain) * scale;
    const float c1 = (__fsqrt_rn(r1.p1) * gain) * scale;
    const int hh = c1 > c0 ? 1 : 0;  // [c,-c] only when strictly larger (correlate.rs:140)
    const float cb = hh ? c1 : c0;
    if (bin == 0 || cb >= best) {  // last maximum across bins (correlate.rs:151-154)
      best = cb;
      bbin = bin;
      bh = hh;
    }
  }
  const RowOut w = r[2 * bbin + bh];
  float second = w.p2 >= 0.f ? (__fsqrt_rn(w.p2) * gain) * scale : -1.0f;
  if (!(second > 0.0f)) second = 1e-6f;  // correlate.rs:172-174
  const PeakSlot s = slots[prn];
  tga_peak pk;
  pk.best_peak = best;
  pk.second_peak = second;
  pk.best_freq_idx = static_cast<uint64_t>(bbin);
  pk.best_codephase = static_cast<uint64_t>(w.i1);
  out[s.base + static_cast<long long>(ant) * s.stride + s.local] = pk;
  if (win) win[gid] = 2 * bbin + bh;
  if (nan) atomicOr(nan_flag, 1);
}

inline unsigned grid_for(long long total, int block) {
  return static_cast<unsigned>((total + block - 1) / block);
}

}  // namespace

// ---- launchers -------------------------------------------------------------------------------
void launch_ant_sums(const uint8_t* packed, size_t row_bytes, int n_ant, int64_t L, int fmt,
                     long long* partial, double* means, int* nan_flag, hipStream_t st) {
  hipLaunchKernelGGL(k_ant_partials, dim3(n_ant * kAntSplit), dim3(256), 0, st, packed,
                     row_bytes, static_cast<long long>(L), fmt, partial);
  TGA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_ant_means, dim3(n_ant), dim3(kAntSplit), 0, st, partial,
                     static_cast<long long>(L), means, nan_flag);
  TGA_HIP(hipGetLastError());
}

bool fused_supported(size_t P, size_t N, int cw) {
  // the rows kernel's masked second peak needs outputs of one lane more than 2 cw apart
  return P == static_cast<size_t>(kFusedP) && N == 2 * P && cw >= 0 && cw <= 16;
}

std::vector<float2> fused_twiddles() {
  std::vector<float2> t(256);
  for (int i = 0; i < 128; ++i) {
    for (int part = 0; part < 2; ++part) {
      const long long e = part == 0 ? i : 128LL * i;
      const double a = -2.0 * kPiD * static_cast<double>(e % kFusedP) / kFusedP;
      t[part * 128 + i] = make_float2(static_cast<float>(std::cos(a)),
                                      static_cast<float>(std::sin(a)));
    }
  }
  return t;
}

void launch_fused_forward(const uint8_t* packed, size_t row_bytes, int fmt, const double* means,
                          const double* fc, int n_ant, int n_bin, double phase_const,
                          const float2* tau, const float2* tw, const float2* carrier, float2* X,
                          hipStream_t st) {
  fused::FwdArgs a{packed, row_bytes, fmt, means, fc, n_bin, phase_const, tau, tw, X, nullptr, 0};
  a.carrier = carrier;
  a.clen = 2LL * fused::P;
  hipLaunchKernelGGL(fused::k_forward<false>, dim3(n_ant * n_bin * 2), dim3(fused::T), 0, st, a);
  TGA_HIP(hipGetLastError());
}

// Grid of a persistent rows kernel (one workgroup per CU: the LDS row takes the CU): the
// device's CU count rounded down to a multiple of `groups`, or one block per item when fewer.
long long persistent_blocks(long long n_items, int groups) {
  static thread_local int cus = 0, dev = -1;
  int d = 0;
  TGA_HIP(hipGetDevice(&d));
  if (d != dev) {
    TGA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d));
    dev = d;
  }
  const long long per_wave = std::max<long long>(groups, (cus / groups) * groups);
  return std::min(n_items, per_wave);
}

void launch_fused_rows(const float2* X, const float2* rep, const float2* tw, int n_ant,
                       int n_bin, int n_prn, int cw, RowOut* rows, hipStream_t st,
                       const FusedRowsConfig& cfg, float* dump) {
  const int groups = n_prn < cfg.groups ? n_prn : cfg.groups;
  const int per = (n_prn + groups - 1) / groups;
  fused::RowArgs a{X, rep, tw, n_bin, n_prn, groups, per, cw, rows};
  a.dump = dump;
  const long long blocks = static_cast<long long>(n_ant) * n_bin * 2 * groups;
  if (dump && cfg.variant == 0)
    hipLaunchKernelGGL((fused::k_rows<0, 0, false, true>), dim3(static_cast<unsigned>(blocks)),
                       dim3(fused::T), 0, st, a);
  else if (dump)
    hipLaunchKernelGGL((fused::k_rows<1, 0, false, true>), dim3(static_cast<unsigned>(blocks)),
                       dim3(fused::T), 0, st, a);
  else if (cfg.variant == 0)
    hipLaunchKernelGGL(fused::k_rows<0>, dim3(static_cast<unsigned>(blocks)), dim3(fused::T), 0,
                       st, a);
  else
    hipLaunchKernelGGL(fused::k_rows<1>, dim3(static_cast<unsigned>(blocks)), dim3(fused::T), 0,
                       st, a);
  TGA_HIP(hipGetLastError());
}

void launch_fused_forward_nc(const uint8_t* packed, size_t row_bytes, int fmt,
                             const double* means, const double* fc, int n_ant, int n_bin,
                             int epochs, double phase_const, const float2* tw,
                             const float2* carrier, float2* X, hipStream_t st) {
  fused::FwdArgs a{packed, row_bytes, fmt, means, fc, n_bin, phase_const, nullptr, tw, X,
                   nullptr, epochs};
  a.carrier = carrier;
  a.clen = static_cast<long long>(epochs) * fused::P;
  hipLaunchKernelGGL((fused::k_forward<false, true>), dim3(n_ant * n_bin * epochs),
                     dim3(fused::T), 0, st, a);
  TGA_HIP(hipGetLastError());
}

void launch_fused_rows_nc(const float2* X, const float2* rep, const float2* tw, int n_ant,
                          int n_bin, int n_prn, int epochs, int cw, RowOut* rows, int groups,
                          hipStream_t st) {
  groups = n_prn < groups ? n_prn : groups;
  const int per = (n_prn + groups - 1) / groups;
  fused::RowArgs a{X, rep, tw, n_bin, n_prn, groups, per, cw, rows};
  a.epochs = epochs;
  const long long blocks = static_cast<long long>(n_ant) * n_bin * groups;
  hipLaunchKernelGGL((fused::k_rows<0, 0, true>), dim3(static_cast<unsigned>(blocks)),
                     dim3(fused::T), 0, st, a);
  TGA_HIP(hipGetLastError());
}

void launch_fused_forward16(const uint8_t* packed, size_t row_bytes, int fmt,
                            const double* means, const double* fc, int n_ant, int n_bin,
                            double phase_const, const float2* tau, const float2* tw,
                            const float2* carrier, uint32_t* X16, hipStream_t st) {
  fused::FwdArgs a{packed, row_bytes, fmt, means, fc, n_bin, phase_const, tau, tw, nullptr, X16,
                   0};
  a.carrier = carrier;
  a.clen = 2LL * fused::P;
  hipLaunchKernelGGL(fused::k_forward<true>, dim3(n_ant * n_bin * 2), dim3(fused::T), 0, st, a);
  TGA_HIP(hipGetLastError());
}

void launch_fused_rows16(const uint32_t* X16, const uint32_t* rep16, const float2* tw,
                         int n_ant, int n_bin, int n_prn, int cw, RowOut* rows, int groups,
                         hipStream_t st, float* dump) {
  groups = n_prn < groups ? n_prn : groups;
  const int per = (n_prn + groups - 1) / groups;
  fused::Row16Args a{X16, rep16, tw, n_bin, n_prn, groups, per, cw, rows};
  a.n_items = static_cast<long long>(n_ant) * n_bin * 2 * groups;
  a.dump = dump;
  const long long blocks = persistent_blocks(a.n_items, groups);
  if (dump)
    hipLaunchKernelGGL((fused::k_rows16<0, true>), dim3(static_cast<unsigned>(blocks)),
                       dim3(fused::T16), 0, st, a);
  else
    hipLaunchKernelGGL(fused::k_rows16<0>, dim3(static_cast<unsigned>(blocks)),
                       dim3(fused::T16), 0, st, a);
  TGA_HIP(hipGetLastError());
}

void launch_carrier_dec(const double* fc, int n_bin, int R0, bool fold, double phase_const,
                        bool phase64, float2* out, hipStream_t st) {
  const int nh = fold ? 2 : 1;
  const long long total = static_cast<long long>(n_bin) * nh * R0 * kFusedP;
  if (phase64)
    hipLaunchKernelGGL(fused::k_carrier_dec<true>, dim3(grid_for(total, 256)), dim3(256), 0, st,
                       fc, phase_const, R0, nh, total, out);
  else
    hipLaunchKernelGGL(fused::k_carrier_dec<false>, dim3(grid_for(total, 256)), dim3(256), 0, st,
                       fc, phase_const, R0, nh, total, out);
  TGA_HIP(hipGetLastError());
}

std::vector<float2> long_twiddles(size_t P) {
  std::vector<float2> t(1024 + 160);
  for (int j = 0; j < 1024 + 160; ++j) {
    const long long e = j < 1024 ? j : 1024LL * (j - 1024);
    const double a = -2.0 * kPiD * static_cast<double>(e) / static_cast<double>(P);
    t[j] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
  }
  return t;
}

void launch_fwd_long(const uint8_t* packed, size_t row_bytes, int fmt, const double* means,
                     const float2* carrier, const float2* tau, const float2* tw,
                     const float2* twl, int R0, bool fold, int n_ant, int n_bin, float* lev,
                     float2* Y, float2* X, hipStream_t st) {
  const long long nlev = static_cast<long long>(n_ant) * (fold ? 2 : 1) * R0 * kFusedP;
  hipLaunchKernelGGL(fused::k_levels_dec, dim3(grid_for(nlev, 256)), dim3(256), 0, st, packed,
                     row_bytes, fmt, means, R0, fold ? 2 : 1, nlev, lev);
  TGA_HIP(hipGetLastError());
  fused::LongFwdArgs a{lev, carrier, tau, tw, n_ant, n_bin, Y};
  const int nblk = n_bin * R0 * n_ant * (fold ? 2 : 1);
  const dim3 g1(static_cast<unsigned>(8 * ((nblk + 7) / 8)));
  if (R0 == 4 && fold)
    hipLaunchKernelGGL((fused::k_fwd_long<4, true>), g1, dim3(fused::T), 0, st, a);
  else if (R0 == 4)
    hipLaunchKernelGGL((fused::k_fwd_long<4, false>), g1, dim3(fused::T), 0, st, a);
  else if (R0 == 10 && fold)
    hipLaunchKernelGGL((fused::k_fwd_long<10, true>), g1, dim3(fused::T), 0, st, a);
  else if (R0 == 10)
    hipLaunchKernelGGL((fused::k_fwd_long<10, false>), g1, dim3(fused::T), 0, st, a);
  else
    fail(TGA_ERR_UNSUPPORTED, "long-period forward supports 4 and 10 segments");
  TGA_HIP(hipGetLastError());
  const long long n_pairs = static_cast<long long>(n_ant) * n_bin * (fold ? 2 : 1) * (kFusedP / 2);
  if (R0 == 4)
    hipLaunchKernelGGL(fused::k_fwd_twist<4>, dim3(grid_for(n_pairs, 256)), dim3(256), 0, st, Y,
                       twl, n_pairs, X);
  else
    hipLaunchKernelGGL(fused::k_fwd_twist<10>, dim3(grid_for(n_pairs, 256)), dim3(256), 0, st, Y,
                       twl, n_pairs, X);
  TGA_HIP(hipGetLastError());
}

void launch_carrier(const double* fc, int n_bin, long long clen, double phase_const,
                    bool phase64, float2* out, hipStream_t st) {
  const long long total = static_cast<long long>(n_bin) * clen;
  if (phase64)
    hipLaunchKernelGGL(fused::k_carrier<true>, dim3(grid_for(total, 256)), dim3(256), 0, st, fc,
                       phase_const, clen, total, out);
  else
    hipLaunchKernelGGL(fused::k_carrier<false>, dim3(grid_for(total, 256)), dim3(256), 0, st, fc,
                       phase_const, clen, total, out);
  TGA_HIP(hipGetLastError());
}

void launch_to_pfa(const float2* in, float2* out, long long n, hipStream_t st) {
  hipLaunchKernelGGL(fused::k_to_pfa, dim3(grid_for(n, 256)), dim3(256), 0, st, in, out, n);
  TGA_HIP(hipGetLastError());
}

void launch_to_f16(const float2* in, uint32_t* out, long long n, float scale, hipStream_t st) {
  hipLaunchKernelGGL(fused::k_to_f16, dim3(grid_for(n, 256)), dim3(256), 0, st, in, out, n,
                     scale);
  TGA_HIP(hipGetLastError());
}

// Autotuner spectra: value i is a hash of i mapped to [-1, 1) (f32 words, or f16 pairs), so the
// timed kernels see data-dependent work like a real search rather than all-zero rows.
__global__ void k_fill_tune(uint32_t* __restrict__ p, long long n, int f16) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t h = static_cast<uint32_t>(i) * 0x9E3779B1u ^ static_cast<uint32_t>(i >> 32);
  h ^= h >> 15;
  h *= 0x85EBCA77u;
  h ^= h >> 13;
  const float a = static_cast<float>(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
  const float b = static_cast<float>(h & 0xffu) * (2.0f / 256.0f) - 1.0f;
  p[i] = f16 ? fused::pack_f16(make_float2(a, b), 1.0f) : __float_as_uint(a);
}

void launch_fill_tune(void* p, size_t bytes, bool f16, hipStream_t st) {
  const long long n = static_cast<long long>(bytes / 4);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_tune, dim3(grid_for(n, 256)), dim3(256), 0, st,
                     static_cast<uint32_t*>(p), n, f16 ? 1 : 0);
  TGA_HIP(hipGetLastError());
}

void launch_fold_obs(const uint8_t* packed, size_t row_bytes, int fmt, const double* means,
                     const double* fc, int n_ant, int n_bin, int P, double phase_const,
                     const float2* tau, float2* out, bool phase64, hipStream_t st) {
  const long long total = static_cast<long long>(n_ant) * n_bin * 2 * P;
  if (phase64)
    hipLaunchKernelGGL(k_fold_obs<true>, dim3(grid_for(total, 256)), dim3(256), 0, st, packed,
                       row_bytes, fmt, means, fc, n_bin, P, total, phase_const, tau, out);
  else
    hipLaunchKernelGGL(k_fold_obs<false>, dim3(grid_for(total, 256)), dim3(256), 0, st, packed,
                       row_bytes, fmt, means, fc, n_bin, P, total, phase_const, tau, out);
  TGA_HIP(hipGetLastError());
}

void launch_wipe_obs(const uint8_t* packed, size_t row_bytes, int fmt, const double* means,
                     const double* fc, int n_ant, int n_bin, int N, double phase_const,
                     float2* out, bool phase64, hipStream_t st) {
  const long long total = static_cast<long long>(n_ant) * n_bin * N;
  if (phase64)
    hipLaunchKernelGGL(k_wipe_obs<true>, dim3(grid_for(total, 256)), dim3(256), 0, st, packed,
                       row_bytes, fmt, means, fc, n_bin, N, total, phase_const, out);
  else
    hipLaunchKernelGGL(k_wipe_obs<false>, dim3(grid_for(total, 256)), dim3(256), 0, st, packed,
                       row_bytes, fmt, means, fc, n_bin, N, total, phase_const, out);
  TGA_HIP(hipGetLastError());
}

void launch_wipe_signal(const float* sig, const double* pp, const double* fc, int n_bin, int N,
                        float2* out, hipStream_t st) {
  const long long total = static_cast<long long>(n_bin) * N;
  hipLaunchKernelGGL(k_wipe_signal, dim3(grid_for(total, 256)), dim3(256), 0, st, sig, pp, fc,
                     N, total, out);
  TGA_HIP(hipGetLastError());
}

void launch_mul_rows(const float2* X, const float2* rep, int M, int n_prn, int n_bin,
                     bool folded, long long r0, long long r1, float2* Z, hipStream_t st) {
  const long long total = (r1 - r0) * M;
  hipLaunchKernelGGL(k_mul_rows, dim3(grid_for(total, 256)), dim3(256), 0, st, X, rep, M,
                     n_prn, n_bin, folded ? 1 : 0, r0, total, Z);
  TGA_HIP(hipGetLastError());
}

void launch_row_reduce(const float2* Z, int M, int P, int cw, long long n_rows, RowOut* out,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_row_reduce, dim3(static_cast<unsigned>(n_rows)), dim3(kRedT), 0, st, Z,
                     M, P, cw, out);
  TGA_HIP(hipGetLastError());
}

void launch_select(const RowOut* rows, const PeakSlot* slots, int n_ant, int n_prn, int n_bin,
                   float gain, float scale, tga_peak* out, int* nan_flag, hipStream_t st,
                   int* win) {
  const int total = n_ant * n_prn;
  hipLaunchKernelGGL(k_select, dim3(grid_for(total, 128)), dim3(128), 0, st, rows, slots, n_ant,
                     n_prn, n_bin, gain, scale, out, nan_flag, win);
  TGA_HIP(hipGetLastError());
}

bool big_supported(size_t P, size_t N, bool folded) {
  return (folded ? N == 2 * P : N == P) && (P == 4 * static_cast<size_t>(kFusedP) ||
                                             P == 10 * static_cast<size_t>(kFusedP));
}

std::vector<float2> big_twiddles(size_t P) {
  std::vector<float2> t(kFusedP);
  for (int s = 0; s < kFusedP; ++s) {
    // natural index of prime-factor position s (fused::pfa_nat)
    const int n1 = s / 496, b = s - 496 * n1, n2 = b / 31, n3 = b - 31 * n2;
    const long long e = (496LL * n1 + 1023LL * n2 + 528LL * n3) % kFusedP;
    const double a = -2.0 * kPiD * static_cast<double>(e) / static_cast<double>(P);
    t[s] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
  }
  return t;
}

void launch_big_pre(const float2* X, const float2* rep, const float2* twb, const int4* rowtab,
                    int R0, long long r0, long long r1, void* Y, bool f16,
                    unsigned long long* rowkey, int* rownan, int* rowmap, hipStream_t st) {
  const long long nr = r1 - r0;  // rows of this chunk, 32 blocks each
  const dim3 grid(static_cast<unsigned>(nr * 32));
  if (R0 == 4 && !f16)
    hipLaunchKernelGGL((k_big_pre<4, false>), grid, dim3(256), 0, st, X, rep, twb, rowtab, r0, nr,
                       Y, rowkey, rownan, rowmap);
  else if (R0 == 4)
    hipLaunchKernelGGL((k_big_pre<4, true>), grid, dim3(256), 0, st, X, rep, twb, rowtab, r0, nr,
                       Y, rowkey, rownan, rowmap);
  else if (R0 == 10 && !f16)
    hipLaunchKernelGGL((k_big_pre<10, false>), grid, dim3(256), 0, st, X, rep, twb, rowtab, r0,
                       nr, Y, rowkey, rownan, rowmap);
  else if (R0 == 10)
    hipLaunchKernelGGL((k_big_pre<10, true>), grid, dim3(256), 0, st, X, rep, twb, rowtab, r0,
                       nr, Y, rowkey, rownan, rowmap);
  else
    fail(TGA_ERR_UNSUPPORTED, "long-period split supports 4 and 10 sub-rows");
  TGA_HIP(hipGetLastError());
}

void launch_big_rowtab(int4* tab, long long n, int n_prn, int n_bin, int n_ant, int hs, int tile,
                       hipStream_t st) {
  if (n < 1) return;
  hipLaunchKernelGGL(k_big_rowtab, dim3(grid_for(n, 256)), dim3(256), 0, st, tab, n, n_prn,
                     n_bin, n_ant, hs, tile);
  TGA_HIP(hipGetLastError());
}

std::vector<float2> half_twiddles(size_t P) {
  const int R0 = static_cast<int>(P / kFusedP);
  std::vector<float2> t(2 * static_cast<size_t>(fused::kHalfS));
  for (int s = 0; s < fused::kHalfS; ++s) {
    // natural index of position s of a half sub-row (fused::half_nat)
    const int c1 = s / 248, b = s - 248 * c1, c2 = b / 31, c3 = b - 31 * c2;
    const long long m = (496LL * c1 + 7161LL * c2 + 528LL * c3) % fused::kHalfS;
    for (int k = 0; k < 2; ++k) {
      const long long e = (k == 0 ? m : m * R0) % static_cast<long long>(P);
      const double a = -2.0 * kPiD * static_cast<double>(e) / static_cast<double>(P);
      t[k * fused::kHalfS + s] =
          make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
  }
  return t;
}

void launch_half_pre(const float2* X, const float2* rep, const float2* twh, const int4* rowtab,
                     int R0, long long r0, long long r1, float2* Y, unsigned long long* rowkey,
                     int* rownan, int* rowmap, hipStream_t st) {
  const long long nr = r1 - r0;
  const dim3 grid(static_cast<unsigned>(nr * 32));
  if (R0 == 4)
    hipLaunchKernelGGL((fused::k_half_pre<4>), grid, dim3(256), 0, st, X, rep, twh, rowtab, r0, nr,
                       Y, rowkey, rownan, rowmap);
  else if (R0 == 10)
    hipLaunchKernelGGL((fused::k_half_pre<10>), grid, dim3(256), 0, st, X, rep, twh, rowtab, r0,
                       nr, Y, rowkey, rownan, rowmap);
  else
    fail(TGA_ERR_UNSUPPORTED, "long-period split supports 4 and 10 sub-rows");
  TGA_HIP(hipGetLastError());
}

void launch_half_sub(const float2* tw, int R0, int n_prn, int n_bin, int n_ant, int hs, int tile,
                     long long r0, long long r1, const float2* Y, unsigned long long* rowkey,
                     int* rownan, const int* rowmap, RowOut* rows, hipStream_t st,
                     float* dump) {
  const long long nr = r1 - r0;
  fused::RowArgs a{};
  a.hs = hs;
  a.n_ant = n_ant;
  a.tile = tile;
  a.tw = tw;
  a.n_prn = n_prn;
  a.n_bin = n_bin;
  a.Y = Y;
  a.n_sub = nr * 2 * R0;
  a.r0 = r0;
  a.rowkey = rowkey;
  a.rownan = rownan;
  a.rowmap = rowmap;
  a.dump = dump;
  const unsigned blocks = static_cast<unsigned>(a.n_sub < 4096 ? a.n_sub : 4096);
  const dim3 grid(blocks), block(fused::TH);
  if (R0 == 4 && dump)
    hipLaunchKernelGGL((fused::k_halfrows<8, true>), grid, block, 0, st, a);
  else if (R0 == 4)
    hipLaunchKernelGGL((fused::k_halfrows<8>), grid, block, 0, st, a);
  else if (R0 == 10 && dump)
    hipLaunchKernelGGL((fused::k_halfrows<20, true>), grid, block, 0, st, a);
  else if (R0 == 10)
    hipLaunchKernelGGL((fused::k_halfrows<20>), grid, block, 0, st, a);
  else
    fail(TGA_ERR_UNSUPPORTED, "long-period split supports 4 and 10 sub-rows");
  TGA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_big_rowout, dim3(grid_for(nr, 256)), dim3(256), 0, st, rowkey, rownan,
                     r0, nr, n_prn, n_bin, n_ant, hs, tile, rows);
  TGA_HIP(hipGetLastError());
}

template <int R0>
void launch_subrows(bool dump, unsigned blocks, hipStream_t st, const fused::RowArgs& a) {
  if (dump)
    hipLaunchKernelGGL((fused::k_subrows<R0, true>), dim3(blocks), dim3(fused::T), 0, st, a);
  else
    hipLaunchKernelGGL((fused::k_subrows<R0, false>), dim3(blocks), dim3(fused::T), 0, st, a);
}

void launch_big_sub(const float2* tw, int R0, int n_prn, int n_bin, int n_ant, int hs, int tile,
                    long long r0, long long r1, const void* Y, bool f16, unsigned long long* rowkey,
                    int* rownan, const int* rowmap, RowOut* rows, hipStream_t st, int variant,
                    float* dump) {
  const long long nr = r1 - r0;
  const long long n_sub = nr * R0;
  const unsigned blocks = static_cast<unsigned>(n_sub < 2048 ? n_sub : 2048);
  if (f16) {
    fused::Row16Args a{};
    a.tw = tw;
    a.n_prn = n_prn;
    a.n_bin = n_bin;
    a.groups = 1;
    a.per_group = 1;
    a.Y = static_cast<const uint32_t*>(Y);
    a.n_sub = n_sub;
    a.r0 = r0;
    a.rowkey = rowkey;
    a.rownan = rownan;
    a.hs = hs;
    a.n_ant = n_ant;
    a.tile = tile;
    a.dump = dump;
    a.rowmap = rowmap;
    if (dump && R0 == 4)
      hipLaunchKernelGGL((fused::k_rows16<4, true>), dim3(blocks), dim3(fused::T16), 0, st, a);
    else if (dump)
      hipLaunchKernelGGL((fused::k_rows16<10, true>), dim3(blocks), dim3(fused::T16), 0, st, a);
    else if (R0 == 4)
      hipLaunchKernelGGL((fused::k_rows16<4>), dim3(blocks), dim3(fused::T16), 0, st, a);
    else
      hipLaunchKernelGGL((fused::k_rows16<10>), dim3(blocks), dim3(fused::T16), 0, st, a);
  } else {
    fused::RowArgs a{};
    a.hs = hs;
    a.n_ant = n_ant;
    a.tile = tile;
    a.tw = tw;
    a.cw = 0;
    a.n_prn = n_prn;
    a.n_bin = n_bin;
    a.Y = static_cast<const float2*>(Y);
    a.n_sub = n_sub;
    a.r0 = r0;
    a.rowkey = rowkey;
    a.rownan = rownan;
    a.rowmap = rowmap;
    a.dump = dump;
    // the surface form runs the plan's radix-31 implementation, so its search is the same
    if (variant == 2 && R0 == 4)
      launch_subrows<4>(dump != nullptr, blocks, st, a);
    else if (variant == 2)
      launch_subrows<10>(dump != nullptr, blocks, st, a);
    else if (dump && R0 == 4 && variant == 0)
      hipLaunchKernelGGL((fused::k_rows<0, 4, false, true>), dim3(blocks), dim3(fused::T), 0, st,
                         a);
    else if (dump && R0 == 4)
      hipLaunchKernelGGL((fused::k_rows<1, 4, false, true>), dim3(blocks), dim3(fused::T), 0, st,
                         a);
    else if (dump && variant == 0)
      hipLaunchKernelGGL((fused::k_rows<0, 10, false, true>), dim3(blocks), dim3(fused::T), 0,
                         st, a);
    else if (dump)
      hipLaunchKernelGGL((fused::k_rows<1, 10, false, true>), dim3(blocks), dim3(fused::T), 0,
                         st, a);
    else if (R0 == 4 && variant == 0)
      hipLaunchKernelGGL((fused::k_rows<0, 4>), dim3(blocks), dim3(fused::T), 0, st, a);
    else if (R0 == 4)
      hipLaunchKernelGGL((fused::k_rows<1, 4>), dim3(blocks), dim3(fused::T), 0, st, a);
    else if (variant == 0)
      hipLaunchKernelGGL((fused::k_rows<0, 10>), dim3(blocks), dim3(fused::T), 0, st, a);
    else
      hipLaunchKernelGGL((fused::k_rows<1, 10>), dim3(blocks), dim3(fused::T), 0, st, a);
  }
  TGA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_big_rowout, dim3(grid_for(nr, 256)), dim3(256), 0, st, rowkey, rownan,
                     r0, nr, n_prn, n_bin, n_ant, hs, tile, rows);
  TGA_HIP(hipGetLastError());
}

std::vector<float2> four_step_twiddles(size_t P) {
  const int n1 = static_cast<int>(P / fused::k4N2);
  std::vector<float2> t(P);
  for (int k1 = 0; k1 < n1; ++k1)
    for (int n2 = 0; n2 < fused::k4N2; ++n2) {
      const long long e = static_cast<long long>(n2) * k1;  // < P
      const double a = -2.0 * kPiD * static_cast<double>(e) / static_cast<double>(P);
      t[static_cast<size_t>(k1) * fused::k4N2 + n2] =
          make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
  return t;
}

std::vector<float2> radix31_table() {
  std::vector<float2> t(31);
  for (int m = 0; m < 31; ++m) {
    const double a = 2.0 * kPiD * m / 31.0;
    t[m] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
  }
  return t;
}

void launch_big_4s(const float2* X, const float2* rep, const float2* tw4, const float2* c31,
                   int R0, int n_prn, int n_bin, int n_ant, int hs, int tile, long long r0,
                   long long r1, float2* Y, unsigned long long* rowkey, int* rownan,
                   RowOut* rows, hipStream_t st, float* dump) {
  const long long nr = r1 - r0;
  if (nr <= 0) return;
  const dim3 g1(static_cast<unsigned>(nr * (fused::k4N2 / fused::k4Cols)));
  const long long n_pairs = nr * R0 * 33 / 2;
  const dim3 g2(static_cast<unsigned>(std::min<long long>((n_pairs + 3) / 4, 32768)));
  if (R0 == 4) {
    hipLaunchKernelGGL(fused::k_4s_pre<4>, g1, dim3(fused::k4TA), 0, st, X, rep, tw4, n_prn,
                       n_bin, n_ant, hs, tile, r0, nr, Y, rowkey, rownan);
    if (dump)
      hipLaunchKernelGGL((fused::k_4s_rows<4, true>), g2, dim3(fused::k4TB), 0, st, Y, c31,
                         n_prn, n_bin, n_ant, hs, tile, r0, n_pairs, rowkey, rownan, dump);
    else
      hipLaunchKernelGGL((fused::k_4s_rows<4, false>), g2, dim3(fused::k4TB), 0, st, Y, c31,
                         n_prn, n_bin, n_ant, hs, tile, r0, n_pairs, rowkey, rownan, dump);
  } else if (R0 == 10) {
    hipLaunchKernelGGL(fused::k_4s_pre<10>, g1, dim3(fused::k4TA), 0, st, X, rep, tw4, n_prn,
                       n_bin, n_ant, hs, tile, r0, nr, Y, rowkey, rownan);
    if (dump)
      hipLaunchKernelGGL((fused::k_4s_rows<10, true>), g2, dim3(fused::k4TB), 0, st, Y, c31,
                         n_prn, n_bin, n_ant, hs, tile, r0, n_pairs, rowkey, rownan, dump);
    else
      hipLaunchKernelGGL((fused::k_4s_rows<10, false>), g2, dim3(fused::k4TB), 0, st, Y, c31,
                         n_prn, n_bin, n_ant, hs, tile, r0, n_pairs, rowkey, rownan, dump);
  } else {
    fail(TGA_ERR_UNSUPPORTED, "the four-step split supports 4 and 10 sub-rows");
  }
  TGA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_big_rowout, dim3(grid_for(nr, 256)), dim3(256), 0, st, rowkey, rownan,
                     r0, nr, n_prn, n_bin, n_ant, hs, tile, rows);
  TGA_HIP(hipGetLastError());
}

__global__ void k_surface_finish(float* __restrict__ v, long long n, float gain, float scale) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) v[i] = (__fsqrt_rn(v[i]) * gain) * scale;
}

void launch_surface_finish(float* v, long long n, float gain, float scale, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_surface_finish, dim3(grid_for(n, 256)), dim3(256), 0, st, v, n, gain,
                     scale);
  TGA_HIP(hipGetLastError());
}

void launch_big_second_sub(const float2* X, const float2* rep, const float2* tw,
                           const float2* twb, const PeakSlot* slots, int R0, int cw, int n_ant,
                           int n_prn, int n_bin, bool folded, const int* win, float gain,
                           float scale, float2* Zw, unsigned int* m2key, tga_peak* out,
                           hipStream_t st, const uint32_t* X16, const uint32_t* rep16) {
  const int nw = n_ant * n_prn;
  const dim3 g1(static_cast<unsigned>(nw) * 32);
  if (X16 != nullptr)
    hipLaunchKernelGGL(k_win_pre16, g1, dim3(256), 0, st, X16, rep16, n_prn, n_bin, win, nw, Zw,
                       m2key);
  else if (R0 == 1)
    hipLaunchKernelGGL(k_win_pre<1>, g1, dim3(256), 0, st, X, rep, twb, n_prn, n_bin,
                       folded ? 1 : 0, win, nw, Zw, m2key);
  else if (R0 == 4)
    hipLaunchKernelGGL(k_win_pre<4>, g1, dim3(256), 0, st, X, rep, twb, n_prn, n_bin,
                       folded ? 1 : 0, win, nw, Zw, m2key);
  else
    hipLaunchKernelGGL(k_win_pre<10>, g1, dim3(256), 0, st, X, rep, twb, n_prn, n_bin,
                       folded ? 1 : 0, win, nw, Zw, m2key);
  TGA_HIP(hipGetLastError());
  fused::RowArgs a{};
  a.tw = tw;
  a.cw = cw;
  a.n_prn = n_prn;
  a.n_bin = n_bin;
  a.Y = Zw;
  a.n_sub = static_cast<long long>(nw) * R0;
  a.peaks = out;
  a.slots = slots;
  a.m2key = m2key;
  const unsigned blocks = static_cast<unsigned>(a.n_sub < 2048 ? a.n_sub : 2048);
  if (R0 == 1)
    hipLaunchKernelGGL((fused::k_subrows<1, false, true>), dim3(blocks), dim3(fused::T), 0, st, a);
  else if (R0 == 4)
    hipLaunchKernelGGL((fused::k_subrows<4, false, true>), dim3(blocks), dim3(fused::T), 0, st, a);
  else
    hipLaunchKernelGGL((fused::k_subrows<10, false, true>), dim3(blocks), dim3(fused::T), 0, st,
                       a);
  TGA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_win_second, dim3(grid_for(nw, 256)), dim3(256), 0, st, m2key, nw, n_prn,
                     slots, gain, scale, out);
  TGA_HIP(hipGetLastError());
}

void launch_big_second(const float2* X, const float2* rep, const PeakSlot* slots, int M, int P,
                       int cw, int n_ant, int n_prn, int n_bin, bool folded, const int* win,
                       float gain, float scale, float2* Z, float2* Zs, tga_peak* out,
                       hipStream_t st, bool natural) {
  const long long nw = static_cast<long long>(n_ant) * n_prn;
  const long long total = nw * M;
  hipLaunchKernelGGL(k_mul_win, dim3(grid_for(total, 256)), dim3(256), 0, st, X, rep, M, n_prn,
                     n_bin, folded ? 1 : 0, win, total, Z, natural ? 1 : 0);
  TGA_HIP(hipGetLastError());
  float2* z = fft_forward(fft_plan(static_cast<size_t>(M)), Z, Zs, static_cast<size_t>(nw), st);
  hipLaunchKernelGGL(k_second_win, dim3(static_cast<unsigned>(nw)), dim3(kRedT), 0, st, z, M, P,
                     cw, n_prn, slots, gain, scale, out);
  TGA_HIP(hipGetLastError());
}

}  // namespace tga
