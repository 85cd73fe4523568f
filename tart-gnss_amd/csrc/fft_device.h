// Device-side DFT building blocks for the mixed-radix Stockham FFTs (gfx950).
//
// All transforms are forward and unnormalised: Y[k] = sum_n v[n] exp(-2 pi i n k / R).
// Inverse transforms are obtained by the caller through conj(FFT(conj(.))), which leaves
// |.| unchanged, so only the forward direction exists here.
//
// Codelets:
//   dft<1>               identity (one-segment rows)
//   dft<2>, dft<4>       butterflies
//   dft<8>, dft<16>      two-level splits (2x4, 4x4) with constant inner twiddles
//   dft<3>, dft<5>, dft<7>, dft<11>, dft<13>, dft<31>
//                        odd radices by the symmetric-pair method: with a_j = v_j + v_{R-j},
//                        b_j = v_j - v_{R-j}, Y_k = v_0 + sum a_j cos(2 pi jk/R)
//                        - i sum b_j sin(2 pi jk/R) and Y_{R-k} its mirror, (R-1)^2 real FMAs.
// Constants are computed at compile time from Taylor series (no runtime trig).
#pragma once

#include <hip/hip_runtime.h>

namespace tga {

constexpr double kPiD = 3.14159265358979323846264338327950288;

// sin/cos for |x| <= pi by Taylor series; used only in constant expressions.
constexpr double cx_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 40; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 40; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}

// cos(2 pi m / R) and sin(2 pi m / R), m = 0..R-1, rounded to float.
template <int R>
struct Roots {
  float c[R];
  float s[R];
  constexpr Roots() : c(), s() {
    for (int m = 0; m < R; ++m) {
      double a = 2.0 * kPiD * m / R;
      if (a > kPiD) a -= 2.0 * kPiD;
      c[m] = static_cast<float>(cx_cos(a));
      s[m] = static_cast<float>(cx_sin(a));
    }
  }
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// conj(a) * b
__device__ __forceinline__ float2 cconjmul(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// v * exp(-2 pi i e / R) with e a compile-time exponent; trivial angles are exact.
template <int R, int E>
__device__ __forceinline__ float2 rot(float2 v) {
  constexpr int e = ((E % R) + R) % R;
  if constexpr (e == 0) {
    return v;
  } else if constexpr (4 * e == R) {  // -i
    return make_float2(v.y, -v.x);
  } else if constexpr (2 * e == R) {  // -1
    return make_float2(-v.x, -v.y);
  } else if constexpr (4 * e == 3 * R) {  // +i
    return make_float2(-v.y, v.x);
  } else {
    constexpr Roots<R> W{};
    const float c = W.c[e], s = W.s[e];
    // (v.x + i v.y)(c - i s)
    return make_float2(v.x * c + v.y * s, v.y * c - v.x * s);
  }
}

template <int R>
struct Dft;

template <>
struct Dft<1> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <>
struct Dft<2> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <>
struct Dft<4> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    float2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = make_float2(b.x + d.y, b.y - d.x);  // b - i d
    v[3] = make_float2(b.x - d.y, b.y + d.x);  // b + i d
  }
};

// R = R1 * R2 split: n = n1 + R1 n2, k = R2 k1 + k2.
template <int R1, int R2>
struct DftSplit {
  static constexpr int R = R1 * R2;
  __device__ __forceinline__ static void run(float2* v) {
    float2 u[R1][R2];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) u[n1][n2] = v[n1 + R1 * n2];
      Dft<R2>::run(u[n1]);
    }
    twiddle_all(u);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      float2 w[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) w[n1] = u[n1][k2];
      Dft<R1>::run(w);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) v[R2 * k1 + k2] = w[k1];
    }
  }
  template <int N1 = 0, int K2 = 0>
  __device__ __forceinline__ static void twiddle_all(float2 (&u)[R1][R2]) {
    if constexpr (N1 < R1) {
      if constexpr (K2 < R2) {
        u[N1][K2] = rot<R, N1 * K2>(u[N1][K2]);
        twiddle_all<N1, K2 + 1>(u);
      } else {
        twiddle_all<N1 + 1, 0>(u);
      }
    }
  }
};

template <>
struct Dft<8> {
  __device__ __forceinline__ static void run(float2* v) { DftSplit<2, 4>::run(v); }
};
template <>
struct Dft<16> {
  __device__ __forceinline__ static void run(float2* v) { DftSplit<4, 4>::run(v); }
};

template <int R>
struct DftOdd {
  static_assert(R % 2 == 1 && R >= 3, "odd radix");
  __device__ __forceinline__ static void run(float2* v) {
    constexpr Roots<R> W{};
    constexpr int H = (R - 1) / 2;
    float2 a[H], b[H];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      a[j - 1] = cadd(v[j], v[R - j]);
      b[j - 1] = csub(v[j], v[R - j]);
    }
    const float2 x0 = v[0];
    float2 y0 = x0;
#pragma unroll
    for (int j = 0; j < H; ++j) y0 = cadd(y0, a[j]);
    v[0] = y0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      float ar = x0.x, ai = x0.y, br = 0.f, bi = 0.f;
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const int m = (j * k) % R;
        ar = fmaf(a[j - 1].x, W.c[m], ar);
        ai = fmaf(a[j - 1].y, W.c[m], ai);
        br = fmaf(b[j - 1].x, W.s[m], br);
        bi = fmaf(b[j - 1].y, W.s[m], bi);
      }
      v[k] = make_float2(ar + bi, ai - br);      // A - i B
      v[R - k] = make_float2(ar - bi, ai + br);  // A + i B
    }
  }
};

// R = R1 * R2 with gcd(R1, R2) = 1 (Good-Thomas): input n = (R2 n1 + R1 n2) mod R and output
// k = (E1 k1 + E2 k2) mod R with k mod R1 = k1, k mod R2 = k2 make the R-point DFT an R1 x R2
// DFT with no twiddle factors between the stages; both index maps are register renamings.
constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 0;
}
template <int R1, int R2>
struct DftPfa {
  static constexpr int R = R1 * R2;
  static constexpr int E1 = R2 * inv_mod(R2 % R1, R1);
  static constexpr int E2 = R1 * inv_mod(R1 % R2, R2);
  static_assert(E1 % R1 == 1 && E1 % R2 == 0 && E2 % R2 == 1 && E2 % R1 == 0, "coprime radices");
  __device__ __forceinline__ static void run(float2* v) {
    float2 u[R1][R2];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) u[n1][n2] = v[(R2 * n1 + R1 * n2) % R];
      Dft<R2>::run(u[n1]);
    }
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      float2 w[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) w[n1] = u[n1][k2];
      Dft<R1>::run(w);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) v[(E1 * k1 + E2 * k2) % R] = w[k1];
    }
  }
};

template <> struct Dft<3> { __device__ __forceinline__ static void run(float2* v) { DftOdd<3>::run(v); } };
template <> struct Dft<5> { __device__ __forceinline__ static void run(float2* v) { DftOdd<5>::run(v); } };
template <> struct Dft<7> { __device__ __forceinline__ static void run(float2* v) { DftOdd<7>::run(v); } };
template <> struct Dft<11> { __device__ __forceinline__ static void run(float2* v) { DftOdd<11>::run(v); } };
template <> struct Dft<13> { __device__ __forceinline__ static void run(float2* v) { DftOdd<13>::run(v); } };
template <> struct Dft<31> { __device__ __forceinline__ static void run(float2* v) { DftOdd<31>::run(v); } };
// 33 = 3 x 11, 10 = 2 x 5 and 20 = 4 x 5 in registers, prime-factor form
template <> struct Dft<33> { __device__ __forceinline__ static void run(float2* v) { DftPfa<3, 11>::run(v); } };
template <> struct Dft<10> { __device__ __forceinline__ static void run(float2* v) { DftPfa<2, 5>::run(v); } };
template <> struct Dft<20> { __device__ __forceinline__ static void run(float2* v) { DftPfa<4, 5>::run(v); } };

// Twiddle W_n^e from the two-level table: tw1[e % split] * tw2[e / split].
__device__ __forceinline__ float2 twiddle2(const float2* tw1, const float2* tw2, int split,
                                           int e) {
  const int hi = e / split;
  const int lo = e - hi * split;
  return cmul(tw1[lo], tw2[hi]);
}

}  // namespace tga
