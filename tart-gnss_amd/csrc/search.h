// Kernel launchers of the acquisition search (search.hip), called by the host plan code.
#pragma once

#include "internal.h"

namespace tga {

// Per-row result of one (antenna, PRN, bin, hypothesis) correlation row, before scaling:
// p1 = max |z|^2 over code phases, i1 its index (last maximum on ties), p2 = max |z|^2 more
// than `cw` samples (ring distance modulo the code period) from i1 (-1 if none),
// nan = 1 if any |z|^2 was NaN.
struct RowOut {
  float p1;
  float p2;
  int32_t i1;
  int32_t nan;
};

// Maps the global PRN index of a group to its tga_peak slot: out = base + ant * stride + local.
struct PeakSlot {
  int32_t base;
  int32_t stride;
  int32_t local;
  int32_t pad;
};

// ---- decoding / de-meaning --------------------------------------------------------------
// Per-antenna means of the decoded levels (exact integer sums over kAntSplit slices per
// antenna, partial: n_ant * kAntSplit scratch); also zeroes *nan_flag (when non-null).
constexpr int kAntSplit = 64;
void launch_ant_sums(const uint8_t* packed, size_t row_bytes, int n_ant, int64_t L, int fmt,
                     long long* partial, double* means, int* nan_flag, hipStream_t st);

// ---- fused path, code period P = 16368 (16.368 MHz C/A) --------------------------------
constexpr int kFusedP = 16368;
bool fused_supported(size_t P, size_t N, int cw);
// X[(ant*n_bin + bin)*2 + h][P]: FFT of the folded sum (h=0) / twisted difference (h=1).
// carrier: the plan's carrier table (launch_carrier) of 2 P samples per bin.
void launch_fused_forward(const uint8_t* packed, size_t row_bytes, int fmt, const double* means,
                          const double* fc, int n_ant, int n_bin, double phase_const,
                          const float2* tau, const float2* tw, const float2* carrier, float2* X,
                          hipStream_t st);
// Carrier table out[bin][n] for n < clen: (cos, sin) of phi = f32(f64(pc * n) * fc[bin]) evaluated
// in f64 and rounded to f32 (the reference form), or of the f64 phase (phase64, the FP16 form).
// The wipe-off is the same for every antenna, so the forward kernels read it from this table.
void launch_carrier(const double* fc, int n_bin, long long clen, double phase_const,
                    bool phase64, float2* out, hipStream_t st);
// The fused FP32 kernels store X and read the replica spectra in the prime-factor (Good-Thomas)
// order of 16368 = 33 x 16 x 31 (position 496 n1 + 31 n2 + n3 holds natural index
// (496 n1 + 1023 n2 + 528 n3) mod 16368), for which the three radix passes of a correlation row
// need no twiddle factors.  launch_to_pfa reorders n values (rows of 16368) into it.
void launch_to_pfa(const float2* in, float2* out, long long n, hipStream_t st);
// Tunable shape of the rows kernel: `variant` 0 runs the radix-31 pass on the VALU, 1 on the
// matrix cores; `groups` PRN groups per (antenna, bin, hypothesis) (a multiple of the 8 XCDs
// keeps each group's replica spectra within one XCD's L2; fewer groups share each block's
// forward spectrum among more PRNs).
struct FusedRowsConfig {
  int variant = 0;
  int groups = 8;
  int tile = 8 | (8 << 16);  // split groups: tile shape of the row order, TX | TP << 16
};
// rows[((ant*n_prn + prn)*n_bin + bin)*2 + h]
// dump (surface form, here and in launch_fused_rows16 / launch_big_sub): when not null, every
// output power of row (or long row) i is also stored at dump[i * P + sample], P the code period.
void launch_fused_rows(const float2* X, const float2* rep, const float2* tw, int n_ant,
                       int n_bin, int n_prn, int cw, RowOut* rows, hipStream_t st,
                       const FusedRowsConfig& cfg, float* dump = nullptr);
// Non-coherent form (P = 16368): X[(ant*n_bin + bin)*epochs + e][P] are the spectra of the
// one-period epochs; rows[((ant*n_prn + prn)*n_bin + bin)*2] hold the top-2 of the power sums
// over epochs (the h = 1 slot is a row that never wins the selection).
void launch_fused_forward_nc(const uint8_t* packed, size_t row_bytes, int fmt,
                             const double* means, const double* fc, int n_ant, int n_bin,
                             int epochs, double phase_const, const float2* tw,
                             const float2* carrier, float2* X, hipStream_t st);
void launch_fused_rows_nc(const float2* X, const float2* rep, const float2* tw, int n_ant,
                          int n_bin, int n_prn, int epochs, int cw, RowOut* rows, int groups,
                          hipStream_t st);
// FP16 form of the fused path: X and the replica spectra are stored as f16 (re, im) pairs of
// the value times kF16Scale (so |conj(X) R| stays far inside the f16 range for 1- and 2-bit
// input), the radix passes run as f16 GEMMs on the matrix cores with f32 accumulation, and the
// row values are those of the transform times kF16Scale^2 (powers times kF16Scale^4).  X16 and
// rep16 rows are stored in the prime-factor (Good-Thomas) order of 16368 = 33 x 16 x 31 (position
// 496 n1 + 31 n2 + n3 holds natural index (496 n1 + 1023 n2 + 528 n3) mod 16368), for which the
// three radix passes need no twiddle factors.
constexpr float kF16Scale = 1.0f / 256.0f;
// Sub-rows of long rows (P = 4 or 10 x 16368) after the first split stage reach about
// sqrt(R0) |X| |R| ~ 1e6 for 1-bit input; this brings them near 1 in f16.
constexpr float kSplitF16Scale = 1.0f / 1048576.0f;  // 2^-20
void launch_fused_forward16(const uint8_t* packed, size_t row_bytes, int fmt,
                            const double* means, const double* fc, int n_ant, int n_bin,
                            double phase_const, const float2* tau, const float2* tw,
                            const float2* carrier, uint32_t* X16, hipStream_t st);
void launch_fused_rows16(const uint32_t* X16, const uint32_t* rep16, const float2* tw,
                         int n_ant, int n_bin, int n_prn, int cw, RowOut* rows, int groups,
                         hipStream_t st, float* dump = nullptr);
// Fills `bytes` of autotuner spectra with a fixed pseudo-random pattern of values in [-1, 1)
// (f32 words, or f16 pairs with f16 = true).
void launch_fill_tune(void* p, size_t bytes, bool f16, hipStream_t st);
// Forward spectra of long code periods P = R0 x 16368 (R0 = 4 or 10), folded (N = 2P: rows
// (ant, bin, h)) or not (N = P: rows (ant, bin)), in two steps: one 16368-point LDS transform
// per (row, r) of the samples R0 m + r into Y, then the twiddles and R0-point DFTs into X, in
// the segment-wise prime-factor order of the split.  carrier: launch_carrier_dec's table; tau:
// tau_n = exp(-i pi n / P) in the order [r][m] (n = R0 m + r); twl: long_twiddles(P).
void launch_carrier_dec(const double* fc, int n_bin, int R0, bool fold, double phase_const,
                        bool phase64, float2* out, hipStream_t st);
std::vector<float2> long_twiddles(size_t P);
// lev: scratch for the decoded levels, n_ant * (fold ? 2 : 1) * P floats.
void launch_fwd_long(const uint8_t* packed, size_t row_bytes, int fmt, const double* means,
                     const float2* carrier, const float2* tau, const float2* tw,
                     const float2* twl, int R0, bool fold, int n_ant, int n_bin, float* lev,
                     float2* Y, float2* X, hipStream_t st);
// n values = n / 16368 natural-order rows -> f16 pairs times scale in the prime-factor order
void launch_to_f16(const float2* in, uint32_t* out, long long n, float scale, hipStream_t st);
// Twiddle table of the fused kernels: W_P^i (i < 128) then W_P^{128 i} (i < 128).
std::vector<float2> fused_twiddles();

// ---- generic path -------------------------------------------------------------------------
// folded input rows (S / twisted E) for any P, N = 2P: out[(ant*n_bin + bin)*2 + h][P]
// phase64: carrier phase in f64 throughout (the FP16 precision form) instead of the reference's
// phi = f32(pp * fc)
void launch_fold_obs(const uint8_t* packed, size_t row_bytes, int fmt, const double* means,
                     const double* fc, int n_ant, int n_bin, int P, double phase_const,
                     const float2* tau, float2* out, bool phase64, hipStream_t st);
// unfolded wiped rows for N-point correlation: out[ant*n_bin + bin][N]
void launch_wipe_obs(const uint8_t* packed, size_t row_bytes, int fmt, const double* means,
                     const double* fc, int n_ant, int n_bin, int N, double phase_const,
                     float2* out, bool phase64, hipStream_t st);
// level 1: out[bin][N] from f32 samples and f64 phase points
void launch_wipe_signal(const float* sig, const double* pp, const double* fc, int n_bin, int N,
                        float2* out, hipStream_t st);
// Z[r][i] = conj(X[xrow(r)][i]) * rep[rep(r)][i] for rows r in [r0, r1) of the ordering
// r = ((ant*n_prn + prn)*n_bin + bin)*2 + h; folded: xrow = (ant*n_bin+bin)*2+h, else
// xrow = ant*n_bin+bin; rep index = prn*2 + h.
void launch_mul_rows(const float2* X, const float2* rep, int M, int n_prn, int n_bin,
                     bool folded, long long r0, long long r1, float2* Z, hipStream_t st);
// One block per row of Z: argmax / masked second maximum of |z|^2 (ring distance mod P).
void launch_row_reduce(const float2* Z, int M, int P, int cw, long long n_rows, RowOut* out,
                       hipStream_t st);

// ---- selection ----------------------------------------------------------------------------
// Per (ant, prn): last-max over bins of the row peak, [c,-c] only if strictly larger; writes
// tga_peak through `slots`; sets *nan_flag if any row saw a NaN.  corr = sqrt(p)*gain*scale
// with gain 2 for folded rows, 1 for unfolded ones.
// `win` (optional): per (ant, prn) the winning row's bin * 2 + h.
void launch_select(const RowOut* rows, const PeakSlot* slots, int n_ant, int n_prn, int n_bin,
                   float gain, float scale, tga_peak* out, int* nan_flag, hipStream_t st,
                   int* win = nullptr);

// ---- long code periods P = R0 x 16368 (R0 = 4: E1-C, R0 = 10: B1C / L1C at 16.368 MHz) ------
// Folded rows of length P are transformed as R0 twiddled sub-rows of 16368 points by the fused
// LDS kernel; each row keeps only its argmax, and the second peak comes from a recomputation of
// the winning row of every (antenna, PRN) pair.  The forward and replica spectra of these rows
// are stored segment by segment in the prime-factor order of 16368 (launch_to_pfa over the
// whole buffer: position 16368 n0 + t holds natural index 16368 n0 + pfa_nat(t)).
// folded rows (N = 2P) or unfolded single-period rows (N = P, a 10 ms B1C / L1C record)
bool big_supported(size_t P, size_t N, bool folded);
// The pre-pass twiddles W_P^{pfa_nat(s)} for the positions s < 16368 of a segment in the
// prime-factor order.
std::vector<float2> big_twiddles(size_t P);
// Rows [r0, r1) of the tiled enumeration (tiles of TX forward spectra x TP replica spectra, whose
// rows reuse them from the caches, as listed by the row table); results land at the usual row
// index.  launch_big_pre writes the twiddled
// sub-rows of the rows to Y ((r1 - r0) * P complex values) and clears their keys;
// launch_big_sub transforms them and writes the rows' RowOut.
// hs = 1: every (bin, h) row; hs = 2: unfolded rows, h = 0 only (the [c,-c] hypothesis equals
// [c,c] when N <= P, correlate.rs:53-56, and gets a row that never wins the selection).
// f16: Y holds f16 (re, im) pairs scaled by kSplitF16Scale (uint32 per value) and the sub-rows
// run through the matrix-core rows kernel, whose powers then carry kSplitF16Scale^2; otherwise
// Y holds float2 and the f32 rows kernel transforms them.
// tile: TX | TP << 16, rows enumerated in tiles of TX forward spectra x TP replica spectra
void launch_big_pre(const float2* X, const float2* rep, const float2* twb, const int4* rowtab,
                    int R0, long long r0, long long r1, void* Y, bool f16,
                    unsigned long long* rowkey, int* rownan, int* rowmap, hipStream_t st);
// The row table the pre-passes read (rows [0, n) of the tiled order of `tile`): per row the
// forward-spectrum row, the replica row and the RowOut row index, as int4 (x, y, z, 0).
void launch_big_rowtab(int4* tab, long long n, int n_prn, int n_bin, int n_ant, int hs, int tile,
                       hipStream_t st);
// variant: the FP32 sub-row transforms (0: k_rows with the VALU radix-31 codelet, 1: k_rows with
// the matrix-core radix 31 and the next sub-row prefetched into registers, 2: k_subrows, the
// matrix-core form with pass B in place and each lane's best output pair); the FP16 form has one
// (variant 5, the half sub-rows, has its own launchers below).
// rowmap: the row index of each of the chunk's rows, written by launch_big_pre (r1 - r0 ints).
void launch_big_sub(const float2* tw, int R0, int n_prn, int n_bin, int n_ant, int hs,
                    int tile, long long r0, long long r1, const void* Y, bool f16,
                    unsigned long long* rowkey, int* rownan, const int* rowmap, RowOut* rows,
                    hipStream_t st, int variant, float* dump = nullptr);
// Half sub-rows (FP32 split variant 5): the R0 x 16368-point rows as R = 2 R0 sub-rows of
// 8184 = 33 x 8 x 31 points in their CRT order (two transforming workgroups per CU), read from
// the same segment-wise prime-factor spectra.  twh = half_twiddles(P): W_P^m and W_P^{R0 m} for
// the natural index m of each of the 8184 positions.  launch_half_pre writes a chunk's sub-rows
// to Y and clears the rows' keys; launch_half_sub transforms them and writes the rows' RowOut;
// dump as launch_big_sub's.
std::vector<float2> half_twiddles(size_t P);
void launch_half_pre(const float2* X, const float2* rep, const float2* twh, const int4* rowtab,
                     int R0, long long r0, long long r1, float2* Y, unsigned long long* rowkey,
                     int* rownan, int* rowmap, hipStream_t st);
void launch_half_sub(const float2* tw, int R0, int n_prn, int n_bin, int n_ant, int hs, int tile,
                     long long r0, long long r1, const float2* Y, unsigned long long* rowkey,
                     int* rownan, const int* rowmap, RowOut* rows, hipStream_t st,
                     float* dump = nullptr);
// surface form: v[i] = (sqrt(v[i]) * gain) * scale in place (the row values of k_select)
void launch_surface_finish(float* v, long long n, float gain, float scale, hipStream_t st);
// second peaks of the winning rows of sub-row split groups (spectra in the segment-wise
// prime-factor order; R0 = 1: the fused 16368-sample rows, read from the f16 spectra X16 and
// rep16 when those are given): their sub-rows into Zw (n_ant * n_prn * R0 * 16368 complex
// values), the masked sub-row transforms into m2key (n_ant * n_prn), then the peaks'
// second_peak
void launch_big_second_sub(const float2* X, const float2* rep, const float2* tw,
                           const float2* twb, const PeakSlot* slots, int R0, int cw, int n_ant,
                           int n_prn, int n_bin, bool folded, const int* win, float gain,
                           float scale, float2* Zw, unsigned int* m2key, tga_peak* out,
                           hipStream_t st, const uint32_t* X16 = nullptr,
                           const uint32_t* rep16 = nullptr);
// second peaks of the winning rows; Z, Zs hold n_ant * n_prn * M complex values; natural: X and
// rep in natural order (the four-step split) instead of the segment-wise prime-factor order
void launch_big_second(const float2* X, const float2* rep, const PeakSlot* slots, int M, int P,
                       int cw, int n_ant, int n_prn, int n_bin, bool folded, const int* win,
                       float gain, float scale, float2* Z, float2* Zs, tga_peak* out,
                       hipStream_t st, bool natural = false);
// Four-step split (FP32, variant 3): long rows in natural order as N1 x 496 (N1 = R0 x 33):
// k_4s_pre writes the N1-point column transforms, twiddled, as rows of 496 into Y ((r1 - r0) * P
// values) and clears the rows' keys; k_4s_rows transforms the rows of 496 and folds each pair's
// argmax into the row keys; then the rows' RowOut.  tw4 = four_step_twiddles(P), c31 =
// radix31_table(); dump as launch_big_sub's.
std::vector<float2> four_step_twiddles(size_t P);
std::vector<float2> radix31_table();
void launch_big_4s(const float2* X, const float2* rep, const float2* tw4, const float2* c31,
                   int R0, int n_prn, int n_bin, int n_ant, int hs, int tile, long long r0,
                   long long r1, float2* Y, unsigned long long* rowkey, int* rownan,
                   RowOut* rows, hipStream_t st, float* dump = nullptr);

}  // namespace tga
