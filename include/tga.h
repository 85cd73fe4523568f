/*
 * tga.h -- C ABI of the MI355X (gfx950) acquisition-search library.
 *
 * The library replaces the CPU code-phase/Doppler search of tart-gnss-acquire with HIP
 * kernels.  Every entry point takes plain pointers and sizes, borrows caller memory only
 * for the duration of the call, returns an int status (TGA_OK or a negative TGA_ERR_*),
 * never aborts, and leaves a human-readable message for tga_last_error() (thread-local).
 *
 * Three levels, mirroring the reference's Rust API:
 *   level 1  tga_correlate_code      <- correlate::correlate_code   (src/correlate.rs:91-183)
 *   level 2  tga_acquire_single      <- acquire_full / acquire_*_single (acquisition.rs:176-248,
 *                                       galileo.rs:133-206, beidou.rs:85-158, l1c.rs:94-167,
 *                                       sbas.rs:132-206, qzss.rs:135-209)
 *   level 3  tga_plan_* / tga_search <- acquire_all_{gps,galileo,beidou,sbas,l1c,qzss}
 *                                       (acquisition.rs:261-396 and siblings), batched over
 *                                       every antenna, PRN, Doppler bin and hypothesis of every
 *                                       system in one pass per observation.
 * plus replica generation, ACR C/N0 and the RFI report.
 */
#ifndef TGA_H
#define TGA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------------------
 * Each negative code maps to one panic site of the reference, so a Rust shim can re-raise
 * the reference's own panic message. */
#define TGA_OK 0
#define TGA_ERR_INVALID_ARGUMENT (-1) /* null pointer, zero length, inconsistent sizes      */
#define TGA_ERR_EMPTY_BINS (-2)       /* correlate.rs:103 "frequency bin list must be non-empty" */
#define TGA_ERR_PRN_RANGE (-3)        /* acquisition.rs:131-134, sbas.rs:112-116, qzss.rs:112-116,
                                         galileo.rs:100-103, beidou_codes.rs:146-149,
                                         l1c_codes.rs:176-180 */
#define TGA_ERR_ANTENNA_RANGE (-4)    /* observation.rs:141-146                            */
#define TGA_ERR_NAN (-5)              /* correlate.rs:138,153 partial_cmp().unwrap() on NaN */
#define TGA_ERR_UNSUPPORTED (-6)      /* a shape this build does not implement              */
#define TGA_ERR_DEVICE (-7)           /* a HIP runtime error                                */
#define TGA_ERR_NO_DEVICE (-8)        /* no gfx950 device visible                           */
#define TGA_ERR_OUT_OF_MEMORY (-9)

/* GNSS systems, in the key order of CombinedOutput (lib.rs:31-45). */
#define TGA_SYS_GPS 0     /* GPS L1 C/A, PRN 1..38      (acquisition.rs:69)  */
#define TGA_SYS_GALILEO 1 /* Galileo E1-C, PRN 1..50    (galileo_codes.rs:10) */
#define TGA_SYS_BEIDOU 2  /* BeiDou B1Cp, PRN 1..63     (beidou_codes.rs:15)  */
#define TGA_SYS_SBAS 3    /* SBAS L1 C/A, PRN 120..158  (sbas.rs:21,30)       */
#define TGA_SYS_L1C 4     /* GPS L1Cp, PRN 1..63        (l1c_codes.rs:14)     */
#define TGA_SYS_QZSS 5    /* QZSS L1 C/A, PRN 184..206  (qzss.rs:21,30)       */
#define TGA_NUM_SYSTEMS 6

/* Sample formats of a packed observation row. */
#define TGA_INPUT_1BIT 1 /* MSB-first bits, 1 -> +1, 0 -> -1 (observation.rs:140-151,232-240)  */
#define TGA_INPUT_2BIT 2 /* MSB-first bit pairs (sign, magnitude): sign 1 -> +, magnitude 1 -> 3,
                            so the levels are {-3,-1,+1,+3}; an extension, the reference reads
                            1-bit data only */

/* ACR tables (acr_tables.rs:23,87,163,247). */
#define TGA_ACR_GPS_L1_CA 0 /* GPS, SBAS and QZSS C/A */
#define TGA_ACR_GALILEO_E1C 1
#define TGA_ACR_BEIDOU_B1C 2
#define TGA_ACR_GPS_L1C 3

/* Result of one (antenna, PRN) search -- mirrors CorrelationPeak (correlate.rs:27-38):
 * best_peak = max |IFFT(X . conj C_h)| / sqrt(N) over bins, hypotheses and code phases;
 * second_peak = largest value more than one chip (ceil(fs/1.023 MHz) samples, ring distance
 * modulo the code period) away from the peak in the winning row, or 1e-6. */
typedef struct tga_peak {
  float best_peak;
  float second_peak;
  uint64_t best_freq_idx;  /* index into the bin list the caller gave */
  uint64_t best_codephase; /* level 1: sample index in [0, N); level 3: index modulo the code
                              period, which is all acquire_all_* uses (acquisition.rs:237) */
} tga_peak;

/* Thread-local description of the last error on the calling thread ("" if none). */
const char* tga_last_error(void);
/* Library version string. */
const char* tga_version(void);
/* Number of visible HIP devices (0 on a host without one; never an error). */
int tga_device_count(int* count);

/* ---- replicas (level 2 building blocks) ------------------------------------------------ */

/* Number of samples-per-period chips of a system's primary code: 1023 (C/A), 4092 (E1-C),
 * 20460 (B1C/L1C after the BOC(1,1) [+c,-c] split). */
int tga_code_length(int32_t system, size_t* chips);

/* Chips (+1/-1) of one PRN: generate_ca_code (acquisition.rs:130-137), generate_sbas_ca_code
 * (sbas.rs:111-118), generate_qzss_ca_code (qzss.rs), generate_e1c_code (galileo.rs:99-105),
 * generate_b1c_pilot_code (beidou_codes.rs:145-158), generate_l1c_pilot_code
 * (l1c_codes.rs:175-189).  `cap` must be at least tga_code_length(). */
int tga_code_chips(int32_t system, int32_t prn, int8_t* out, size_t cap);

/* The replica resampled to exactly num_samples samples: c[n] = chips[floor(n/(spc/L)) mod L]
 * (acquisition.rs:146-165, galileo.rs:112-122, beidou_codes.rs:166-177,
 * l1c_codes.rs:197-208). */
int tga_code_resampled(int32_t system, int32_t prn, double samples_per_code,
                       size_t num_samples, double* out);

/* code_spectra (correlate.rs:47-68): conj(FFT_n(code)) and conj(FFT_n(code with samples
 * >= min(period, n) negated)), interleaved (re, im) float pairs, computed on the GPU. */
int tga_code_spectra(const double* code, size_t n, size_t period, float* code_complex,
                     float* code_complex_alt);

/* ---- level 1 --------------------------------------------------------------------------- */

/* correlate_code (correlate.rs:91-183), same arguments without the FFT plans:
 *   signal      num_samples de-meaned f32 samples
 *   phasepoints num_samples values TAU*i/fs (f64)
 *   code_complex, code_complex_alt  num_samples complex (re, im) float pairs, already
 *               conjugated, as code_spectra returns them
 *   fc          n_fc absolute bin frequencies (Hz)
 * Same semantics: f32(pp*fc) phase, e^{+j phase} wipe-off, last-max argmax in a row,
 * [c,-c] only when strictly larger, last max across bins, masked second peak.  Any
 * num_samples >= 1 is accepted.  Reentrant and thread-safe (rayon calls it concurrently). */
int tga_correlate_code(const float* signal, const double* phasepoints,
                       const float* code_complex, const float* code_complex_alt,
                       const double* fc, size_t n_fc, size_t num_samples,
                       size_t code_period_samples, double sampling_freq, tga_peak* out);

/* tga_correlate_code without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory -- code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Same semantics and kernels as tga_correlate_code,
 * but every argument is DEVICE memory the caller owns (d_code_complex / d_code_complex_alt:
 * num_samples complex (re, im) float pairs, already conjugated, as code_spectra returns them),
 * the peak is written to d_out (device) and d_nan (device, may be NULL) receives 1 where
 * tga_correlate_code would return TGA_ERR_NAN, else 0.  The call only enqueues work on
 * `stream` (a hipStream_t, NULL for the null stream) and returns; it never synchronises.
 * Working memory is a fixed 22 MiB device array inside the library, which each GPU the library
 * runs on holds until the process ends; calls that overlap on one GPU share it and race (one
 * call at a time per GPU).  Limits, else TGA_ERR_UNSUPPORTED: 2 <= num_samples <= 327360 with
 * prime factors in {2, 3, 5, 7, 11, 13, 31} (every 16.368 MHz and 4.092 MHz reference length),
 * n_fc <= 4096. */
int tga_correlate_code_noalloc(const float* d_signal, const double* d_phasepoints,
                               const float* d_code_complex, const float* d_code_complex_alt,
                               const double* d_fc, size_t n_fc, size_t num_samples,
                               size_t code_period_samples, double sampling_freq, tga_peak* d_out,
                               int32_t* d_nan, void* stream);

/* ---- level 2 --------------------------------------------------------------------------- */

/* One PRN on one antenna (acquire_full and siblings): 300 Hz grid
 * fc[i] = center - band + i*2*band/(n-1), n = round(2*band/300)+1 (single bin if n == 1),
 * replica resampled to n_samples, code_spectra, correlate_code, then
 * codephase_frac = (best_codephase mod P)/P and frequency = fc[best] - center. */
typedef struct tga_acq_result {
  int32_t prn;
  int32_t pad;
  double signal_strength;
  double second_peak;
  double codephase_frac;
  double frequency;
} tga_acq_result;

int tga_acquire_single(int32_t system, const float* signal, const double* phasepoints,
                       size_t n_samples, double sampling_freq, double center_freq,
                       double search_band, int32_t prn, size_t samples_per_period,
                       tga_acq_result* out);

/* ---- level 3: batched search plans ----------------------------------------------------- */

/* One system of a plan. */
typedef struct tga_system_desc {
  int32_t system;      /* TGA_SYS_* */
  int32_t n_prn;       /* number of PRNs; 0 with prns == NULL selects the full range */
  const int32_t* prns; /* PRN numbers in the reference's numbering */
  int32_t n_fc;        /* number of Doppler bins (>= 1) */
  const double* fc;    /* absolute bin centre frequencies (Hz), e.g. IF + Doppler */
  int64_t num_samples; /* coherent length N; 0 selects the reference's min(2P, record) */
  /* 0: the reference form (N-sample coherent correlation, [c,c] / [c,-c] hypotheses).
   * E >= 1: non-coherent form, an extension with no counterpart in the reference: E
   * consecutive one-period epochs (E P <= record_samples), |y_e|^2 summed over them with y_e
   * the unnormalised inverse-transform correlation of epoch e; best_peak = sqrt(sum) / sqrt(P)
   * (the reference form's scale for E = 1), second peak masked as in the reference form.  Needs the
   * fused P = 16368 path (TGA_ERR_UNSUPPORTED otherwise); ACR C/N0 tables are not calibrated
   * for it. */
  int32_t nc_epochs;
  /* TGA_REPLICA_REFERENCE (0): the reference's replica for the system.  TGA_REPLICA_E1C_BOC11
   * (1), Galileo only, an extension with no counterpart in the reference (whose E1-C replica is
   * the plain 4092-chip code, galileo.rs:112-122): the E1-C chips with the BOC(1,1) subcarrier,
   * each chip c as the half-chips [+c, -c] (8184 per 4 ms period), the usual receiver model of
   * the E1-C pilot's CBOC(6,1,1/11) modulation.  Its secondary code (CS25, one chip per 4 ms
   * period) needs no wipe inside an N = 2P search: the [c,c] / [c,-c] hypotheses cover both
   * signs of the second period.  ACR C/N0 uses the E1-C table, calibrated for the plain replica
   * on plain-chip signals. */
  int32_t replica;
} tga_system_desc;
#define TGA_REPLICA_REFERENCE 0
#define TGA_REPLICA_E1C_BOC11 1

/* Arithmetic form of the correlation rows.
 *   TGA_PRECISION_FP32  the reference form: f32 spectra and f32 transforms, results within
 *                       1e-5 of the reference's peaks (the parity path).
 *   TGA_PRECISION_FP16  an extension with no counterpart in the reference: groups served by
 *                       the fused P = 16368 kernels store the spectra as f16 and run the
 *                       radix-33/16/31 passes as f16 GEMMs on the matrix cores with f32
 *                       accumulation; the carrier phase stays in f64 (the FP32 form rounds
 *                       it to f32 like the reference);
 *                       other groups run the FP32 form.  A row whose values leave the f16
 *                       range reports TGA_ERR_NAN. */
#define TGA_PRECISION_FP32 0
#define TGA_PRECISION_FP16 1

typedef struct tga_plan_desc {
  double sampling_freq;   /* Hz; 16.368e6 for TART */
  int32_t n_ant;          /* antenna rows per observation */
  int32_t input_format;   /* TGA_INPUT_1BIT or TGA_INPUT_2BIT */
  int64_t record_samples; /* samples per antenna row (the whole record: de-meaning uses all) */
  int32_t n_systems;
  int32_t precision;      /* TGA_PRECISION_FP32 (0) or TGA_PRECISION_FP16 */
  const tga_system_desc* systems;
} tga_plan_desc;

typedef struct tga_plan tga_plan;

/* Builds replica spectra, FFT tables and device buffers for every system.  The plan owns all
 * device memory; one plan must not be used by two threads at once. */
int tga_plan_create(const tga_plan_desc* desc, tga_plan** plan);
int tga_plan_destroy(tga_plan* plan);

/* Bytes per packed antenna row: ceil(record_samples * bits / 8). */
int tga_plan_row_bytes(const tga_plan* plan, size_t* row_bytes);
/* Number of tga_peak records a search writes: sum over systems of n_ant * n_prn.
 * Layout: systems in plan order, then antenna, then PRN in the order given. */
int tga_plan_peak_count(const tga_plan* plan, size_t* count);
/* PRN list of system `sys_index` of the plan (after defaulting). */
int tga_plan_prns(const tga_plan* plan, int32_t sys_index, int32_t* prns, size_t cap,
                  size_t* n_prn);

/* Search one observation held in host memory: packed[n_ant][row_bytes]. */
int tga_search(tga_plan* plan, const uint8_t* packed, size_t n_ant, size_t row_bytes,
               tga_peak* out);

/* Search one observation already resident in device memory (d_packed is a device pointer
 * to n_ant * row_bytes bytes).  `stream` is a hipStream_t or NULL for the plan's stream.
 * Peaks are copied to `out` in host memory before the call returns. */
int tga_search_device(tga_plan* plan, const void* d_packed, void* stream, tga_peak* out);

/* hipGraph replay, on by default.  A plan's first search launches its kernels one by one; the
 * second captures the whole search (kernel nodes only, reading the plan's own input buffer)
 * into a graph; every later search is one graph launch, preceded by a device-to-device copy
 * of d_packed into that buffer when d_packed is another address.  While timing is on, and for
 * plans with a Bluestein transform, searches run eagerly.  enable = 0 drops the graph.
 * No counterpart in the reference (its searches are rayon loops on the CPU). */
int tga_plan_set_graph(tga_plan* plan, int32_t enable);

/* Per-kernel timing with HIP events recorded around each launch on the launch stream.
 * kernel: 0 = correlate rows (the dominant kernel), 1 = forward spectra, 2 = select. */
int tga_plan_set_timing(tga_plan* plan, int32_t enable);
int tga_plan_get_timing(tga_plan* plan, int32_t kernel, double* total_ms, int64_t* launches,
                        double* bytes_per_launch);
/* Matrix-core arithmetic of one search's correlation rows (timing slot 0), counted from the
 * kernels each execution group runs after autotuning: f32_flops for v_mfma_f32_16x16x4f32
 * (2048 FLOP each: the radix-31 pass of the FP32 rows and sub-rows, the four-step rows) and
 * f16_flops for v_mfma_f32_16x16x32_f16 (16384 FLOP each: every pass of the FP16 form).  Divided
 * by the slot's time it gives the matrix cores' achieved rate in the DFT-as-GEMM stages.
 * No counterpart in the reference (a measurement aid). */
int tga_plan_rows_matrix_flops(const tga_plan* plan, double* f32_flops, double* f16_flops);
/* Short names of the kernels behind a timing slot, as rocprof lists them, joined by '+' over
 * the plan's execution groups (e.g. "k_rows+k_big_pre<4>+k_rows<0,4>" for a C/A group beside
 * an E1-C split group); "" for a slot out of range. */
const char* tga_plan_kernel_name(const tga_plan* plan, int32_t kernel);
/* Describes the plan's execution path for each system ("fused16368", "generic", ...). */
int tga_plan_describe(const tga_plan* plan, char* buf, size_t cap);

/* Correlation surface of a search (no counterpart in the reference, which forms the same
 * values, |IFFT(X . code_complex)|[i] / sqrt(N) per bin and hypothesis in correlate.rs:118-146,
 * and keeps only their maxima): for stating the error of a precision form against an FP64
 * truth.  Searches d_packed (device memory, as tga_search_device, eagerly) and writes every
 * row's values at code phases i < P -- the reference's surface at i and i + P for N = 2P --
 * to d_surface (device memory), system by system in plan order as [ant][prn][bin][h][P]
 * floats on the scale of tga_peak::best_peak (the winning row's maximum equals its best_peak).
 * Unfolded rows (N = P) have no h = 1 row: zeros.  out (host, optional) receives the search's
 * peaks as tga_search_device would.  Covers plans whose groups run the fused (P = 16368) or
 * split (P = 4 or 10 x 16368) kernels, else TGA_ERR_UNSUPPORTED; cap < the size
 * tga_plan_surface_size reports is TGA_ERR_INVALID_ARGUMENT.  Allocates a staging buffer of the
 * same size. */
/* The carrier table a fused (16368-sample) group's forward kernel multiplies the samples by,
 * copied to host memory: (cos, sin) pairs [bin][sample] of phi = f32(f64(pp[n]) * fc) through
 * f64 sincos rounded to f32 (correlate.rs:110-115), or of the f64 phase in the FP16 form; n_floats
 * receives 2 x bins x samples (samples = 2P, or P x epochs in the non-coherent form); out may be
 * NULL to query the size.  A diagnostic with no counterpart in the reference.  Other groups:
 * TGA_ERR_UNSUPPORTED. */
int tga_plan_carrier(const tga_plan* plan, int32_t sys_index, float* out, size_t cap,
                     size_t* n_floats);
/* Floats of system sys_index's surface, or of the whole plan's for sys_index = -1. */
int tga_plan_surface_size(const tga_plan* plan, int32_t sys_index, size_t* n_floats);
int tga_plan_surface(tga_plan* plan, const void* d_packed, void* stream, float* d_surface,
                     size_t cap, tga_peak* out);

/* ---- ACR C/N0 (acr.rs:49-84; gate acquisition.rs:345-352) ------------------------------ */

/* estimate_cn0: *valid = 0 for None (r_a <= 1 or not finite). */
int tga_estimate_cn0(double r_a, int32_t table, double* cn0, int32_t* valid);
/* The gate plus lookup used by acquire_all_*: valid only if v_s > 0 and v_m > v_s. */
int tga_acr_cn0(double v_m, double v_s, int32_t table, double* cn0, int32_t* valid);

/* ---- RFI report (rfi.rs:83-184) -------------------------------------------------------- */

typedef struct tga_rfi_line {
  double frequency;
  double excess_db;
} tga_rfi_line;

typedef struct tga_rfi_channel {
  int32_t antenna;
  int32_t dead;          /* rfi.rs:146-156: variance < 1e-9 */
  double acf_lag1;       /* NaN when dead */
  double acf_lag3;
  double acf_lag16;
  double period_samples; /* fs / frequency of the strongest line; NaN when no line */
  int32_t n_lines;       /* 0..3 */
  int32_t pad;
  tga_rfi_line lines[3]; /* strongest first */
} tga_rfi_channel;

/* analyze_channel for every antenna of a packed observation (1- or 2-bit), on the GPU:
 * f64 FFT of the whole de-meaned record, |X|^2 below n/2, median floor, local maxima
 * > 15 dB, 10 kHz clustering, at most 3 lines, ACF at lags 1, 3 and 16.  record_samples
 * must exceed 16 (the largest ACF lag). */
int tga_rfi(const uint8_t* packed, size_t n_ant, size_t row_bytes, int64_t record_samples,
            int32_t input_format, double sampling_freq, tga_rfi_channel* out);
/* The same report for an observation already in device memory (d_packed: n_ant * row_bytes
 * bytes of device memory), e.g. the buffer a tga_search_device call searches. */
int tga_rfi_device(const void* d_packed, size_t n_ant, size_t row_bytes, int64_t record_samples,
                   int32_t input_format, double sampling_freq, tga_rfi_channel* out);

/* The --rfi report as JSON (serde_json::to_string_pretty of RfiReport, rfi.rs:73-77, written
 * at main.rs:468): {"antenna_numbers": [...], "channels": [{antenna, dead, acf_lag1,
 * acf_lag3, acf_lag16, period_samples, lines: [{frequency, excess_db}]}]}, None (NaN in the
 * struct) as null.  tga_rfi_report_json serialises channels a tga_rfi / tga_rfi_device call
 * returned (host only, no GPU); tga_rfi_json runs tga_rfi on a packed observation first
 * (replaces rfi::run + the serialisation, main.rs:439-478).  The string is library-owned:
 * release it with tga_free_string. */
int tga_rfi_report_json(const tga_rfi_channel* channels, size_t n_ant, char** json);
int tga_rfi_json(const uint8_t* packed, size_t n_ant, size_t row_bytes, int64_t record_samples,
                 int32_t input_format, double sampling_freq, char** json);

/* ---- host orchestration and JSON (acquire_all_* + lib.rs CombinedOutput) ---------------- */

/* Runs the batched search of `plan` on one packed observation and writes the reference's JSON
 * (serde_json::to_string_pretty of CombinedOutput, main.rs:893, lib.rs:31-45) into a
 * library-owned string released with tga_free_string.
 *   center_freq[s]     IF subtracted from the bin frequency to report the Doppler, per plan
 *                      system (4.092e6 for every system of the reference)
 *   antenna_numbers    n_ant labels written to "antenna_numbers" (the antenna filter)
 *   cn0                non-zero adds "cn0_acr" (acquisition.rs:341-357)
 *   filter_phase_mad, filter_freq_mad
 *                      MAD filters of the CLI (main.rs:242-255); negative or NaN = off
 * Per SV: strengths, phases ((best_codephase mod P)/P), freqs, optional cn0_acr, and the
 * median/MAD fields when n_ant > 1 (acquisition.rs:365-377); results sorted by sv label. */
int tga_acquire_all_json(tga_plan* plan, const uint8_t* packed, size_t n_ant,
                         size_t row_bytes, const double* center_freq,
                         const int64_t* antenna_numbers, int32_t cn0, double filter_phase_mad,
                         double filter_freq_mad, char** json);
/* The same JSON from peaks a previous tga_search / tga_search_device returned for `plan`. */
int tga_peaks_json(tga_plan* plan, const tga_peak* peaks, const double* center_freq,
                   const int64_t* antenna_numbers, int32_t cn0, double filter_phase_mad,
                   double filter_freq_mad, char** json);

/* ACR C/N0 of every peak of a previous tga_search for `plan` (acquisition.rs:341-357 per
 * (antenna, PRN)): cn0[i] for peaks[i], from its system's table, with the reference's gate
 * (v_s > 0 and v_m > v_s) and lookup; NaN where the reference has None.  Host-only, f64. */
int tga_peaks_cn0(const tga_plan* plan, const tga_peak* peaks, double* cn0);

/* ---- antenna-relative C/N0 report (--test-antennas) -------------------------------------- */

/* antenna_test::run (antenna_test.rs:105-222) and its JSON (serde_json::to_string_pretty of
 * AntennaTestOutput, main.rs:848-880), written into a library-owned string.
 * Rows are the cn0_acr vectors of CombinedOutput in its order (gps, galileo, beidou, sbas,
 * l1c, qzss; each sorted by label): sv[r] is the label, cn0[r * n_ant + a] the estimate of
 * antenna column a, NaN where that antenna has none (Option::None); a row of NaN is a
 * satellite without cn0_acr.  A satellite enters the reference set when some antenna reaches
 * min_cn0 (44.0 in the reference CLI); each antenna column is scored by the median of its
 * estimates over that set; columns without any estimate stay unscored (null fields).
 * Needs no GPU. */
int tga_antenna_report_json(size_t n_sv, const char* const* sv, const double* cn0,
                            size_t n_ant, const int64_t* antenna_numbers, double min_cn0,
                            char** json);
/* The same report from peaks a previous tga_search returned for `plan`, with the C/N0
 * estimates forced on (main.rs:181-182) and the MAD filters applied first (main.rs:820-845). */
int tga_peaks_antenna_report_json(tga_plan* plan, const tga_peak* peaks,
                                  const double* center_freq, const int64_t* antenna_numbers,
                                  double filter_phase_mad, double filter_freq_mad,
                                  double min_cn0, char** json);
/* Search one packed observation with `plan`, then tga_peaks_antenna_report_json. */
int tga_acquire_antenna_report_json(tga_plan* plan, const uint8_t* packed, size_t n_ant,
                                    size_t row_bytes, const double* center_freq,
                                    const int64_t* antenna_numbers, double filter_phase_mad,
                                    double filter_freq_mad, double min_cn0, char** json);
void tga_free_string(char* s);
/* The JSON text of one f64 (ryu shortest round-trip digits in serde_json's layout). */
int tga_format_f64(double x, char* buf, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* TGA_H */
