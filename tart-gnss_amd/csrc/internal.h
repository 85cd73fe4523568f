// Internal declarations shared by the host and device translation units of libtga.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "tga.h"

namespace tga {

// An error carrying a TGA_ERR_* status; the C ABI layer turns it into a return code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] void fail(int code, const std::string& msg);
void check_hip(hipError_t e, const char* what, const char* file, int line);
#define TGA_HIP(x) ::tga::check_hip((x), #x, __FILE__, __LINE__)

// ---- replica generation (codes.cpp) ----------------------------------------------------
struct SystemInfo {
  int id;
  const char* key;         // JSON key in CombinedOutput (lib.rs:31-45)
  const char* label_fmt;   // sv label, e.g. "GPS%02d"
  int prn_first, prn_last; // inclusive reference range
  double period_s;         // primary code period
  size_t chips;            // chips per period (BOC half-chips for B1C/L1C)
  int acr_table;
};
const SystemInfo& system_info(int system);
void check_prn(int system, int prn);
std::vector<int8_t> code_chips(int system, int prn);
// chips of a plan replica: code_chips for TGA_REPLICA_REFERENCE, the BOC(1,1) half-chips of
// the E1-C code for TGA_REPLICA_E1C_BOC11 (Galileo only)
std::vector<int8_t> replica_chips(int system, int prn, int replica);
// c[n] = chips[floor(n / (spc / L)) mod L], f64 index arithmetic
void resample(const std::vector<int8_t>& chips, double samples_per_code, size_t n, double* out);
size_t samples_per_period(int system, double fs);

// ---- ACR (acr.cpp) ---------------------------------------------------------------------
bool estimate_cn0(double r_a, int table, double* cn0);

// ---- device helpers ----------------------------------------------------------------------
int current_device_checked();

// A device buffer owned by RAII.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  void alloc(size_t b);
  void release();
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// ---- generic batched FFT (fft_generic.hip) ----------------------------------------------
// Forward unnormalised DFT of `batch` contiguous rows of length n (row stride n complex),
// Stockham passes through global memory.  Returns the buffer holding the result (`data` or
// `scratch`).
struct FftPlan;
const FftPlan* fft_plan(size_t n);  // cached, thread-safe
float2* fft_forward(const FftPlan* plan, float2* data, float2* scratch, size_t batch,
                    hipStream_t stream);
size_t fft_plan_size(const FftPlan* plan);
// Bluestein plans allocate work buffers and drain the stream per call: not capturable.
bool fft_plan_bluestein(const FftPlan* plan);
// Twiddle table W_n^e = tw1[e % split] * tw2[e / split] (device pointers) used by kernels.
struct TwiddleView {
  const float2* tw1;
  const float2* tw2;
  int split;
};
TwiddleView fft_twiddles(const FftPlan* plan);

// Heap-free form of fft_forward for lengths whose every factor has a register codelet (2, 3,
// 4, 5, 7, 8, 11, 13, 16, 31): fft_codelet_passes writes the pass order of fft_plan(n) into
// radix[0..cap) and returns the count (-1 for any other length); the twiddle table of
// fft_twiddle_count(n) float2 entries is computed on the GPU by launch_fft_twiddles into
// caller-provided device memory; fft_forward_tables runs the passes and returns the buffer
// holding the result.  Results equal fft_forward's for the same length.
int fft_codelet_passes(size_t n, int* radix, int cap);
int fft_twiddle_count(size_t n);
void launch_fft_twiddles(size_t n, float2* tw, hipStream_t st);
float2* fft_forward_tables(size_t n, const int* radix, int n_radix, const float2* tw,
                           float2* data, float2* scratch, size_t batch, hipStream_t st);

// Double precision variant for the RFI spectrum (rfi.rs:96-99 uses an f64 FFT).
struct FftPlanD;
const FftPlanD* fft_plan_d(size_t n);
double2* fft_forward_d(const FftPlanD* plan, double2* data, double2* scratch, size_t batch,
                       hipStream_t stream);

}  // namespace tga
