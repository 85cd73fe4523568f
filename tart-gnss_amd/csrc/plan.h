// Batched search plan (level 3) shared by plan.cpp and the JSON host layer.
#pragma once

#include <array>
#include <string>
#include <utility>
#include <vector>

#include "internal.h"
#include "search.h"

namespace tga {

constexpr double kPi = 3.14159265358979323846;

struct Group {
  size_t P = 0, N = 0, M = 0;  // code period, coherent length, transform length
  std::vector<double> fc;
  bool folded = false, fused = false, fp16 = false;
  int nc = 0;        // non-coherent form: one-period epochs summed (0 = reference form)
  bool big = false;  // P = big_r0 x 16368: sub-row split through the fused LDS kernel
  int big_r0 = 0;
  bool big16 = false;  // FP16 form of the split: f16 sub-rows, matrix-core rows kernel
  std::vector<std::vector<double>> codes;  // replicas, plan creation only
  std::vector<PeakSlot> slots;
  int n_prn = 0, n_bin = 0;
  size_t n_xrow = 0;
  long long n_rows = 0, chunk_rows = 0;
  float scale = 1.f;
  int cw = 1;
  DevBuf d_fc, d_slots, d_tau, d_rep, d_X, d_Xs, d_rows, d_Z, d_Zs, d_tw;
  // carrier (cos, sin) per (bin, sample), built at plan creation: fused groups [bin][sample],
  // sub-row split groups in the decimated order of launch_fwd_long
  DevBuf d_carrier;
  DevBuf d_rep16, d_X16;  // FP16 form: f16 (re, im) pairs scaled by kF16Scale
  DevBuf d_Y, d_rowkey, d_rownan, d_rowmap, d_twb, d_win, d_Zw, d_Zws;  // long-period split
  DevBuf d_twh;  // half sub-row twiddles (FP32 split variant 5)
  DevBuf d_rowtab;  // split groups: (X row, replica row, RowOut row) per enumerated row
  DevBuf d_tw4, d_c31;  // four-step split: W_P^{n2 k1} as [k1][n2], radix-31 cos / sin
  DevBuf d_twl;  // sub-row split: W_P^e two-level table of the long-period forward
  DevBuf d_m2;   // masked maximum (f32 bits) of each winning row (win_sub groups)
  DevBuf d_lev;  // sub-row split: decoded levels of the long-period forward
  float2* X = nullptr;  // where the forward spectra of the last search landed
  FusedRowsConfig rows_cfg;  // rows-kernel shape picked by the autotuner
  std::vector<std::pair<FusedRowsConfig, double>> tune;  // candidates timed, ms per launch
  double row_bytes_alg = 0, fwd_bytes_alg = 0;
  size_t dump_off = 0;  // surface form: first float of the group's rows in the staging buffer
  // FP32 split groups on the four-step form (variant 3) keep their spectra in natural order
  bool natural() const { return big && !big16 && rows_cfg.variant == 3; }
  // second peaks from a masked transform of each winning row (launch_big_second_sub): the
  // sub-row split groups and the fused rows of the reference form (one segment)
  bool win_sub() const { return (big && !natural()) || (fused && nc == 0); }
};

struct SysEntry {
  int system = 0;
  std::vector<int> prns;
  int group = 0;
  int first_prn = 0;
  size_t peak_offset = 0;
};

struct Timer {
  hipEvent_t start = nullptr, stop = nullptr;
  bool pending = false;
  double total_ms = 0.0;
  long long launches = 0;
  double bytes = 0.0;
};

struct Plan {
  double fs = 0, phase_const = 0;
  int n_ant = 0, fmt = 1, precision = TGA_PRECISION_FP32;
  long long L = 0;
  size_t row_bytes = 0, n_peaks = 0;
  std::vector<SysEntry> systems;
  std::vector<Group> groups;
  DevBuf d_means, d_partial, d_input, d_peaks, d_nan;
  hipStream_t stream = nullptr;
  bool timing = false;
  std::array<std::string, 3> kernel_names;  // kernels behind each timing slot
  std::array<Timer, 3> timers;
  // hipGraph replay: the first search of a plan runs eagerly, the second captures the whole
  // search (reading d_input) into `gexec`, later ones replay it.  With timing on, the timer
  // records are captured as event-record nodes (graph_slots: the timers the graph sets); a
  // change of the timing switch captures again; a runtime without event nodes in captures
  // runs timed searches eagerly.  Plans with a Bluestein transform (which allocates per
  // call) stay eager.
  bool graph_enabled = true, graphable = true, warmed = false;
  bool capturing = false, graph_timed = false;
  bool timed_graphs = true;  // cleared when the runtime refuses event nodes in a capture
  std::array<bool, 3> graph_slots{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long graph_launches = 0;

  ~Plan();
  void build(const tga_plan_desc& d);
  // dump (surface form): when not null, every row's output powers also go to
  // dump + g.dump_off + row * P (group layout, see surface())
  void search(const uint8_t* d_in, hipStream_t st, float* dump = nullptr);
  // Correlation surface of every row of a search of d_in, system by system in plan order:
  // [ant][prn][bin][h][P] floats on the scale of tga_peak::best_peak.
  size_t surface_floats(int sys_index = -1) const;
  void surface(const uint8_t* d_in, float* d_out, hipStream_t st);
  // Enqueue one search of d_in on st: eager, or a copy into d_input plus a graph replay.
  void run(const uint8_t* d_in, hipStream_t st);
  void drop_graph();
  void finish(tga_peak* out, hipStream_t st);
  std::string describe() const;
  // matrix-core FLOP of the rows stage of one search (f32 and f16 MFMAs)
  void rows_matrix_flops(double& f32, double& f16) const;
  void time_record(hipEvent_t ev, int slot, hipStream_t st, const char* what);
  void time_begin(int slot, hipStream_t st);
  void time_end(int slot, hipStream_t st, double bytes);
  void time_collect();
  void autotune(Group& g);
  void autotune_split(Group& g);
  void rep_to_pfa(Group& g, long long n);
  void alloc_winners(Group& g);
  // the rows of a sub-row split group (variants 0, 1, 2, 5), chunk by chunk, on st
  void split_rows(Group& g, hipStream_t st, float* dump);
};

// Error-to-status wrapper shared by the C ABI translation units.
int plan_guard_impl(void (*fn)(void*), void* ctx);
template <class F>
int plan_guard(F&& f) {
  return plan_guard_impl([](void* c) { (*static_cast<F*>(c))(); }, &f);
}

}  // namespace tga
