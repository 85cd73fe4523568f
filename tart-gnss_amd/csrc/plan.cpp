// Level-3 batched search plans: every antenna, PRN, Doppler bin and hypothesis of every
// system of one observation in a handful of launches (replaces the rayon loops of
// acquire_all_* -- acquisition.rs:261-396 and its five siblings).
//
// Systems that share (code period P, coherent length N, bin grid) share one group: one
// forward-spectrum buffer and one rows launch over all their PRNs (GPS, SBAS and QZSS on the
// reference's common grid form one group of 100 PRNs).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "internal.h"
#include "plan.h"
#include "search.h"

namespace tga {

namespace {

// Thrown out of a capture whose HIP runtime refuses external event records.
struct NoTimedCapture {};

template <class T>
void upload(DevBuf& b, const std::vector<T>& h, hipStream_t st) {
  b.alloc(h.size() * sizeof(T));
  TGA_HIP(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
  TGA_HIP(hipStreamSynchronize(st));  // the host vector may be a temporary
}

std::vector<float2> tau_table(size_t P) {
  std::vector<float2> t(P);
  for (size_t n = 0; n < P; ++n) {
    const double a = kPi * static_cast<double>(n) / static_cast<double>(P);
    t[n] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(-std::sin(a)));
  }
  return t;
}

// Kernels of one group behind a timing slot (0 rows, 1 forward spectra, 2 selection).
std::string group_kernels(const Group& g, int slot) {
  const std::string r0 = std::to_string(g.big_r0);
  switch (slot) {
    case 0:
      return g.nc > 0  ? "k_rows<NC>"
             : g.fp16  ? "k_rows16"
             : g.fused ? "k_rows"
             : g.big16 ? "k_big_pre<" + r0 + ",f16>+k_rows16<" + r0 + ">"
             : g.natural() ? "k_4s_pre<" + r0 + ">+k_4s_rows<" + r0 + ">"
             : g.big && g.rows_cfg.variant == 2 ? "k_big_pre<" + r0 + ">+k_subrows<" + r0 + ">"
             : g.big && g.rows_cfg.variant == 5
                 ? "k_half_pre<" + r0 + ">+k_halfrows<" + std::to_string(2 * g.big_r0) + ">"
             : g.big   ? "k_big_pre<" + r0 + ">+k_rows<" + std::to_string(g.rows_cfg.variant) +
                             "," + r0 + ">"
                       : "k_mul_rows+k_pass+k_row_reduce";
    case 1:
      return g.fused ? "k_forward"
             : g.big && !g.natural()
                 ? "k_levels_dec+k_fwd_long<" + r0 + ">+k_fwd_twist<" + r0 + ">"
             : g.folded ? "k_fold_obs+k_pass"
                        : "k_wipe_obs+k_pass";
    default:
      return g.win_sub() ? (g.fp16 ? std::string("k_select+k_win_pre16")
                                   : "k_select+k_win_pre<" + std::string(g.big ? r0 : "1") +
                                         ">") +
                               "+k_subrows<" + std::string(g.big ? r0 : "1") +
                               ",mask>+k_win_second"
             : g.big ? "k_select+k_mul_win+k_pass+k_second_win"
                     : "k_select";
  }
}

}  // namespace

Plan::~Plan() {
  for (auto& t : timers) {
    if (t.start) (void)hipEventDestroy(t.start);
    if (t.stop) (void)hipEventDestroy(t.stop);
  }
  drop_graph();
  if (stream) (void)hipStreamDestroy(stream);
}

void Plan::drop_graph() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  gexec = nullptr;
  graph = nullptr;
}

void Plan::run(const uint8_t* d_in, hipStream_t st) {
  if (st == nullptr) st = stream;
  if (!graph_enabled || !graphable || !warmed || (timing && !timed_graphs)) {
    search(d_in, st);
    warmed = true;  // lazily built transform plans now exist: nothing allocates any more
    return;
  }
  if (gexec && graph_timed != timing) drop_graph();
  if (!gexec) {
    TGA_HIP(hipStreamSynchronize(stream));
    graph_slots.fill(false);
    capturing = true;
    TGA_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    bool refused = false;
    try {
      search(d_input.as<uint8_t>(), stream);
    } catch (...) {
      capturing = false;
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      try {
        throw;
      } catch (const NoTimedCapture&) {
        refused = true;
      }
    }
    if (refused) {
      // this runtime records no event nodes in a capture: timed searches run eagerly
      timed_graphs = false;
      search(d_in, st);
      return;
    }
    capturing = false;
    TGA_HIP(hipStreamEndCapture(stream, &graph));
    TGA_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    graph_timed = timing;
  }
  if (d_in != d_input.as<uint8_t>())
    TGA_HIP(hipMemcpyAsync(d_input.p, d_in, static_cast<size_t>(n_ant) * row_bytes,
                           hipMemcpyDeviceToDevice, st));
  TGA_HIP(hipGraphLaunch(gexec, st));
  ++graph_launches;
  for (size_t k = 0; k < timers.size(); ++k)
    if (graph_timed && graph_slots[k]) timers[k].pending = true;
}

void Plan::build(const tga_plan_desc& d) {
  if (!(d.sampling_freq > 0.0) || !std::isfinite(d.sampling_freq))
    fail(TGA_ERR_INVALID_ARGUMENT, "sampling_freq must be positive");
  if (d.n_ant < 1) fail(TGA_ERR_INVALID_ARGUMENT, "n_ant must be >= 1");
  if (d.record_samples < 1) fail(TGA_ERR_INVALID_ARGUMENT, "record_samples must be >= 1");
  if (d.input_format != TGA_INPUT_1BIT && d.input_format != TGA_INPUT_2BIT)
    fail(TGA_ERR_INVALID_ARGUMENT, "input_format must be TGA_INPUT_1BIT or TGA_INPUT_2BIT");
  if (d.n_systems < 1 || d.systems == nullptr)
    fail(TGA_ERR_INVALID_ARGUMENT, "a plan needs at least one system");
  if (d.precision != TGA_PRECISION_FP32 && d.precision != TGA_PRECISION_FP16)
    fail(TGA_ERR_INVALID_ARGUMENT, "precision must be TGA_PRECISION_FP32 or TGA_PRECISION_FP16");
  precision = d.precision;
  current_device_checked();
  fs = d.sampling_freq;
  n_ant = d.n_ant;
  L = d.record_samples;
  fmt = d.input_format;
  const int bits = fmt == TGA_INPUT_1BIT ? 1 : 2;
  row_bytes = static_cast<size_t>((L * bits + 7) / 8);
  phase_const = (2.0 * kPi) / fs;  // TAU / fs (acquisition.rs:289)
  TGA_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));

  // ---- systems -> groups ----------------------------------------------------------------
  size_t peak_off = 0;
  for (int s = 0; s < d.n_systems; ++s) {
    const tga_system_desc& sd = d.systems[s];
    const SystemInfo& info = system_info(sd.system);
    SysEntry e;
    e.system = sd.system;
    if (sd.n_prn == 0 && sd.prns == nullptr) {
      for (int p = info.prn_first; p <= info.prn_last; ++p) e.prns.push_back(p);
    } else {
      if (sd.n_prn < 1 || sd.prns == nullptr)
        fail(TGA_ERR_INVALID_ARGUMENT, "n_prn / prns inconsistent");
      e.prns.assign(sd.prns, sd.prns + sd.n_prn);
    }
    for (int p : e.prns) check_prn(sd.system, p);
    if (sd.n_fc < 1 || sd.fc == nullptr)
      fail(TGA_ERR_EMPTY_BINS, "frequency bin list must be non-empty");
    std::vector<double> fc(sd.fc, sd.fc + sd.n_fc);
    for (double f : fc)
      if (!std::isfinite(f)) fail(TGA_ERR_INVALID_ARGUMENT, "bin frequencies must be finite");
    const size_t P = samples_per_period(sd.system, fs);
    if (P == 0) fail(TGA_ERR_INVALID_ARGUMENT, "sampling rate below one sample per period");
    const int nc = sd.nc_epochs;
    if (nc < 0) fail(TGA_ERR_INVALID_ARGUMENT, "nc_epochs must be >= 0");
    if (nc > 0) {
      if (static_cast<long long>(nc) * static_cast<long long>(P) > L)
        fail(TGA_ERR_INVALID_ARGUMENT, "nc_epochs periods exceed the record length");
      if (P != static_cast<size_t>(kFusedP) || std::ceil(fs / 1.023e6) > 16)
        fail(TGA_ERR_UNSUPPORTED, "the non-coherent form needs 16368-sample code periods");
    }
    size_t N = nc > 0 ? P
                      : sd.num_samples > 0 ? static_cast<size_t>(sd.num_samples)
                                           : std::min<size_t>(2 * P, static_cast<size_t>(L));
    if (N > static_cast<size_t>(L))
      fail(TGA_ERR_INVALID_ARGUMENT, "num_samples exceeds the record length");
    if (N > (1u << 26)) fail(TGA_ERR_UNSUPPORTED, "coherent length too large");
    e.peak_offset = peak_off;
    peak_off += static_cast<size_t>(n_ant) * e.prns.size();
    // replica of N samples (acquisition.rs:219); folding needs c[n+P] == c[n]
    std::vector<std::vector<double>> codes;
    bool fold = nc == 0 && N == 2 * P;
    if (sd.replica != TGA_REPLICA_REFERENCE && sd.replica != TGA_REPLICA_E1C_BOC11)
      fail(TGA_ERR_INVALID_ARGUMENT, "unknown replica " + std::to_string(sd.replica));
    if (sd.replica == TGA_REPLICA_E1C_BOC11 && sd.system != TGA_SYS_GALILEO)
      fail(TGA_ERR_UNSUPPORTED, "the BOC(1,1) replica is defined for Galileo E1-C only");
    for (int p : e.prns) {
      std::vector<double> c(N);
      resample(replica_chips(sd.system, p, sd.replica), static_cast<double>(P), N, c.data());
      if (fold)
        for (size_t n = 0; n < P; ++n)
          if (c[n + P] != c[n]) { fold = false; break; }
      codes.push_back(std::move(c));
    }
    // find or create the group
    Group* g = nullptr;
    for (auto& gg : groups)
      if (gg.P == P && gg.N == N && gg.fc == fc && gg.folded == fold && gg.nc == nc) g = &gg;
    if (!g) {
      groups.emplace_back();
      g = &groups.back();
      g->P = P;
      g->N = N;
      g->fc = fc;
      g->folded = fold;
      g->nc = nc;
      g->fused = nc > 0 ||
                 (fold && fused_supported(P, N, static_cast<int>(std::ceil(fs / 1.023e6))));
      g->big = nc == 0 && !g->fused && big_supported(P, N, fold);
      g->big_r0 = g->big ? static_cast<int>(P / kFusedP) : 0;
      g->M = fold ? P : N;
    }
    e.group = static_cast<int>(g - groups.data());
    e.first_prn = static_cast<int>(g->codes.size());
    for (auto& c : codes) g->codes.push_back(std::move(c));
    for (size_t i = 0; i < e.prns.size(); ++i) {
      PeakSlot sl;
      sl.base = static_cast<int32_t>(e.peak_offset);
      sl.stride = static_cast<int32_t>(e.prns.size());
      sl.local = static_cast<int32_t>(i);
      sl.pad = 0;
      g->slots.push_back(sl);
    }
    systems.push_back(std::move(e));
  }
  n_peaks = peak_off;

  // ---- device state per group -------------------------------------------------------------
  d_means.alloc(static_cast<size_t>(n_ant) * sizeof(double));
  d_partial.alloc(static_cast<size_t>(n_ant) * kAntSplit * sizeof(long long));
  d_input.alloc(static_cast<size_t>(n_ant) * row_bytes);
  d_peaks.alloc(n_peaks * sizeof(tga_peak));
  d_nan.alloc(sizeof(int));
  for (auto& g : groups) {
    const int n_prn = static_cast<int>(g.codes.size());
    const int n_bin = static_cast<int>(g.fc.size());
    const size_t M = g.M;
    g.scale = 1.0f / std::sqrt(static_cast<float>(g.N));  // correlate.rs:123
    g.cw = static_cast<int>(std::ceil(fs / 1.023e6));     // correlate.rs:157
    upload(g.d_fc, g.fc, stream);
    upload(g.d_slots, g.slots, stream);
    // replica time series -> spectra (plan creation only)
    std::vector<float2> h(static_cast<size_t>(n_prn) * 2 * M);
    std::vector<float2> tau;
    if (g.folded) {
      tau = tau_table(g.P);
      upload(g.d_tau, tau, stream);
    }
    for (int p = 0; p < n_prn; ++p) {
      const std::vector<double>& c = g.codes[p];
      float2* h0 = h.data() + (static_cast<size_t>(p) * 2) * M;
      float2* h1 = h0 + M;
      for (size_t n = 0; n < M; ++n) {
        const float cv = static_cast<float>(c[n]);
        h0[n] = make_float2(cv, 0.f);
        if (g.folded) {
          h1[n] = make_float2(cv * tau[n].x, cv * tau[n].y);
        } else {
          const float a = n < std::min(g.P, g.N) ? cv : -cv;  // correlate.rs:53-56
          h1[n] = make_float2(a, 0.f);
        }
      }
    }
    DevBuf a;
    upload(a, h, stream);
    DevBuf sc(h.size() * sizeof(float2));
    float2* r = fft_forward(fft_plan(M), a.as<float2>(), sc.as<float2>(),
                            static_cast<size_t>(n_prn) * 2, stream);
    TGA_HIP(hipStreamSynchronize(stream));
    g.d_rep = std::move(r == a.as<float2>() ? a : sc);
    g.codes.clear();
    g.codes.shrink_to_fit();
    g.fp16 = g.fused && g.nc == 0 && precision == TGA_PRECISION_FP16;
    g.big16 = g.big && precision == TGA_PRECISION_FP16;
    // forward spectra and per-row results
    g.n_xrow = static_cast<size_t>(n_ant) * n_bin * (g.nc > 0 ? g.nc : g.folded ? 2 : 1);
    g.n_rows = static_cast<long long>(n_ant) * n_prn * n_bin * 2;
    g.d_rows.alloc(static_cast<size_t>(g.n_rows) * sizeof(RowOut));
    g.n_prn = n_prn;
    g.n_bin = n_bin;
    if (g.fused) {
      // the forward kernels' carrier table: one (cos, sin) per (bin, sample) for all antennas
      const long long clen = static_cast<long long>(g.nc > 0 ? g.nc : 2) * kFusedP;
      g.d_carrier.alloc(static_cast<size_t>(n_bin) * clen * sizeof(float2));
      launch_carrier(g.d_fc.as<double>(), n_bin, clen, phase_const, g.fp16,
                     g.d_carrier.as<float2>(), stream);
      TGA_HIP(hipStreamSynchronize(stream));
    }
    if (g.fp16) {
      const long long nr = static_cast<long long>(n_prn) * 2 * static_cast<long long>(M);
      g.d_rep16.alloc(static_cast<size_t>(nr) * sizeof(uint32_t));
      launch_to_f16(g.d_rep.as<float2>(), g.d_rep16.as<uint32_t>(), nr, kF16Scale, stream);
      TGA_HIP(hipStreamSynchronize(stream));
      g.d_rep.release();
      g.d_X16.alloc(g.n_xrow * M * sizeof(uint32_t));
      upload(g.d_tw, fused_twiddles(), stream);
      alloc_winners(g);
      autotune(g);
    } else if (g.fused) {  // FP32 fused rows, reference or non-coherent form
      rep_to_pfa(g, static_cast<long long>(n_prn) * 2 * static_cast<long long>(M));
      g.d_X.alloc(g.n_xrow * M * sizeof(float2));
      upload(g.d_tw, fused_twiddles(), stream);
      if (g.win_sub()) alloc_winners(g);
      autotune(g);
    } else if (g.big) {
      g.d_X.alloc(g.n_xrow * M * sizeof(float2));
      g.d_Xs.alloc(g.n_xrow * M * sizeof(float2));
      upload(g.d_tw, fused_twiddles(), stream);
      upload(g.d_twb, big_twiddles(g.P), stream);
      if (!g.big16) {
        upload(g.d_tw4, four_step_twiddles(g.P), stream);
        upload(g.d_c31, radix31_table(), stream);
        upload(g.d_twh, half_twiddles(g.P), stream);
      }
      // sub-rows of a chunk of rows per pre-pass / transform pair, on the plan's stream (a
      // second stream for the pre-pass of the next chunk measured no faster: the two kernels
      // then slow each other down).  Larger chunks amortise each launch pair's ramp and tail
      // (per all24 search: 1000 ms at 2 GiB, 894 ms at 16 GiB, 885 ms at 32 GiB of sub-rows
      // at the time; later 826.5 / 817.9 / 816.9 ms at 32 / 48 / 64 GiB), so a chunk holds up
      // to 48 GiB, at most a sixth of the free device memory; TGA_SPLIT_CHUNK_MB sets it
      // explicitly.
      const size_t y_elem = g.big16 ? sizeof(uint32_t) : sizeof(float2);
      const char* env_mb = std::getenv("TGA_SPLIT_CHUNK_MB");
      size_t free_b = 0, total_b = 0;
      TGA_HIP(hipMemGetInfo(&free_b, &total_b));
      const long long chunk_bytes =
          env_mb ? std::max(1, std::atoi(env_mb)) * (1LL << 20)
                 : std::max<long long>(256LL << 20, std::min<long long>(48LL << 30, free_b / 6));
      const long long cap = std::max<long long>(
          1, chunk_bytes / (static_cast<long long>(M) * static_cast<long long>(y_elem)));
      g.chunk_rows = std::min<long long>(cap, g.n_rows);
      g.d_Y.alloc(static_cast<size_t>(g.chunk_rows) * M * y_elem);
      g.d_rowkey.alloc(static_cast<size_t>(g.n_rows) * sizeof(unsigned long long));
      g.d_rownan.alloc(static_cast<size_t>(g.n_rows) * sizeof(int));
      g.d_rowmap.alloc(static_cast<size_t>(g.chunk_rows) * sizeof(int));
      g.d_rowtab.alloc(static_cast<size_t>(g.n_rows / (g.folded ? 1 : 2)) * sizeof(int4));
      if (g.n_rows > (1LL << 31) - 1)
        fail(TGA_ERR_UNSUPPORTED, "a long-period group holds more than 2^31 - 1 rows");
      alloc_winners(g);
      autotune_split(g);
      if (g.natural()) g.d_Zws.alloc(g.d_Zw.bytes);  // the generic transforms' ping-pong
      // the sub-row forms read their spectra segment by segment in the prime-factor order of
      // 16368; the four-step form reads them in natural order
      if (!g.natural()) {
        rep_to_pfa(g, static_cast<long long>(n_prn) * 2 * static_cast<long long>(M));
        // their forward spectra come from the two-step LDS forward (launch_fwd_long): the
        // carrier in its decimated order (f64 phase in the FP16 form), tau as [r][m]
        g.d_carrier.alloc(static_cast<size_t>(n_bin) * (g.folded ? 2 : 1) * g.P * sizeof(float2));
        launch_carrier_dec(g.d_fc.as<double>(), n_bin, g.big_r0, g.folded, phase_const, g.big16,
                           g.d_carrier.as<float2>(), stream);
        if (g.folded) {
          std::vector<float2> td(g.P);
          for (int r = 0; r < g.big_r0; ++r)
            for (size_t m = 0; m < kFusedP; ++m) td[r * kFusedP + m] = tau[g.big_r0 * m + r];
          upload(g.d_tau, td, stream);
        }
        upload(g.d_twl, long_twiddles(g.P), stream);
        g.d_lev.alloc(static_cast<size_t>(n_ant) * (g.folded ? 2 : 1) * g.P * sizeof(float));
        TGA_HIP(hipStreamSynchronize(stream));
      }
    } else {
      g.d_X.alloc(g.n_xrow * M * sizeof(float2));
      g.d_Xs.alloc(g.n_xrow * M * sizeof(float2));
      // rows are processed in chunks of at most ~32 Mi complex values (256 MiB) per buffer
      const long long cap = std::max<long long>(1, (32LL << 20) / static_cast<long long>(M));
      g.chunk_rows = std::min<long long>(cap, g.n_rows);
      g.d_Z.alloc(static_cast<size_t>(g.chunk_rows) * M * sizeof(float2));
      g.d_Zs.alloc(static_cast<size_t>(g.chunk_rows) * M * sizeof(float2));
    }
    // algorithmic bytes of the rows stage: each row reads its forward spectrum and replica
    // spectrum once (M complex f32 each)
    const double elem = g.fp16 ? 4.0 : 8.0;  // bytes per complex value
    // non-coherent rows: one (X, R) pair per epoch, half as many rows (one hypothesis)
    // unfolded split rows: only h = 0 is transformed (the hypotheses coincide)
    const double per_row = g.nc > 0 ? 0.5 * g.nc : (g.big && !g.folded) ? 0.5 : 1.0;
    g.row_bytes_alg =
        static_cast<double>(g.n_rows) * per_row * 2.0 * static_cast<double>(M) * elem;
    // forward: each (antenna, bin, fold) row writes its spectrum; the fused kernels also read
    // the carrier table of the row's samples (2P values, one period per non-coherent epoch)
    const double car_row = g.fused ? (g.nc > 0 ? 1.0 : 2.0) * static_cast<double>(g.P) * 8.0 : 0.0;
    g.fwd_bytes_alg = static_cast<double>(g.n_xrow) * (static_cast<double>(M) * elem + car_row);
  }
  for (auto& g : groups)
    if (fft_plan_bluestein(fft_plan(g.M))) graphable = false;
  // the kernels behind each timing slot, as rocprof lists them, joined over the groups
  for (int slot = 0; slot < 3; ++slot) {
    std::string names;
    for (const auto& g : groups) {
      const std::string n = group_kernels(g, slot);
      if (names.find(n) == std::string::npos) names += (names.empty() ? "" : "+") + n;
    }
    kernel_names[slot] = names;
  }
  TGA_HIP(hipStreamSynchronize(stream));
  for (auto& t : timers) {
    TGA_HIP(hipEventCreate(&t.start));
    TGA_HIP(hipEventCreate(&t.stop));
    // a timer event captured as an event-record node must have been recorded once before
    TGA_HIP(hipEventRecord(t.start, stream));
    TGA_HIP(hipEventRecord(t.stop, stream));
  }
  TGA_HIP(hipStreamSynchronize(stream));
}

// Times every candidate once after a warm-up launch, then the candidates within 3 % of the
// best twice more (one timing varies by about 1 %), and keeps the fastest: g.tune holds each
// candidate's best time, g.rows_cfg the winner.
template <typename Launch>
void tune_pick(Group& g, const std::vector<FusedRowsConfig>& cands, hipStream_t stream,
               Launch&& launch) {
  hipEvent_t e0, e1;
  TGA_HIP(hipEventCreate(&e0));
  TGA_HIP(hipEventCreate(&e1));
  auto time_one = [&](const FusedRowsConfig& c) {
    TGA_HIP(hipEventRecord(e0, stream));
    launch(c);
    TGA_HIP(hipEventRecord(e1, stream));
    TGA_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    TGA_HIP(hipEventElapsedTime(&ms, e0, e1));
    return static_cast<double>(ms);
  };
  double best = 1e30;
  for (const FusedRowsConfig& c : cands) {
    launch(c);
    const double ms = time_one(c);
    g.tune.push_back({c, ms});
    best = std::min(best, ms);
  }
  for (auto& t : g.tune)
    if (t.second <= 1.03 * best)
      for (int k = 0; k < 2; ++k) t.second = std::min(t.second, time_one(t.first));
  best = 1e30;
  for (const auto& t : g.tune)
    if (t.second < best) {
      best = t.second;
      g.rows_cfg = t.first;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

// Buffers of the winners' masked second peaks: the winning row of each (antenna, PRN) pair,
// its masked maximum and its spectrum products (one row of M complex values per pair).
void Plan::alloc_winners(Group& g) {
  const size_t nw = static_cast<size_t>(n_ant) * g.n_prn;
  g.d_win.alloc(nw * sizeof(int));
  g.d_m2.alloc(nw * sizeof(unsigned int));
  g.d_Zw.alloc(nw * g.M * sizeof(float2));
}

// Rows-kernel autotuner: times every candidate shape on the group's real sizes, with
// pseudo-random spectra, and keeps the fastest (tune_pick).  Candidates: the radix-31
// implementation of the f32 kernel (VALU codelet or matrix cores) and the number of XCD-local
// PRN groups (8, 16 or 32: more groups give more, shorter workgroups, each over fewer PRNs).
// TGA_FUSED_VARIANT=0|1 pins the radix-31 implementation (tests use it to cover both) and
// TGA_ROWS_GROUPS=n the grouping.
// The fused FP32 rows kernels read the replica spectra in the prime-factor order of 16368.
void Plan::rep_to_pfa(Group& g, long long n) {
  DevBuf t(static_cast<size_t>(n) * sizeof(float2));
  launch_to_pfa(g.d_rep.as<float2>(), t.as<float2>(), n, stream);
  TGA_HIP(hipStreamSynchronize(stream));
  g.d_rep = std::move(t);
}

void Plan::autotune(Group& g) {
  const char* pin_v = std::getenv("TGA_FUSED_VARIANT");
  const char* pin_g = std::getenv("TGA_ROWS_GROUPS");
  std::vector<int> grs;
  if (pin_g) {
    grs.push_back(std::max(1, std::atoi(pin_g)));
  } else {
    for (int gr : {1, 2, 4, 8, 16, 32})  // groupings beyond the PRN count only add empty groups
      if (gr <= 8 || gr / 2 < g.n_prn) grs.push_back(gr);
  }
  std::vector<FusedRowsConfig> cands;
  // the FP16 and non-coherent kernels have one radix-31 implementation
  const bool one_variant = g.fp16 || g.nc > 0;
  for (int v = 0; v < (one_variant ? 1 : 2); ++v) {
    if (pin_v && !one_variant && std::atoi(pin_v) != v) continue;
    for (int gr : grs) {
      FusedRowsConfig c;
      c.variant = v;
      c.groups = gr;
      cands.push_back(c);
    }
  }
  g.rows_cfg = cands[0];
  g.tune.clear();
  if (cands.size() == 1) return;
  DevBuf& xs = g.fp16 ? g.d_X16 : g.d_X;
  launch_fill_tune(xs.p, xs.bytes, g.fp16, stream);
  auto launch = [&](const FusedRowsConfig& c) {
    if (g.nc > 0)
      launch_fused_rows_nc(g.d_X.as<float2>(), g.d_rep.as<float2>(), g.d_tw.as<float2>(), n_ant,
                           g.n_bin, g.n_prn, g.nc, g.cw, g.d_rows.as<RowOut>(), c.groups, stream);
    else if (g.fp16)
      launch_fused_rows16(g.d_X16.as<uint32_t>(), g.d_rep16.as<uint32_t>(), g.d_tw.as<float2>(),
                          n_ant, g.n_bin, g.n_prn, g.cw, g.d_rows.as<RowOut>(), c.groups, stream);
    else
      launch_fused_rows(g.d_X.as<float2>(), g.d_rep.as<float2>(), g.d_tw.as<float2>(), n_ant,
                        g.n_bin, g.n_prn, g.cw, g.d_rows.as<RowOut>(), stream, c);
  };
  tune_pick(g, cands, stream, launch);
}

// Split-path autotuner: times the first chunk of rows (both kernels of the form, after one
// warm-up) for every candidate on the group's real sizes, with pseudo-random spectra, and keeps
// the fastest.  Candidates: the FP32 form (variants 0, 1 and 2: R0 sub-rows of 16368 through the
// LDS transform with the VALU or the matrix-core radix 31 (k_rows: 0, 1), or with the
// matrix-core radix 31, pass B in place and only each lane's best output pair kept (k_subrows:
// 2; 2 R0 half sub-rows of 8184 points with two transforming workgroups per CU: 5); variant 3:
// the four-step N1 x 496 form; the FP16 form has one) and the
// tile shape of the row order (TX forward spectra x TP replica spectra whose rows run together,
// so their spectra are fetched once into L2).  TGA_SPLIT_VARIANT=0..5 and TGA_SPLIT_TILE=TXxTP
// pin them.  The
// choice fixes the spectra's layout: natural order for variant 3, else the segment-wise
// prime-factor order (the replicas are reordered after tuning).
void Plan::autotune_split(Group& g) {
  const char* pin_v = std::getenv("TGA_SPLIT_VARIANT");
  const char* pin_t = std::getenv("TGA_SPLIT_TILE");
  std::vector<int> variants;
  if (g.big16)
    variants = {0};
  else if (pin_v) {
    const int v = std::atoi(pin_v);
    if (v < 0 || v > 5 || v == 4)
      fail(TGA_ERR_INVALID_ARGUMENT, "TGA_SPLIT_VARIANT must be 0, 1, 2, 3 or 5");
    variants = {v};
  } else {
    variants = {5, 2, 1, 0, 3};
  }
  std::vector<int> tiles;
  if (pin_t) {
    int tx = 0, tp = 0;
    if (std::sscanf(pin_t, "%dx%d", &tx, &tp) != 2 || tx < 1 || tp < 1 || tx > 4096 ||
        tp > 4096)
      fail(TGA_ERR_INVALID_ARGUMENT, "TGA_SPLIT_TILE must read TXxTP (e.g. 8x8)");
    tiles = {tx | (tp << 16)};
  } else {
    for (auto [tx, tp] : {std::pair{8, 8}, {4, 16}, {16, 4}, {2, 32}, {32, 2}})
      tiles.push_back(tx | (tp << 16));
  }
  std::vector<FusedRowsConfig> cands;
  for (int v : variants)
    for (int t : tiles) {
      // the four-step form is timed on two shapes unless pinned (its step 1 reads whole tiles of
      // columns, its step 2 none: the shape matters less)
      if (v == 3 && !pin_t && t != (8 | (8 << 16)) && t != (4 | (16 << 16))) continue;
      FusedRowsConfig c;
      c.variant = v;
      c.groups = 1;
      c.tile = t;
      cands.push_back(c);
    }
  g.rows_cfg = cands[0];
  g.tune.clear();
  const int hs = g.folded ? 1 : 2;
  // the row table of the chosen order, every enumerated row (the pre-passes read it)
  auto build_rowtab = [&] {
    if (!g.natural())
      launch_big_rowtab(g.d_rowtab.as<int4>(), g.n_rows / hs, g.n_prn, g.n_bin, n_ant, hs,
                        g.rows_cfg.tile, stream);
  };
  if (cands.size() == 1) {
    build_rowtab();
    return;
  }
  launch_fill_tune(g.d_X.p, g.d_X.bytes, false, stream);
  const long long r1 = std::min(g.n_rows / hs, g.chunk_rows);
  auto launch = [&](const FusedRowsConfig& c) {
    if (c.variant == 3) {
      launch_big_4s(g.d_X.as<float2>(), g.d_rep.as<float2>(), g.d_tw4.as<float2>(),
                    g.d_c31.as<float2>(), g.big_r0, g.n_prn, g.n_bin, n_ant, hs, c.tile, 0, r1,
                    g.d_Y.as<float2>(), g.d_rowkey.as<unsigned long long>(),
                    g.d_rownan.as<int>(), g.d_rows.as<RowOut>(), stream);
      return;
    }
    // the candidate's row order for the timed chunk (a short kernel beside the chunk's pair)
    launch_big_rowtab(g.d_rowtab.as<int4>(), r1, g.n_prn, g.n_bin, n_ant, hs, c.tile, stream);
    if (c.variant == 5) {
      launch_half_pre(g.d_X.as<float2>(), g.d_rep.as<float2>(), g.d_twh.as<float2>(),
                      g.d_rowtab.as<int4>(), g.big_r0, 0, r1, g.d_Y.as<float2>(),
                      g.d_rowkey.as<unsigned long long>(), g.d_rownan.as<int>(),
                      g.d_rowmap.as<int>(), stream);
      launch_half_sub(g.d_tw.as<float2>(), g.big_r0, g.n_prn, g.n_bin, n_ant, hs, c.tile, 0, r1,
                      g.d_Y.as<float2>(), g.d_rowkey.as<unsigned long long>(),
                      g.d_rownan.as<int>(), g.d_rowmap.as<int>(), g.d_rows.as<RowOut>(),
                      stream);
      return;
    }
    launch_big_pre(g.d_X.as<float2>(), g.d_rep.as<float2>(), g.d_twb.as<float2>(),
                   g.d_rowtab.as<int4>(), g.big_r0, 0, r1, g.d_Y.p, g.big16,
                   g.d_rowkey.as<unsigned long long>(), g.d_rownan.as<int>(),
                   g.d_rowmap.as<int>(), stream);
    launch_big_sub(g.d_tw.as<float2>(), g.big_r0, g.n_prn, g.n_bin, n_ant, hs, c.tile, 0, r1,
                   g.d_Y.p, g.big16, g.d_rowkey.as<unsigned long long>(), g.d_rownan.as<int>(),
                   g.d_rowmap.as<int>(), g.d_rows.as<RowOut>(), stream, c.variant);
  };
  tune_pick(g, cands, stream, launch);
  build_rowtab();
}

// While a graph is being captured the records become event-record nodes of the graph, and
// every replay of it sets the timers again.
void Plan::time_record(hipEvent_t ev, int slot, hipStream_t st, const char* what) {
  const hipError_t e = hipEventRecordWithFlags(
      ev, st, capturing ? hipEventRecordExternal : hipEventRecordDefault);
  if (e == hipSuccess) return;
  (void)hipGetLastError();
  if (capturing && e == hipErrorInvalidValue) throw NoTimedCapture{};
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(st, &cs);
  fail(TGA_ERR_DEVICE, std::string("hipEventRecordWithFlags(") + what + " of timer " +
                           std::to_string(slot) + (capturing ? ", external" : ", default") +
                           ", stream capture status " + std::to_string(static_cast<int>(cs)) +
                           "): " + hipGetErrorString(e));
}

void Plan::time_begin(int slot, hipStream_t st) {
  if (!timing) return;
  time_record(timers[slot].start, slot, st, "start");
}

void Plan::time_end(int slot, hipStream_t st, double bytes) {
  if (!timing) return;
  time_record(timers[slot].stop, slot, st, "stop");
  timers[slot].bytes = bytes;
  if (capturing)
    graph_slots[slot] = true;
  else
    timers[slot].pending = true;
}

void Plan::time_collect() {
  if (!timing) return;
  for (auto& t : timers) {
    if (!t.pending) continue;
    float ms = 0.f;
    TGA_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
    t.total_ms += ms;
    t.launches += 1;
    t.pending = false;
  }
}

// Row-value gain of k_select: folded rows carry half the 2P-point gain; the FP16 forms' values
// carry kF16Scale^2 (fused) or kSplitF16Scale (split).
static float select_gain(const Group& g) {
  const float fold = g.folded ? 2.0f : 1.0f;
  return fold / (g.fp16 ? kF16Scale * kF16Scale : g.big16 ? kSplitF16Scale : 1.0f);
}

// Rows of a sub-row split group, chunk by chunk on st: the pre-pass writes the chunk's sub-rows,
// the transform reduces them to the rows' keys and RowOut.  (Running the next chunk's pre-pass
// on a second stream into a second buffer beside this chunk's transform was measured slower in
// every form: the two kernels share the CUs and slow each other down.)
void Plan::split_rows(Group& g, hipStream_t st, float* dump) {
  const int hs = g.folded ? 1 : 2;
  const long long n_enum = g.n_rows / hs;
  const bool half = g.rows_cfg.variant == 5 && !g.big16;
  float* dmp = dump ? dump + g.dump_off : nullptr;
  auto pre = [&](long long r0, long long r1, hipStream_t s) {
    void* Y = g.d_Y.p;
    int* rowmap = g.d_rowmap.as<int>();
    if (half)
      launch_half_pre(g.X, g.d_rep.as<float2>(), g.d_twh.as<float2>(), g.d_rowtab.as<int4>(),
                      g.big_r0, r0, r1, static_cast<float2*>(Y),
                      g.d_rowkey.as<unsigned long long>(), g.d_rownan.as<int>(), rowmap, s);
    else
      launch_big_pre(g.X, g.d_rep.as<float2>(), g.d_twb.as<float2>(), g.d_rowtab.as<int4>(),
                     g.big_r0, r0, r1, Y, g.big16, g.d_rowkey.as<unsigned long long>(),
                     g.d_rownan.as<int>(), rowmap, s);
  };
  auto sub = [&](long long r0, long long r1, hipStream_t s) {
    void* Y = g.d_Y.p;
    const int* rowmap = g.d_rowmap.as<int>();
    if (half)
      launch_half_sub(g.d_tw.as<float2>(), g.big_r0, g.n_prn, g.n_bin, n_ant, hs,
                      g.rows_cfg.tile, r0, r1, static_cast<const float2*>(Y),
                      g.d_rowkey.as<unsigned long long>(), g.d_rownan.as<int>(), rowmap,
                      g.d_rows.as<RowOut>(), s, dmp);
    else
      launch_big_sub(g.d_tw.as<float2>(), g.big_r0, g.n_prn, g.n_bin, n_ant, hs,
                     g.rows_cfg.tile, r0, r1, Y, g.big16, g.d_rowkey.as<unsigned long long>(),
                     g.d_rownan.as<int>(), rowmap, g.d_rows.as<RowOut>(), s,
                     g.rows_cfg.variant, dmp);
  };
  const long long n_chunks = (n_enum + g.chunk_rows - 1) / g.chunk_rows;
  auto bounds = [&](long long c) {
    return std::pair<long long, long long>{c * g.chunk_rows,
                                           std::min(n_enum, (c + 1) * g.chunk_rows)};
  };
  for (long long c = 0; c < n_chunks; ++c) {
    const auto [r0, r1] = bounds(c);
    pre(r0, r1, st);
    sub(r0, r1, st);
  }
}

void Plan::search(const uint8_t* d_in, hipStream_t st, float* dump) {
  if (st == nullptr) st = stream;
  launch_ant_sums(d_in, row_bytes, n_ant, L, fmt, d_partial.as<long long>(),
                  d_means.as<double>(), d_nan.as<int>(), st);
  double rows_bytes = 0.0, fwd_bytes = 0.0;
  for (auto& g : groups) {
    rows_bytes += g.row_bytes_alg;
    fwd_bytes += g.fwd_bytes_alg;
  }
  // forward spectra of every group
  time_begin(1, st);
  for (auto& g : groups) {
    const int M = static_cast<int>(g.M);
    if (g.nc > 0) {
      launch_fused_forward_nc(d_in, row_bytes, fmt, d_means.as<double>(), g.d_fc.as<double>(),
                              n_ant, g.n_bin, g.nc, phase_const, g.d_tw.as<float2>(),
                              g.d_carrier.as<float2>(), g.d_X.as<float2>(), st);
      g.X = g.d_X.as<float2>();
    } else if (g.fp16) {
      launch_fused_forward16(d_in, row_bytes, fmt, d_means.as<double>(), g.d_fc.as<double>(),
                             n_ant, g.n_bin, phase_const, g.d_tau.as<float2>(),
                             g.d_tw.as<float2>(), g.d_carrier.as<float2>(),
                             g.d_X16.as<uint32_t>(), st);
    } else if (g.fused) {
      launch_fused_forward(d_in, row_bytes, fmt, d_means.as<double>(), g.d_fc.as<double>(),
                           n_ant, g.n_bin, phase_const, g.d_tau.as<float2>(),
                           g.d_tw.as<float2>(), g.d_carrier.as<float2>(), g.d_X.as<float2>(), st);
      g.X = g.d_X.as<float2>();
    } else if (g.big && !g.natural()) {
      launch_fwd_long(d_in, row_bytes, fmt, d_means.as<double>(), g.d_carrier.as<float2>(),
                      g.d_tau.as<float2>(), g.d_tw.as<float2>(), g.d_twl.as<float2>(), g.big_r0,
                      g.folded, n_ant, g.n_bin, g.d_lev.as<float>(), g.d_Xs.as<float2>(),
                      g.d_X.as<float2>(), st);
      g.X = g.d_X.as<float2>();
    } else {
      if (g.folded)
        launch_fold_obs(d_in, row_bytes, fmt, d_means.as<double>(), g.d_fc.as<double>(), n_ant,
                        g.n_bin, M, phase_const, g.d_tau.as<float2>(), g.d_X.as<float2>(),
                        g.big16, st);
      else
        launch_wipe_obs(d_in, row_bytes, fmt, d_means.as<double>(), g.d_fc.as<double>(), n_ant,
                        g.n_bin, M, phase_const, g.d_X.as<float2>(), g.big16, st);
      g.X = fft_forward(fft_plan(g.M), g.d_X.as<float2>(), g.d_Xs.as<float2>(), g.n_xrow, st);
      if (g.big && !g.natural()) {  // segment-wise prime-factor order of the long rows
        float2* dst = g.X == g.d_X.as<float2>() ? g.d_Xs.as<float2>() : g.d_X.as<float2>();
        launch_to_pfa(g.X, dst, static_cast<long long>(g.n_xrow) * M, st);
        g.X = dst;
      }
    }
  }
  time_end(1, st, fwd_bytes);
  // correlation rows of every group
  time_begin(0, st);
  for (auto& g : groups) {
    const int M = static_cast<int>(g.M);
    if (g.nc > 0) {
      launch_fused_rows_nc(g.X, g.d_rep.as<float2>(), g.d_tw.as<float2>(), n_ant, g.n_bin,
                           g.n_prn, g.nc, g.cw, g.d_rows.as<RowOut>(), g.rows_cfg.groups, st);
    } else if (g.fp16) {
      launch_fused_rows16(g.d_X16.as<uint32_t>(), g.d_rep16.as<uint32_t>(), g.d_tw.as<float2>(),
                          n_ant, g.n_bin, g.n_prn, g.cw, g.d_rows.as<RowOut>(),
                          g.rows_cfg.groups, st, dump ? dump + g.dump_off : nullptr);
    } else if (g.fused) {
      launch_fused_rows(g.X, g.d_rep.as<float2>(), g.d_tw.as<float2>(), n_ant, g.n_bin, g.n_prn,
                        g.cw, g.d_rows.as<RowOut>(), st, g.rows_cfg,
                        dump ? dump + g.dump_off : nullptr);
    } else if (g.big) {
      const int hs = g.folded ? 1 : 2;
      const long long n_enum = g.n_rows / hs;
      if (g.natural()) {
        for (long long r0 = 0; r0 < n_enum; r0 += g.chunk_rows)
          launch_big_4s(g.X, g.d_rep.as<float2>(), g.d_tw4.as<float2>(), g.d_c31.as<float2>(),
                        g.big_r0, g.n_prn, g.n_bin, n_ant, hs, g.rows_cfg.tile, r0,
                        std::min(n_enum, r0 + g.chunk_rows), g.d_Y.as<float2>(),
                        g.d_rowkey.as<unsigned long long>(), g.d_rownan.as<int>(),
                        g.d_rows.as<RowOut>(), st, dump ? dump + g.dump_off : nullptr);
        continue;
      }
      split_rows(g, st, dump);
    } else {
      const FftPlan* fp = fft_plan(g.M);
      for (long long r0 = 0; r0 < g.n_rows; r0 += g.chunk_rows) {
        const long long r1 = std::min(g.n_rows, r0 + g.chunk_rows);
        launch_mul_rows(g.X, g.d_rep.as<float2>(), M, g.n_prn, g.n_bin, g.folded, r0, r1,
                        g.d_Z.as<float2>(), st);
        float2* z = fft_forward(fp, g.d_Z.as<float2>(), g.d_Zs.as<float2>(),
                                static_cast<size_t>(r1 - r0), st);
        launch_row_reduce(z, M, static_cast<int>(g.P), g.cw, r1 - r0,
                          g.d_rows.as<RowOut>() + r0, st);
      }
    }
  }
  time_end(0, st, rows_bytes);
  time_begin(2, st);
  for (auto& g : groups) {
    // the winners' second peaks of split groups and fused FP32 rows are recomputed in f32
    const float fold = g.folded ? 2.0f : 1.0f;
    const float gain = select_gain(g);
    launch_select(g.d_rows.as<RowOut>(), g.d_slots.as<PeakSlot>(), n_ant, g.n_prn, g.n_bin,
                  gain, g.scale, d_peaks.as<tga_peak>(), d_nan.as<int>(), st,
                  g.big || g.win_sub() ? g.d_win.as<int>() : nullptr);
    if (g.win_sub())
      launch_big_second_sub(g.X, g.d_rep.as<float2>(), g.d_tw.as<float2>(),
                            g.d_twb.as<float2>(), g.d_slots.as<PeakSlot>(),
                            g.big ? g.big_r0 : 1, g.cw,
                            n_ant, g.n_prn, g.n_bin, g.folded, g.d_win.as<int>(),
                            g.fp16 ? gain : fold, g.scale,
                            g.d_Zw.as<float2>(), g.d_m2.as<unsigned int>(),
                            d_peaks.as<tga_peak>(), st,
                            g.fp16 ? g.d_X16.as<uint32_t>() : nullptr,
                            g.fp16 ? g.d_rep16.as<uint32_t>() : nullptr);
    else if (g.big)
      launch_big_second(g.X, g.d_rep.as<float2>(), g.d_slots.as<PeakSlot>(),
                        static_cast<int>(g.M), static_cast<int>(g.P), g.cw, n_ant, g.n_prn,
                        g.n_bin, g.folded, g.d_win.as<int>(), fold, g.scale, g.d_Zw.as<float2>(),
                        g.d_Zws.as<float2>(), d_peaks.as<tga_peak>(), st, g.natural());
  }
  time_end(2, st, 0.0);
}

size_t Plan::surface_floats(int sys_index) const {
  size_t n = 0;
  for (size_t i = 0; i < systems.size(); ++i) {
    const SysEntry& e = systems[i];
    const Group& g = groups[e.group];
    if (sys_index < 0 || static_cast<size_t>(sys_index) == i)
      n += static_cast<size_t>(n_ant) * e.prns.size() * g.n_bin * 2 * g.P;
  }
  return n;
}

void Plan::surface(const uint8_t* d_in, float* d_out, hipStream_t st) {
  if (st == nullptr) st = stream;
  size_t tot = 0;
  for (auto& g : groups) {
    if (g.nc > 0 || !(g.fused || g.fp16 || g.big))
      fail(TGA_ERR_UNSUPPORTED,
           "the surface form covers the fused (16368) and split (4 or 10 x 16368) groups only");
    g.dump_off = tot;
    tot += static_cast<size_t>(n_ant) * g.n_prn * g.n_bin * 2 * g.P;
  }
  DevBuf stage(tot * sizeof(float));
  // unfolded long rows have no h = 1 row: those stay zero
  TGA_HIP(hipMemsetAsync(stage.p, 0, tot * sizeof(float), st));
  search(d_in, st, stage.as<float>());
  for (auto& g : groups)
    launch_surface_finish(stage.as<float>() + g.dump_off,
                          static_cast<long long>(n_ant) * g.n_prn * g.n_bin * 2 * g.P,
                          select_gain(g), g.scale, st);
  // group layout [ant][group prn][bin][h][P] -> system by system [ant][prn][bin][h][P]
  size_t o = 0;
  for (const auto& e : systems) {
    const Group& g = groups[e.group];
    const size_t per_prn = static_cast<size_t>(g.n_bin) * 2 * g.P;
    const size_t width = e.prns.size() * per_prn;
    TGA_HIP(hipMemcpy2DAsync(d_out + o, width * sizeof(float),
                             stage.as<float>() + g.dump_off + e.first_prn * per_prn,
                             g.n_prn * per_prn * sizeof(float), width * sizeof(float), n_ant,
                             hipMemcpyDeviceToDevice, st));
    o += width * n_ant;
  }
  TGA_HIP(hipStreamSynchronize(st));
}

void Plan::finish(tga_peak* out, hipStream_t st) {
  if (st == nullptr) st = stream;
  int nan = 0;
  TGA_HIP(hipMemcpyAsync(out, d_peaks.p, n_peaks * sizeof(tga_peak), hipMemcpyDeviceToHost, st));
  TGA_HIP(hipMemcpyAsync(&nan, d_nan.p, sizeof(int), hipMemcpyDeviceToHost, st));
  TGA_HIP(hipStreamSynchronize(st));
  time_collect();
  if (nan)
    fail(TGA_ERR_NAN, precision == TGA_PRECISION_FP16
                          ? "non-finite correlation value (NaN input or f16 range exceeded)"
                          : "called `Option::unwrap()` on a `None` value (NaN in the correlation)");
}

// Matrix-core instructions per transform, from the kernels' tile loops: the FP32 radix-31 pass
// is 16 v_mfma_f32_16x16x4f32 per tile of 16 butterflies (33 tiles per 16368 points, 17 per 8184
// including the half tile, 1 per 496-point four-step row); the FP16 rows kernel runs radix 33 as
// 31 tiles x 5 x 3, radix 16 as 64 tiles x 2 and radix 31 as 33 tiles x 8
// v_mfma_f32_16x16x32_f16.  The non-coherent form and the generic engine use no matrix cores.
void Plan::rows_matrix_flops(double& f32, double& f16) const {
  constexpr double kF32 = 2.0 * 16 * 16 * 4, kF16 = 2.0 * 16 * 16 * 32;
  constexpr double kRows31 = 33 * 16, kHalf31 = 17 * 16, kFourStep31 = 16;
  constexpr double kRows16 = 31 * 15 + 64 * 2 + 33 * 8;
  f32 = f16 = 0.0;
  for (const auto& g : groups) {
    if (g.nc > 0) continue;
    const double rows = static_cast<double>(g.n_rows);
    if (g.fp16) {
      f16 += rows * kRows16 * kF16;
    } else if (g.fused) {
      if (g.rows_cfg.variant == 1) f32 += rows * kRows31 * kF32;
    } else if (g.big) {
      const double long_rows = rows / (g.folded ? 1.0 : 2.0);  // unfolded: h = 0 rows only
      const double r0 = g.big_r0;
      if (g.big16)
        f16 += long_rows * r0 * kRows16 * kF16;
      else if (g.rows_cfg.variant == 1 || g.rows_cfg.variant == 2)
        f32 += long_rows * r0 * kRows31 * kF32;
      else if (g.rows_cfg.variant == 5)
        f32 += long_rows * 2.0 * r0 * kHalf31 * kF32;
      else if (g.rows_cfg.variant == 3)
        f32 += long_rows * (static_cast<double>(g.P) / 496.0) * kFourStep31 * kF32;
    }
  }
}

std::string Plan::describe() const {
  std::string s;
  for (const auto& e : systems) {
    const Group& g = groups[e.group];
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "%s: P=%zu N=%zu bins=%zu prns=%zu group=%d path=%s rows_variant=%d "
                  "groups=%d tune_ms=",
                  system_info(e.system).key, g.P, g.N, g.fc.size(), e.prns.size(), e.group,
                  g.nc > 0 ? "fused16368-noncoherent"
                  : g.fp16 ? "fused16368-fp16"
                  : g.fused ? "fused16368"
                  : g.natural() ? (g.big_r0 == 4 ? "split132x496" : "split330x496")
                  : g.big ? (g.big_r0 == 4 ? (g.big16 ? "split4x16368-fp16" : "split4x16368")
                                           : (g.big16 ? "split10x16368-fp16" : "split10x16368"))
                          : (g.folded ? "generic-folded" : "generic-unfolded"),
                  g.rows_cfg.variant, g.rows_cfg.groups);
    s += buf;
    // every candidate the autotuner timed: v<variant>g<groups>:<ms> (fused groups) or
    // v<variant>t<TX>x<TP>:<ms> (split groups)
    for (size_t k = 0; k < g.tune.size(); ++k) {
      const FusedRowsConfig& c = g.tune[k].first;
      if (g.big)
        std::snprintf(buf, sizeof buf, "%sv%dt%dx%d:%.3f", k ? "," : "", c.variant,
                      c.tile & 0xffff, c.tile >> 16, g.tune[k].second);
      else
        std::snprintf(buf, sizeof buf, "%sv%dg%d:%.3f", k ? "," : "", c.variant, c.groups,
                      g.tune[k].second);
      s += buf;
    }
    if (g.big) {
      std::snprintf(buf, sizeof buf, " tile=%dx%d", g.rows_cfg.tile & 0xffff,
                    g.rows_cfg.tile >> 16);
      s += buf;
    }
    s += g.tune.empty() ? "-\n" : "\n";
  }
  s += !graphable         ? "graph=off (Bluestein transform)"
       : !graph_enabled   ? "graph=off"
       : gexec            ? "graph=replay launches=" + std::to_string(graph_launches)
                          : "graph=pending (captured at the second search)";
  s += timed_graphs ? "\n" : " timed=eager (no event nodes in captures)\n";
  return s;
}

}  // namespace tga

using namespace tga;

extern "C" {

int tga_plan_create(const tga_plan_desc* desc, tga_plan** plan) {
  return plan_guard([&] {
    if (!desc || !plan) fail(TGA_ERR_INVALID_ARGUMENT, "desc and plan must not be null");
    *plan = nullptr;
    auto p = std::make_unique<Plan>();
    p->build(*desc);
    *plan = reinterpret_cast<tga_plan*>(p.release());
  });
}

int tga_plan_destroy(tga_plan* plan) {
  return plan_guard([&] { delete reinterpret_cast<Plan*>(plan); });
}

int tga_plan_row_bytes(const tga_plan* plan, size_t* row_bytes) {
  return plan_guard([&] {
    if (!plan || !row_bytes) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    *row_bytes = reinterpret_cast<const Plan*>(plan)->row_bytes;
  });
}

int tga_plan_peak_count(const tga_plan* plan, size_t* count) {
  return plan_guard([&] {
    if (!plan || !count) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    *count = reinterpret_cast<const Plan*>(plan)->n_peaks;
  });
}

int tga_plan_prns(const tga_plan* plan, int32_t sys_index, int32_t* prns, size_t cap,
                  size_t* n_prn) {
  return plan_guard([&] {
    if (!plan || !n_prn) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    const Plan* p = reinterpret_cast<const Plan*>(plan);
    if (sys_index < 0 || sys_index >= static_cast<int>(p->systems.size()))
      fail(TGA_ERR_INVALID_ARGUMENT, "system index out of range");
    const auto& v = p->systems[sys_index].prns;
    *n_prn = v.size();
    if (prns) {
      if (cap < v.size()) fail(TGA_ERR_INVALID_ARGUMENT, "capacity too small");
      std::copy(v.begin(), v.end(), prns);
    }
  });
}

int tga_search(tga_plan* plan, const uint8_t* packed, size_t n_ant, size_t row_bytes,
               tga_peak* out) {
  return plan_guard([&] {
    if (!plan || !packed || !out) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    Plan* p = reinterpret_cast<Plan*>(plan);
    if (n_ant != static_cast<size_t>(p->n_ant))
      fail(TGA_ERR_ANTENNA_RANGE, "antenna count " + std::to_string(n_ant) +
                                      " does not match the plan's " + std::to_string(p->n_ant));
    if (row_bytes != p->row_bytes)
      fail(TGA_ERR_INVALID_ARGUMENT, "row_bytes " + std::to_string(row_bytes) +
                                         " does not match the plan's " +
                                         std::to_string(p->row_bytes));
    TGA_HIP(hipMemcpyAsync(p->d_input.p, packed, n_ant * row_bytes, hipMemcpyHostToDevice,
                           p->stream));
    p->run(p->d_input.as<uint8_t>(), p->stream);
    p->finish(out, p->stream);
  });
}

int tga_search_device(tga_plan* plan, const void* d_packed, void* stream, tga_peak* out) {
  return plan_guard([&] {
    if (!plan || !d_packed || !out) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    Plan* p = reinterpret_cast<Plan*>(plan);
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : p->stream;
    p->run(static_cast<const uint8_t*>(d_packed), st);
    p->finish(out, st);
  });
}

int tga_plan_surface_size(const tga_plan* plan, int32_t sys_index, size_t* n_floats) {
  return plan_guard([&] {
    if (!plan || !n_floats) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    const Plan* p = reinterpret_cast<const Plan*>(plan);
    if (sys_index < -1 || sys_index >= static_cast<int32_t>(p->systems.size()))
      fail(TGA_ERR_INVALID_ARGUMENT, "system index out of range");
    *n_floats = p->surface_floats(sys_index);
  });
}

int tga_plan_surface(tga_plan* plan, const void* d_packed, void* stream, float* d_surface,
                     size_t cap, tga_peak* out) {
  return plan_guard([&] {
    if (!plan || !d_packed || !d_surface) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    Plan* p = reinterpret_cast<Plan*>(plan);
    const size_t need = p->surface_floats();
    if (cap < need)
      fail(TGA_ERR_INVALID_ARGUMENT, "surface buffer holds " + std::to_string(cap) +
                                         " floats, the plan's rows need " + std::to_string(need));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : p->stream;
    p->surface(static_cast<const uint8_t*>(d_packed), d_surface, st);
    if (out) p->finish(out, st);
  });
}

int tga_plan_carrier(const tga_plan* plan, int32_t sys_index, float* out, size_t cap,
                     size_t* n_floats) {
  return plan_guard([&] {
    if (!plan || !n_floats) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    const Plan* p = reinterpret_cast<const Plan*>(plan);
    if (sys_index < 0 || sys_index >= static_cast<int32_t>(p->systems.size()))
      fail(TGA_ERR_INVALID_ARGUMENT, "system index out of range");
    const Group& g = p->groups[p->systems[sys_index].group];
    if (!g.fused || !g.d_carrier.p)
      fail(TGA_ERR_UNSUPPORTED, "only fused (16368-sample) groups keep a carrier table");
    *n_floats = g.d_carrier.bytes / sizeof(float);
    if (out) {
      if (cap < *n_floats) fail(TGA_ERR_INVALID_ARGUMENT, "carrier buffer too small");
      TGA_HIP(hipMemcpy(out, g.d_carrier.p, g.d_carrier.bytes, hipMemcpyDeviceToHost));
    }
  });
}

int tga_plan_set_timing(tga_plan* plan, int32_t enable) {
  return plan_guard([&] {
    if (!plan) fail(TGA_ERR_INVALID_ARGUMENT, "null plan");
    Plan* p = reinterpret_cast<Plan*>(plan);
    p->timing = enable != 0;
    for (auto& t : p->timers) {
      t.total_ms = 0;
      t.launches = 0;
      t.pending = false;
    }
  });
}

int tga_plan_set_graph(tga_plan* plan, int32_t enable) {
  return plan_guard([&] {
    if (!plan) fail(TGA_ERR_INVALID_ARGUMENT, "null plan");
    Plan* p = reinterpret_cast<Plan*>(plan);
    TGA_HIP(hipStreamSynchronize(p->stream));
    p->graph_enabled = enable != 0;
    if (!p->graph_enabled) p->drop_graph();
  });
}

int tga_plan_get_timing(tga_plan* plan, int32_t kernel, double* total_ms, int64_t* launches,
                        double* bytes_per_launch) {
  return plan_guard([&] {
    if (!plan || !total_ms || !launches) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    Plan* p = reinterpret_cast<Plan*>(plan);
    if (kernel < 0 || kernel >= static_cast<int>(p->timers.size()))
      fail(TGA_ERR_INVALID_ARGUMENT, "timing slot out of range");
    *total_ms = p->timers[kernel].total_ms;
    *launches = p->timers[kernel].launches;
    if (bytes_per_launch) *bytes_per_launch = p->timers[kernel].bytes;
  });
}

int tga_plan_rows_matrix_flops(const tga_plan* plan, double* f32_flops, double* f16_flops) {
  return plan_guard([&] {
    if (!plan || !f32_flops || !f16_flops) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    reinterpret_cast<const Plan*>(plan)->rows_matrix_flops(*f32_flops, *f16_flops);
  });
}

const char* tga_plan_kernel_name(const tga_plan* plan, int32_t kernel) {
  const Plan* p = reinterpret_cast<const Plan*>(plan);
  if (!p || kernel < 0 || kernel > 2) return "";
  return p->kernel_names[kernel].c_str();
}

int tga_plan_describe(const tga_plan* plan, char* buf, size_t cap) {
  return plan_guard([&] {
    if (!plan || !buf || cap == 0) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    std::string s = reinterpret_cast<const Plan*>(plan)->describe();
    std::strncpy(buf, s.c_str(), cap - 1);
    buf[cap - 1] = 0;
  });
}

}  // extern "C"
