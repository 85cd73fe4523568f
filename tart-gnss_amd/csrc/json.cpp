// Host orchestration of acquire_all_* and the reference's JSON output.
//
// One batched search of the plan replaces the rayon loops of acquire_all_{gps,galileo,beidou,
// sbas,l1c,qzss} (acquisition.rs:261-396 and siblings); this file maps its peaks to the
// reference's per-SV records and writes CombinedOutput (lib.rs:31-45) the way
// serde_json::to_string_pretty does (main.rs:893): two-space indent, keys in struct order,
// Option::None fields skipped, None inside vectors as null, f64 through shortest round-trip
// digits laid out like ryu (integral values with ".0", plain decimals for 1e-5 <= |x| < 1e16,
// otherwise d.ddde<exp>).
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <string>

#include "internal.h"
#include "plan.h"

namespace tga {

// f64 -> the text serde_json (ryu) prints.
std::string json_f64(double x) {
  if (!std::isfinite(x)) return "null";
  if (x == 0.0) return std::signbit(x) ? "-0.0" : "0.0";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, std::fabs(x), std::chars_format::scientific);
  std::string s(buf, r.ptr);
  const size_t e = s.find('e');
  std::string digits;
  for (size_t i = 0; i < e; ++i)
    if (s[i] != '.') digits += s[i];
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  const int exp10 = std::atoi(s.c_str() + e + 1);  // value = d.ddd * 10^exp10
  const int length = static_cast<int>(digits.size());
  const int kk = exp10 + 1;  // 10^(kk-1) <= v < 10^kk
  const int k = kk - length;  // value = digits * 10^k
  std::string out = std::signbit(x) ? "-" : "";
  if (0 <= k && kk <= 16) {
    out += digits + std::string(k, '0') + ".0";
  } else if (0 < kk && kk <= 16) {
    out += digits.substr(0, kk) + "." + digits.substr(kk);
  } else if (-5 < kk && kk <= 0) {
    out += "0." + std::string(-kk, '0') + digits;
  } else if (length == 1) {
    out += digits + "e" + std::to_string(kk - 1);
  } else {
    out += digits.substr(0, 1) + "." + digits.substr(1) + "e" + std::to_string(kk - 1);
  }
  return out;
}

namespace {

// stats.rs:10-30
double median(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  const size_t mid = v.size() / 2;
  return v.size() % 2 == 0 ? (v[mid - 1] + v[mid]) / 2.0 : v[mid];
}

double mad(const std::vector<double>& v, double med) {
  std::vector<double> d(v.size());
  for (size_t i = 0; i < v.size(); ++i) d[i] = std::fabs(v[i] - med);
  return median(d);
}

struct SvRecord {
  std::string sv;
  std::vector<double> strengths, phases, freqs;
  bool has_cn0 = false;
  std::vector<double> cn0;
  std::vector<bool> cn0_ok;
  bool has_stats = false;
  double phase_median = 0, phase_mad = 0, freq_median = 0, freq_mad = 0;
};

void copy_out(const std::string& s, char** json) {
  char* out = new char[s.size() + 1];
  std::memcpy(out, s.c_str(), s.size() + 1);
  *json = out;
}

struct SystemOut {
  const char* key;
  std::vector<SvRecord> recs;
};

class Writer {
 public:
  std::string out;
  void raw(const std::string& s) { out += s; }
  void nl(int depth) {
    out += '\n';
    out.append(static_cast<size_t>(2 * depth), ' ');
  }
  void key(const char* k, int depth, bool first) {
    if (!first) out += ',';
    nl(depth);
    out += '"';
    out += k;
    out += "\": ";
  }
  void str(const std::string& s) {
    out += '"';
    for (char c : s) {
      if (c == '"' || c == '\\') out += '\\';
      out += c;
    }
    out += '"';
  }
  template <class F>
  void array(size_t n, int depth, F&& elem) {
    if (n == 0) {
      out += "[]";
      return;
    }
    out += '[';
    for (size_t i = 0; i < n; ++i) {
      if (i) out += ',';
      nl(depth + 1);
      elem(i);
    }
    nl(depth);
    out += ']';
  }
};

}  // namespace

// The per-SV records of every plan system, in CombinedOutput's field order (lib.rs:33-44),
// after the MAD filters and the sort by label: the data acquire_all_* hands to the JSON
// writer and to antenna_test::run.
std::vector<SystemOut> collect_outputs(Plan& p, const tga_peak* peaks, const double* center,
                                       bool cn0, double filt_phase, double filt_freq) {
  const int order[TGA_NUM_SYSTEMS] = {TGA_SYS_GPS, TGA_SYS_GALILEO, TGA_SYS_BEIDOU,
                                      TGA_SYS_SBAS, TGA_SYS_L1C, TGA_SYS_QZSS};
  std::vector<SystemOut> outs;
  for (int sysid : order) {
    int si = -1;
    for (size_t i = 0; i < p.systems.size(); ++i)
      if (p.systems[i].system == sysid) {
        if (si >= 0) fail(TGA_ERR_UNSUPPORTED, "a system appears twice in the plan");
        si = static_cast<int>(i);
      }
    if (si < 0) continue;
    const SysEntry& e = p.systems[si];
    const Group& g = p.groups[e.group];
    const SystemInfo& info = system_info(sysid);
    const double P = static_cast<double>(g.P);
    SystemOut so;
    so.key = info.key;
    for (size_t j = 0; j < e.prns.size(); ++j) {
      SvRecord r;
      char lab[32];
      std::snprintf(lab, sizeof lab, info.label_fmt, e.prns[j]);
      r.sv = lab;
      std::vector<double> second;
      for (int a = 0; a < p.n_ant; ++a) {
        const tga_peak& pk = peaks[e.peak_offset + static_cast<size_t>(a) * e.prns.size() + j];
        r.strengths.push_back(static_cast<double>(pk.best_peak));
        second.push_back(static_cast<double>(pk.second_peak));
        // acquisition.rs:237-239
        r.phases.push_back(static_cast<double>(pk.best_codephase % g.P) / P);
        r.freqs.push_back(g.fc[pk.best_freq_idx] - center[si]);
      }
      if (cn0) {  // acquisition.rs:341-357
        for (int a = 0; a < p.n_ant; ++a) {
          double c = 0;
          bool ok = false;
          const double vm = r.strengths[a], vs = second[a];
          if (vs > 0.0 && vm > vs) ok = estimate_cn0((vm / vs) * (vm / vs), info.acr_table, &c);
          r.cn0.push_back(c);
          r.cn0_ok.push_back(ok);
          r.has_cn0 = r.has_cn0 || ok;
        }
      }
      if (p.n_ant > 1) {  // acquisition.rs:365-377
        r.has_stats = true;
        r.phase_median = median(r.phases);
        r.phase_mad = mad(r.phases, r.phase_median);
        r.freq_median = median(r.freqs);
        r.freq_mad = mad(r.freqs, r.freq_median);
      }
      // MAD filters (main.rs:242-255): a missing MAD always passes
      if (r.has_stats && filt_phase >= 0 && !(r.phase_mad <= filt_phase)) continue;
      if (r.has_stats && filt_freq >= 0 && !(r.freq_mad <= filt_freq)) continue;
      so.recs.push_back(std::move(r));
    }
    std::stable_sort(so.recs.begin(), so.recs.end(),
                     [](const SvRecord& a, const SvRecord& b) { return a.sv < b.sv; });
    outs.push_back(std::move(so));
  }
  return outs;
}

std::string combined_json(const std::vector<SystemOut>& outs, const int64_t* antenna_numbers,
                          size_t n_ant) {
  Writer w;
  w.raw("{");
  bool first_sys = true;
  for (const SystemOut& so : outs) {
    const std::vector<SvRecord>& recs = so.recs;
    w.key(so.key, 1, first_sys);
    first_sys = false;
    w.raw("{");
    w.key("antenna_numbers", 2, true);
    w.array(n_ant, 2, [&](size_t i) { w.raw(std::to_string(antenna_numbers[i])); });
    w.key("results", 2, false);
    w.array(recs.size(), 2, [&](size_t i) {
      const SvRecord& r = recs[i];
      const int d = 4;
      w.raw("{");
      w.key("sv", d, true);
      w.str(r.sv);
      auto vec = [&](const char* k, const std::vector<double>& v) {
        w.key(k, d, false);
        w.array(v.size(), d, [&](size_t q) { w.raw(json_f64(v[q])); });
      };
      vec("strengths", r.strengths);
      vec("phases", r.phases);
      vec("freqs", r.freqs);
      if (r.has_cn0) {
        w.key("cn0_acr", d, false);
        w.array(r.cn0.size(), d, [&](size_t q) {
          w.raw(r.cn0_ok[q] ? json_f64(r.cn0[q]) : std::string("null"));
        });
      }
      if (r.has_stats) {
        w.key("phase_median", d, false);
        w.raw(json_f64(r.phase_median));
        w.key("phase_mad", d, false);
        w.raw(json_f64(r.phase_mad));
        w.key("freq_median", d, false);
        w.raw(json_f64(r.freq_median));
        w.key("freq_mad", d, false);
        w.raw(json_f64(r.freq_mad));
      }
      w.nl(d - 1);
      w.raw("}");
    });
    w.nl(1);
    w.raw("}");
  }
  if (first_sys) {
    w.raw("}");
  } else {
    w.nl(0);
    w.raw("}");
  }
  return w.out;
}

// RfiReport (rfi.rs:73-77) as serde_json::to_string_pretty writes it (main.rs:468): every
// field of RfiChannel present in struct order (rfi.rs:54-70), Option::None as null (NaN in
// tga_rfi_channel), lines as {frequency, excess_db} objects, antenna_numbers the channels'
// antenna indices in order (rfi.rs:188-191).
std::string rfi_report_json(const tga_rfi_channel* ch, size_t n) {
  Writer w;
  auto opt = [](double v) { return std::isnan(v) ? std::string("null") : json_f64(v); };
  w.raw("{");
  w.key("antenna_numbers", 1, true);
  w.array(n, 1, [&](size_t i) { w.raw(std::to_string(ch[i].antenna)); });
  w.key("channels", 1, false);
  w.array(n, 1, [&](size_t i) {
    const tga_rfi_channel& c = ch[i];
    const int d = 3;
    w.raw("{");
    w.key("antenna", d, true);
    w.raw(std::to_string(c.antenna));
    w.key("dead", d, false);
    w.raw(c.dead ? "true" : "false");
    w.key("acf_lag1", d, false);
    w.raw(opt(c.acf_lag1));
    w.key("acf_lag3", d, false);
    w.raw(opt(c.acf_lag3));
    w.key("acf_lag16", d, false);
    w.raw(opt(c.acf_lag16));
    w.key("period_samples", d, false);
    w.raw(opt(c.period_samples));
    w.key("lines", d, false);
    w.array(static_cast<size_t>(c.n_lines), d, [&](size_t q) {
      w.raw("{");
      w.key("frequency", d + 2, true);
      w.raw(json_f64(c.lines[q].frequency));
      w.key("excess_db", d + 2, false);
      w.raw(json_f64(c.lines[q].excess_db));
      w.nl(d + 1);
      w.raw("}");
    });
    w.nl(d - 1);
    w.raw("}");
  });
  w.nl(0);
  w.raw("}");
  return w.out;
}

std::string acquire_all_json(Plan& p, const tga_peak* peaks, const double* center,
                             const int64_t* antenna_numbers, bool cn0, double filt_phase,
                             double filt_freq) {
  return combined_json(collect_outputs(p, peaks, center, cn0, filt_phase, filt_freq),
                       antenna_numbers, static_cast<size_t>(p.n_ant));
}

// antenna_test::run (antenna_test.rs:105-222) over rows (sv label, per-antenna C/N0 or NaN
// for "no estimate"), in CombinedOutput order, and the report as serde_json::to_string_pretty
// writes AntennaTestOutput (main.rs:880): every field present, None as null.
//   1. reference set: rows with some antenna at C/N0 >= min_cn0 (a row without any estimate,
//      cn0_acr None, never qualifies);
//   2. per antenna column, the median of its estimates over the reference set; columns with
//      no estimate at all are unscored;
//   3. rank scored antennas by median, descending, ties by antenna number; offset = median -
//      best median; n_sats = estimates >= min_cn0 in the column; antennas listed by number.
std::string antenna_report_json(const std::vector<std::string>& svs,
                                const std::vector<std::vector<double>>& cn0,
                                const int64_t* antenna_numbers, size_t n_ant, double min_cn0) {
  std::vector<size_t> cand;
  for (size_t r = 0; r < svs.size(); ++r) {
    bool any = false;
    for (size_t a = 0; a < n_ant; ++a) any = any || (!std::isnan(cn0[r][a]) && cn0[r][a] >= min_cn0);
    if (any) cand.push_back(r);
  }
  struct Stat {
    int64_t antenna;
    size_t n_sats;
    bool scored;
    double median, offset;
    size_t rank;
  };
  std::vector<Stat> stats;
  if (!cand.empty()) {
    std::vector<std::vector<double>> col(n_ant);
    for (size_t r : cand)
      for (size_t a = 0; a < n_ant; ++a)
        if (!std::isnan(cn0[r][a])) col[a].push_back(cn0[r][a]);
    std::vector<std::pair<size_t, double>> ranked;  // (column, median)
    for (size_t a = 0; a < n_ant; ++a)
      if (!col[a].empty()) ranked.emplace_back(a, median(col[a]));
    std::stable_sort(ranked.begin(), ranked.end(), [&](const auto& x, const auto& y) {
      if (x.second != y.second) return x.second > y.second;
      return antenna_numbers[x.first] < antenna_numbers[y.first];
    });
    const double best = ranked[0].second;
    for (size_t a = 0; a < n_ant; ++a) {
      Stat s{antenna_numbers[a], 0, false, 0.0, 0.0, 0};
      for (double c : col[a]) s.n_sats += c >= min_cn0;
      stats.push_back(s);
    }
    for (size_t k = 0; k < ranked.size(); ++k) {
      Stat& s = stats[ranked[k].first];
      s.scored = true;
      s.median = ranked[k].second;
      s.offset = ranked[k].second - best;
      s.rank = k + 1;
    }
    std::stable_sort(stats.begin(), stats.end(),
                     [](const Stat& x, const Stat& y) { return x.antenna < y.antenna; });
  }
  Writer w;
  w.raw("{");
  w.key("antenna_numbers", 1, true);
  w.array(n_ant, 1, [&](size_t i) { w.raw(std::to_string(antenna_numbers[i])); });
  w.key("min_cn0_db_hz", 1, false);
  w.raw(json_f64(min_cn0));
  w.key("n_reference_satellites", 1, false);
  w.raw(std::to_string(cand.size()));
  w.key("reference_satellites", 1, false);
  w.array(cand.size(), 1, [&](size_t i) { w.str(svs[cand[i]]); });
  w.key("antennas", 1, false);
  w.array(stats.size(), 1, [&](size_t i) {
    const Stat& s = stats[i];
    w.raw("{");
    w.key("antenna", 3, true);
    w.raw(std::to_string(s.antenna));
    w.key("n_sats", 3, false);
    w.raw(std::to_string(s.n_sats));
    w.key("median_cn0_db_hz", 3, false);
    w.raw(s.scored ? json_f64(s.median) : std::string("null"));
    w.key("offset_db", 3, false);
    w.raw(s.scored ? json_f64(s.offset) : std::string("null"));
    w.key("rank", 3, false);
    w.raw(s.scored ? std::to_string(s.rank) : std::string("null"));
    w.nl(2);
    w.raw("}");
  });
  w.key("per_satellite", 1, false);
  w.array(cand.size(), 1, [&](size_t i) {
    w.raw("{");
    w.key("sv", 3, true);
    w.str(svs[cand[i]]);
    w.key("cn0_db_hz", 3, false);
    w.array(n_ant, 3, [&](size_t a) { w.raw(json_f64(cn0[cand[i]][a])); });
    w.nl(2);
    w.raw("}");
  });
  w.nl(0);
  w.raw("}");
  return w.out;
}

}  // namespace tga

using namespace tga;

extern "C" {

int tga_format_f64(double x, char* buf, size_t cap) {
  return plan_guard([&] {
    if (!buf || cap == 0) fail(TGA_ERR_INVALID_ARGUMENT, "null buffer");
    const std::string s = json_f64(x);
    if (s.size() + 1 > cap) fail(TGA_ERR_INVALID_ARGUMENT, "buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
  });
}

int tga_acquire_all_json(tga_plan* plan, const uint8_t* packed, size_t n_ant, size_t row_bytes,
                         const double* center_freq, const int64_t* antenna_numbers, int32_t cn0,
                         double filter_phase_mad, double filter_freq_mad, char** json) {
  return plan_guard([&] {
    if (!plan || !packed || !center_freq || !antenna_numbers || !json)
      fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    *json = nullptr;
    Plan* p = reinterpret_cast<Plan*>(plan);
    std::vector<tga_peak> peaks(p->n_peaks);
    if (int rc = tga_search(plan, packed, n_ant, row_bytes, peaks.data()); rc != TGA_OK)
      fail(rc, tga_last_error());
    copy_out(acquire_all_json(*p, peaks.data(), center_freq, antenna_numbers, cn0 != 0,
                              filter_phase_mad, filter_freq_mad),
             json);
  });
}

int tga_peaks_json(tga_plan* plan, const tga_peak* peaks, const double* center_freq,
                   const int64_t* antenna_numbers, int32_t cn0, double filter_phase_mad,
                   double filter_freq_mad, char** json) {
  return plan_guard([&] {
    if (!plan || !peaks || !center_freq || !antenna_numbers || !json)
      fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    *json = nullptr;
    Plan* p = reinterpret_cast<Plan*>(plan);
    copy_out(acquire_all_json(*p, peaks, center_freq, antenna_numbers, cn0 != 0,
                              filter_phase_mad, filter_freq_mad),
             json);
  });
}

int tga_peaks_cn0(const tga_plan* plan, const tga_peak* peaks, double* cn0) {
  return plan_guard([&] {
    if (!plan || !peaks || !cn0) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    const Plan* p = reinterpret_cast<const Plan*>(plan);
    for (const SysEntry& e : p->systems) {
      const int table = system_info(e.system).acr_table;
      const size_t n = static_cast<size_t>(p->n_ant) * e.prns.size();
      for (size_t i = e.peak_offset; i < e.peak_offset + n; ++i) {
        const double vm = peaks[i].best_peak, vs = peaks[i].second_peak;
        double c = 0.0;
        const bool ok = vs > 0.0 && vm > vs && estimate_cn0((vm / vs) * (vm / vs), table, &c);
        cn0[i] = ok ? c : std::nan("");
      }
    }
  });
}

int tga_antenna_report_json(size_t n_sv, const char* const* sv, const double* cn0,
                            size_t n_ant, const int64_t* antenna_numbers, double min_cn0,
                            char** json) {
  return plan_guard([&] {
    if (!json || (n_sv && (!sv || (n_ant && !cn0))) || (n_ant && !antenna_numbers))
      fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    *json = nullptr;
    std::vector<std::string> svs(n_sv);
    std::vector<std::vector<double>> rows(n_sv, std::vector<double>(n_ant));
    for (size_t r = 0; r < n_sv; ++r) {
      if (!sv[r]) fail(TGA_ERR_INVALID_ARGUMENT, "null sv label");
      svs[r] = sv[r];
      for (size_t a = 0; a < n_ant; ++a) rows[r][a] = cn0[r * n_ant + a];
    }
    copy_out(antenna_report_json(svs, rows, antenna_numbers, n_ant, min_cn0), json);
  });
}

int tga_peaks_antenna_report_json(tga_plan* plan, const tga_peak* peaks,
                                  const double* center_freq, const int64_t* antenna_numbers,
                                  double filter_phase_mad, double filter_freq_mad,
                                  double min_cn0, char** json) {
  return plan_guard([&] {
    if (!plan || !peaks || !center_freq || !antenna_numbers || !json)
      fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    *json = nullptr;
    Plan* p = reinterpret_cast<Plan*>(plan);
    const size_t n_ant = static_cast<size_t>(p->n_ant);
    std::vector<std::string> svs;
    std::vector<std::vector<double>> rows;
    // --test-antennas forces the C/N0 estimates (main.rs:181-182)
    for (const SystemOut& so : collect_outputs(*p, peaks, center_freq, true, filter_phase_mad,
                                               filter_freq_mad))
      for (const SvRecord& r : so.recs) {
        if (!r.has_cn0) continue;  // cn0_acr None
        svs.push_back(r.sv);
        rows.emplace_back(n_ant);
        for (size_t a = 0; a < n_ant; ++a)
          rows.back()[a] = r.cn0_ok[a] ? r.cn0[a] : std::nan("");
      }
    copy_out(antenna_report_json(svs, rows, antenna_numbers, n_ant, min_cn0), json);
  });
}

int tga_acquire_antenna_report_json(tga_plan* plan, const uint8_t* packed, size_t n_ant,
                                    size_t row_bytes, const double* center_freq,
                                    const int64_t* antenna_numbers, double filter_phase_mad,
                                    double filter_freq_mad, double min_cn0, char** json) {
  return plan_guard([&] {
    if (!plan || !packed || !json) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    *json = nullptr;
    Plan* p = reinterpret_cast<Plan*>(plan);
    std::vector<tga_peak> peaks(p->n_peaks);
    if (int rc = tga_search(plan, packed, n_ant, row_bytes, peaks.data()); rc != TGA_OK)
      fail(rc, tga_last_error());
    if (int rc = tga_peaks_antenna_report_json(plan, peaks.data(), center_freq,
                                               antenna_numbers, filter_phase_mad,
                                               filter_freq_mad, min_cn0, json);
        rc != TGA_OK)
      fail(rc, tga_last_error());
  });
}

int tga_rfi_report_json(const tga_rfi_channel* channels, size_t n_ant, char** json) {
  return plan_guard([&] {
    if (!json || (n_ant && !channels)) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    *json = nullptr;
    for (size_t i = 0; i < n_ant; ++i)
      if (channels[i].n_lines < 0 || channels[i].n_lines > 3)
        fail(TGA_ERR_INVALID_ARGUMENT, "n_lines must be in 0..3");
    copy_out(rfi_report_json(channels, n_ant), json);
  });
}

int tga_rfi_json(const uint8_t* packed, size_t n_ant, size_t row_bytes, int64_t record_samples,
                 int32_t input_format, double sampling_freq, char** json) {
  return plan_guard([&] {
    if (!json) fail(TGA_ERR_INVALID_ARGUMENT, "null argument");
    *json = nullptr;
    std::vector<tga_rfi_channel> ch(n_ant);
    if (int rc = tga_rfi(packed, n_ant, row_bytes, record_samples, input_format,
                         sampling_freq, ch.data());
        rc != TGA_OK)
      fail(rc, tga_last_error());
    copy_out(rfi_report_json(ch.data(), n_ant), json);
  });
}

void tga_free_string(char* s) { delete[] s; }

}  // extern "C"
