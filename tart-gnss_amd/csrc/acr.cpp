// ACR C/N0 estimation (acr.rs:49-84) on the host in f64.
//
// r_A = (V_m / V_s)^2 is mapped to dB-Hz by linear interpolation in ln(r_A) over the
// Monte-Carlo tables (acr_tables.rs), clamped to the first/last entry outside the table and
// to [20, 70] dB-Hz inside it.
#include <algorithm>
#include <cmath>

#include "internal.h"
#include "tables_gen.h"  // kAcrTables[4], kAcrSizes[4]

namespace tga {

bool estimate_cn0(double r_a, int table, double* cn0) {
  if (table < 0 || table > 3) fail(TGA_ERR_INVALID_ARGUMENT, "unknown ACR table");
  if (!(r_a > 1.0) || !std::isfinite(r_a)) return false;  // acr.rs:50-52
  const AcrPoint* t = kAcrTables[table];
  const int n = kAcrSizes[table];
  if (r_a <= t[0].r) { *cn0 = t[0].cn0; return true; }
  if (r_a >= t[n - 1].r) { *cn0 = t[n - 1].cn0; return true; }
  // first entry with r >= r_a (partition_point of r < r_a)
  const AcrPoint* hi = std::lower_bound(t, t + n, r_a,
                                        [](const AcrPoint& p, double v) { return p.r < v; });
  int idx = static_cast<int>(hi - t);
  if (idx == 0) { *cn0 = t[0].cn0; return true; }
  if (idx >= n) { *cn0 = t[n - 1].cn0; return true; }
  const double lr = std::log(r_a), l0 = std::log(t[idx - 1].r), l1 = std::log(t[idx].r);
  const double f = (lr - l0) / (l1 - l0);
  double c = t[idx - 1].cn0 + f * (t[idx].cn0 - t[idx - 1].cn0);
  *cn0 = std::min(std::max(c, 20.0), 70.0);
  return true;
}

}  // namespace tga
