"""Antenna-relative C/N0 report, --test-antennas (oracle restatement; test infrastructure).

Follows antenna_test.rs:105-222 and its serialisation (main.rs:848-880):
  * reference set: SVs, over CombinedOutput's fields in order (lib.rs:33-44), whose cn0_acr is
    present and has some antenna >= min_cn0 (antenna_test.rs:115-131);
  * antenna_numbers from the first output present (antenna_test.rs:117-119);
  * per antenna column, its estimates over the reference set (antenna_test.rs:160-169); the
    median (stats.rs:10-30) of each non-empty column, ranked by median descending, ties by
    antenna number (antenna_test.rs:173-179); offset = median - best; n_sats = estimates
    >= min_cn0 in the column (antenna_test.rs:181-199); columns with no estimate unscored
    (null median, offset and rank, n_sats 0; antenna_test.rs:201-212); antennas sorted by
    number (antenna_test.rs:213);
  * an empty reference set gives empty antennas and per_satellite (antenna_test.rs:147-156).
"""
from . import jsonfmt as _jsonfmt
from . import stats as _stats

KEY_ORDER = ["gps", "galileo", "beidou", "sbas", "l1c", "qzss"]


def run(combined, min_cn0):
    """combined: {key: {"antenna_numbers": [...], "results": [{"sv", "cn0_acr"?...}]}}."""
    antenna_numbers, candidates = [], []
    for key in KEY_ORDER:
        out = combined.get(key)
        if out is None:
            continue
        if not antenna_numbers:
            antenna_numbers = list(out["antenna_numbers"])
        for r in out["results"]:
            cn0s = r.get("cn0_acr")
            if cn0s is None:
                continue
            if any(c is not None and c >= min_cn0 for c in cn0s):
                candidates.append((r["sv"], list(cn0s)))
    report = {"antenna_numbers": antenna_numbers, "min_cn0_db_hz": float(min_cn0),
              "n_reference_satellites": len(candidates),
              "reference_satellites": [sv for sv, _ in candidates],
              "antennas": [], "per_satellite": []}
    if not candidates:
        return report
    report["per_satellite"] = [{"sv": sv, "cn0_db_hz": c} for sv, c in candidates]
    cols = [[] for _ in antenna_numbers]
    for _, cn0s in candidates:
        for col, c in enumerate(cn0s):
            if c is not None:
                cols[col].append(float(c))
    scored = [(a, _stats.median(v)) for a, v in zip(antenna_numbers, cols) if v]
    scored.sort(key=lambda t: (-t[1], t[0]))
    best = scored[0][1]
    stats = []
    for rank, (ant, med) in enumerate(scored):
        col = antenna_numbers.index(ant)
        stats.append({"antenna": ant, "n_sats": sum(c >= min_cn0 for c in cols[col]),
                      "median_cn0_db_hz": med, "offset_db": med - best, "rank": rank + 1})
    for ant, v in zip(antenna_numbers, cols):
        if not v:
            stats.append({"antenna": ant, "n_sats": 0, "median_cn0_db_hz": None,
                          "offset_db": None, "rank": None})
    stats.sort(key=lambda s: s["antenna"])
    report["antennas"] = stats
    return report


def report_json(report):
    """serde_json::to_string_pretty(&AntennaTestOutput) (main.rs:880)."""
    return _jsonfmt.dumps_pretty(report)
