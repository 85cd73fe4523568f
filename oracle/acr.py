"""ACR C/N0 estimation (oracle restatement of acr.rs:49-84 and the ratio gate of
acquisition.rs:341-357).  Tables are read from the data file the library also embeds."""
import math
import os

import numpy as np

_TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "tart-gnss_amd", "csrc", "tables", "acr_tables.txt")


def load_tables(path=_TABLE):
    tables = {}
    cur = None
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            if line.startswith("table "):
                _, name, count = line.split()
                cur = tables.setdefault(name, [])
                continue
            r, c = line.split()
            cur.append((float(r), float(c)))
    return tables


TABLES = load_tables()


def estimate_cn0(r_a, table):
    """acr.rs:49-84: None if r_a <= 1 or not finite; clamp to the table ends; otherwise
    interpolate linearly in ln(r_a) between partition_point brackets, clamp to [20, 70]."""
    if isinstance(table, str):
        table = TABLES[table]
    if not math.isfinite(r_a) or r_a <= 1.0:
        return None
    rs = np.array([t[0] for t in table])
    if r_a <= table[0][0]:
        return table[0][1]
    if r_a >= table[-1][0]:
        return table[-1][1]
    idx = int(np.searchsorted(rs, r_a, side="left"))
    if idx == 0:
        return table[0][1]
    if idx >= len(table):
        return table[-1][1]
    r_lo, c_lo = table[idx - 1]
    r_hi, c_hi = table[idx]
    t = (math.log(r_a) - math.log(r_lo)) / (math.log(r_hi) - math.log(r_lo))
    return min(max(c_lo + t * (c_hi - c_lo), 20.0), 70.0)


def acr_from_peaks(v_m, v_s, table):
    """acquisition.rs:345-352: only if v_s > 0 and v_m > v_s, r = (v_m / v_s)^2."""
    if v_s > 0.0 and v_m > v_s:
        return estimate_cn0((v_m / v_s) ** 2, table)
    return None
