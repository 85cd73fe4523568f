"""ACR C/N0 calibration tables by Monte Carlo on the GPU search.

Restates examples/gen_acr_tables.rs (gen_acr_tables.rs:1-230) on the batched plan: for each
C/N0 on a grid, TRIALS two-period signals at 16.368 MHz (random code phase and carrier phase,
carrier at 100 kHz, unit-variance noise, amplitude A = 2 sqrt(C/N0 / fs), gen_acr_tables.rs
:60-64) are one-bit quantized and de-meaned like a TART observation, searched for PRN 1 in the
single bin at 100 kHz, and reduced to

    r_A = mean(V_m^2) / mean(V_s^2)          (peak and second-peak powers, paper eq. 31)

with the reference's monotonicity fix (gen_acr_tables.rs:170-175).  Every trial of a C/N0
point is one antenna row of one plan, so a point costs one batched search instead of TRIALS
calls of acquire_full.  The pseudo-random streams are numpy's, not the reference's xorshift, so
tables agree with acr_tables.rs within Monte Carlo error, not digit for digit.

    PYTHONPATH=tart-gnss_amd python -m tga.acr_tables gps --trials 100   # prints a table in the reference's format
"""
import argparse
import math
import sys

import numpy as np

from . import api
from . import synth

FS = 16.368e6
F_CARRIER = 100e3
SPECS = {
    # system: (table name, code period at 16.368 MHz, C/N0 low, high) -- gen_acr_tables.rs:186-229
    "gps": ("GPS_L1_CA_ACR_TABLE", 16368, 33.0, 62.0),
    "galileo": ("GALILEO_E1C_ACR_TABLE", 65472, 27.0, 62.0),
    "beidou": ("BEIDOU_B1C_ACR_TABLE", 163680, 23.0, 62.0),
    "l1c": ("GPS_L1C_ACR_TABLE", 163680, 23.0, 62.0),
}


def trial_rows(system, cn0_db, trials, rng, prn=1):
    """Packed one-bit rows [trials][2P/8] of the generator's synth() (gen_acr_tables.rs:66-80)
    and the code phases the search should find."""
    _, P, _, _ = SPECS[system]
    n = 2 * P
    code = api.code_resampled(system, prn, float(P), n)
    a = 2.0 * math.sqrt(10.0 ** (cn0_db / 10.0) / FS)
    delays = rng.integers(0, P, trials)
    thetas = rng.uniform(0.0, 2.0 * math.pi, trials)
    i = np.arange(n)
    carrier_w = 2.0 * math.pi * F_CARRIER / FS
    rows = np.empty((trials, n // 8), dtype=np.uint8)
    for t in range(trials):
        cont = code[(i + delays[t]) % P] * (a * np.cos(carrier_w * i + thetas[t]))
        cont += rng.standard_normal(n)
        rows[t] = synth.pack_1bit(cont)  # v >= 0 -> 1: quantize_bipolar_demean's sign rule
    return rows, (P - delays) % P


def ratio(system, cn0_db, trials, rng, plan=None):
    """r_A at one C/N0 from `trials` single-bin searches batched as antenna rows."""
    _, P, _, _ = SPECS[system]
    rows, want = trial_rows(system, cn0_db, trials, rng)
    if plan is None:
        plan = api.Plan(FS, trials, 2 * P, [api.SystemSearch(system, [F_CARRIER], [1])],
                        input_format=1)
    pk = plan.search(rows)[0][:, 0]
    vm = pk["best_peak"].astype(np.float64)
    vs = pk["second_peak"].astype(np.float64)
    found = float(np.mean(pk["best_codephase"].astype(np.int64) % P == want))
    return float(np.sum(vm * vm) / np.sum(vs * vs)), found, plan


# The reference's end-to-end C/N0 tests (acquisition.rs:551-595, galileo.rs:465-510,
# beidou.rs:414-459, l1c.rs:426-470): trials without / with the sign flip, code-phase step per
# trial (delay = t * step mod P).
E2E = {"gps": (80, 120, 137), "galileo": (60, 120, 251), "beidou": (50, 120, 379),
       "l1c": (50, 120, 431)}


def flip_rows(system, cn0_db, trials, flip, rng, prn=1):
    """Packed one-bit rows [trials][2P/8] of testutil.rs:60-81 synth_signal_cn0: code delayed by
    t * step mod P samples on a cos carrier at 100 kHz (phase 0), amplitude 2 sqrt(C/N0 / fs),
    unit Gaussian noise; with `flip` the sign flips at the code epoch between the two periods,
    absolute sample P + (P - delay) mod P.  Quantized like quantize_bipolar_demean
    (simulator.rs:237-244: v >= 0 -> +1; the plan de-means the record)."""
    _, P, _, _ = SPECS[system]
    n = 2 * P
    code = api.code_resampled(system, prn, float(P), n)
    a = 2.0 * math.sqrt(10.0 ** (cn0_db / 10.0) / FS)
    i = np.arange(n)
    carrier = a * np.cos(2.0 * math.pi * F_CARRIER * i / FS)
    delays = (np.arange(trials) * E2E[system][2]) % P
    rows = np.empty((trials, n // 8), dtype=np.uint8)
    for t in range(trials):
        d = int(delays[t])
        cont = code[(i + d) % P] * carrier
        if flip:
            cont[P + (P - d) % P:] *= -1.0
        cont += rng.standard_normal(n)
        rows[t] = synth.pack_1bit(cont.astype(np.float32))
    return rows, (P - delays) % P


def end_to_end(system, cn0_db=45.0, flip=False, trials=None, seed=None):
    """Median ACR C/N0 estimate of the reference's end-to-end test: per trial r_A =
    (signal_strength / second_peak)^2 of a single-bin search (PRN 1 at 100 kHz, every trial one
    antenna row of one plan) through estimate_cn0 with the system's table; trials whose ratio
    falls outside the table are dropped.  Returns (median, estimates used, detection rate).  The
    noise is numpy's, not the reference's xorshift stream."""
    name, P, _, _ = SPECS[system]
    if trials is None:
        trials = E2E[system][1 if flip else 0]
    rng = np.random.default_rng(seed if seed is not None else (P + 7 * int(flip)))
    rows, want = flip_rows(system, cn0_db, trials, flip, rng)
    plan = api.Plan(FS, trials, 2 * P, [api.SystemSearch(system, [F_CARRIER], [1])],
                    input_format=1)
    pk = plan.search(rows)[0][:, 0]
    plan.close()
    r_a = (pk["best_peak"].astype(np.float64) / pk["second_peak"].astype(np.float64)) ** 2
    ests = [e for e in (api.estimate_cn0(float(r), name) for r in r_a) if e is not None]
    found = float(np.mean(pk["best_codephase"].astype(np.int64) % P == want))
    return (float(np.median(ests)) if ests else None), len(ests), found


def generate(system, trials=100, seed=2026, step=0.5):
    """[(r_A, C/N0)] over the reference's C/N0 grid for `system`."""
    name, P, lo, hi = SPECS[system]
    rng = np.random.default_rng(seed)
    out, prev, plan = [], 0.0, None
    cn0 = lo
    while cn0 <= hi + 1e-9:
        r_a, _, plan = ratio(system, cn0, trials, rng, plan)
        if r_a <= prev:  # strict monotonicity, gen_acr_tables.rs:170-175
            r_a = prev + 0.001
        prev = r_a
        out.append((r_a, cn0))
        cn0 += step
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("system", choices=sorted(SPECS))
    ap.add_argument("--trials", type=int, default=100)
    ap.add_argument("--seed", type=int, default=2026)
    args = ap.parse_args(argv)
    name = SPECS[args.system][0]
    print(f"// {name} ({args.trials} trials per point, batched GPU search)")
    print(f"pub static {name}: &[(f64, f64)] = &[")
    for r_a, cn0 in generate(args.system, args.trials, args.seed):
        print(f"    ({r_a:.4f}, {cn0:.1f}),")
    print("];")
    return 0


if __name__ == "__main__":
    sys.exit(main())
