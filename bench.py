"""Benchmark: batched code-phase/Doppler search throughput on one MI355X.

Default workload (``--workload all24``): all six systems in one plan, 24 antennas, 2-bit real
samples at 16.368 MHz, 40 ms records (654720 samples); seeded signals: 8 GPS, 2 SBAS and 1 QZSS
at C/N0 uniform in [38, 50] dB-Hz and 6 each of Galileo E1-C, BeiDou B1C and GPS L1C at
[40, 48] dB-Hz, per-SV Doppler and per-antenna carrier offsets from a 3 m disk; search GPS +
SBAS + QZSS L1 C/A 38 + 39 + 23 PRNs x 41 bins (+-10 kHz, 500 Hz), E1-C 50 SVs x 81 bins
(+-5 kHz, 125 Hz), B1C and L1C 63 SVs x 201 bins each (+-5 kHz, 50 Hz), N = 2P (reference
form, two hypotheses): 1.0746e11 cells per search, FP32, plus ACR C/N0 on the host.  The FP16
matrix-core form (f16 spectra or f16 sub-rows, radix passes as f16 GEMMs, f64 carrier phase)
then runs the same observation and is reported under extra.fp16_mfma with its difference to
the FP32 peaks.  extra.mfma (and extra.fp16_mfma.mfma) gives the matrix cores' achieved rate in
the rows stage: the plan's MFMA FLOP per search over the stage's mean time, against the dense
f32 / f16 peaks.

Other workloads (same JSON line):
  ca24     GPS + SBAS + QZSS L1 C/A only (the C/A part of all24 at 10 ms: 1.611e9 cells).
  ca1      GPS L1 C/A, 1 antenna, 1-bit, 10 ms: 4 signals at 45 dB-Hz, PRN 1-32 x 21 bins
           (+-5 kHz, 500 Hz).
  ca1nc    ca1 with 1 ms coherent x 10 non-coherent epochs (the non-coherent form; same cells).
  pilots24 Galileo E1-C (P = 65472), BeiDou B1C and GPS L1C (P = 163680), 24 antennas, 1-bit,
           40 ms: 6 signals per system at 40-48 dB-Hz; E1-C 50 SVs x 81 bins (125 Hz), B1C and
           L1C 63 SVs x 201 bins (50 Hz), N = 2P.
  stream64 64 independent 24-antenna observations (2-bit, 10 ms) resident in HBM, each a random
           4-12 SV mix from all six systems at 35-50 dB-Hz plus 3 CW tones per antenna at
           -10..+10 dB; step k searches observation k mod 64 (all systems, the grids of all24),
           computes its ACR C/N0 and its RFI report; extra.rfi counts the tones found as report
           lines.  About 1 s per step: use --steps 16 for a short run.

A step = one tga_search_device call on an observation already resident in HBM: decode +
de-mean, forward spectra, all correlation rows, selection, copy of the peaks to the host, and
the f64 ACR lookups.  After the warm-up every step replays the plan's search as one hipGraph
whose timed stages carry event-record nodes (the kernel times of the roofline).  Metric: search
cells/s = sum(n_ant * n_prn * n_bin * P) / t.  The roofline block covers the rows stage of the
search (every correlation-row kernel of every group, HIP events around the stage on the plan's
stream); its algorithmic bytes are 2 M S_c per row (M = P folded, S_c = 8 B).

Prints one JSON line (see the repository's measurement contract).
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "tart-gnss_amd"))
sys.path.insert(0, REPO)

METRIC = "search cells/s (ant×PRN×Doppler×phase): FP16-MFMA vs FP32 GPU vs CPU"
FS = 16.368e6
IF = 4.092e6
HBM_PEAK_GBS = 8000.0
# MI355X dense matrix-core peaks: f32 (v_mfma_f32_16x16x4f32) and f16 (v_mfma_f32_16x16x32_f16)
MFMA_PEAK_TFLOPS = {"f32": 157.3, "f16": 2500.0}


def mfma_block(plan, stage_s):
    """Matrix-core rate of the rows stage: the plan's MFMA FLOP per search (counted from the
    kernels its groups run) over the stage's measured mean time, per MFMA type."""
    f32, f16 = plan.rows_matrix_flops()
    out = {}
    for kind, flops in (("f32", f32), ("f16", f16)):
        if flops > 0:
            rate = flops / stage_s / 1e12
            out[kind] = {"flops_per_launch": flops, "achieved_tflops": rate,
                         "peak_tflops": MFMA_PEAK_TFLOPS[kind],
                         "frac": rate / MFMA_PEAK_TFLOPS[kind]}
    return out
PERIOD_S = {"gps": 0.001, "sbas": 0.001, "qzss": 0.001, "galileo": 0.004, "beidou": 0.010,
            "l1c": 0.010}
PRNS = {"gps": range(1, 39), "sbas": range(120, 159), "qzss": range(184, 207),
        "galileo": range(1, 51), "beidou": range(1, 64), "l1c": range(1, 64)}


def grid(half, step):
    n = int(round(2 * half / step)) + 1
    return [IF - half + step * i for i in range(n)]


def sv(rng, system, prn, cn0_lo, cn0_hi, dop):
    P = int(FS * PERIOD_S[system])
    return dict(system=system, prn=int(prn), cn0=float(rng.uniform(cn0_lo, cn0_hi)),
                doppler=float(rng.uniform(-dop, dop)), delay=int(rng.integers(0, P)),
                phase=float(rng.uniform(0, 6.28)), az=float(rng.uniform(0, 360)),
                el=float(rng.uniform(15, 90)))


def workload(name, rng):
    """-> dict(desc, n_ant, L, fmt, signals, systems=[(system, prns, fc)])."""
    if name == "ca24":
        sig = [sv(rng, "gps", p, 38, 50, 9000) for p in rng.choice(range(1, 33), 8, replace=False)]
        sig += [sv(rng, "sbas", p, 38, 50, 9000)
                for p in rng.choice(range(120, 159), 2, replace=False)]
        sig += [sv(rng, "qzss", p, 38, 50, 9000)
                for p in rng.choice(range(184, 207), 1, replace=False)]
        fc = grid(10000.0, 500.0)
        return dict(desc="GPS+SBAS+QZSS L1 C/A, 24 antennas, 2-bit, 16.368 MHz, 10 ms; "
                         "100 PRNs x 41 bins (+-10 kHz, 500 Hz), N=2P=32736, FP32 + ACR",
                    n_ant=24, L=163680, fmt=2, signals=sig,
                    systems=[(s, list(PRNS[s]), fc) for s in ("gps", "sbas", "qzss")])
    if name == "ca1":
        sig = [sv(rng, "gps", p, 45, 45, 5000) for p in rng.choice(range(1, 33), 4, replace=False)]
        return dict(desc="GPS L1 C/A, 1 antenna, 1-bit, 16.368 MHz, 10 ms; PRN 1-32 x 21 bins "
                         "(+-5 kHz, 500 Hz), N=2P=32736, FP32 + ACR",
                    n_ant=1, L=163680, fmt=1, signals=sig,
                    systems=[("gps", list(range(1, 33)), grid(5000.0, 500.0))])
    if name == "ca1nc":
        w = workload("ca1", rng)
        w["desc"] = ("GPS L1 C/A, 1 antenna, 1-bit, 16.368 MHz, 10 ms; PRN 1-32 x 21 bins "
                     "(+-5 kHz, 500 Hz), 1 ms coherent x 10 non-coherent, FP32")
        w["nc_epochs"] = 10
        return w
    if name == "pilots24":
        sig = []
        for s, hi in (("galileo", 51), ("beidou", 64), ("l1c", 64)):
            sig += [sv(rng, s, p, 40, 48, 5000) for p in rng.choice(range(1, hi), 6, replace=False)]
        return dict(desc="Galileo E1-C (P=65472) + BeiDou B1C + GPS L1C (P=163680), 24 antennas, "
                         "1-bit, 16.368 MHz, 40 ms; 50 x 81 bins (125 Hz) + 2 x 63 x 201 bins "
                         "(50 Hz), N=2P, FP32 + ACR",
                    n_ant=24, L=654720, fmt=1, signals=sig,
                    systems=[("galileo", list(PRNS["galileo"]), grid(5000.0, 125.0)),
                             ("beidou", list(PRNS["beidou"]), grid(5000.0, 50.0)),
                             ("l1c", list(PRNS["l1c"]), grid(5000.0, 50.0))])
    if name == "stream64":
        # 64 independent observations, each a random 4-12 SV mix from every system at 35-50
        # dB-Hz and 3 CW tones per antenna at -10..+10 dB against the noise
        pool = [(s, p) for s in ("gps", "sbas", "qzss", "galileo", "beidou", "l1c")
                for p in PRNS[s]]
        stream = []
        for _ in range(64):
            pick = rng.choice(len(pool), int(rng.integers(4, 13)), replace=False)
            sig = [sv(rng, pool[i][0], pool[i][1], 35, 50, 5000) for i in pick]
            tones = [[(float(rng.uniform(0.2e6, 8.0e6)),
                       math.sqrt(2.0 * 10.0 ** (rng.uniform(-10.0, 10.0) / 10.0)))
                      for _ in range(3)] for _ in range(24)]
            stream.append((sig, tones))
        return dict(desc="stream of 64 observations: all six systems (C/A N=2P, E1-C N=2P, "
                         "B1C/L1C N=P), 24 antennas, 2-bit, 16.368 MHz, 10 ms, 4-12 SVs at "
                         "35-50 dB-Hz and 3 CW tones per antenna; C/A +-10 kHz / 500 Hz, E1-C "
                         "+-5 kHz / 125 Hz, B1C+L1C +-5 kHz / 50 Hz, FP32 + ACR + RFI report per "
                         "observation",
                    n_ant=24, L=163680, fmt=2, signals=stream[0][0], stream=stream,
                    systems=all_systems())
    if name == "all24":
        sig = [sv(rng, "gps", p, 38, 50, 5000) for p in rng.choice(range(1, 33), 8, replace=False)]
        sig += [sv(rng, "sbas", p, 38, 50, 5000)
                for p in rng.choice(range(120, 159), 2, replace=False)]
        sig += [sv(rng, "qzss", p, 38, 50, 5000)
                for p in rng.choice(range(184, 207), 1, replace=False)]
        for s, hi in (("galileo", 51), ("beidou", 64), ("l1c", 64)):
            sig += [sv(rng, s, p, 40, 48, 5000) for p in rng.choice(range(1, hi), 6, replace=False)]
        return dict(desc="all six systems, 24 antennas, 2-bit, 16.368 MHz, 40 ms; C/A 100 PRNs x "
                         "41 bins (+-10 kHz, 500 Hz), E1-C 50 x 81 (125 Hz), B1C + L1C 2 x 63 x 201 "
                         "(50 Hz), N=2P, FP32 + ACR (FP16 form beside it)",
                    n_ant=24, L=654720, fmt=2, signals=sig, systems=all_systems())
    raise SystemExit(f"unknown workload {name}")


def all_systems():
    """C/A grid of the 24-antenna C/A workload, pilot grids of the pilots workload: bins as fine
    as each coherent length needs (first sinc null 500 / 125 / 100 Hz off)."""
    step = {"gps": 500.0, "sbas": 500.0, "qzss": 500.0, "galileo": 125.0, "beidou": 50.0,
            "l1c": 50.0}
    half = {"gps": 10000.0, "sbas": 10000.0, "qzss": 10000.0, "galileo": 5000.0,
            "beidou": 5000.0, "l1c": 5000.0}
    return [(s, list(PRNS[s]), grid(half[s], step[s]))
            for s in ("gps", "sbas", "qzss", "galileo", "beidou", "l1c")]


class DeviceInput:
    """The observation in device memory, through the HIP runtime the library itself links (the
    first one loaded), so the process holds one runtime and searches replay as graphs with
    event-timed stages."""

    def __init__(self, packed):
        import ctypes as C
        import tga
        tga.lib()
        self._C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self._p = C.c_void_p()
        n = C.c_size_t(packed.nbytes)
        self._check(self.hip.hipMalloc(C.byref(self._p), n), "hipMalloc")
        self._check(self.hip.hipMemcpy(self._p, packed.ctypes.data_as(C.c_void_p), n, 1),
                    "hipMemcpy")  # 1 = hipMemcpyHostToDevice
        self.sync()

    @property
    def ptr(self):
        return self._p.value

    def sync(self):
        self._check(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")

    @staticmethod
    def _check(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")


def surface_errors(dev, L, fmt, wsystems, signals, levels0):
    """Correlation-surface error of the FP16 and FP32 forms against an FP64 truth, per system:
    a one-antenna plan of the workload's systems over antenna 0 of the resident observation
    (per system the first injected PRN and one more, the 3 central bins) through
    tga_plan_surface, and the truth |IFFT(FFT(x e^{j 2 pi f n / fs}) conj(FFT(code)))| / sqrt(N)
    of the de-meaned record in float64 (numpy).  Error of a row: max |diff| / the row's peak;
    the worst row (both hypotheses, code phases 0..P-1) of each system."""
    import ctypes as C
    import tga
    x = levels0.astype(np.float64)
    x -= x.sum() / len(x)
    sub = []
    for s, prns, fc in wsystems:
        inj = [g["prn"] for g in signals if g["system"] == s][:1]
        pr = inj + [p for p in prns if p not in inj][:2 - len(inj)]
        mid = len(fc) // 2
        sub.append((s, pr, fc[max(mid - 1, 0):mid + 2]))
    searches = [tga.SystemSearch(s, fc, pr) for s, pr, fc in sub]
    out = {}
    for prec in ("fp16", "fp32"):
        plan = tga.Plan(FS, 1, L, searches, input_format=fmt, precision=prec)
        n = plan.surface_size()
        buf = C.c_void_p()
        dev._check(dev.hip.hipMalloc(C.byref(buf), C.c_size_t(4 * n)), "hipMalloc")
        try:
            plan.surface_device(dev.ptr, buf.value, n)  # antenna 0 is the first packed row
            flat = np.empty(n, dtype=np.float32)
            dev._check(dev.hip.hipMemcpy(flat.ctypes.data_as(C.c_void_p), buf,
                                         C.c_size_t(4 * n), 2), "hipMemcpy")  # 2 = D2H
        finally:
            dev.hip.hipFree(buf)
        err, off = {}, 0
        for si, (s, pr, fc) in enumerate(sub):
            k = plan.surface_size(si)
            P = k // (len(pr) * len(fc) * 2)
            surf = flat[off:off + k].reshape(len(pr), len(fc), 2, P)
            off += k
            N = min(2 * P, L)
            ph = (2.0 * math.pi / FS) * np.arange(N, dtype=np.float64)
            worst = 0.0
            for b, f in enumerate(fc):
                X = np.fft.fft(x[:N] * np.exp(1j * ph * f))
                for j, prn in enumerate(pr):
                    code = tga.code_resampled(s, prn, float(P), N)
                    alt = code.copy()
                    alt[P:] *= -1.0
                    for h, c in enumerate((code, alt) if N == 2 * P else (code,)):
                        y = np.fft.ifft(X * np.conj(np.fft.fft(c))) * N
                        t = np.abs(y[:P]) / math.sqrt(N)
                        g = surf[j, b, h].astype(np.float64)
                        worst = max(worst, float(np.abs(g - t).max() / t.max()))
            err[s] = worst
        plan.close()
        out[prec] = err
    return out


def load_traffic(path, kernel, workload_name):
    """HBM bytes per launch (per search for a rows stage of several kernels) of `kernel`
    measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE) on this workload, from the committed
    summary; None if absent (the bytes of a launch depend on its rows)."""
    try:
        with open(path) as f:
            d = json.load(f)
        return float(d["workloads"][workload_name]["kernels"][kernel]["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="all24",
                    choices=["ca24", "ca1", "ca1nc", "pilots24", "all24", "stream64"])
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU sample")
    ap.add_argument("--seed", type=int, default=0x5EED_2026 + 2)
    ap.add_argument("--traffic", default=os.path.join(REPO, "profiles", "pmc_traffic.json"))
    ap.add_argument("--no-fp16", action="store_true",
                    help="skip the FP16 matrix-core form (reported beside the FP32 value)")
    ap.add_argument("--no-surface", action="store_true",
                    help="skip the correlation-surface error of the FP16 / FP32 forms")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="first run one elementwise copy of MIB MiB (reads and writes MIB MiB) "
                         "so hardware byte counters can be scaled against a known size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the search "
              "shards by antenna with no collective, see the design notes", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy:
        import torch  # the calibration copy below is a torch kernel; torch loads first
    import tga
    from tga import synth

    global np
    import numpy as np

    assert tga.device_count() > 0, "no GPU visible"
    rng = np.random.default_rng(args.seed)
    w = workload(args.workload, rng)
    n_ant, L, fmt, signals = w["n_ant"], w["L"], w["fmt"], w["signals"]
    stream = w.get("stream")
    t0 = time.perf_counter()
    if stream:  # the observations one after another: [n_obs * n_ant, row_bytes]
        packed = np.concatenate([synth.synth_observation(n_ant, L, FS, sig, fmt=fmt,
                                                         seed=args.seed + k, tones=tones)
                                 for k, (sig, tones) in enumerate(stream)])
    else:
        packed = synth.synth_observation(n_ant, L, FS, signals, fmt=fmt, seed=args.seed)
    t_synth = time.perf_counter() - t0
    n_obs = len(stream) if stream else 1
    obs_bytes = n_ant * packed.shape[1]
    nc = w.get("nc_epochs", 0)
    systems = [tga.SystemSearch(s, fc, prns, nc_epochs=nc) for s, prns, fc in w["systems"]]
    t0 = time.perf_counter()
    plan = tga.Plan(FS, n_ant, L, systems, input_format=fmt)
    t_plan = time.perf_counter() - t0
    periods = [int(FS * PERIOD_S[s]) for s, _, _ in w["systems"]]
    cells = sum(n_ant * len(prns) * len(fc) * P
                for (s, prns, fc), P in zip(w["systems"], periods))
    n_prn = sum(len(prns) for _, prns, _ in w["systems"])

    dev = DeviceInput(packed)
    sync = dev.sync
    ptr = dev.ptr
    if args.calib_copy:
        src = torch.ones(args.calib_copy * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 2.0, out=dst)  # one elementwise kernel: reads and writes MIB MiB
        torch.cuda.synchronize()
        del src, dst

    state = {"k": 0, "rfi": [], "peaks": []}

    def step(pl):
        k = state["k"] % n_obs  # stream: the next observation in HBM
        state["k"] += 1
        p = ptr + k * obs_bytes
        peaks = pl.search_device(p)
        # ACR C/N0 per (antenna, PRN) in f64 on the host (acquisition.rs:341-357), natively
        cn0 = pl.cn0(peaks)
        if stream:  # the narrowband RFI report of the same observation (rfi.rs:142-184)
            state["rfi"].append((k, tga.rfi_device(p, n_ant, packed.shape[1], L, fmt, FS)))
            state["peaks"].append((k, peaks))
        return peaks, int(sum(np.count_nonzero(~np.isnan(c)) for c in cn0))

    def timed(pl):
        """Warm-up, then exactly args.steps timed steps; returns (seconds, last peaks, cn0)."""
        state["k"] = 0
        state["rfi"] = []
        state["peaks"] = []
        for _ in range(args.warmup):
            step(pl)
        pl.set_timing(True)
        sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            pk, n_ok = step(pl)
        sync()
        dt = time.perf_counter() - t0
        return dt, pk, n_ok

    dt, peaks, n_cn0 = timed(plan)
    stream_peaks = state["peaks"][-args.steps:]
    rfi = None
    if stream:  # injected CW tones found as report lines (within 10 kHz), over the timed steps
        hit = tot = dead = 0
        for k, chans in state["rfi"][-args.steps:]:
            for a, ch in enumerate(chans):
                dead += ch["dead"]
                for f, _ in stream[k][1][a]:
                    tot += 1
                    hit += any(abs(ln["frequency"] - f) <= 10e3 for ln in ch["lines"])
        rfi = {"tones_found": f"{hit}/{tot}", "dead_channels": dead,
               "observations": len(state["rfi"][-args.steps:])}
    ms_step = dt * 1000.0 / args.steps
    value = cells * args.steps / dt

    rows_ms, rows_n, rows_bytes = plan.timing(0)
    fwd_ms, fwd_n, _ = plan.timing(1)
    sel_ms, sel_n, _ = plan.timing(2)
    avg_rows_s = rows_ms / 1000.0 / max(rows_n, 1)
    achieved = rows_bytes / avg_rows_s / 1e9
    traffic = load_traffic(args.traffic, plan.kernel_name(0), args.workload)

    fp16 = None
    desc = plan.describe()
    if not args.no_fp16 and not nc and ("fused16368" in desc or "split" in desc):
        # the FP16 matrix-core form on the same observation, compared with the FP32 peaks
        plan16 = tga.Plan(FS, n_ant, L, systems, input_format=fmt, precision="fp16")
        dt16, peaks16, _ = timed(plan16)
        r_ms, r_n, r_bytes = plan16.timing(0)
        avg16 = r_ms / 1000.0 / max(r_n, 1)
        rel, same = [], 0
        for a32, a16 in zip(peaks, peaks16):
            b32 = a32["best_peak"].astype(np.float64)
            rel.append(np.abs(a16["best_peak"].astype(np.float64) - b32) / b32)
            same += int(np.sum((a32["best_codephase"] == a16["best_codephase"]) &
                               (a32["best_freq_idx"] == a16["best_freq_idx"])))
        rel = np.concatenate([r.ravel() for r in rel])
        fp16 = {"value": cells * args.steps / dt16, "unit": "cells/s",
                "ms_per_step": dt16 * 1000.0 / args.steps, "kernel": plan16.kernel_name(0),
                "kernel_ms": avg16 * 1000.0, "alg_bytes_per_launch": r_bytes,
                "achieved_GBs": r_bytes / avg16 / 1e9,
                "frac": r_bytes / avg16 / 1e9 / HBM_PEAK_GBS,
                "speedup_vs_fp32": (cells * args.steps / dt16) / value,
                "best_peak_rel_diff_vs_fp32": {"max": float(rel.max()),
                                               "median": float(np.median(rel))},
                "same_bin_and_codephase_as_fp32": f"{same}/{rel.size}",
                "mfma": mfma_block(plan16, avg16),
                "plan": plan16.describe().splitlines()[0]}

    surface = None
    if not args.no_surface and not nc and not stream:
        errs = surface_errors(dev, L, fmt, w["systems"], signals,
                              synth.unpack_levels(packed[:1], L, fmt)[0])
        surface = {"fp16_vs_fp64": errs["fp16"], "fp32_vs_fp64": errs["fp32"],
                   "measure": "worst row of max |diff| / row peak over code phases 0..P-1",
                   "sample": "antenna 0; per system the first injected PRN and one more, the 3 "
                             "central bins, both hypotheses"}

    # detections: every injected SV found at its code phase on most antennas
    sysidx = {s: i for i, (s, _, _) in enumerate(w["systems"])}

    def detections(sigs, pk):
        n = 0
        for s in sigs:
            si = sysidx[s["system"]]
            P = periods[si]
            j = list(plan.prns[si]).index(s["prn"])
            cp = (P - s["delay"]) % P
            hits = sum(abs(int(pk[si][a, j]["best_codephase"]) - cp) <= 1 for a in range(n_ant))
            n += hits >= n_ant // 2
        return n

    if stream:  # every observation of the timed steps
        found = sum(detections(stream[k][0], pk) for k, pk in stream_peaks)
        signals = [s for k, _ in stream_peaks for s in stream[k][0]]
    else:
        found = detections(signals, peaks)

    cpu = None
    if not args.no_cpu_baseline:
        # each system of the workload on antenna 0 for an equal share of the budget (whole
        # PRNs, at least one), then the workload's own mix projected from the per-system rates:
        # value = total cells / sum over systems of (system cells / system rate)
        from oracle import cpu_port
        workers = min(16, os.cpu_count() or 1)
        levels = synth.unpack_levels(packed[:n_ant], L, fmt)
        share = args.cpu_budget / len(w["systems"])
        per_sys, t_all, cells_all, t_proj = [], 0.0, 0.0, 0.0
        for s, prns, fc in w["systems"]:
            c, t, d = cpu_port.time_sample(levels[0], FS, s, prns, fc, workers, budget_s=share,
                                           epochs=nc)
            rate = c / t
            full = n_ant * len(prns) * len(fc) * int(FS * PERIOD_S[s])
            per_sys.append(f"{s} {d} PRN {t:.1f} s {rate:.3g} cells/s")
            t_all += t
            cells_all += full
            t_proj += full / rate
        # the same port on one thread, over the first system of antenna 0, for a few seconds
        c1, t1, d1 = cpu_port.time_sample(levels[0], FS, w["systems"][0][0], w["systems"][0][1],
                                          w["systems"][0][2], 1, budget_s=args.cpu_budget / 4,
                                          epochs=nc)
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f
                              if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu = {"value": cells_all / t_proj, "unit": "cells/s", "cores": workers, "kind": "port",
               "sample": ("antenna 0, every system for an equal share of "
                          f"{args.cpu_budget:.0f} s ({'; '.join(per_sys)}), "
                          + (f"{nc} one-period epochs per bin (1 replica FFT per PRN, "
                             f"{nc} forward + {nc} inverse FFTs per bin)" if nc else
                             "N=2P, reference work structure (2 replica + 3 FFTs per bin per "
                             "PRN)")
                          + f", complex64 pocketfft over {workers} threads, {t_all:.1f} s; value "
                          "= the workload's cells over the time the per-system rates project"),
               "value_1_core": c1 / t1 if t1 > 0 else None,
               "sample_1_core": f"{d1} PRN searches of {w['systems'][0][0]} on antenna 0, "
                                f"{t1:.1f} s",
               "host": {"nproc": os.cpu_count(), "model": model}}

    out = {
        "metric": METRIC, "value": value, "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w["desc"], "name": args.workload, "n_ant": n_ant,
                   "n_prn": n_prn, "n_bins": [len(fc) for _, _, fc in w["systems"]],
                   "code_periods": periods, "cells": cells,
                   "form": (f"non-coherent ({nc} one-period epochs, powers summed)" if nc
                            else "reference (dual hypothesis, folded P-point rows)"),
                   "plan": plan.describe().strip().splitlines()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": plan.kernel_name(0), "kernel_ms": avg_rows_s * 1000.0,
                     "alg_bytes_per_launch": rows_bytes},
        "cpu_baseline": cpu,
        "extra": {"forward_ms": fwd_ms / max(fwd_n, 1), "select_ms": sel_ms / max(sel_n, 1),
                  "mfma": mfma_block(plan, avg_rows_s),
                  "plan_create_s": t_plan, "synth_s": t_synth, "cn0_valid": n_cn0,
                  "injected_found": f"{found}/{len(signals)}", "rfi": rfi,
                  "observations_per_s": args.steps / dt if stream else None,
                  "graph": plan.describe().strip().splitlines()[-1], "fp16_mfma": fp16,
                  "surface_error": surface},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
