"""One plan holding every system of CombinedOutput (lib.rs:31-45) -- what acquire_all_gps,
_galileo, _beidou, _sbas, _l1c and _qzss compute one after another in the reference's --all
mode (main.rs:725-810) -- searched in one batched pass per observation.  At 16.368 MHz the plan
has three execution groups with different paths:

  * GPS + SBAS + QZSS L1 C/A, P = 16368: the fused LDS kernels;
  * Galileo E1-C, P = 65472 = 4 x 16368: the long-period split (sub-rows of 16368);
  * BeiDou B1C + GPS L1C, P = 163680 = 10 x 16368: the split with ten sub-rows per row.

Each group writes its peaks at its own offsets of the shared peak array, runs its own carrier
table / forward spectra, and the split groups' chunked pre-pass and transform launches join the
fused group's launches on the plan's streams.  Checked here:

  * 40 ms records (N = 2P for every system), 1- and 2-bit input: every (antenna, PRN) of every
    system against the oracle (acquire_all_* semantics: de-mean over the whole record,
    truncate to N, correlate_code), the parity bar of tests/parity.py;
  * the FP16 precision form of the same plan against the oracle's FP64 truth (best peaks within
    1e-3, same bin and code phase outside near-ties, detections equal to the FP32 form);
  * hipGraph replays of three different observations in mixed order, bit-identical to eager
    searches, through host and device input;
  * a stream of three 10 ms 2-bit observations with three CW tones per antenna resident in device
    memory: search_device of each (C/A folded, E1-C folded, B1C/L1C unfolded N = P rows) against
    the oracle, and rfi_device at L = 163680 against oracle/rfi.py.
"""
import numpy as np
import pytest
import torch

import tga
from tga import synth
from oracle import codes as ocodes
from oracle import correlate as oc
from oracle import rfi as orfi
from parity import assert_peak_matches, near_ties, oracle_search

pytestmark = pytest.mark.gpu

FS = 16.368e6
IF = 4.092e6
PERIOD = {"gps": 16368, "sbas": 16368, "qzss": 16368, "galileo": 65472, "beidou": 163680,
          "l1c": 163680}
# bin spacing of each coherent length (first sinc null 500 / 125 / 50 Hz off at N = 2P)
STEP = {"gps": 500.0, "sbas": 500.0, "qzss": 500.0, "galileo": 125.0, "beidou": 50.0,
        "l1c": 50.0}
PRNS = {"gps": [3, 17, 25], "sbas": [131, 138], "qzss": [194, 195], "galileo": [11, 27],
        "beidou": [19, 30], "l1c": [7, 40]}
ORDER = ["gps", "galileo", "beidou", "sbas", "l1c", "qzss"]  # CombinedOutput order


def fc_of(system):
    s = STEP[system]
    return [IF - s, IF, IF + s]


def signals(seed):
    """One injected SV per system (its first PRN) at 46-52 dB-Hz, Doppler on the grid."""
    rng = np.random.default_rng(seed)
    out = []
    for s in ORDER:
        out.append(dict(system=s, prn=PRNS[s][0], cn0=float(rng.uniform(46.0, 52.0)),
                        doppler=float(rng.choice([-1.0, 0.0, 1.0])) * STEP[s],
                        delay=int(rng.integers(0, PERIOD[s])), phase=float(rng.uniform(0, 6.28)),
                        az=float(rng.uniform(0, 360)), el=float(rng.uniform(20, 90))))
    return out


def make_plan(n_ant, L, fmt, precision="fp32"):
    return tga.Plan(FS, n_ant, L, [tga.SystemSearch(s, fc_of(s), PRNS[s]) for s in ORDER],
                    input_format=fmt, precision=precision)


def check_all_against_oracle(plan, peaks, packed, L, fmt):
    levels = synth.unpack_levels(packed, L, fmt)
    n = 0
    for si, s in enumerate(ORDER):
        P = PERIOD[s]
        N = min(2 * P, L)
        fc = fc_of(s)
        for j, prn in enumerate(PRNS[s]):
            code = ocodes.code_resampled(s, prn, float(P), N)
            for a in range(plan.n_ant):
                ref, sig, pp, cc, ca = oracle_search(levels[a], FS, P, N, fc, code)
                assert_peak_matches(peaks[si][a, j], ref, P, near_ties(sig, pp, cc, ca, fc, N, P),
                                    what=f"{s} {prn} ant {a}")
                n += 1
    return n


# code-phase tolerance of a detection (samples): the correlation top of a 16-sample C/A chip
# over 2 ms is flat to within the noise at 46 dB-Hz; the pilots integrate 8-20 ms
TOL = {"gps": 4, "sbas": 4, "qzss": 4, "galileo": 2, "beidou": 1, "l1c": 1}


def assert_detected(peaks, sigs, n_ant):
    for s in sigs:
        si = ORDER.index(s["system"])
        P = PERIOD[s["system"]]
        for a in range(n_ant):
            pk = peaks[si][a, 0]
            d = (int(pk["best_codephase"]) - (P - s["delay"])) % P
            assert min(d, P - d) <= TOL[s["system"]], (s, a, pk)
            assert fc_of(s["system"])[int(pk["best_freq_idx"])] - IF == s["doppler"], (s, a)


@pytest.mark.parametrize("fmt", [1, 2])
def test_all_systems_one_plan_matches_oracle(gpu, fmt):
    L = 4 * 163680  # 40 ms: N = 2P for every system
    sigs = signals(40 + fmt)
    packed = synth.synth_observation(2, L, FS, sigs, fmt=fmt, seed=400 + fmt)
    plan = make_plan(2, L, fmt)
    desc = plan.describe()
    print(desc, flush=True)
    # the split groups run the sub-row form or the four-step form, whichever the autotuner picks
    assert "fused16368 " in desc
    assert "split4x16368 " in desc or "split132x496 " in desc
    assert "split10x16368 " in desc or "split330x496 " in desc
    assert plan.n_peaks == 2 * sum(len(PRNS[s]) for s in ORDER)
    peaks = plan.search(packed)
    # the peak array is laid out system by system, antenna-major inside each system
    assert [p.shape for p in peaks] == [(2, len(PRNS[s])) for s in ORDER]
    assert check_all_against_oracle(plan, peaks, packed, L, fmt) == plan.n_peaks
    assert_detected(peaks, sigs, 2)


FP16_RTOL = 1e-3
FP16_TIE = 2e-3


def test_all_systems_fp16_form_matches_fp64_truth(gpu):
    L = 4 * 163680
    fmt = 2
    sigs = signals(77)
    packed = synth.synth_observation(2, L, FS, sigs, fmt=fmt, seed=77)
    p16 = make_plan(2, L, fmt, "fp16")
    desc = p16.describe()
    for path in ("fused16368-fp16", "split4x16368-fp16", "split10x16368-fp16"):
        assert path in desc
    got16 = p16.search(packed)
    got32 = make_plan(2, L, fmt).search(packed)
    levels = synth.unpack_levels(packed, L, fmt)
    worst = {}
    for si, s in enumerate(ORDER):
        P = PERIOD[s]
        N = 2 * P
        fc = fc_of(s)
        for j, prn in enumerate(PRNS[s]):
            code = ocodes.code_resampled(s, prn, float(P), N)
            cc, ca = oc.code_spectra(code, P, mode="fp64")
            pp = oc.phasepoints(FS, N)
            for a in range(2):
                lv = levels[a].astype(np.float64)
                sig = lv - lv.sum() / len(lv)
                ref = oc.correlate_code(sig[:N], pp, cc, ca, fc, N, P, FS, mode="fp64")
                vals = sorted((v for _, v in ref["rows"]), reverse=True)
                tie = len(vals) > 1 and vals[1] >= vals[0] * (1 - FP16_TIE)
                g = got16[si][a, j]
                rel = abs(float(g["best_peak"]) - ref["best_peak"]) / ref["best_peak"]
                worst[s] = max(worst.get(s, 0.0), rel)
                what = f"fp16 {s} {prn} ant {a}: {g} vs {ref['best_peak']}"
                assert rel <= FP16_RTOL, what
                if not tie:
                    assert int(g["best_freq_idx"]) == ref["best_freq_idx"], what
    print("fp16 worst best_peak relative error vs fp64 per system:",
          {k: f"{v:.2e}" for k, v in worst.items()})
    # detections of the injected SVs: identical to the FP32 form and to the truth
    assert_detected(got16, sigs, 2)
    assert_detected(got32, sigs, 2)


def test_all_systems_graph_replays_match_eager(gpu):
    L = 4 * 163680
    obs = [synth.synth_observation(1, L, FS, signals(90 + k), fmt=1, seed=900 + k)
           for k in range(3)]
    plan = make_plan(1, L, 1)
    assert "graph=pending" in plan.describe()
    plan.set_graph(False)
    want = [plan.search(o) for o in obs]
    plan.set_graph(True)
    same = lambda a, b: all(np.array_equal(x[f], y[f]) for x, y in zip(a, b)
                            for f in x.dtype.names)
    assert not same(want[0], want[1])
    for k, i in enumerate([0, 1, 2, 0, 2, 1]):
        assert same(plan.search(obs[i]), want[i]), (k, i)
    assert "graph=replay launches=6" in plan.describe()
    dev = [torch.from_numpy(o).cuda() for o in obs]
    torch.cuda.synchronize()
    for i in (2, 0, 1):
        assert same(plan.search_device(dev[i].data_ptr()), want[i]), ("device", i)


def test_stream_of_10ms_observations_with_tones(gpu):
    # 10 ms records: C/A and E1-C keep N = 2P (folded), B1C / L1C have N = P = L (unfolded rows,
    # single hypothesis, correlate.rs:53-56); three CW tones per antenna at -10..+10 dB
    L, n_ant, fmt = 163680, 2, 2
    rng = np.random.default_rng(5)
    stream = []
    for k in range(3):
        tones = [[(float(rng.uniform(0.3e6, 7.9e6)),
                   float(np.sqrt(2.0 * 10.0 ** (rng.uniform(-10.0, 10.0) / 10.0))))
                  for _ in range(3)] for _ in range(n_ant)]
        sigs = signals(500 + k)
        stream.append((sigs, tones, synth.synth_observation(n_ant, L, FS, sigs, fmt=fmt,
                                                            seed=600 + k, tones=tones)))
    plan = tga.Plan(FS, n_ant, L, [tga.SystemSearch(s, fc_of(s), PRNS[s]) for s in ORDER],
                    input_format=fmt)
    desc = plan.describe()
    assert ("split10x16368" in desc or "split330x496" in desc) and "N=163680" in desc
    dev = torch.from_numpy(np.concatenate([o for _, _, o in stream])).cuda()
    torch.cuda.synchronize()
    row_bytes = stream[0][2].shape[1]
    found = total = 0
    for k, (sigs, tones, packed) in enumerate(stream):
        ptr = dev.data_ptr() + k * n_ant * row_bytes
        peaks = plan.search_device(ptr)
        print(f"observation {k}: searched", flush=True)
        check_all_against_oracle(plan, peaks, packed, L, fmt)
        chans = tga.rfi_device(ptr, n_ant, row_bytes, L, fmt, FS)
        levels = synth.unpack_levels(packed, L, fmt)
        for a in range(n_ant):
            ref = orfi.analyze_channel(a, levels[a], FS)
            got = chans[a]
            assert got["dead"] == ref["dead"]
            assert [l["frequency"] for l in got["lines"]] == [l["frequency"] for l in ref["lines"]]
            for g, r in zip(got["lines"], ref["lines"]):
                assert g["excess_db"] == pytest.approx(r["excess_db"], rel=1e-9)
            for f in ("acf_lag1", "acf_lag3", "acf_lag16", "period_samples"):
                assert (got[f] is None) == (ref[f] is None)
                if ref[f] is not None:
                    assert got[f] == pytest.approx(ref[f], rel=1e-9, abs=1e-12)
            for f, _ in tones[a]:
                total += 1
                found += any(abs(l["frequency"] - f) <= 10e3 for l in got["lines"])
        assert tga.rfi_report_json(chans) == orfi.report_json(chans)
    print(f"CW tones found as report lines: {found}/{total}")
    assert found >= total // 2
