"""GPU tests of the FP16 precision form of the fused P = 16368 path (an extension with no
counterpart in the reference): f16 spectra, radix passes as f16 GEMMs on the matrix cores with
f32 accumulation, f64 carrier phase.  It is checked against the oracle's FP64 truth
(oracle/correlate.py mode="fp64", exact phase, no f32 rounding anywhere):

  * best_peak and second_peak within FP16_RTOL relative of the FP64 values;
  * bin and code phase identical unless the FP64 candidates are within FP16_TIE of each other;
  * every injected signal detected at the same bin and code phase as the FP32 form.
"""
import numpy as np
import pytest

import tga
from tga import synth
from oracle import codes as ocodes
from oracle import correlate as oc

pytestmark = pytest.mark.gpu

FS = 16.368e6
IF = 4.092e6
P = 16368
FP16_RTOL = 1e-3
FP16_TIE = 2e-3


def fp64_truth(levels, fc, code):
    lv = np.asarray(levels, dtype=np.float64)
    sig = lv - lv.sum() / len(lv)
    N = 2 * P
    pp = oc.phasepoints(FS, N)
    cc, ca = oc.code_spectra(code, P, mode="fp64")
    ref = oc.correlate_code(sig[:N], pp, cc, ca, fc, N, P, FS, mode="fp64")
    # nearest competitor of the winner, for the tie rule
    vals = sorted((v for _, v in ref["rows"]), reverse=True)
    bin_tie = len(vals) > 1 and vals[1] >= vals[0] * (1 - FP16_TIE)
    return ref, bin_tie


def run(fmt, n_ant, signals, systems, seed, precision):
    L = 163680
    packed = synth.synth_observation(n_ant, L, FS, signals, fmt=fmt, seed=seed)
    plan = tga.Plan(FS, n_ant, L, [tga.SystemSearch(n, fc, prns) for n, prns, fc in systems],
                    input_format=fmt, precision=precision)
    return plan, plan.search(packed), packed, L


SIGNALS = [dict(system="gps", prn=4, cn0=48.0, doppler=2300.0, delay=1234, phase=0.3),
           dict(system="gps", prn=19, cn0=42.0, doppler=-4100.0, delay=16000, phase=1.1),
           dict(system="sbas", prn=133, cn0=46.0, doppler=900.0, delay=77, phase=2.2),
           dict(system="qzss", prn=195, cn0=50.0, doppler=-600.0, delay=8000, phase=0.0)]
FC = [IF - 5000.0 + 500.0 * i for i in range(21)]
SYSTEMS = [("gps", [4, 19, 7, 30], FC), ("sbas", [133, 140], FC), ("qzss", [195], FC)]


@pytest.mark.parametrize("fmt", [1, 2])
def test_fp16_matches_fp64_truth(gpu, fmt):
    plan, peaks, packed, L = run(fmt, 2, SIGNALS, SYSTEMS, 91 + fmt, "fp16")
    assert "fused16368-fp16" in plan.describe()
    levels = synth.unpack_levels(packed, L, fmt)
    worst = 0.0
    for si, (name, prns, fc) in enumerate(SYSTEMS):
        for a in range(2):
            for j, prn in enumerate(prns):
                code = ocodes.code_resampled(name, prn, float(P), 2 * P)
                ref, bin_tie = fp64_truth(levels[a], fc, code)
                got = peaks[si][a, j]
                rel = abs(float(got["best_peak"]) - ref["best_peak"]) / ref["best_peak"]
                rel2 = abs(float(got["second_peak"]) - ref["second_peak"]) / ref["second_peak"]
                worst = max(worst, rel)
                what = f"{name} {prn} ant {a}: {got} vs {ref['best_peak']}, {ref['second_peak']}"
                assert rel <= FP16_RTOL, what
                if not bin_tie:
                    assert int(got["best_freq_idx"]) == ref["best_freq_idx"], what
                    assert rel2 <= FP16_RTOL, what
    print(f"fp16 worst best_peak relative error vs fp64: {worst:.2e}")


def test_fp16_detections_equal_fp32(gpu):
    _, p16, _, _ = run(2, 3, SIGNALS, SYSTEMS, 17, "fp16")
    _, p32, _, _ = run(2, 3, SIGNALS, SYSTEMS, 17, "fp32")
    sysidx = {n: i for i, (n, _, _) in enumerate(SYSTEMS)}
    for s in SIGNALS:
        si = sysidx[s["system"]]
        j = SYSTEMS[si][1].index(s["prn"])
        for a in range(3):
            g16, g32 = p16[si][a, j], p32[si][a, j]
            assert int(g16["best_freq_idx"]) == int(g32["best_freq_idx"])
            assert int(g16["best_codephase"]) == int(g32["best_codephase"])
            # 16 samples per chip: noise moves a weak signal's argmax along the peak plateau
            d = abs(int(g16["best_codephase"]) - (P - s["delay"]) % P)
            assert min(d, P - d) <= 2
            assert abs(float(g16["best_peak"]) / float(g32["best_peak"]) - 1) < FP16_RTOL


def test_fp16_strong_signal_stays_in_range(gpu):
    # 65 dB-Hz on 2-bit samples: the largest partial sums of the passes stay inside f16
    sig = [dict(system="gps", prn=11, cn0=65.0, doppler=1000.0, delay=4321, phase=0.5)]
    systems = [("gps", [11, 12], FC)]
    _, p16, _, _ = run(2, 1, sig, systems, 5, "fp16")
    _, p32, _, _ = run(2, 1, sig, systems, 5, "fp32")
    g16, g32 = p16[0][0, 0], p32[0][0, 0]
    assert int(g16["best_codephase"]) == (P - 4321) % P
    assert int(g16["best_freq_idx"]) == int(g32["best_freq_idx"])
    assert abs(float(g16["best_peak"]) / float(g32["best_peak"]) - 1) < FP16_RTOL


def test_fp16_precision_errors(gpu):
    with pytest.raises(tga.TgaError):
        tga.Plan(FS, 1, 2 * P, [tga.SystemSearch("gps", FC, [1])], precision=7)
    # a group the fused kernels do not serve runs the FP32 form
    plan = tga.Plan(FS, 1, P + 100, [tga.SystemSearch("gps", FC, [1])], precision="fp16")
    assert "fp16" not in plan.describe()
