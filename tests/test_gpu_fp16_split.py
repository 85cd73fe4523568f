"""FP16 precision form of the long-period split path (an extension with no counterpart in the
reference): E1-C (P = 4 x 16368) and B1C / L1C (P = 10 x 16368) rows whose first split stage
(conj(X) . R and the R0-point DFT) runs in f32 and whose sub-rows are stored in f16 for the
matrix-core rows kernel; the carrier phase stays f64.  Checked like the fused FP16 form,
against the oracle's FP64 truth (oracle/correlate.py mode="fp64"):

  * best_peak within FP16_RTOL relative of the FP64 value;
  * bin identical unless the FP64 candidates are within FP16_TIE of each other; where the code
    phase is the FP64 one, second_peak (recomputed in f32 on the winning row) within 2e-5 --
    a code phase one sample off moves the masked window, and with it the second peak;
  * injected signals detected at the same bin and code phase as the FP32 split path.
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
FP16_RTOL = 1e-3
FP16_TIE = 2e-3


def grid(center, band, step):
    n = int(round(2 * band / step)) + 1
    return [center - band + i * step for i in range(n)]


def fp64_truth(levels, name, prn, fc, P, N):
    lv = np.asarray(levels, dtype=np.float64)
    sig = lv - lv.sum() / len(lv)
    pp = oc.phasepoints(FS, N)
    code = ocodes.code_resampled(name, prn, float(P), N)
    cc, ca = oc.code_spectra(code, P, mode="fp64")
    ref = oc.correlate_code(sig[:N], pp, cc, ca, fc, N, P, FS, mode="fp64")
    vals = sorted((v for _, v in ref["rows"]), reverse=True)
    return ref, len(vals) > 1 and vals[1] >= vals[0] * (1 - FP16_TIE)


def plans(L, n_ant, fmt, signals, systems, seed):
    packed = synth.synth_observation(n_ant, L, FS, signals, fmt=fmt, seed=seed)
    out = {}
    for prec in ("fp16", "fp32"):
        plan = tga.Plan(FS, n_ant, L, [tga.SystemSearch(n, fc, p) for n, p, fc in systems],
                        input_format=fmt, precision=prec)
        out[prec] = (plan, plan.search(packed))
    return out, packed


CASES = {
    # name: (P, L, n_ant, fmt, signals, systems, seed, path)
    "e1c": (65472, 2 * 65472 + 500, 2, 1,
            [dict(system="galileo", prn=11, cn0=50.0, doppler=500.0, delay=40000),
             dict(system="galileo", prn=27, cn0=47.0, doppler=-1000.0, delay=123)],
            [("galileo", [3, 11, 27], grid(IF, 1000.0, 500.0))], 21, "split4x16368-fp16"),
    "b1c": (163680, 2 * 163680, 1, 2,
            [dict(system="beidou", prn=19, cn0=48.0, doppler=-500.0, delay=99999)],
            [("beidou", [19, 20], grid(IF, 1000.0, 500.0))], 19, "split10x16368-fp16"),
    "l1c": (163680, 2 * 163680, 1, 1,
            [dict(system="l1c", prn=7, cn0=48.0, doppler=500.0, delay=5)],
            [("l1c", [7, 8], grid(IF, 500.0, 500.0))], 7, "split10x16368-fp16"),
    "b1c_unfolded": (163680, 163680, 2, 1,
                     [dict(system="beidou", prn=30, cn0=52.0, doppler=1000.0, delay=12345)],
                     [("beidou", [30, 31], grid(IF, 1000.0, 500.0))], 12, "split10x16368-fp16"),
}


@pytest.mark.parametrize("case", list(CASES))
def test_fp16_split_matches_fp64_truth(gpu, case):
    P, L, n_ant, fmt, signals, systems, seed, path = CASES[case]
    res, packed = plans(L, n_ant, fmt, signals, systems, seed)
    plan16, p16 = res["fp16"]
    assert path in plan16.describe()
    r0 = P // 16368
    assert plan16.kernel_name(0) == f"k_big_pre<{r0},f16>+k_rows16<{r0}>"
    levels = synth.unpack_levels(packed, L, fmt)
    N = min(2 * P, L)
    worst = 0.0
    for si, (name, prns, fc) in enumerate(systems):
        for a in range(n_ant):
            for j, prn in enumerate(prns):
                ref, tie = fp64_truth(levels[a], name, prn, fc, P, N)
                got = p16[si][a, j]
                rel = abs(float(got["best_peak"]) - ref["best_peak"]) / ref["best_peak"]
                rel2 = abs(float(got["second_peak"]) - ref["second_peak"]) / ref["second_peak"]
                worst = max(worst, rel)
                what = f"{case} {name} {prn} ant {a}: {got} vs {ref['best_peak']}"
                assert rel <= FP16_RTOL, what
                if not tie:
                    assert int(got["best_freq_idx"]) == ref["best_freq_idx"], what
                    if int(got["best_codephase"]) == ref["best_codephase"]:
                        assert rel2 <= 2e-5, what
    print(f"{case}: fp16 split worst best_peak relative error vs fp64 {worst:.2e}")

    # detections: the FP32 split path's bin and code phase
    plan32, p32 = res["fp32"]
    sysidx = {n: i for i, (n, _, _) in enumerate(systems)}
    for s in signals:
        si = sysidx[s["system"]]
        j = systems[si][1].index(s["prn"])
        for a in range(n_ant):
            g16, g32 = p16[si][a, j], p32[si][a, j]
            assert int(g16["best_freq_idx"]) == int(g32["best_freq_idx"])
            assert int(g16["best_codephase"]) == int(g32["best_codephase"])
            assert int(g16["best_codephase"]) == (P - s["delay"]) % P
            # against the FP32 form the carrier phase differs too (f64 here, the reference's
            # f32-rounded phi there: about 0.03 rad of rounding at 5e5 rad over 20 ms)
            assert abs(float(g16["best_peak"]) / float(g32["best_peak"]) - 1) < 5e-3
            # both recomputed in f32 on the same winning row; the FP16 form's spectra carry
            # the carrier phase in f64, the FP32 form's the reference's f32-rounded phase
            assert float(g16["second_peak"]) == pytest.approx(float(g32["second_peak"]),
                                                              rel=5e-3)


def test_fp16_split_strong_signal_stays_in_range(gpu):
    # 62 dB-Hz over 20 ms on 2-bit samples: the sub-rows and pass outputs stay inside f16
    sig = [dict(system="beidou", prn=5, cn0=62.0, doppler=0.0, delay=777, phase=0.2)]
    systems = [("beidou", [5], [IF])]
    res, _ = plans(2 * 163680, 1, 2, sig, systems, 3)
    g16, g32 = res["fp16"][1][0][0, 0], res["fp32"][1][0][0, 0]
    assert int(g16["best_codephase"]) == (163680 - 777) % 163680
    assert abs(float(g16["best_peak"]) / float(g32["best_peak"]) - 1) < FP16_RTOL
