"""Correlation-surface error per system (tga_plan_surface): every row of one six-system plan --
the array correlate.rs:118-146 forms per bin and hypothesis, |IFFT(X . code_complex)| / sqrt(N),
whose maximum the search keeps -- against the oracle, at code phases 0..P-1 (the rows of
N = 2P searches are P-periodic).

Error of a row: max_i |S_gpu[i] - S_ref[i]| / max_i S_ref[i] (relative to the row's peak, the
value best_peak reports), and the worst over every (antenna, PRN, bin, hypothesis) of a system.
  * FP16 form against the FP64 truth (f64 carrier phase, f64 transforms): below 1e-3;
  * FP32 form against the oracle's FP32 arithmetic (the reference's f32 carrier rounding,
    correlate.rs:110-115): below 1e-5, the peak-metric bar of the FP32 path;
  * the winning row's surface holds best_peak at best_codephase mod P, and the surface form's
    search returns the same peaks as the plain search (bit for bit).
"""
import numpy as np
import pytest

import tga
from tga import synth
from oracle import codes as ocodes
from oracle import correlate as oc
from test_gpu_multisystem import FS, ORDER, PERIOD, PRNS, fc_of, make_plan, signals

pytestmark = pytest.mark.gpu

L = 4 * 163680  # 40 ms: N = 2P for every system
N_ANT = 2
BOUND = {"fp16": 1e-3, "fp32": 1e-5}


def truth(levels, system, mode):
    """Oracle surfaces [n_ant, n_prn, n_bin, 2, P] of one system."""
    P = PERIOD[system]
    N = 2 * P
    fc = fc_of(system)
    pp = oc.phasepoints(FS, N)
    out = np.empty((N_ANT, len(PRNS[system]), len(fc), 2, P))
    for a in range(N_ANT):
        lv = levels[a].astype(np.float64)
        sig = lv - lv.sum() / len(lv)  # de-mean over the whole record, then truncate
        sig = sig[:N] if mode == "fp64" else sig[:N].astype(np.float32)
        for j, prn in enumerate(PRNS[system]):
            code = ocodes.code_resampled(system, prn, float(P), N)
            cc, ca = oc.code_spectra(code, P, mode=mode)
            out[a, j] = oc.surfaces(sig, pp, cc, ca, fc, N, P, mode=mode)
    return out


@pytest.mark.parametrize("precision,split", [("fp16", None), ("fp32", None), ("fp32", "5")],
                         ids=["fp16", "fp32", "fp32-half-subrows"])
def test_surface_error_per_system(gpu, precision, split, monkeypatch):
    # split = "5": the FP32 split groups pinned to the half sub-rows (k_halfrows' surface form)
    if split is not None:
        monkeypatch.setenv("TGA_SPLIT_VARIANT", split)
    packed = synth.synth_observation(N_ANT, L, FS, signals(77), fmt=2, seed=77)
    plan = make_plan(N_ANT, L, 2, precision)
    print(plan.describe(), flush=True)
    if split is not None:
        assert plan.describe().count(f"rows_variant={split}") == 3  # E1-C, B1C, L1C
    surf, peaks = plan.surface(packed)
    assert [s.shape for s in surf] == [(N_ANT, len(PRNS[s]), 3, 2, PERIOD[s]) for s in ORDER]
    plain = plan.search(packed)
    for si in range(len(ORDER)):
        for f in peaks[si].dtype.names:
            assert np.array_equal(peaks[si][f], plain[si][f]), (ORDER[si], f)
    levels = synth.unpack_levels(packed, L, 2)
    mode = "fp64" if precision == "fp16" else "fp32"
    worst, rms = {}, {}
    for si, s in enumerate(ORDER):
        ref = truth(levels, s, mode)
        got = surf[si].astype(np.float64)
        peak = ref.max(axis=-1)
        err = np.abs(got - ref).max(axis=-1) / peak
        worst[s] = float(err.max())
        rms[s] = float((np.sqrt(((got - ref) ** 2).mean(axis=-1)) / peak).max())
        # the winning row's surface holds the reported peak at the reported code phase
        P = PERIOD[s]
        for a in range(N_ANT):
            for j in range(len(PRNS[s])):
                pk = peaks[si][a, j]
                b = int(pk["best_freq_idx"])
                row = surf[si][a, j, b]
                h = int(np.argmax(row.max(axis=-1) == np.float32(pk["best_peak"])))
                assert row[h].max() == np.float32(pk["best_peak"]), (s, a, j)
                assert row[h, int(pk["best_codephase"]) % P] == row[h].max(), (s, a, j)
    print(f"{precision} surface error vs {mode} oracle, worst row (max |diff| / row peak):",
          {k: f"{v:.2e}" for k, v in worst.items()})
    print(f"{precision} surface error, worst row RMS / row peak:",
          {k: f"{v:.2e}" for k, v in rms.items()}, flush=True)
    for s in ORDER:
        assert worst[s] < BOUND[precision], (s, worst[s])


def test_surface_rejects_short_buffer_and_unsupported_groups(gpu):
    import ctypes as C
    import torch
    from tga._lib import lib
    packed = synth.synth_observation(1, 2 * 16368, FS, [], fmt=1, seed=3)
    plan = tga.Plan(FS, 1, 2 * 16368, [tga.SystemSearch("gps", [4.092e6], [1])])
    n = plan.surface_size()
    assert n == 1 * 1 * 1 * 2 * 16368
    dev = torch.from_numpy(packed).cuda()
    buf = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(tga.TgaError):
        plan.surface_device(dev.data_ptr(), buf.data_ptr(), n - 1)
    # a generic-path group (P = 2046 at 2.046 MHz) has no surface form
    g = tga.Plan(2.046e6, 1, 2 * 2046, [tga.SystemSearch("gps", [0.0], [1])])
    small = synth.synth_observation(1, 2 * 2046, 2.046e6, [], fmt=1, seed=4)
    with pytest.raises(tga.TgaError):
        g.surface(small)
    assert lib().tga_plan_surface_size(None, -1, C.byref(C.c_size_t())) != 0
