"""Pins the oracle's correlate_code (CPU).

Independent checks: numpy FFT against a direct DFT matrix; the FFT correlation against an
O(N^2) circular correlation written from the definition; tie rules on constructed rows; the
reference's recovery identity best_codephase % P == (P - delay) % P (testutil.rs:45-47,
correlate.rs:213-274, acquisition.rs:425-449); the contrast test (acquisition.rs:451-471);
and the two-period folding identity the GPU path relies on.
"""
import math

import numpy as np
import pytest

from oracle import codes
from oracle import correlate as oc


def xorshift_uniform(seed, n):
    out = np.empty(n)
    x = seed
    for i in range(n):
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        out[i] = (x >> 11) / float(1 << 53)
    return out


def synth(code, period, delay, f0, fs, noise, seed):
    """testutil.rs:21-40 (vectorised; same xorshift stream)."""
    n = len(code)
    i = np.arange(n)
    u = xorshift_uniform(seed, n)
    return (code[(i + delay) % period] * np.cos(2 * math.pi * f0 * i / fs)
            + noise * (2 * u - 1)).astype(np.float32)


@pytest.mark.parametrize("n", [1, 2, 7, 31, 64, 341, 528, 1023])
def test_fft_matches_direct_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    W = np.exp(-2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
    assert np.allclose(np.fft.fft(x), W @ x, rtol=1e-10, atol=1e-9)


def brute_corr(iq, code):
    """|y[m]| = N |sum_n iq[(n + m) mod N] * code[n]|: the unnormalised forward and inverse
    transforms of the reference each carry a factor N^(1/2) relative to the unitary DFT."""
    N = len(iq)
    return N * np.array([abs(np.dot(np.roll(iq, -m), code)) for m in range(N)])


@pytest.mark.parametrize("N,P", [(64, 32), (62, 31), (90, 45), (33, 33), (40, 16)])
def test_correlate_matches_brute_force(N, P):
    rng = np.random.default_rng(N * 100 + P)
    fs = 1.0e6
    sig = rng.standard_normal(N).astype(np.float32)
    pp = oc.phasepoints(fs, N)
    code = rng.choice([-1.0, 1.0], size=N)
    cc, ca = oc.code_spectra(code, P)
    fc = [0.0, 1234.5, -7000.0]
    res = oc.correlate_code(sig, pp, cc, ca, fc, N, P, fs)
    alt = code.copy()
    alt[min(P, N):] *= -1
    scale = 1.0 / math.sqrt(N)
    best = None
    for fi, f in enumerate(fc):
        iq = oc.wipe(sig, pp, f, N)
        rows = [brute_corr(iq, code.astype(np.float32).astype(np.float64)) * scale,
                brute_corr(iq, alt.astype(np.float32).astype(np.float64)) * scale]
        v0, v1 = rows[0].max(), rows[1].max()
        h = 1 if v1 > v0 * (1 + 1e-9) else 0
        v = max(v0, v1) if h else v0
        if best is None or v >= best[1] * (1 - 1e-9):
            best = (fi, v, rows[h])
    assert res["best_freq_idx"] == best[0]
    assert abs(res["best_peak"] - best[1]) <= 1e-5 * best[1]
    row = best[2]
    assert res["best_codephase"] == int(np.flatnonzero(row >= row.max() * (1 - 1e-9))[-1])


def test_tie_rules():
    # in-row: last maximum; across bins: last maximum; hypotheses: [c,c] wins ties
    v = np.array([1.0, 3.0, 2.0, 3.0, 0.5], dtype=np.float32)
    assert oc.last_argmax(v)[0] == 3
    rows = [(1, np.float32(2.0), None), (0, np.float32(2.0), None), (4, np.float32(1.0), None)]
    assert oc.select_bins(rows) == 1
    with pytest.raises(oc.OraclePanic):
        oc.last_argmax(np.array([1.0, np.nan], dtype=np.float32))
    with pytest.raises(oc.OraclePanic):
        oc.correlate_code(np.zeros(4, np.float32), np.zeros(4), np.ones(4), np.ones(4), [], 4,
                          4, 1e6)


def test_second_peak_mask():
    P = 100
    corr = np.zeros(200, dtype=np.float32)
    corr[10] = 5.0   # main peak
    corr[27] = 4.0   # ring distance 17 > cw=16: eligible
    corr[26] = 4.5   # distance 16: masked (strict >)
    corr[199] = 4.7  # index 199 -> 99 mod P, distance 11: masked
    assert oc.second_peak(corr, 10, P, 16.368e6) == np.float32(4.0)
    # nothing eligible -> 1e-6 floor
    assert oc.second_peak(np.zeros(10, np.float32), 0, 10, 16.368e6) == 1e-6


@pytest.mark.parametrize("delay", [0, 50, 128, 200])
def test_recovers_delay(delay):
    """correlate.rs:213-240."""
    period = 256
    code = codes.resample(codes.gps_ca(1), float(period), period)
    fs = 1.023e6
    sig = synth(code, period, delay, 0.0, fs, 0.05, 42)
    cc, _ = oc.code_spectra(code, period)
    r = oc.correlate_code(sig, oc.phasepoints(fs, period), cc, cc, [0.0], period, period, fs)
    assert r["best_codephase"] % period == (period - delay) % period
    assert r["best_freq_idx"] == 0 and r["best_peak"] > 0


def test_acquire_full_recovers_delay_and_doppler():
    """acquisition.rs:425-449."""
    fs = 2.046e6
    period = int(fs / 1000.0)
    code = codes.resample(codes.gps_ca(1), float(period), 2 * period)
    sig = synth(code, period, 500, 1200.0, fs, 0.05, 1)
    r = oc.acquire_single("gps", sig, oc.phasepoints(fs, len(sig)), fs, 0.0, 3000.0, 1, period)
    assert round(r["codephase_frac"] * period) == (period - 500) % period
    assert abs(abs(r["frequency"]) - 1200.0) <= 250.0


def test_contrast():
    """acquisition.rs:451-471."""
    fs = 2.046e6
    period = int(fs / 1000.0)
    code = codes.resample(codes.gps_ca(1), float(period), 2 * period)
    n = len(code)
    inj = synth(code, period, 123, 0.0, fs, 0.05, 2)
    noise = (np.sin(np.arange(n) * 3.3) * 0.05).astype(np.float32)
    pp = oc.phasepoints(fs, n)
    a = oc.acquire_single("gps", inj, pp, fs, 0.0, 3000.0, 1, period)
    b = oc.acquire_single("gps", noise, pp, fs, 0.0, 3000.0, 1, period)
    assert a["signal_strength"] > 10 * b["signal_strength"]


@pytest.mark.parametrize("N,P", [(64, 32), (62, 31), (33, 33)])
def test_surfaces_match_brute_force(N, P):
    # oracle.surfaces is the array the GPU's surface form is checked against: every row, both
    # hypotheses, phases 0..P-1, against the direct circular correlation
    rng = np.random.default_rng(N + 7 * P)
    fs = 1.0e6
    sig = rng.standard_normal(N)
    pp = oc.phasepoints(fs, N)
    code = rng.choice([-1.0, 1.0], size=N)
    cc, ca = oc.code_spectra(code, P, mode="fp64")
    fc = [0.0, 2500.0]
    surf = oc.surfaces(sig, pp, cc, ca, fc, N, P, mode="fp64")
    assert surf.shape == (2, 2, P)
    alt = code.copy()
    alt[min(P, N):] *= -1
    for b, f in enumerate(fc):
        iq = oc.wipe(sig, pp, f, N, "fp64")
        for h, c in enumerate((code, alt)):
            want = brute_corr(iq, c)[:P] / math.sqrt(N)
            assert np.allclose(surf[b, h], want, rtol=1e-10, atol=1e-10), (b, h)
        rows = oc.bin_rows(sig, pp, cc, ca, fc, N, mode="fp64")
        assert rows[b][1] == pytest.approx(surf[b].max(), rel=1e-12)


@pytest.mark.parametrize("flip", [False, True])
def test_folding_identity(flip):
    # The GPU's P-point folded rows must equal the reference's 2P-point rows, up to m -> m mod P.
    fs = 2.046e6
    P = 2046
    code = codes.resample(codes.gps_ca(5), float(P), 2 * P)
    sig = synth(code, P, 333, 700.0, fs, 0.3, 9)
    if flip:
        sig[P + (P - 333) % P:] *= -1
    pp = oc.phasepoints(fs, 2 * P)
    cc, ca = oc.code_spectra(code, P, mode="fp64")
    fc = [0.0, 650.0, 700.0]
    ref = oc.bin_rows(sig, pp, cc, ca, fc, 2 * P, mode="fp64")
    fold = oc.folded_rows(sig, pp, code, fc, P, fs, mode="fp64")
    for bi in range(len(fc)):
        full = []
        for h, rep in enumerate((cc, ca)):
            X = np.fft.fft(oc.wipe(sig, pp, fc[bi], 2 * P, "fp64"))
            y = np.abs(np.fft.ifft(X * rep) * 2 * P) / math.sqrt(2 * P)
            assert np.allclose(y[:P], y[P:], rtol=1e-9, atol=1e-9)
            assert np.allclose(y[:P], fold[bi][h], rtol=1e-9, atol=1e-9)
            full.append(y)
        assert ref[bi][1] == pytest.approx(max(full[0].max(), full[1].max()), rel=1e-12)


def test_noncoherent_one_epoch_is_the_single_period_correlation():
    # E = 1: the non-coherent search of one period equals correlate_code with N = P, where
    # both hypotheses coincide (correlate.rs:53-56)
    P, fs = 1023, 1.023e6
    code = codes.code_resampled("gps", 5, float(P), P)
    rng = np.random.default_rng(3)
    sig = (np.roll(code, -100) + rng.standard_normal(P)).astype(np.float32)
    pp = oc.phasepoints(fs, P)
    fc = [-500.0, 0.0, 500.0]
    cc, ca = oc.code_spectra(code, P)
    ref = oc.correlate_code(sig, pp, cc, ca, fc, P, P, fs)
    nc = oc.noncoherent(sig, pp, code, fc, P, 1, fs)
    assert nc["best_freq_idx"] == ref["best_freq_idx"]
    assert nc["best_codephase"] == ref["best_codephase"]
    assert nc["best_peak"] == pytest.approx(float(ref["best_peak"]), rel=1e-6)
    assert nc["second_peak"] == pytest.approx(float(ref["second_peak"]), rel=1e-6)


def test_noncoherent_matches_brute_force():
    # E = 3 epochs of P = 64 against sum_e |sum_n x_e[n] c[(n - m) mod P]|^2 written out; the
    # unnormalised inverse transform carries a factor P, so the strength is sqrt(P S)
    P, E, fs = 64, 3, 64000.0
    rng = np.random.default_rng(11)
    code = np.where(rng.random(P) < 0.5, -1.0, 1.0)
    sig = rng.standard_normal(E * P).astype(np.float32)
    pp = oc.phasepoints(fs, E * P)
    f = 1000.0
    x = oc.wipe(sig, pp, f, E * P, "fp64").reshape(E, P)
    S = np.zeros(P)
    for m in range(P):
        for e in range(E):
            S[m] += abs(sum(x[e, n] * code[(n - m) % P] for n in range(P))) ** 2
    nc = oc.noncoherent(sig, pp, code, [f], P, E, fs, mode="fp64")
    want = np.sqrt(S * P)
    assert nc["best_codephase"] == int(np.flatnonzero(want == want.max())[-1])
    assert nc["best_peak"] == pytest.approx(want.max(), rel=1e-12)


def test_noncoherent_recovers_weak_signal():
    # 10 one-period epochs detect a delay that a single period misses more often
    P, fs, E = 1023, 1.023e6, 10
    code = codes.code_resampled("gps", 9, float(P), P)
    rng = np.random.default_rng(8)
    n = np.arange(E * P)
    sig = (0.35 * code[(n + 300) % P] + rng.standard_normal(E * P)).astype(np.float32)
    pp = oc.phasepoints(fs, E * P)
    nc = oc.noncoherent(sig, pp, code, [0.0], P, E, fs)
    assert nc["best_codephase"] == (P - 300) % P
    assert nc["best_peak"] > 3 * nc["second_peak"] / 2
