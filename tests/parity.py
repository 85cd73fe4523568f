"""Parity helpers shared by the GPU tests: compare a GPU peak with the oracle's result.

Bar: bin index and code phase bit-exact unless the oracle itself shows a near-tie (the top
two candidates within NEAR_TIE relative, where FP32 rounding of any FFT may reorder them);
best_peak and second_peak within PEAK_RTOL relative of the oracle's value.
"""
import numpy as np

from oracle import correlate as oc

PEAK_RTOL = 1e-5
NEAR_TIE = 2e-5


def oracle_search(levels, fs, P, N, fc, code_n):
    """Reference semantics of acquire_all_*: de-mean over the whole record (f64), truncate to
    N, f32 signal, phasepoints, code_spectra of the N-sample replica, correlate_code."""
    lv = np.asarray(levels, dtype=np.float64)
    mean = lv.sum() / len(lv)
    sig = (lv - mean)[:N].astype(np.float32)
    pp = oc.phasepoints(fs, N)
    cc, ca = oc.code_spectra(code_n, P)
    return oc.correlate_code(sig, pp, cc, ca, fc, N, P, fs), sig, pp, cc, ca


def near_ties(sig, pp, cc, ca, fc, N, P):
    """(bin_tie, phase_tie): whether the oracle's winning (bin, hypothesis) or its code phase
    is within NEAR_TIE of a different candidate."""
    scale = 1.0 / np.sqrt(N)
    cand = []  # (value, bin, h, corr)
    for bi, f in enumerate(fc):
        X = np.fft.fft(oc.wipe(sig, pp, f, N))
        for h, rep in enumerate((cc, ca)):
            corr = np.abs(np.fft.ifft(X * np.asarray(rep, np.complex128)) * N) * scale
            cand.append((corr.max(), bi, h, corr))
    cand.sort(key=lambda c: -c[0])
    top = cand[0]
    bin_tie = len(cand) > 1 and cand[1][0] >= top[0] * (1 - NEAR_TIE) and \
        (cand[1][1], cand[1][2]) != (top[1], top[2])
    corr = top[3]
    m = int(np.argmax(corr)) % P
    folded = np.zeros(P)
    np.maximum.at(folded, np.arange(len(corr)) % P, corr)
    folded[m] = 0.0
    phase_tie = folded.max() >= top[0] * (1 - NEAR_TIE)
    return bin_tie, phase_tie


def _get(o, k):
    return getattr(o, k) if hasattr(o, k) else o[k]


def assert_peak_matches(gpu, ref, P, ties=(False, False), what=""):
    gpu = {k: _get(gpu, k) for k in ("best_peak", "second_peak", "best_freq_idx",
                                     "best_codephase")}
    bp, sp = float(gpu["best_peak"]), float(gpu["second_peak"])
    rbp, rsp = float(ref["best_peak"]), float(ref["second_peak"])
    assert abs(bp - rbp) <= PEAK_RTOL * abs(rbp), f"{what} best_peak {bp} vs {rbp}"
    if not ties[0]:
        assert int(gpu["best_freq_idx"]) == int(ref["best_freq_idx"]), f"{what} bin"
        if not ties[1]:
            assert int(gpu["best_codephase"]) % P == int(ref["best_codephase"]) % P, \
                f"{what} codephase {int(gpu['best_codephase'])} vs {int(ref['best_codephase'])}"
            assert abs(sp - rsp) <= PEAK_RTOL * abs(rsp) + 1e-12, f"{what} second {sp} vs {rsp}"
