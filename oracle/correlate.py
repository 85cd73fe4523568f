"""Code-phase / Doppler search (oracle restatement of correlate.rs and the per-PRN wrappers).

``correlate_code`` follows correlate.rs:91-183 step by step:
  * carrier wipe-off, phase = f32(pp[n] * fc) with pp in f64, e^{+j phase}, f32 products
    (correlate.rs:110-115);
  * forward FFT, per hypothesis product with the caller's conjugated replica spectrum,
    unnormalised inverse FFT, |.| scaled by f32(1/sqrt(N)) (correlate.rs:118-134);
  * in-row argmax taking the LAST maximum, the [c,-c] hypothesis replacing [c,c] only when
    strictly larger, the LAST maximum across bins (correlate.rs:135-154);
  * second peak: max of the winning row outside a ring window of ceil(fs/1.023e6) samples
    around best_codephase mod P (strict >), floor 1e-6 (correlate.rs:156-175).
The FFTs run in float64 on the float32-rounded inputs, so the oracle is the exact result of
the reference's f32-specified inputs; rustfft's own f32 rounding (about 1e-6 relative at the
peak) is the only difference to the reference and is inside the parity tolerance.

``mode="fp64"`` skips every f32 rounding (FP64 truth for the precision-form error reports).
"""
import math

import numpy as np

from . import codes as _codes


class OraclePanic(ValueError):
    """Raised where the reference panics (empty bin list, NaN comparison)."""


def rust_round(x):
    """f64::round: halves away from zero (Python's round() is half-to-even)."""
    r = math.floor(abs(x))
    if abs(x) - r >= 0.5:
        r += 1
    return int(math.copysign(r, x))


def phasepoints(fs, n):
    """testutil.rs:10-13 / acquisition.rs:289-291: TAU/fs * i in f64."""
    pc = (2.0 * math.pi) / fs
    return pc * np.arange(n, dtype=np.float64)


def code_spectra(code, period, mode="fp32"):
    """correlate.rs:47-68: conj(FFT([c,c])) and conj(FFT([c,-c])), second part negated from
    index min(period, len)."""
    code = np.asarray(code, dtype=np.float64)
    alt = code.copy()
    alt[min(period, len(alt)):] *= -1.0
    if mode == "fp32":
        code = code.astype(np.float32).astype(np.float64)
        alt = alt.astype(np.float32).astype(np.float64)
        cc = np.conj(np.fft.fft(code)).astype(np.complex64)
        ca = np.conj(np.fft.fft(alt)).astype(np.complex64)
        return cc, ca
    return np.conj(np.fft.fft(code)), np.conj(np.fft.fft(alt))


def wipe(signal, pp, freq, num_samples, mode="fp32"):
    """correlate.rs:110-115."""
    n = min(len(signal), len(pp), num_samples)
    iq = np.zeros(num_samples, dtype=np.complex128)
    if mode == "fp32":
        s = np.asarray(signal[:n], dtype=np.float32)
        ph = (np.asarray(pp[:n], dtype=np.float64) * float(freq)).astype(np.float32)
        ph64 = ph.astype(np.float64)
        c = np.cos(ph64).astype(np.float32)
        sn = np.sin(ph64).astype(np.float32)
        iq[:n] = (c * s).astype(np.float64) + 1j * (sn * s).astype(np.float64)
    else:
        s = np.asarray(signal[:n], dtype=np.float64)
        ph = np.asarray(pp[:n], dtype=np.float64) * float(freq)
        iq[:n] = s * np.exp(1j * ph)
    return iq


def last_argmax(v):
    """Iterator::max_by semantics: the last index among equal maxima; NaN panics."""
    if np.isnan(v).any():
        raise OraclePanic("called `Option::unwrap()` on a `None` value (NaN in correlation)")
    i = int(np.flatnonzero(v == v.max())[-1])
    return i, v[i]


def bin_rows(signal, pp, cc, cc_alt, fc, num_samples, mode="fp32"):
    """Per bin: (peak_idx, peak_val, corr) of the winning hypothesis (correlate.rs:107-146)."""
    N = num_samples
    if mode == "fp32":
        scale = np.float32(1.0) / np.sqrt(np.float32(N))
    else:
        scale = 1.0 / math.sqrt(N)
    out = []
    for freq in fc:
        X = np.fft.fft(wipe(signal, pp, freq, N, mode))
        best = None
        for rep in (cc, cc_alt):
            y = np.fft.ifft(X * np.asarray(rep, dtype=np.complex128)) * N
            if mode == "fp32":
                corr = (np.abs(y).astype(np.float32) * scale).astype(np.float32)
            else:
                corr = np.abs(y) * scale
            idx, val = last_argmax(corr)
            if best is None or val > best[1]:
                best = (idx, val, corr)
        out.append(best)
    return out


def surfaces(signal, pp, cc, cc_alt, fc, num_samples, P, mode="fp64"):
    """Every row's correlation values at code phases 0..P-1, [n_bin, 2, P]: corr =
    |IFFT(X . rep)| / sqrt(N) for rep = cc, cc_alt (correlate.rs:118-131, the array whose last
    maximum bin_rows keeps).  For N = 2P the rows are P-periodic (the replicas are [c,c] and
    [c,-c]), so phases P..2P-1 repeat these."""
    N = num_samples
    out = np.empty((len(fc), 2, P), dtype=np.float64)
    for b, freq in enumerate(fc):
        X = np.fft.fft(wipe(signal, pp, freq, N, mode))
        for h, rep in enumerate((cc, cc_alt)):
            y = np.fft.ifft(X * np.asarray(rep, dtype=np.complex128)) * N
            out[b, h] = np.abs(y[:P]) / math.sqrt(N)
    return out


def second_peak(best_corr, best_codephase, P, fs):
    """correlate.rs:156-175."""
    cw = int(math.ceil(fs / 1.023e6))
    main = best_codephase % P
    idx = np.arange(len(best_corr)) % P
    d = np.abs(idx - main)
    dist = np.minimum(d, P - d)
    mask = dist > cw
    sp = -np.inf
    if mask.any():
        sp = best_corr[mask].max()
    if not sp > 0.0:
        sp = 1e-6
    return sp


def select_bins(rows):
    """max_by over bins in bin order (last maximum wins) -- correlate.rs:150-154."""
    vals = np.array([r[1] for r in rows])
    if np.isnan(vals).any():
        raise OraclePanic("NaN peak across bins")
    fi, _ = last_argmax(vals)
    return fi


def correlate_code(signal, pp, cc, cc_alt, fc, num_samples, code_period_samples, fs,
                   mode="fp32"):
    """correlate.rs:91-183 -> dict(best_peak, second_peak, best_freq_idx, best_codephase)."""
    if len(fc) == 0:
        raise OraclePanic("frequency bin list must be non-empty")
    rows = bin_rows(signal, pp, cc, cc_alt, fc, num_samples, mode)
    fi = select_bins(rows)
    idx, val, corr = rows[fi]
    sp = second_peak(corr, idx, code_period_samples, fs)
    if mode == "fp32":
        sp = np.float32(sp)
    return {"best_peak": val, "second_peak": sp, "best_freq_idx": fi, "best_codephase": idx,
            "rows": [(r[0], r[1]) for r in rows]}


def folded_rows(signal, pp, code_period, fc, P, fs, mode="fp64"):
    """Two-period folding identity used by the GPU path (not a reference function).

    For N = 2P and replicas [c,c] / [c,-c]: the [c,c] correlation equals the P-point cyclic
    correlation of x[0:P] + x[P:2P]; the [c,-c] correlation equals, in magnitude, the P-point
    cyclic correlation of (x[0:P] - x[P:2P]) e^{-j pi n/P} with c[n] e^{-j pi n/P}.  The
    unnormalised 2P-point transforms carry twice the gain of the P-point ones (FFT_2P([c,c])
    at bin 2k is 2 FFT_P(c) at k), hence the factor 2.  Returns, per bin, the two magnitude
    rows (length P) on the reference scale 1/sqrt(2P)."""
    c = np.asarray(code_period, dtype=np.float64)[:P]
    tw = np.exp(-1j * np.pi * np.arange(P) / P)
    RS = np.fft.fft(c)
    RE = np.fft.fft(c * tw)
    scale = 1.0 / math.sqrt(2 * P)
    rows = []
    for freq in fc:
        x = wipe(signal, pp, freq, 2 * P, mode)
        s = x[:P] + x[P:]
        e = (x[:P] - x[P:]) * tw
        zs = np.fft.ifft(np.fft.fft(s) * np.conj(RS)) * P
        ze = np.fft.ifft(np.fft.fft(e) * np.conj(RE)) * P
        rows.append((2.0 * np.abs(zs) * scale, 2.0 * np.abs(ze) * scale))
    return rows


def acquire_single(system, signal, pp, fs, center_freq, search_band, prn, period,
                   mode="fp32"):
    """Per-PRN wrapper (acquisition.rs:176-248 and the five siblings): 300 Hz grid,
    resampled replica of exactly len(signal) samples, correlate_code, result mapping."""
    n = len(signal)
    nb = rust_round(2.0 * search_band / 300.0) + 1
    if nb == 1:
        fc = [center_freq]
    else:
        fc = [center_freq - search_band + i * (2.0 * search_band) / (nb - 1.0) for i in range(nb)]
    code = _codes.code_resampled(system, prn, float(period), n)
    cc, ca = code_spectra(code, period, mode)
    pk = correlate_code(signal, pp, cc, ca, fc, n, period, fs, mode)
    cp = pk["best_codephase"] % period
    return {"prn": prn, "signal_strength": float(pk["best_peak"]),
            "second_peak": float(pk["second_peak"]), "codephase_frac": cp / float(period),
            "frequency": fc[pk["best_freq_idx"]] - center_freq, "peak": pk}


def noncoherent(signal, pp, code_period, fc, P, epochs, fs, mode="fp32"):
    """Non-coherent search (an extension, not a reference function): per bin,
    S[m] = sum over epochs e < E of |y_e[m]|^2, y_e the unnormalised P-point circular
    correlation (the inverse transform without 1/P, like rustfft's) of the wiped epoch
    x[eP:(e+1)P] (carrier phase from the absolute sample index, correlate.rs:110-115 rounding
    in fp32 mode) with one code period.  Strength sqrt(S) / sqrt(P), which for E = 1 is the
    reference form's scale; last maximum in a row and across bins; second peak by the rule of
    correlate.rs:156-175.  Returns the correlate_code dict."""
    c = np.asarray(code_period, dtype=np.float64)[:P]
    if mode == "fp32":
        C = np.conj(np.fft.fft(c.astype(np.float32).astype(np.float64)))
        scale = np.float32(1.0) / np.sqrt(np.float32(P))
    else:
        C = np.conj(np.fft.fft(c))
        scale = 1.0 / math.sqrt(P)
    rows = []
    for f in fc:
        x = wipe(signal, pp, f, epochs * P, mode).reshape(epochs, P)
        y = np.fft.ifft(np.fft.fft(x, axis=1) * C[None, :], axis=1) * P
        S = (np.abs(y) ** 2).sum(axis=0)
        corr = np.sqrt(S)
        corr = (corr.astype(np.float32) * scale).astype(np.float32) if mode == "fp32" \
            else corr * scale
        idx, val = last_argmax(corr)
        rows.append((idx, val, corr))
    fi = select_bins(rows)
    idx, val, corr = rows[fi]
    sp = second_peak(corr, idx, P, fs)
    if mode == "fp32":
        sp = np.float32(sp)
    return {"best_peak": val, "second_peak": sp, "best_freq_idx": fi, "best_codephase": idx,
            "rows": [(r[0], r[1]) for r in rows]}
