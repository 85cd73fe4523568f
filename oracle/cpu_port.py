"""CPU port of the reference's search with its work structure, timed as the CPU baseline.

Mirrors acquire_all_* -> acquire_full -> correlate_code (acquisition.rs:281-317,
correlate.rs:91-146): per (antenna, PRN) two N-point replica FFTs (code_spectra), then per
Doppler bin one wipe-off, one forward FFT and two inverse FFTs with |.| and argmax, N = 2P,
all in complex64 (the reference's f32 rustfft) with scipy's pocketfft.  The bins of one PRN
are transformed as one batch over ``workers`` threads, standing in for rayon's parallel
iterator over bins (correlate.rs:106).  It is the oracle's timing port, not the checker.
"""
import math
import time

import numpy as np
import scipy.fft as sfft

from . import codes


def search_prn(sig32, pp, code, fc, N, P, workers):
    """One (antenna, PRN) search: returns (best_peak, best_bin, best_codephase)."""
    c = code.astype(np.float32)
    alt = c.copy()
    alt[min(P, N):] *= -1
    cc = np.conj(sfft.fft(c.astype(np.complex64), workers=workers))
    ca = np.conj(sfft.fft(alt.astype(np.complex64), workers=workers))
    ph = np.multiply.outer(np.asarray(fc), pp).astype(np.float32)
    iq = (np.cos(ph) * sig32 + 1j * (np.sin(ph) * sig32)).astype(np.complex64)
    X = sfft.fft(iq, axis=-1, workers=workers)
    scale = np.float32(1.0 / math.sqrt(N))
    best = (-1.0, 0, 0)
    for rep in (cc, ca):
        y = sfft.ifft(X * rep, axis=-1, workers=workers, norm="forward")
        corr = np.abs(y) * scale
        idx = corr.argmax(axis=-1)
        val = corr[np.arange(len(fc)), idx]
        b = int(val.argmax())
        if val[b] > best[0]:
            best = (float(val[b]), b, int(idx[b]))
    return best


def search_prn_nc(sig32, pp, code, fc, P, epochs, workers):
    """One (antenna, PRN) non-coherent search in the same complex64 work structure: one P-point
    replica FFT, then per bin one wipe-off of E periods, E forward and E inverse P-point FFTs,
    |y|^2 summed over the epochs and the argmax.  Returns (best_power, best_bin, codephase)."""
    cc = np.conj(sfft.fft(code[:P].astype(np.complex64), workers=workers))
    ph = np.multiply.outer(np.asarray(fc), pp[:epochs * P]).astype(np.float32)
    s = sig32[:epochs * P]
    iq = (np.cos(ph) * s + 1j * (np.sin(ph) * s)).astype(np.complex64)
    X = sfft.fft(iq.reshape(len(fc), epochs, P), axis=-1, workers=workers)
    y = sfft.ifft(X * cc, axis=-1, workers=workers, norm="forward")
    pw = (y.real * y.real + y.imag * y.imag).sum(axis=1)
    idx = pw.argmax(axis=-1)
    val = pw[np.arange(len(fc)), idx]
    b = int(val.argmax())
    return float(val[b]), b, int(idx[b])


def time_sample(levels, fs, system, prns, fc, workers, budget_s=10.0, epochs=0):
    """Run whole PRNs on one antenna until ``budget_s`` elapses; returns
    (cells searched, seconds, PRNs done).  epochs > 0: the non-coherent form over that many
    one-period epochs."""
    P = codes.samples_per_period(system, fs)
    N = epochs * P if epochs > 0 else min(2 * P, len(levels))
    lv = np.asarray(levels, dtype=np.float64)
    sig32 = (lv - lv.sum() / len(lv))[:N].astype(np.float32)
    pp = (2.0 * math.pi / fs) * np.arange(N, dtype=np.float64)
    done = 0
    t0 = time.perf_counter()
    for prn in prns:
        if epochs > 0:
            code = codes.code_resampled(system, prn, float(P), P)
            search_prn_nc(sig32, pp, code, fc, P, epochs, workers)
        else:
            code = codes.code_resampled(system, prn, float(P), N)
            search_prn(sig32, pp, code, fc, N, P, workers)
        done += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return done * len(fc) * P, dt, done
