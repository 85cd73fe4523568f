"""The CPU timing port (oracle/cpu_port.py) does the search it is timed as: its coherent and
non-coherent per-PRN searches find the oracle's bin and code phase on a strong signal, so the
cpu_baseline figure of bench.py prices the same work as the GPU path."""
import math

import numpy as np

from tga import synth
from oracle import codes as ocodes
from oracle import correlate as ocor
from oracle import cpu_port

FS = 16.368e6
IF = 4.092e6
P = 16368


def _signal(L, prn, doppler, delay, seed):
    signals = [dict(system="gps", prn=prn, cn0=50.0, doppler=doppler, delay=delay)]
    packed = synth.synth_observation(1, L, FS, signals, fmt=1, seed=seed)
    lv = synth.unpack_levels(packed, L, 1)[0].astype(np.float64)
    return (lv - lv.sum() / L).astype(np.float32)


def test_coherent_port_matches_oracle():
    L = 2 * P
    sig = _signal(L, 5, 1000.0, 4321, seed=3)
    fc = [IF + d for d in (-1000.0, 0.0, 1000.0, 2000.0)]
    code = ocodes.code_resampled("gps", 5, float(P), L)
    pp = (2.0 * math.pi / FS) * np.arange(L, dtype=np.float64)
    peak, b, cp = cpu_port.search_prn(sig, pp, code, fc, L, P, 1)
    cc, ca = ocor.code_spectra(code, P)
    ref = ocor.correlate_code(sig, ocor.phasepoints(FS, L), cc, ca, fc, L, P, FS)
    assert b == ref["best_freq_idx"] == 2
    assert cp % P == ref["best_codephase"] % P == (P - 4321) % P
    assert abs(peak - float(ref["best_peak"])) <= 1e-4 * float(ref["best_peak"])


def test_noncoherent_port_matches_oracle():
    E = 4
    sig = _signal(E * P, 11, -1500.0, 777, seed=4)
    fc = [IF + d for d in (-2000.0, -1500.0, -1000.0)]
    code = ocodes.code_resampled("gps", 11, float(P), P)
    pp = (2.0 * math.pi / FS) * np.arange(E * P, dtype=np.float64)
    pw, b, cp = cpu_port.search_prn_nc(sig, pp, code, fc, P, E, 1)
    ref = ocor.noncoherent(sig, ocor.phasepoints(FS, E * P), code, fc, P, E, FS)
    assert b == ref["best_freq_idx"] == 1
    assert cp == ref["best_codephase"] == (P - 777) % P
    # the port keeps the power sum; the oracle reports sqrt(S) / sqrt(P)
    assert abs(math.sqrt(pw / P) - float(ref["best_peak"])) <= 1e-4 * float(ref["best_peak"])
