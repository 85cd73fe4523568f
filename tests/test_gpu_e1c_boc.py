"""Galileo E1-C with the BOC(1,1) replica (TGA_REPLICA_E1C_BOC11, an extension: the reference
correlates E1-C with plain chips, galileo.rs:112-122) on E1-C signals that carry the BOC(1,1)
subcarrier and the CS25 secondary code, 4 ms periods at 16.368 MHz (P = 65472), N = 2P.

  * every (antenna, PRN) of the BOC plan against the oracle run with the same replica
    (correlate_code semantics, parity bar of tests/parity.py), 1- and 2-bit input;
  * detections: each injected SV at its code phase in the BOC plan whether its two periods carry
    equal CS25 chips ([c,c] wins) or opposite ones ([c,-c] wins); the reference's plain replica
    on the same signals returns peaks well below the BOC replica's (a plain chip integrates the
    +- half-chips of the subcarrier to zero at the true code phase);
  * the option is refused for other systems.
"""
import numpy as np
import pytest

import tga
from tga import synth
from oracle import codes as ocodes
from parity import assert_peak_matches, near_ties, oracle_search

pytestmark = pytest.mark.gpu

FS = 16.368e6
IF = 4.092e6
P = 65472
N = 2 * P
FC = [IF - 125.0, IF, IF + 125.0]
PRNS = [11, 27, 5]


def sigs():
    # CS25 chips 3, 4 are equal (the second period keeps its sign); chips 1, 2 differ
    return [dict(system="galileo", prn=11, cn0=47.0, doppler=0.0, delay=20001, phase=0.4,
                 boc=True, secondary=3),
            dict(system="galileo", prn=27, cn0=47.0, doppler=125.0, delay=777, phase=2.0,
                 boc=True, secondary=1)]


def plan(n_ant, fmt, replica):
    return tga.Plan(FS, n_ant, N, [tga.SystemSearch("galileo", FC, PRNS, replica=replica)],
                    input_format=fmt)


@pytest.mark.parametrize("fmt", [1, 2])
def test_boc_replica_matches_oracle_and_detects(gpu, fmt):
    packed = synth.synth_observation(2, N, FS, sigs(), fmt=fmt, seed=31 + fmt)
    boc = plan(2, fmt, "boc11").search(packed)[0]
    ref = plan(2, fmt, "reference").search(packed)[0]
    levels = synth.unpack_levels(packed, N, fmt)
    for j, prn in enumerate(PRNS):
        code = ocodes.resample(ocodes.e1c_boc11(prn), float(P), N)
        for a in range(2):
            r, sig, pp, cc, ca = oracle_search(levels[a], FS, P, N, FC, code)
            assert_peak_matches(boc[a, j], r, P, near_ties(sig, pp, cc, ca, FC, N, P),
                                what=f"boc PRN {prn} ant {a}")
    for j, s in enumerate(sigs()):
        for a in range(2):
            pk = boc[a, j]
            d = (int(pk["best_codephase"]) - (P - s["delay"])) % P
            assert min(d, P - d) <= 1, (s["prn"], a, pk)
            assert FC[int(pk["best_freq_idx"])] - IF == s["doppler"], (s["prn"], a)
            ratio = float(pk["best_peak"]) / float(ref[a, j]["best_peak"])
            print(f"PRN {s['prn']} ant {a}: BOC / plain replica peak {ratio:.2f}", flush=True)
            assert ratio > 1.4, (s["prn"], a, ratio)


def test_boc_replica_is_galileo_only(gpu):
    with pytest.raises(tga.TgaError):
        tga.Plan(FS, 1, 2 * 16368, [tga.SystemSearch("gps", [IF], [1], replica="boc11")])
