"""GPU search of real TART data: the committed 2 ms cut of the reference's observation
(tests/golden/tart_2ms.hdf, 24 antennas, 1-bit, 16.368 MHz) searched for GPS PRN 1-32 on the
reference CLI's grid (41 bins of 300 Hz around 4.092 MHz, acquisition.rs:189-198,
main.rs:738-739), checked against the oracle run per (antenna, PRN) with the reference's
semantics, and written as the reference's JSON."""
import json
import os

import numpy as np
import pytest

import tga
from tga import synth
from oracle import codes as ocodes
from parity import assert_peak_matches, near_ties, oracle_search

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "tart_2ms.hdf")
IF = 4.092e6
FC = [IF - 6000.0 + 300.0 * i for i in range(41)]


def test_real_observation_matches_oracle(gpu):
    obs = tga.read_observation(FIXTURE)
    fs, L = obs.sampling_frequency, obs.num_samples
    plan = tga.Plan(fs, obs.num_antenna, L, [tga.SystemSearch("gps", FC, list(range(1, 33)))],
                    input_format=tga.INPUT_1BIT)
    assert "fused16368" in plan.describe()
    peaks = plan.search(obs.data)
    levels = synth.unpack_levels(obs.data, L, 1)
    P = 16368
    for a in (0, 17):
        for j, prn in enumerate(range(1, 33)):
            if prn % 4 != a % 4 + 1:  # a quarter of the PRNs per antenna keeps the oracle quick
                continue
            code = ocodes.code_resampled("gps", prn, float(P), 2 * P)
            ref, sig, pp, cc, ca = oracle_search(levels[a], fs, P, 2 * P, FC, code)
            assert_peak_matches(peaks[0][a, j], ref, P, near_ties(sig, pp, cc, ca, FC, 2 * P, P),
                                what=f"ant {a} prn {prn}")
    text = plan.peaks_json(peaks, IF, list(range(24)), cn0=True)
    d = json.loads(text)
    assert [r["sv"] for r in d["gps"]["results"]] == [f"GPS{p:02d}" for p in range(1, 33)]
    # strongest SV by median strength, for the record
    best = max(d["gps"]["results"], key=lambda r: np.median(r["strengths"]))
    print("strongest:", best["sv"], "phase MAD", best["phase_mad"], "freq", best["freq_median"])
