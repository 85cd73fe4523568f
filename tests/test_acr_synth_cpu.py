"""Host side of the ACR Monte Carlo (tga/acr_tables.py): the trial rows of the reference's
end-to-end C/N0 tests (testutil.rs:60-81 synth_signal_cn0), checked on the CPU where the
signal dominates the noise so every quantized sample's sign is known."""
import math

import numpy as np
import pytest

import tga
from tga import acr_tables, synth


@pytest.mark.parametrize("flip", [False, True])
def test_flip_rows_sign_flips_at_the_second_code_epoch(flip):
    system = "gps"
    P = acr_tables.SPECS[system][1]
    rows, want = acr_tables.flip_rows(system, 200.0, 3, flip, np.random.default_rng(1))
    assert rows.shape == (3, 2 * P // 8)
    code = tga.code_resampled(system, 1, float(P), 2 * P)
    i = np.arange(2 * P)
    carrier = np.cos(2.0 * math.pi * acr_tables.F_CARRIER * i / acr_tables.FS)
    step = acr_tables.E2E[system][2]
    for t in range(3):
        d = (t * step) % P
        assert want[t] == (P - d) % P
        clean = code[(i + d) % P] * carrier
        if flip:
            clean[P + (P - d) % P:] *= -1.0
        bits = synth.unpack_levels(rows[t:t + 1], 2 * P, 1)[0] > 0
        strong = np.abs(clean) > 1e-2  # amplitude 5000: noise cannot reach these signs
        assert np.array_equal(bits[strong], clean[strong] > 0), t


def test_end_to_end_trial_counts_follow_the_reference():
    # acquisition.rs:569, galileo.rs:484, beidou.rs:433, l1c.rs:444
    assert acr_tables.E2E == {"gps": (80, 120, 137), "galileo": (60, 120, 251),
                              "beidou": (50, 120, 379), "l1c": (50, 120, 431)}
