"""GPU Monte Carlo of the ACR calibration (examples/gen_acr_tables.rs restated on the batched
plan, tga/acr_tables.py) against the reference's committed tables (acr_tables.rs): the ratio
r_A measured at a C/N0 must map back to that C/N0 through estimate_cn0 (acr.rs:49-84) within
the Monte Carlo spread.  This is the statistical end-to-end check of the search's peak and
second-peak powers on the production path (16.368 MHz, one-bit samples); compare the
reference's own estimate tests (acquisition.rs:551-595, galileo.rs:465-510, beidou.rs:414-459,
l1c.rs:425-470)."""
import numpy as np
import pytest

import tga
from tga import acr_tables

pytestmark = pytest.mark.gpu

TABLE = {"gps": "GPS_L1_CA_ACR_TABLE", "galileo": "GALILEO_E1C_ACR_TABLE",
         "beidou": "BEIDOU_B1C_ACR_TABLE", "l1c": "GPS_L1C_ACR_TABLE"}


# (system, C/N0, trials, tolerance in dB, detection-rate bounds): at 40 dB-Hz a 2 ms GPS search
# locks on noise in about half the trials; longer codes integrate longer and detect lower
@pytest.mark.parametrize("system,cn0,trials,tol,det", [
    ("gps", 40.0, 400, 0.7, (0.2, 0.8)), ("gps", 46.0, 400, 0.6, (0.9, 1.0)),
    ("gps", 52.0, 200, 0.6, (0.99, 1.0)), ("galileo", 38.0, 200, 0.8, (0.0, 1.0)),
    ("beidou", 34.0, 100, 1.0, (0.9, 1.0))])
def test_monte_carlo_ratio_inverts_to_cn0(gpu, system, cn0, trials, tol, det):
    rng = np.random.default_rng(int(cn0 * 10) + len(system))
    r_a, found, _ = acr_tables.ratio(system, cn0, trials, rng)
    est = tga.estimate_cn0(r_a, TABLE[system])
    assert est is not None and abs(est - cn0) <= tol, (system, cn0, r_a, est)
    assert det[0] <= found <= det[1], (system, cn0, found)


def test_generated_table_is_monotonic(gpu):
    # a short stretch of the GPS grid, generated like gen_acr_tables.rs and printed format
    rng = np.random.default_rng(5)
    prev, plan = 0.0, None
    for cn0 in (44.0, 48.0, 52.0):
        r_a, _, plan = acr_tables.ratio("gps", cn0, 100, rng, plan)
        assert r_a > prev
        prev = r_a


# The reference's end-to-end estimate tests on all four calibrated systems, both hypotheses:
# no flip -> median within +-1 dB of the injected 45 dB-Hz; a sign flip at the code epoch
# ([c,-c] wins) -> median in [-4, -0.5] dB below it (the flipped window is biased low by 0 to
# -6 dB depending on where the flip lands, median -2.5 dB).
@pytest.mark.parametrize("flip", [False, True])
@pytest.mark.parametrize("system", ["gps", "galileo", "beidou", "l1c"])
def test_cn0_end_to_end_with_and_without_flip(gpu, system, flip):
    med, n, found = acr_tables.end_to_end(system, 45.0, flip)
    lo, hi = (-4.0, -0.5) if flip else (-1.0, 1.0)
    print(f"{system} flip={flip}: median {med} from {n} estimates, detected {found:.2f}",
          flush=True)
    assert n > 0, (system, flip)
    assert lo <= med - 45.0 <= hi, (system, flip, med, n)

