"""--test-antennas on the GPU path: Plan.test_antennas_json (search, forced C/N0 estimates,
MAD filters, antenna_test::run) against the oracle's report built from the same peaks, and the
reference's end-to-end ranking test (antenna_test.rs:589-704) restated: three antennas see GPS
PRNs 1, 7 and 20 through one-bit quantization, antenna 2 with 8x the noise, and the report must
rank it last by more than 5 dB while the two quiet antennas agree within 1 dB."""
import json

import numpy as np
import pytest

import tga
from oracle import acquire as oacq
from oracle import antenna_test as oat

pytestmark = pytest.mark.gpu

FS = 16.368e6
GPS_IF = 4.092e6
SAMPLES_PER_MS = 16368
CARRIER = GPS_IF - 2000.0  # antenna_test.rs:596-599
PRNS = [1, 7, 20]


def xorshift_uniform(seed, n):
    """The test module's XorShift (antenna_test.rs:601-618): n draws in [0, 1)."""
    m = (1 << 64) - 1
    x = seed | 1
    out = np.empty(n)
    for i in range(n):
        x ^= (x << 13) & m
        x ^= x >> 7
        x ^= (x << 17) & m
        out[i] = (x >> 11) / float(1 << 53)
    return out


def observation_with_gains(n_ant, gains, noise_levels):
    """observation_with_gains (antenna_test.rs:624-656), packed MSB-first."""
    n = 2 * SAMPLES_PER_MS
    i = np.arange(n)
    carrier = np.cos(2.0 * np.pi * CARRIER * (i / FS))
    clean = np.zeros(n)
    for k, prn in enumerate(PRNS):
        chips = tga.code_chips("gps", prn).astype(np.float64)
        code = chips[np.arange(SAMPLES_PER_MS) // 16]
        clean += code[(i + 137 * (k + 1)) % SAMPLES_PER_MS] * carrier
    rows = []
    for j in range(n_ant):
        amp = noise_levels[min(j, len(noise_levels) - 1)]
        v = clean * gains[j] + amp * (2.0 * xorshift_uniform(1000 + j, n) - 1.0)
        rows.append(np.packbits((v >= 0.0).astype(np.uint8)))  # simulator.rs:228-230
    return np.stack(rows)


@pytest.fixture(scope="module")
def setup(gpu):
    packed = observation_with_gains(3, [1.0] * 3, [2.0, 2.0, 16.0])
    fc = tga.search_frequencies(GPS_IF, 6000.0)
    plan = tga.Plan(FS, 3, 2 * SAMPLES_PER_MS, [tga.SystemSearch("gps", fc, PRNS)])
    return plan, packed, fc


def oracle_report(plan, peaks, fc, min_cn0, fpm=None, ffm=None):
    pk = peaks[0]
    dicts = [[{k: pk[a, j][k] for k in pk.dtype.names} for j in range(pk.shape[1])]
             for a in range(pk.shape[0])]
    combined = {"gps": oacq.system_output("gps", PRNS, dicts, fc, GPS_IF, SAMPLES_PER_MS,
                                          [0, 1, 2], True, fpm, ffm)}
    return oat.report_json(oat.run(json.loads(oacq.combined_json(combined)), min_cn0))


def test_end_to_end_ranking_recovers_injected_gains(setup):
    plan, packed, _ = setup
    rep = json.loads(plan.test_antennas_json(packed, GPS_IF, min_cn0=40.0))
    assert rep["antenna_numbers"] == [0, 1, 2]
    assert rep["n_reference_satellites"] == len(PRNS)
    assert rep["reference_satellites"] == ["GPS01", "GPS07", "GPS20"]
    st = rep["antennas"]
    best = min(st, key=lambda s: s["rank"])
    assert best["antenna"] != 2, st
    weak = st[2]
    assert weak["rank"] == 3 and weak["offset_db"] < -5.0, st
    m0, m1 = st[0]["median_cn0_db_hz"], st[1]["median_cn0_db_hz"]
    assert abs(m0 - m1) < 1.0 and m0 - weak["median_cn0_db_hz"] > 5.0, st


@pytest.mark.parametrize("min_cn0,fpm,ffm", [(40.0, None, None), (44.0, None, None),
                                             (55.0, None, None), (40.0, 0.0, None),
                                             (40.0, None, 0.0)])
def test_report_matches_oracle_on_the_same_peaks(setup, min_cn0, fpm, ffm):
    plan, packed, fc = setup
    peaks = plan.search(packed)
    got = plan.peaks_test_antennas_json(peaks, GPS_IF, min_cn0=min_cn0, filter_phase_mad=fpm,
                                        filter_freq_mad=ffm)
    assert got == oracle_report(plan, peaks, fc, min_cn0, fpm, ffm)
    # the same report from the CombinedOutput text the library writes
    text = plan.peaks_json(peaks, GPS_IF, cn0=True, filter_phase_mad=fpm, filter_freq_mad=ffm)
    assert got == tga.antenna_test_json(text, min_cn0)


def test_acquire_form_equals_peaks_form(setup):
    plan, packed, _ = setup
    peaks = plan.search(packed)
    assert (plan.test_antennas_json(packed, GPS_IF, antenna_numbers=[4, 9, 2])
            == plan.peaks_test_antennas_json(peaks, GPS_IF, antenna_numbers=[4, 9, 2]))
