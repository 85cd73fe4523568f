"""Pins of the oracle's RFI report, ACR C/N0 lookup and robust statistics against closed
forms and the reference's own known answers, chosen so that a plausible restatement mistake
fails one of them:

RFI (rfi.rs:83-168)
  * |X_k|^2 of a tone on FFT bin k is (A n / 2)^2, and the power spectrum equals a direct
    O(n^2) DFT on a ragged length;
  * the floor is sorted[half/2] (not the element beside it, not an average): a designed
    spectrum of powers 2^j gives excess 10 log10(P / 2^10) exactly;
  * local maxima are P[k] > P[k-1] and P[k] >= P[k+1] on 1 <= k < half-1: a plateau reports
    its left edge, and the end bins never report;
  * clustering keeps a line only more than 10 kHz (strictly) from every kept line, at most
    three, strongest first, equal excesses in ascending frequency (stable sort);
  * the ACF is normalised by n - lag: closed forms on period-4 sequences;
  * the dead-channel variance is the population variance (divided by n) against 1e-9;
  * the reference's square-wave / white-noise / weak-vs-strong / GPS / dead tests.
ACR (acr.rs:49-84, tests acr.rs:103-169)
  * interpolation is linear in ln r: r at the geometric mean of two table points gives the
    arithmetic mean of their C/N0 (linear-in-r interpolation does not);
  * exact table points, below-first, saturation, cut-off, monotone sweep, the KATs;
  * the [20, 70] clamp applies to interpolated values only (a synthetic table).
stats (stats.rs:32-61): the five cases.
"""
import math

import numpy as np
import pytest

from oracle import acr as oacr
from oracle import rfi as orfi
from oracle import stats as ostats

FS = 16.368e6


# ---- RFI --------------------------------------------------------------------------------

def test_tone_on_bin_power_closed_form():
    n, k0, amp, theta = 4096, 123, 0.7, 0.4
    i = np.arange(n)
    z = amp * np.cos(2 * np.pi * k0 * i / n + theta)
    p = orfi.power_spectrum(z)
    assert len(p) == n // 2
    assert p[k0] == pytest.approx((amp * n / 2) ** 2, rel=1e-12)
    others = np.delete(p, k0)
    assert others.max() < 1e-18 * p[k0]


def test_power_spectrum_is_a_direct_dft():
    rng = np.random.default_rng(5)
    n = 90  # ragged: 2 * 3^2 * 5
    z = rng.standard_normal(n)
    k = np.arange(n // 2)[:, None]
    m = np.arange(n)[None, :]
    X = (z[None, :] * np.exp(-2j * np.pi * k * m / n)).sum(axis=1)
    np.testing.assert_allclose(orfi.power_spectrum(z), np.abs(X) ** 2, rtol=1e-10, atol=1e-10)


def designed(half, spikes, base=None):
    """Powers rising monotonically with k (no local maxima of their own) with spikes."""
    p = np.array(base if base is not None else 1.0 + np.arange(half) / half, dtype=np.float64)
    for k, v in spikes.items():
        p[k] = v
    return p


def test_floor_is_sorted_half_over_two():
    # 19 background powers 2^0 .. 2^18 rising with k around one spike: half = 20, so the
    # floor is the 11th smallest, 2^10 (2^9 or 2^11 or their mean would shift the excess)
    half = 20
    base = [2.0 ** j for j in range(half - 1)]
    p = np.array(base[:7] + [1e9] + base[7:], dtype=np.float64)
    lines = orfi.lines_from_powers(p, 2 * half, 2 * half * 1000.0)
    assert len(lines) == 1
    assert lines[0]["frequency"] == 7000.0
    assert lines[0]["excess_db"] == pytest.approx(10 * math.log10(1e9 / 1024.0), abs=1e-12)


def test_local_maximum_rule_on_plateaus_and_ends():
    half = 64
    # a plateau reports its left edge only (strict on the left, non-strict on the right):
    # bins 10-11 -> 10, bins 29-31 -> 29; bins 0 and half-1 lie outside the scanned range
    # even when they are the largest
    p = designed(half, {10: 1e6, 11: 1e6, 29: 2e6, 30: 2e6, 31: 2e6, 0: 1e9, half - 1: 1e9})
    lines = orfi.lines_from_powers(p, 2 * half, 2 * half * 1000.0)
    assert sorted(l["frequency"] for l in lines) == [10000.0, 29000.0]
    # a peak equal to its left neighbour but above its right one is not a maximum
    p = designed(half, {20: 5e6, 21: 5e6})
    p[22] = 1.0
    lines = orfi.lines_from_powers(p, 2 * half, 2 * half * 1000.0)
    assert [l["frequency"] for l in lines] == [20000.0]


def test_clustering_is_strictly_more_than_10_khz_and_capped_at_three():
    half = 400  # bin_hz = 1000 Hz exactly, so k * bin_hz is exact
    p = designed(half, {100: 1e12, 110: 1e11, 120: 1e10, 200: 1e9, 300: 1e8})
    lines = orfi.lines_from_powers(p, 2 * half, 2 * half * 1000.0)
    # 110 lies exactly 10 kHz from 100 and is absorbed; 120 lies 10 kHz from the absorbed 110
    # but 20 kHz from the kept 100 and stays; 300 is beyond the cap of three
    assert [l["frequency"] for l in lines] == [100000.0, 120000.0, 200000.0]
    assert [l["excess_db"] for l in lines] == sorted((l["excess_db"] for l in lines),
                                                     reverse=True)


def test_equal_excess_keeps_ascending_frequency():
    half = 400
    p = designed(half, {300: 1e9, 50: 1e9, 170: 1e9})
    lines = orfi.lines_from_powers(p, 2 * half, 2 * half * 1000.0)
    assert [l["frequency"] for l in lines] == [50000.0, 170000.0, 300000.0]


def test_short_record_has_no_lines():
    assert orfi.detect_lines(np.cos(np.arange(7.0)), FS) == []


def test_acf_closed_forms_period_four():
    # [1, 1, -1, -1] x 16: zero mean; lag 1 sums 63 terms alternating +1/-1 from +1 -> +1,
    # lag 3 sums 61 terms alternating from -1 -> -1, lag 16 = one full period shift -> n - 16
    b = np.tile([1.0, 1.0, -1.0, -1.0], 16)
    ch = orfi.analyze_channel(3, b, FS)
    assert not ch["dead"] and ch["antenna"] == 3
    assert ch["acf_lag1"] == pytest.approx(1 / 63, rel=1e-14)
    assert ch["acf_lag3"] == pytest.approx(-1 / 61, rel=1e-14)
    assert ch["acf_lag16"] == pytest.approx(1.0, rel=1e-14)
    # [3, 1, -1, -1] x 16 (2-bit levels): mean 0.5, z = [2.5, 0.5, -1.5, -1.5]; per period the
    # lag-1 products sum to -1 (1.25 - 0.75 + 2.25 - 3.75), 63 terms = 15 periods + 2.75;
    # lag 3: 15 periods + (-3.75) over 61 terms; lag 16: 12 periods of sum z^2 = 11 over 48
    b = np.tile([3.0, 1.0, -1.0, -1.0], 16)
    ch = orfi.analyze_channel(0, b, FS)
    assert ch["acf_lag1"] == pytest.approx(-12.25 / 63, rel=1e-14)
    assert ch["acf_lag3"] == pytest.approx(-18.75 / 61, rel=1e-14)
    assert ch["acf_lag16"] == pytest.approx(132 / 48, rel=1e-14)


def test_dead_channel_uses_population_variance():
    # n = 20 zeros with one sample eps: var = eps^2 (n-1) / n^2, dead below 1e-9, i.e. for
    # eps < 1.4510e-4 (dividing by n - 1 instead would move the threshold to 1.4142e-4)
    for eps, dead in ((1.43e-4, True), (1.47e-4, False)):
        b = np.zeros(20)
        b[0] = eps
        ch = orfi.analyze_channel(0, b, FS)
        assert ch["dead"] is dead, eps
    ch = orfi.analyze_channel(4, np.full(100, -1.0), FS)
    assert ch == {"antenna": 4, "dead": True, "acf_lag1": None, "acf_lag3": None,
                  "acf_lag16": None, "period_samples": None, "lines": []}


def test_period_samples_from_strongest_line():
    n, fs = 4096, 4.096e6  # 1000 Hz bins
    i = np.arange(n)
    z = np.cos(2 * np.pi * 37 * i / n) + 0.01 * np.cos(2 * np.pi * 900 * i / n)
    ch = orfi.analyze_channel(0, z, fs)
    assert ch["lines"][0]["frequency"] == 37 * (fs / n)
    assert ch["period_samples"] == fs / (37 * (fs / n))


N_REF = 64 * 16368  # rfi.rs:182


def square_243(n):
    per = FS / 2.43e6
    i = np.arange(n, dtype=np.float64)
    return np.where(np.fmod(i, per) < per / 2.0, 1.0, -1.0)


def bipolar(sig):
    """simulator.rs:237-244 (one-bit quantise; analyze_channel de-means again)."""
    return np.where(sig >= 0.0, 1.0, -1.0)


@pytest.fixture(scope="module")
def ref_channels():
    rng = np.random.default_rng(9)
    noise = rng.standard_normal((3, N_REF))
    i = np.arange(N_REF)
    from oracle import codes as ocodes
    code = ocodes.code_resampled("gps", 1, 16368.0, 2 * 16368)
    a = 2.0 * np.sqrt(10 ** 4.0 / FS)
    gps = code[(i + 137) % 16368] * np.cos(2 * np.pi * 100e3 * i / FS) * a + noise[2]
    return {"strong": orfi.analyze_channel(0, bipolar(2.0 * square_243(N_REF) + noise[0]), FS),
            "weak": orfi.analyze_channel(0, bipolar(0.1 * square_243(N_REF) + noise[0]), FS),
            "white": orfi.analyze_channel(0, bipolar(noise[1]), FS),
            "gps": orfi.analyze_channel(0, bipolar(gps), FS)}


def test_reference_rfi_assertions_on_the_oracle(ref_channels):
    sq = ref_channels["strong"]
    top = sq["lines"][0]  # rfi.rs:214-233
    assert abs(top["frequency"] - 2.43e6) < 1e3 and top["excess_db"] > 15.0
    assert abs(sq["period_samples"] - FS / 2.43e6) < 0.2
    assert sq["acf_lag1"] > 0.3 and sq["acf_lag3"] < -0.4 and sq["acf_lag16"] < 0.0
    assert ref_channels["white"]["lines"] == []  # rfi.rs:236-246
    weak = ref_channels["weak"]["lines"]  # rfi.rs:249-268
    assert weak and top["excess_db"] > weak[0]["excess_db"] + 10.0
    assert ref_channels["gps"]["lines"] == []  # rfi.rs:271-287


# ---- ACR --------------------------------------------------------------------------------

TABLES = sorted(oacr.TABLES)


@pytest.mark.parametrize("name", TABLES)
def test_acr_log_space_interpolation(name):
    t = oacr.TABLES[name]
    checked = 0
    for (r0, c0), (r1, c1) in zip(t, t[1:]):
        for w in (0.5, 0.25, 0.8):
            r = math.exp((1 - w) * math.log(r0) + w * math.log(r1))
            want = c0 + w * (c1 - c0)
            if r0 < r < r1 and 20.0 <= want <= 70.0:
                assert oacr.estimate_cn0(r, t) == pytest.approx(want, abs=1e-9), (r0, r1, w)
                checked += 1
        # linear-in-r interpolation differs visibly at the bracket's geometric mean
        rg = math.sqrt(r0 * r1)
        lin = c0 + (rg - r0) / (r1 - r0) * (c1 - c0)
        if abs(lin - 0.5 * (c0 + c1)) > 1e-3:
            assert abs(oacr.estimate_cn0(rg, t) - lin) > 1e-4
    assert checked > 20


@pytest.mark.parametrize("name", TABLES)
def test_acr_reference_unit_tests(name):
    t = oacr.TABLES[name]
    assert len(t) >= 10 and t[0][0] > 1.0  # acr.rs:116-125
    for a, b in zip(t, t[1:]):
        assert a[0] < b[0] and a[1] < b[1]
    for r in (0.5, 1.0, float("nan"), float("inf")):  # acr.rs:128-134
        assert oacr.estimate_cn0(r, t) is None
    for r, c in t[::7]:  # acr.rs:137-148: exact table points round-trip
        assert oacr.estimate_cn0(r, t) == c
    assert oacr.estimate_cn0(0.5 * (1.0 + t[0][0]), t) == t[0][1]
    assert oacr.estimate_cn0(t[-1][0] * 10.0, t) == t[-1][1]  # acr.rs:151-158
    prev, r = t[0][1], t[0][0]  # acr.rs:171-184
    while r < t[-1][0]:
        r *= 1.01
        c = oacr.estimate_cn0(r, t)
        assert c >= prev - 1e-9
        prev = c


def test_acr_known_answers_and_gate():
    t = oacr.TABLES["GPS_L1_CA_ACR_TABLE"]
    assert abs(oacr.estimate_cn0(4.1725, t) - 45.0) < 0.3  # acr.rs:161-168
    assert abs(oacr.estimate_cn0(11.4497, t) - 50.0) < 0.3
    # the gate of acquisition.rs:345-352: v_s > 0 and v_m > v_s, r = (v_m / v_s)^2
    assert oacr.acr_from_peaks(3.0, 0.0, t) is None
    assert oacr.acr_from_peaks(2.0, 2.0, t) is None
    assert oacr.acr_from_peaks(2.0 * math.sqrt(4.1725), 2.0, t) == oacr.estimate_cn0(4.1725, t)


def test_acr_clamp_applies_to_interpolation_only():
    t = [(2.0, 10.0), (4.0, 80.0)]
    assert oacr.estimate_cn0(1.5, t) == 10.0   # at or below the first point: as tabulated
    assert oacr.estimate_cn0(2.0001, t) == 20.0  # interpolated ~10.01 -> clamped
    assert oacr.estimate_cn0(3.9999, t) == 70.0  # interpolated ~79.99 -> clamped
    assert oacr.estimate_cn0(5.0, t) == 80.0   # at or above the last point: as tabulated


# ---- stats ------------------------------------------------------------------------------

def test_stats_reference_cases():
    assert ostats.median([1.0, 3.0, 2.0]) == 2.0  # stats.rs:37-40
    assert ostats.median([1.0, 4.0, 2.0, 3.0]) == 2.5  # stats.rs:42-45
    assert ostats.median([7.0]) == 7.0  # stats.rs:47-50
    assert ostats.mad([1.0, 2.0, 3.0, 4.0, 5.0], 3.0) == 1.0  # stats.rs:52-56
    assert ostats.mad([2.0, 2.0, 2.0], 2.0) == 0.0  # stats.rs:58-61
    with pytest.raises(ValueError):
        ostats.median([])
