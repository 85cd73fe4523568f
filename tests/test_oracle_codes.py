"""Pins the oracle's spreading-code generators (CPU).

Checks against data independent of the oracle's formulas: the reference's own KAT vectors
(tests/golden/code_kats.json, from sbas.rs:436-442, qzss.rs:437-441, beidou_codes.rs:225-253,
l1c_codes.rs:256-289), the IS-GPS-200 first-10-chip table for GPS PRN 1..32, Gold-code
correlation properties and the Legendre residue count.
"""
import json
import os

import numpy as np
import pytest

from oracle import codes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "code_kats.json")

# IS-GPS-200 Table 3-Ia "First 10 Chips (Octal)" for PRN 1..32 (standard XOR convention).
IS_GPS_200_FIRST10 = ["1440", "1620", "1710", "1744", "1133", "1455", "1131", "1454", "1626",
                      "1504", "1642", "1750", "1764", "1772", "1775", "1776", "1156", "1467",
                      "1633", "1715", "1746", "1763", "1063", "1706", "1743", "1761", "1770",
                      "1774", "1127", "1453", "1625", "1712"]


def kats():
    with open(GOLDEN) as f:
        return json.load(f)


def chips10_octal(chips):
    """+1 -> 0, -1 -> 1; 10 bits as 1 + 3 + 3 + 3 (the reference's helper, sbas.rs:417-426)."""
    b = [0 if v == 1 else 1 for v in chips[:10]]
    return "%d%d%d%d" % (b[0], 4 * b[1] + 2 * b[2] + b[3], 4 * b[4] + 2 * b[5] + b[6],
                         4 * b[7] + 2 * b[8] + b[9])


def chips24_octal(bits):
    return "".join(str(4 * bits[3 * i] + 2 * bits[3 * i + 1] + bits[3 * i + 2]) for i in range(8))


def test_sbas_first10_kat():
    k = kats()["sbas_first10_octal"]
    for i, exp in enumerate(k["values"]):
        assert chips10_octal(codes.sbas_ca(k["prn_base"] + i)) == exp, i


def test_qzss_first10_kat():
    k = kats()["qzss_first10_octal"]
    for i, exp in enumerate(k["values"]):
        assert chips10_octal(codes.qzss_ca(k["prn_base"] + i)) == exp, i


def test_gps_first10_is_gps_200_inverted():
    # The reference's chips are -(g1*g2): the bitwise complement of IS-GPS-200's XOR chips
    # (sbas.rs:430-434), so its octal is 1777 XOR the table value.
    for prn, tab in enumerate(IS_GPS_200_FIRST10, start=1):
        exp = "%04o" % (0o1777 ^ int(tab, 8))
        assert chips10_octal(codes.gps_ca(prn)) == exp, prn


def test_gps_delay_table_quirks():
    # PRN 34 and 37 share delay 950 (acquisition.rs:86), so their codes are identical.
    assert np.array_equal(codes.gps_ca(34), codes.gps_ca(37))
    assert not np.array_equal(codes.gps_ca(34), codes.gps_ca(35))


@pytest.mark.parametrize("prn", [1, 7, 19, 32, 36])
def test_gold_code_correlation_values(prn):
    # A balanced Gold code of length 1023: periodic autocorrelation 1023 at lag 0 and only
    # {-65, -1, 63} elsewhere; cross-correlations against other PRNs take the same 3 values.
    c = codes.gps_ca(prn).astype(np.int64)
    ac = np.array([np.dot(c, np.roll(c, k)) for k in range(1023)])
    assert ac[0] == 1023
    assert set(np.unique(ac[1:])).issubset({-65, -1, 63})
    other = codes.gps_ca(prn % 32 + 1).astype(np.int64)
    cc = np.array([np.dot(c, np.roll(other, k)) for k in range(0, 1023, 7)])
    assert set(np.unique(cc)).issubset({-65, -1, 63})


@pytest.mark.parametrize("fam,gen", [("b1cp", codes.b1cp_bits), ("l1cp", codes.l1cp_bits)])
def test_weil_first_last24_kat(fam, gen):
    k = kats()
    for prn in range(1, 64):
        bits = gen(prn)
        assert len(bits) == 10230
        assert chips24_octal(bits[:24]) == k[f"{fam}_first24_octal"]["values"][prn - 1], prn
        assert chips24_octal(bits[-24:]) == k[f"{fam}_last24_octal"]["values"][prn - 1], prn


@pytest.mark.parametrize("p", [10243, 10223])
def test_legendre_residue_count(p):
    seq = codes.legendre(p)
    assert seq[0] == 0 and int(seq.sum()) == (p - 1) // 2


def test_boc_and_bipolar_mapping():
    b = codes.b1cp_bits(1)
    pil = codes.b1c_pilot(1)
    assert len(pil) == 20460
    assert np.array_equal(pil[0::2], 1 - 2 * b.astype(np.int16))
    assert np.array_equal(pil[1::2], -pil[0::2])


def test_e1c_decode():
    c = codes.e1c(1)
    assert len(c) == 4092 and set(np.unique(c)) == {-1, 1}
    # first hex digit of PRN 1 is 'B' = 1011 (galileo_codes.rs data) -> +1 -1 +1 +1
    assert list(c[:4]) == [1, -1, 1, 1]
    for prn in range(1, 51):
        c = codes.e1c(prn)
        # memory codes are nearly balanced
        assert abs(int(c.astype(np.int64).sum())) < 200, prn


@pytest.mark.parametrize("bad", [("gps", 0), ("gps", 39), ("sbas", 1), ("sbas", 200),
                                 ("qzss", 183), ("galileo", 51), ("beidou", 64), ("l1c", 0)])
def test_prn_range_panics(bad):
    with pytest.raises(ValueError):
        codes.chips(*bad)


def test_resample_integer_index():
    # spc/L = 16 exactly at 16.368 MHz: the replica is each chip repeated 16 times.
    ch = codes.gps_ca(3)
    r = codes.resample(ch, 16368.0, 32736)
    assert np.array_equal(r[:16368], np.repeat(ch, 16))
    assert np.array_equal(r[16368:], r[:16368])
    # a partial last period (acquisition.rs:474-484)
    for n in (1000, 1023, 2046, 5049):
        assert len(codes.resample(ch, 1000.0, n)) == n


def test_e1c_boc11_is_the_plain_code_times_a_half_chip_square_wave():
    # resampled at 16 samples per chip: the BOC(1,1) replica equals the plain E1-C replica times
    # +1 on the first 8 samples of every chip and -1 on the last 8
    from oracle import codes as oc
    P = 65472
    n = np.arange(2 * P)
    sq = np.where((n % 16) < 8, 1.0, -1.0)
    for prn in (1, 27, 50):
        boc = oc.resample(oc.e1c_boc11(prn), float(P), 2 * P)
        plain = oc.resample(oc.e1c(prn), float(P), 2 * P)
        assert np.array_equal(boc, plain * sq), prn


def test_synth_boc_code_matches_oracle_replica():
    from oracle import codes as oc
    from tga import synth
    for prn in (3, 44):
        got = synth.e1c_boc11_code(prn, 65472.0, 2 * 65472)
        assert np.array_equal(got, oc.resample(oc.e1c_boc11(prn), 65472.0, 2 * 65472)), prn
