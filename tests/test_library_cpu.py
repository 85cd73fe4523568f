"""libtga.so on the CPU host: it loads, exports every entry point of include/tga.h, and its
host-side pieces (replica generation, resampling, ACR lookup, error paths) agree with the
oracle.  No kernel is launched here."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import tga
from tga import _lib
from oracle import acr as oacr
from oracle import codes as ocodes

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "tga.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tga_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    # and the Python binding prototypes cover exactly those symbols
    assert sorted(_lib.PROTOTYPES) == names


def test_version_and_device_count():
    assert b"gfx950" in _lib.lib().tga_version()
    assert tga.device_count() >= 0


@pytest.mark.parametrize("system", ["gps", "sbas", "qzss", "galileo", "beidou", "l1c"])
def test_chips_match_oracle(system):
    for prn in ocodes.PRN_RANGES[system]:
        assert np.array_equal(tga.code_chips(system, prn), ocodes.chips(system, prn)), prn


@pytest.mark.parametrize("system,spc,n", [("gps", 16368.0, 32736), ("gps", 1000.0, 5049),
                                          ("sbas", 1023.0, 2050), ("galileo", 65472.0, 130944),
                                          ("beidou", 10230.0, 20464), ("l1c", 163680.0, 1000),
                                          ("qzss", 2046.0, 2043)])
def test_resample_matches_oracle(system, spc, n):
    prn = ocodes.PRN_RANGES[system][3]
    assert np.array_equal(tga.code_resampled(system, prn, spc, n),
                          ocodes.code_resampled(system, prn, spc, n))


def test_prn_range_error_codes():
    for system, prn in [("gps", 0), ("gps", 39), ("sbas", 1), ("qzss", 207), ("galileo", 51),
                        ("beidou", 0), ("l1c", 64)]:
        with pytest.raises(tga.TgaError) as e:
            tga.code_chips(system, prn)
        assert e.value.name == "TGA_ERR_PRN_RANGE"
        assert "PRN must be in" in e.value.message


@pytest.mark.parametrize("table", sorted(oacr.TABLES))
def test_estimate_cn0_matches_oracle(table):
    t = oacr.TABLES[table]
    rs = [0.5, 1.0, float("nan"), float("inf"), t[0][0], 0.5 * (1 + t[0][0]), t[-1][0],
          t[-1][0] * 10]
    rs += [r for r, _ in t[::5]]
    rs += list(np.exp(np.linspace(math.log(t[0][0]), math.log(t[-1][0]), 200)))
    for r in rs:
        assert tga.estimate_cn0(r, table) == oacr.estimate_cn0(r, t), r


def test_acr_known_answers():
    # acr.rs:148-155
    assert abs(tga.estimate_cn0(4.1725, "GPS_L1_CA_ACR_TABLE") - 45.0) < 0.3
    assert abs(tga.estimate_cn0(11.4497, "GPS_L1_CA_ACR_TABLE") - 50.0) < 0.3


def test_gpu_entry_points_fail_cleanly_without_device():
    if tga.device_count() > 0:
        pytest.skip("a device is visible; this checks the no-device error path")
    sig = np.zeros(64, np.float32)
    pp = np.zeros(64)
    cc = np.ones(64, np.complex64)
    with pytest.raises(tga.TgaError) as e:
        tga.correlate_code(sig, pp, cc, cc, [0.0], 64, 32, 1e6)
    assert e.value.name == "TGA_ERR_NO_DEVICE"
    with pytest.raises(tga.TgaError) as e:
        tga.correlate_code(sig, pp, cc, cc, [], 64, 32, 1e6)
    assert e.value.name == "TGA_ERR_EMPTY_BINS"
    with pytest.raises(tga.TgaError) as e:
        tga.Plan(16.368e6, 1, 163680, [tga.SystemSearch("gps", [4.092e6])])
    assert e.value.name == "TGA_ERR_NO_DEVICE"
