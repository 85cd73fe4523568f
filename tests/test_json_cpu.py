"""JSON number formatting (CPU): the library's writer and the oracle's against serde_json's
known outputs (ryu shortest digits; main.rs:893 prints with serde_json::to_string_pretty)."""
import math
import struct

import numpy as np
import pytest

import tga
from oracle import jsonfmt

# (value, text serde_json prints) -- hand-written expectations of ryu's layout
KNOWN = [(0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (-1500.0, "-1500.0"), (0.1, "0.1"),
         (0.3, "0.3"), (0.1 + 0.2, "0.30000000000000004"), (2.0 / 3.0, "0.6666666666666666"),
         (100.0, "100.0"), (1e15, "1000000000000000.0"), (1e16, "1e16"),
         (1.2345678901234568e17, "1.2345678901234568e17"), (1e-5, "0.00001"),
         (1.5e-5, "0.000015"), (1e-6, "1e-6"), (1e-7, "1e-7"), (1.5e-7, "1.5e-7"),
         (45.0, "45.0"), (0.5, "0.5"), (123.456, "123.456"), (5e-324, "5e-324"),
         (1.7976931348623157e308, "1.7976931348623157e308"), (float("nan"), "null"),
         (float("inf"), "null"), (12345.678, "12345.678"), (0.0001, "0.0001"),
         (9007199254740993.0, "9007199254740992.0")]


@pytest.mark.parametrize("x,text", KNOWN)
def test_oracle_matches_known_serde_output(x, text):
    assert jsonfmt.fmt_f64(x) == text


@pytest.mark.parametrize("x,text", KNOWN)
def test_library_matches_known_serde_output(x, text):
    assert tga.format_f64(x) == text


def test_library_matches_oracle_on_random_doubles():
    rng = np.random.default_rng(7)
    vals = list(rng.standard_normal(2000) * 10.0 ** rng.integers(-30, 30, 2000))
    vals += [float(np.float32(v)) for v in rng.random(1000) * 1000]  # widened f32 peaks
    vals += [struct.unpack("<d", struct.pack("<Q", int(b)))[0]
             for b in rng.integers(1, 0x7FEFFFFFFFFFFFFF, 1000, dtype=np.int64)]
    for v in vals:
        if math.isfinite(v):
            t = tga.format_f64(v)
            assert t == jsonfmt.fmt_f64(v), v
            assert float(t) == v  # round trip


def test_pretty_layout():
    text = jsonfmt.dumps_pretty({"gps": {"antenna_numbers": [0, 1], "results": []},
                                 "qzss": {"x": [None, 1.0]}})
    assert text == ('{\n  "gps": {\n    "antenna_numbers": [\n      0,\n      1\n    ],\n'
                    '    "results": []\n  },\n  "qzss": {\n    "x": [\n      null,\n'
                    '      1.0\n    ]\n  }\n}')
