"""The --rfi JSON writer (tga_rfi_report_json: serde_json::to_string_pretty of RfiReport,
rfi.rs:73-77, main.rs:468) against the oracle's serde restatement, byte for byte, on channel
records that exercise every field shape: a dead channel (None -> null, no lines), a live
channel without lines (period_samples null), one with three lines, and f64 values whose ryu
layout differs from Python's repr (1e-7, 1.5e20, whole numbers with ".0", -0.0).  Host-only:
no GPU is needed."""
import json

import pytest

import tga
from oracle import rfi as orfi

CHANNELS = [
    {"antenna": 0, "dead": False, "acf_lag1": 0.41234567890123, "acf_lag3": -0.7,
     "acf_lag16": -0.5000000000000001, "period_samples": 16.368e6 / 2430000.0,
     "lines": [{"frequency": 2430000.0, "excess_db": 41.25},
               {"frequency": 4860000.0, "excess_db": 22.0},
               {"frequency": 7290000.0, "excess_db": 15.000000000000002}]},
    {"antenna": 1, "dead": False, "acf_lag1": 1e-7, "acf_lag3": -0.0, "acf_lag16": 1.5e20,
     "period_samples": None, "lines": []},
    {"antenna": 2, "dead": True, "acf_lag1": None, "acf_lag3": None, "acf_lag16": None,
     "period_samples": None, "lines": []},
    {"antenna": 3, "dead": False, "acf_lag1": 0.0, "acf_lag3": 3.0, "acf_lag16": 123456.75,
     "period_samples": 1234.5, "lines": [{"frequency": 13330.0, "excess_db": 99.99}]},
]


def test_rfi_report_json_matches_serde_restatement():
    text = tga.rfi_report_json(CHANNELS)
    assert text == orfi.report_json(CHANNELS)
    d = json.loads(text)
    assert list(d) == ["antenna_numbers", "channels"]
    assert d["antenna_numbers"] == [0, 1, 2, 3]
    assert list(d["channels"][0]) == ["antenna", "dead", "acf_lag1", "acf_lag3", "acf_lag16",
                                      "period_samples", "lines"]
    assert d["channels"][2]["acf_lag1"] is None and d["channels"][2]["lines"] == []
    assert '"frequency": 2430000.0' in text and '"acf_lag1": 1e-7' in text
    assert '"acf_lag16": 1.5e20' in text and '"acf_lag3": -0.0' in text


def test_rfi_report_json_empty_and_errors():
    assert tga.rfi_report_json([]) == orfi.report_json([])
    bad = dict(CHANNELS[0], lines=CHANNELS[0]["lines"] + [{"frequency": 1.0, "excess_db": 1.0}])
    with pytest.raises(ValueError):
        tga.rfi_report_json([bad])
