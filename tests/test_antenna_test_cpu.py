"""--test-antennas report on the CPU: the oracle (oracle/antenna_test.py) against the known
answers of antenna_test.rs's unit tests, and the library's writer (tga_antenna_report_json,
via tga.antenna_test_json) against both, byte for byte against the oracle's serde_json
layout."""
import json
import pathlib

import numpy as np
import pytest

import tga
from oracle import antenna_test as oat

KATS = json.loads((pathlib.Path(__file__).parent / "golden" / "antenna_test_kats.json")
                  .read_text())["cases"]


def _get(doc, path):
    for k in path:
        if k == "#len":
            return len(doc)
        doc = doc[k]
    return doc


def _check(report, text, case):
    for path, want in case["expect"]:
        assert _get(report, path) == want, (case["name"], path)
    for frag in case.get("text_contains", []):
        assert frag in text


@pytest.mark.parametrize("case", KATS, ids=[c["name"] for c in KATS])
def test_oracle_known_answers(case):
    rep = oat.run(case["combined"], case["min_cn0"])
    _check(rep, oat.report_json(rep), case)


@pytest.mark.parametrize("case", KATS, ids=[c["name"] for c in KATS])
def test_library_known_answers(case):
    text = tga.antenna_test_json(case["combined"], case["min_cn0"])
    _check(json.loads(text), text, case)
    assert text == oat.report_json(oat.run(case["combined"], case["min_cn0"]))


def test_oracle_rejects_plausible_mistakes():
    # ascending instead of descending ranking, or a mean instead of the median, fail a KAT
    case = next(c for c in KATS if c["name"] == "basic_ranking_and_offsets")
    rep = oat.run(case["combined"], case["min_cn0"])
    assert rep["antennas"][2]["rank"] == 1 and rep["antennas"][0]["rank"] == 3
    even = next(c for c in KATS if c["name"] == "n_sats_varies_per_antenna")
    assert oat.run(even["combined"], even["min_cn0"])["antennas"][1]["median_cn0_db_hz"] == 30.5


def _random_combined(rng):
    n_ant = int(rng.integers(1, 7))
    ants = sorted(rng.choice(32, n_ant, replace=False).tolist())
    out = {}
    for key, fmt in (("gps", "GPS{:02d}"), ("galileo", "GSAT{:02d}"), ("sbas", "SBAS{:03d}"),
                     ("qzss", "QZS{:02d}")):
        if rng.random() < 0.3:
            continue
        rows = []
        for prn in sorted(rng.choice(40, int(rng.integers(0, 8)), replace=False).tolist()):
            if rng.random() < 0.15:
                rows.append({"sv": fmt.format(prn + 1), "cn0_acr": None})
                continue
            # quantised values so that ties in medians and thresholds occur
            c = [None if rng.random() < 0.2 else float(np.round(rng.uniform(30, 55) * 4) / 4)
                 for _ in range(n_ant)]
            if all(v is None for v in c):
                c = None
            rows.append({"sv": fmt.format(prn + 1), "cn0_acr": c})
        out[key] = {"antenna_numbers": ants, "results": rows}
    return out


def test_library_matches_oracle_on_random_outputs():
    rng = np.random.default_rng(11)
    for _ in range(300):
        comb = _random_combined(rng)
        thr = float(rng.choice([35.0, 40.0, 44.0, 47.5, 52.0]))
        assert tga.antenna_test_json(comb, thr) == oat.report_json(oat.run(comb, thr))


def test_library_reads_the_combined_json_text():
    comb = KATS[0]["combined"]
    assert tga.antenna_test_json(json.dumps(comb), 40.0) == tga.antenna_test_json(comb, 40.0)


def test_misaligned_rows_raise():
    comb = {"gps": {"antenna_numbers": [0, 1], "results": [{"sv": "GPS01", "cn0_acr": [45.0]}]}}
    with pytest.raises(ValueError):
        tga.antenna_test_json(comb, 40.0)
