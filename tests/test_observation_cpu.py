"""TART observation reader (CPU): Observation::from_hdf5 (observation.rs:48-92) on the
committed fixture tests/golden/tart_2ms.hdf (the reference's data file cut to 2 ms per antenna,
see tests/golden/make_tart_fixture.py), and on the full reference file when it is present."""
import os

import numpy as np
import pytest

from tga import observation as ob

HERE = os.path.dirname(__file__)
FIXTURE = os.path.join(HERE, "golden", "tart_2ms.hdf")
FULL = "/root/reference/data/data_2026-03-09_04_25_10.145497.hdf"


def test_fixture_fields():
    o = ob.read_observation(FIXTURE)
    assert o.data.shape == (24, 4092) and o.data.dtype == np.uint8
    assert o.num_antenna == 24 and o.num_samples == 32736
    assert o.sampling_frequency == 16.368e6
    assert o.config["baseband_frequency"] == 4.092e6
    assert o.config["num_antenna"] == 24
    assert o.timestamp == "2026-03-09T04:25:10.145497+00:00"
    # observation.rs:140-151 and 232-240: MSB-first bits, 1 -> +1, 0 -> -1
    first = o.data[3, 0]
    want = [1.0 if (first >> (7 - i)) & 1 else -1.0 for i in range(8)]
    assert list(o.antenna(3)[:8]) == want
    with pytest.raises(IndexError):
        o.antenna(24)


@pytest.mark.skipif(not os.path.exists(FULL), reason="reference data file not present")
def test_full_file_and_fixture_agree():
    f = ob.read_observation(FULL)
    o = ob.read_observation(FIXTURE)
    assert f.data.shape == (24, 524288)
    assert np.array_equal(f.data[:, :4092], o.data)
    assert f.config == o.config and f.timestamp == o.timestamp


def test_reader_errors(tmp_path):
    bad = tmp_path / "x.hdf"
    bad.write_bytes(b"not hdf5 at all")
    with pytest.raises(ValueError):
        ob.read_observation(str(bad))
    h = ob._File(open(FIXTURE, "rb").read())
    with pytest.raises(KeyError):
        h.dataset("nothing")
