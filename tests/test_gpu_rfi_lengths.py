"""The narrowband RFI report (rfi.rs:142-184) at the record lengths the reference meets:
2^22 samples per antenna, the length of the real TART file the --rfi mode reads (one f64 FFT of
the whole record per antenna, rfi.rs:96-99), and 163680 samples (a 10 ms record at 16.368 MHz,
whose f64 transform needs the radix-5 pass).  Every field against oracle/rfi.py on the same
packed input (line frequencies and counts exact, excess and ACF within 1e-9 relative), and the
--rfi JSON (tga_rfi_json) byte for byte against the oracle's serde restatement of the same
channels."""
import numpy as np
import pytest

import tga
from tga import synth
from oracle import rfi as orfi

pytestmark = pytest.mark.gpu

FS = 16.368e6


def square(n, f):
    per = FS / f
    i = np.arange(n, dtype=np.float64)
    return np.where(np.fmod(i, per) < per / 2.0, 1.0, -1.0)


def channels(n, seed):
    rng = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.float64)
    tone = np.cos(2 * np.pi * 1.234567e6 * i / FS + 0.3)
    return np.stack([1.5 * square(n, 2.43e6) + rng.standard_normal(n),  # clock harmonic
                     rng.standard_normal(n),                                # noise only
                     0.3 * tone + 0.2 * np.cos(2 * np.pi * 5.5e6 * i / FS)
                     + rng.standard_normal(n),                              # two CW lines
                     np.ones(n)])                                           # dead channel


def check(got, ref, what):
    assert got["antenna"] == ref["antenna"] and got["dead"] == ref["dead"], what
    assert [l["frequency"] for l in got["lines"]] == [l["frequency"] for l in ref["lines"]], what
    for g, r in zip(got["lines"], ref["lines"]):
        assert g["excess_db"] == pytest.approx(r["excess_db"], rel=1e-9), what
    for f in ("acf_lag1", "acf_lag3", "acf_lag16", "period_samples"):
        assert (got[f] is None) == (ref[f] is None), (what, f)
        if ref[f] is not None:
            assert got[f] == pytest.approx(ref[f], rel=1e-9, abs=1e-12), (what, f)


@pytest.mark.parametrize("n,fmt", [(1 << 22, 1), (163680, 2)], ids=["2^22-1bit", "163680-2bit"])
def test_rfi_record_lengths(gpu, n, fmt):
    sig = channels(n, 7 if fmt == 1 else 8)
    packed = synth.pack_1bit(sig) if fmt == 1 else synth.pack_2bit(sig, threshold=1.0)
    got = tga.rfi(packed, n, fmt, FS)
    levels = synth.unpack_levels(packed, n, fmt)
    for a in range(levels.shape[0]):
        check(got[a], orfi.analyze_channel(a, levels[a], FS), f"n={n} ant {a}")
    assert abs(got[0]["lines"][0]["frequency"] - 2.43e6) < 1e3
    assert got[1]["lines"] == []
    assert any(abs(l["frequency"] - 1.234567e6) < 2 * FS / n for l in got[2]["lines"])
    assert got[3]["dead"]
    text = tga.rfi_json(packed, n, fmt, FS)
    assert text == orfi.report_json(got)
