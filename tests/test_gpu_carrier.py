"""The carrier table of the fused forward kernels (tga_plan_carrier), made once per plan and
multiplied into every antenna's samples, pinned value by value against the wipe-off of the
reference (correlate.rs:110-115, acquisition.rs:289-291): phi = f32(f64(TAU / fs * n) * fc), then
cos / sin evaluated in f64 and rounded to f32 -- the oracle's fp32 wipe.  The FP16 form keeps
the phase in f64 (cos / sin of f64(TAU / fs * n) * fc rounded to f32).

Both sides round an f64 cos / sin to f32; two f64 implementations that differ by an ulp can
round differently only when the f64 value sits within an ulp of an f32 rounding boundary (about
one value in 10^8), so the bar is bit equality with at most one 1-ulp difference per 10^5
values.  The FP32 and FP16 tables differ in a large fraction of their values at these lengths,
so a switch of the FP32 form to the f64 phase (or back) fails the test.
"""
import math

import numpy as np
import pytest

import tga

pytestmark = pytest.mark.gpu

FS = 16.368e6
IF = 4.092e6
FC = [IF - 7000.0, IF - 500.0, IF, IF + 1500.0, IF + 9999.5]


def expected(fs, fc, n_samples, phase64):
    pc = (2.0 * math.pi) / fs
    pp = pc * np.arange(n_samples, dtype=np.float64)  # f64 product, as phasepoints
    out = np.empty((len(fc), n_samples), dtype=np.complex64)
    for b, f in enumerate(fc):
        ph = pp * np.float64(f)
        if not phase64:
            ph = ph.astype(np.float32).astype(np.float64)
        out[b] = np.cos(ph).astype(np.float32) + 1j * np.sin(ph).astype(np.float32)
    return out


def assert_bits(got, want, what):
    g = got.view(np.uint32).astype(np.int64)
    w = want.view(np.uint32).astype(np.int64)
    diff = np.abs(g - w)
    n_off = int(np.count_nonzero(diff))
    assert diff.max() <= 1, (what, int(diff.max()))
    assert n_off <= max(1, g.size // 100000), (what, n_off, g.size)
    return n_off


@pytest.mark.parametrize("precision,nc", [("fp32", 0), ("fp16", 0), ("fp32", 3)])
def test_carrier_table_matches_the_reference_wipe(gpu, precision, nc):
    L = 3 * 16368
    plan = tga.Plan(FS, 1, L, [tga.SystemSearch("gps", FC, [1, 2], nc_epochs=nc)],
                    precision=precision)
    got = plan.carrier(0)
    n_samples = (nc if nc else 2) * 16368
    assert got.shape == (len(FC), n_samples)
    want = expected(FS, FC, n_samples, phase64=precision == "fp16")
    n_off = assert_bits(got, want, (precision, nc))
    print(f"{precision} nc={nc}: {got.size} values, {n_off} one-ulp differences", flush=True)
    # the other phase form differs widely: the test tells them apart
    other = expected(FS, FC, n_samples, phase64=precision != "fp16")
    assert np.count_nonzero(got.view(np.uint32) != other.view(np.uint32)) > got.size // 10


def test_carrier_only_for_fused_groups(gpu):
    plan = tga.Plan(FS, 1, 2 * 65472, [tga.SystemSearch("galileo", [IF], [1])])
    with pytest.raises(tga.TgaError):
        plan.carrier(0)
