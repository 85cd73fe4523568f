"""tga_correlate_code_noalloc (the level-1 correlation with no device allocation: its working
set lives in a fixed device array inside the library) against tga_correlate_code on the same
inputs, through the C ABI: three calls in a row on one GPU -- E1-C at N = 2P = 130944 with 7
bins, then a smaller C/A call (N = 32736, 5 bins, which reuses the array the first call left
behind), then B1C at the largest length N = 327360 (one bin per chunk) -- each to the parity
bar of the level-1 tests (bin and code phase exact, peaks within 1e-5 relative), plus the size
limits it rejects."""
import numpy as np
import pytest
import torch

import tga
from tga import synth

pytestmark = pytest.mark.gpu

FS = 16.368e6
IF = 4.092e6
PEAK = np.dtype([("best_peak", "<f4"), ("second_peak", "<f4"), ("best_freq_idx", "<u8"),
                 ("best_codephase", "<u8")])


def case(system, prn, P, N, fc, delay, seed):
    rng = np.random.default_rng(seed)
    code = tga.code_resampled(system, prn, float(P), N)
    i = np.arange(N)
    x = 0.5 * code[(i + delay) % P] * np.cos(2 * np.pi * fc[len(fc) // 2] * i / FS + 0.3) \
        + rng.standard_normal(N)
    lv = np.where(x >= 0, 1.0, -1.0)
    sig = (lv - lv.mean()).astype(np.float32)
    pp = synth.phasepoints(FS, N)
    cc, ca = tga.code_spectra(code, P)
    return sig, pp, cc, ca, np.asarray(fc, np.float64)


def run_noalloc(sig, pp, cc, ca, fc, N, P):
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda()
           for a in (sig, pp, cc.view(np.float32), ca.view(np.float32), fc)]
    out = torch.zeros(PEAK.itemsize, dtype=torch.uint8, device="cuda")
    nan = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    tga.correlate_code_noalloc(dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(),
                               dev[3].data_ptr(), dev[4].data_ptr(), len(fc), N, P, FS,
                               out.data_ptr(), nan.data_ptr(),
                               torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    pk = np.frombuffer(out.cpu().numpy().tobytes(), dtype=PEAK)[0]
    return pk, int(nan.cpu()[0])


CASES = [  # (system, prn, P, N, bins, delay): large, then smaller, then the largest length
    ("galileo", 11, 65472, 130944, [IF + 125.0 * k for k in range(-3, 4)], 40000),
    ("gps", 5, 16368, 32736, [IF + 500.0 * k for k in range(-2, 3)], 1234),
    ("beidou", 19, 163680, 327360, [IF - 50.0, IF, IF + 50.0], 99999),
]


def test_noalloc_matches_correlate_code(gpu):
    for k, (system, prn, P, N, fc, delay) in enumerate(CASES):
        print(f"call {k}: {system} N={N} bins={len(fc)}", flush=True)
        sig, pp, cc, ca, fcs = case(system, prn, P, N, fc, delay, seed=k)
        want = tga.correlate_code(sig, pp, cc, ca, fcs, N, P, FS)
        got, nan = run_noalloc(sig, pp, cc, ca, fcs, N, P)
        assert nan == 0
        assert int(got["best_freq_idx"]) == want.best_freq_idx, (k, got, want)
        assert int(got["best_codephase"]) == want.best_codephase, (k, got, want)
        assert abs(float(got["best_peak"]) - want.best_peak) <= 1e-5 * want.best_peak
        assert abs(float(got["second_peak"]) - want.second_peak) <= 1e-5 * want.second_peak
        assert int(got["best_codephase"]) % P == (P - delay) % P


def test_noalloc_rejects_sizes_beyond_its_limits(gpu):
    d = torch.zeros(1 << 20, dtype=torch.float64, device="cuda")
    out = torch.zeros(PEAK.itemsize, dtype=torch.uint8, device="cuda")
    p = d.data_ptr()
    for n_fc, N, what in ((3, 327360 * 2, "too long"), (3, 17 * 1000, "prime factor 17"),
                          (4097, 32736, "too many bins")):
        with pytest.raises(tga.TgaError) as e:
            tga.correlate_code_noalloc(p, p, p, p, p, n_fc, N, 16368, FS, out.data_ptr())
        assert e.value.name == "TGA_ERR_UNSUPPORTED", what
    with pytest.raises(tga.TgaError) as e:
        tga.correlate_code_noalloc(p, p, p, p, p, 0, 32736, 16368, FS, out.data_ptr())
    assert e.value.name == "TGA_ERR_EMPTY_BINS"
