"""GPU tests of the host layers around the search: the acquire_all_* JSON (CombinedOutput,
lib.rs:31-45; per-SV records acquisition.rs:304-395; MAD filters main.rs:242-255) and the
narrowband RFI report (rfi.rs:142-184, tests rfi.rs:196-297 restated on the GPU path).

JSON: the library's text is compared byte for byte with the oracle's restatement built from
the same peaks (the peaks themselves are checked against the oracle's search in
test_gpu_plan.py), and the peaks of acquire_all_json are checked against the oracle search.
RFI: every field against oracle/rfi.py on the same packed input (line frequencies and counts
exact, excess and ACF to 1e-9 relative), plus the reference's own assertions."""
import json

import numpy as np
import pytest

import tga
from tga import synth
from oracle import acquire as oacq
from oracle import codes as ocodes
from oracle import rfi as orfi
from parity import assert_peak_matches, near_ties, oracle_search

pytestmark = pytest.mark.gpu

FS = 16.368e6
IF = 4.092e6
FC = [IF - 5000.0 + 500.0 * i for i in range(21)]


def peak_dicts(arr):
    return [[{k: arr[a, j][k] for k in arr.dtype.names} for j in range(arr.shape[1])]
            for a in range(arr.shape[0])]


def oracle_json(plan, peaks, center, antenna_numbers, cn0, fpm=None, ffm=None):
    outs = {}
    for si, s in enumerate(plan.systems):
        P = ocodes.samples_per_period(s.system, FS)
        outs[s.system] = oacq.system_output(s.system, list(plan.prns[si]),
                                            peak_dicts(peaks[si]), list(s.fc), center,
                                            P, antenna_numbers, cn0, fpm, ffm)
    return oacq.combined_json(outs)


@pytest.fixture(scope="module")
def obs3(gpu):
    sig = [dict(system="gps", prn=5, cn0=52.0, doppler=1800.0, delay=4000, phase=0.4),
           dict(system="gps", prn=17, cn0=47.0, doppler=-3100.0, delay=9001, phase=1.0),
           dict(system="sbas", prn=131, cn0=50.0, doppler=700.0, delay=123, phase=2.0),
           dict(system="qzss", prn=194, cn0=49.0, doppler=-1200.0, delay=15000, phase=0.0)]
    L = 2 * 16368 + 777
    packed = synth.synth_observation(3, L, FS, sig, fmt=1, seed=31)
    systems = [tga.SystemSearch("qzss", FC, [193, 194]),
               tga.SystemSearch("gps", FC, [17, 5, 9]),
               tga.SystemSearch("sbas", FC, [131, 120])]
    plan = tga.Plan(FS, 3, L, systems, input_format=1)
    return plan, packed, sig, L


def test_acquire_all_json_matches_oracle(obs3):
    plan, packed, _, L = obs3
    ants = [4, 11, 2]
    text = plan.acquire_all_json(packed, IF, ants, cn0=True)
    peaks = plan.search(packed)
    assert text == oracle_json(plan, peaks, IF, ants, True)
    assert text == plan.peaks_json(peaks, IF, ants, cn0=True)
    d = json.loads(text)
    assert list(d) == ["gps", "sbas", "qzss"]  # CombinedOutput field order
    assert d["gps"]["antenna_numbers"] == ants
    assert [r["sv"] for r in d["gps"]["results"]] == ["GPS05", "GPS09", "GPS17"]
    assert [r["sv"] for r in d["sbas"]["results"]] == ["SBAS120", "SBAS131"]
    rec = d["gps"]["results"][0]
    assert list(rec) == ["sv", "strengths", "phases", "freqs", "cn0_acr", "phase_median",
                         "phase_mad", "freq_median", "freq_mad"]
    assert rec["phase_mad"] < 2e-4 and rec["freq_median"] == 2000.0
    assert all(c is not None and 40 < c < 70 for c in rec["cn0_acr"])
    # the peaks behind the JSON are the reference's peaks
    levels = synth.unpack_levels(packed, L, 1)
    P = 16368
    for si, s in enumerate(plan.systems):
        for j, prn in enumerate(plan.prns[si]):
            code = ocodes.code_resampled(s.system, prn, float(P), 2 * P)
            for a in range(3):
                ref, sg, pp, cc, ca = oracle_search(levels[a], FS, P, 2 * P, FC, code)
                assert_peak_matches(peaks[si][a, j], ref, P,
                                    near_ties(sg, pp, cc, ca, FC, 2 * P, P),
                                    what=f"{s.system} {prn} ant {a}")


def test_json_without_cn0_and_filters(obs3):
    plan, packed, _, _ = obs3
    peaks = plan.search(packed)
    plain = plan.peaks_json(peaks, IF, [0, 1, 2])
    assert plain == oracle_json(plan, peaks, IF, [0, 1, 2], False)
    assert "cn0_acr" not in plain
    for fpm, ffm in ((1e-3, None), (None, 100.0), (0.0, 0.0), (1e-3, 250.0)):
        t = plan.peaks_json(peaks, IF, [0, 1, 2], False, fpm, ffm)
        assert t == oracle_json(plan, peaks, IF, [0, 1, 2], False, fpm, ffm)
    kept = json.loads(plan.peaks_json(peaks, IF, [0, 1, 2], False, 1e-3, 250.0))
    svs = {r["sv"] for s in kept.values() for r in s["results"]}
    assert {"GPS05", "GPS17", "SBAS131", "QZSS194"} <= svs  # true signals agree across antennas


def test_json_single_antenna_has_no_stats(gpu):
    sig = [dict(system="gps", prn=3, cn0=50.0, doppler=0.0, delay=50, phase=0.0)]
    L = 2 * 16368
    packed = synth.synth_observation(1, L, FS, sig, fmt=2, seed=5)
    plan = tga.Plan(FS, 1, L, [tga.SystemSearch("gps", FC, [3, 4])], input_format=2)
    peaks = plan.search(packed)
    text = plan.peaks_json(peaks, IF, [7], cn0=True, filter_phase_mad=0.0)
    assert text == oracle_json(plan, peaks, IF, [7], True, 0.0)
    d = json.loads(text)
    assert len(d["gps"]["results"]) == 2  # a missing MAD passes the filter
    assert "phase_median" not in text and "freq_mad" not in text


def test_json_errors(obs3):
    plan, packed, _, _ = obs3
    with pytest.raises(ValueError):
        plan.peaks_json(plan.search(packed), IF, [0, 1])
    with pytest.raises(tga.TgaError):
        plan.acquire_all_json(packed[:2], IF, [0, 1, 2])


# ---- RFI -----------------------------------------------------------------------------------

N_RFI = 64 * 16368  # 64 ms, the reference's test length


def square_243(n):
    per = FS / 2.43e6
    i = np.arange(n, dtype=np.float64)
    return np.where(np.fmod(i, per) < per / 2.0, 1.0, -1.0)


def gps_like(n, rng):
    period = 16368
    code = ocodes.code_resampled("gps", 1, float(period), 2 * period)
    a = 2.0 * np.sqrt(10 ** 4.0 / FS)
    i = np.arange(n)
    return code[(i + 137) % period] * np.cos(2 * np.pi * 100e3 * i / FS) * a + \
        rng.standard_normal(n)


def rfi_channels():
    rng = np.random.default_rng(2024)
    noise = rng.standard_normal((4, N_RFI))
    return np.stack([2.0 * square_243(N_RFI) + noise[0],   # strong square wave
                     noise[1],                               # white noise
                     0.1 * square_243(N_RFI) + noise[2],   # weak square wave, same noise
                     gps_like(N_RFI, rng),                   # spread-spectrum signal
                     -np.ones(N_RFI),                        # dead channel
                     0.1 * square_243(N_RFI) + noise[0]])  # weak, same noise as channel 0


def assert_channel_matches(got, ref, what):
    assert got["dead"] == ref["dead"], what
    assert len(got["lines"]) == len(ref["lines"]), (what, got["lines"], ref["lines"])
    for g, r in zip(got["lines"], ref["lines"]):
        assert g["frequency"] == r["frequency"], what
        assert g["excess_db"] == pytest.approx(r["excess_db"], rel=1e-9), what
    for k in ("acf_lag1", "acf_lag3", "acf_lag16", "period_samples"):
        if ref[k] is None:
            assert got[k] is None, (what, k)
        else:
            assert got[k] == pytest.approx(ref[k], rel=1e-9, abs=1e-12), (what, k)


@pytest.fixture(scope="module")
def rfi_1bit(gpu):
    sig = rfi_channels()
    packed = synth.pack_1bit(sig)
    return packed, tga.rfi(packed, N_RFI, 1, FS)


def test_rfi_matches_oracle_1bit(rfi_1bit):
    packed, got = rfi_1bit
    levels = synth.unpack_levels(packed, N_RFI, 1)
    for a in range(levels.shape[0]):
        ref = orfi.analyze_channel(a, levels[a], FS)
        assert got[a]["antenna"] == a
        assert_channel_matches(got[a], ref, f"channel {a}")


def test_rfi_reference_assertions(rfi_1bit):
    _, ch = rfi_1bit
    sq, white, weak, gps, dead, weak0 = ch
    # rfi.rs:214-233
    assert not sq["dead"]
    top = sq["lines"][0]
    assert abs(top["frequency"] - 2.43e6) < 1e3 and top["excess_db"] > 15.0
    assert abs(sq["period_samples"] - FS / 2.43e6) < 0.2
    assert sq["acf_lag1"] > 0.3 and sq["acf_lag3"] < -0.4 and sq["acf_lag16"] < 0.0
    # rfi.rs:236-246
    assert not white["dead"] and white["lines"] == []
    # rfi.rs:249-268 (same noise for the weak and strong interferer)
    assert weak0["lines"] and sq["lines"][0]["excess_db"] > weak0["lines"][0]["excess_db"] + 10
    assert weak["lines"]
    # rfi.rs:271-287
    assert gps["lines"] == []
    # rfi.rs:290-296
    assert dead["dead"] and dead["lines"] == [] and dead["acf_lag1"] is None


def test_rfi_2bit_and_ragged_length(gpu):
    rng = np.random.default_rng(77)
    n = 61453  # 7 * 8779: a ragged record whose f64 spectrum runs the Bluestein form
    sig = np.stack([1.5 * square_243(n) + rng.standard_normal(n),
                    rng.standard_normal(n) * 0.8])
    packed = synth.pack_2bit(sig, threshold=1.0)
    got = tga.rfi(packed, n, 2, FS)
    levels = synth.unpack_levels(packed, n, 2)
    for a in range(2):
        assert_channel_matches(got[a], orfi.analyze_channel(a, levels[a], FS), f"2bit {a}")


def test_rfi_errors(gpu):
    pk = np.zeros((2, 8), np.uint8)
    with pytest.raises(tga.TgaError):
        tga.rfi(pk, 16, 1, FS)
    with pytest.raises(tga.TgaError):
        tga.rfi(pk, 65, 1, FS)
    with pytest.raises(tga.TgaError):
        tga.rfi(pk, 64, 3, FS)
    with pytest.raises(tga.TgaError):
        tga.rfi(pk, 64, 1, 0.0)


def test_rfi_device_input_matches_host_input(rfi_1bit):
    # tga_rfi_device on the same packed rows already in device memory (torch only allocates)
    import torch
    packed, host = rfi_1bit
    dev = torch.from_numpy(np.ascontiguousarray(packed)).cuda()
    torch.cuda.synchronize()
    got = tga.rfi_device(dev.data_ptr(), packed.shape[0], packed.shape[1], N_RFI, 1, FS)
    assert got == host
