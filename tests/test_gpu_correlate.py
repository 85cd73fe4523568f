"""GPU parity of the level-1/2 entry points (tga_correlate_code, tga_acquire_single,
tga_code_spectra) against the oracle, plus ports of the reference's own tests of the path:
correlate.rs:213-295, acquisition.rs:425-515, sbas.rs:457-518, qzss.rs:457-518,
galileo.rs:400-463, beidou.rs:351-412, l1c.rs:361-424."""
import math

import numpy as np
import pytest

import tga
from tga import synth
from oracle import codes as ocodes
from oracle import correlate as oc
from parity import assert_peak_matches, near_ties

pytestmark = pytest.mark.gpu


def make_case(system, prn, P, N, fs, delay, f0, noise, seed):
    code = ocodes.code_resampled(system, prn, float(P), N)
    long = ocodes.code_resampled(system, prn, float(P), max(N, P))
    sig = synth.synth_signal(long, P, delay, f0, fs, noise, seed)[:N]
    return code, sig


@pytest.mark.parametrize("N,P,fs,nbins", [(256, 256, 1.023e6, 3), (512, 256, 1.023e6, 5),
                                          (2046, 1023, 1.023e6, 7), (4092, 2046, 2.046e6, 9),
                                          (2043, 2046, 2.046e6, 4), (4099, 2046, 2.046e6, 3),
                                          (32736, 16368, 16.368e6, 5),
                                          # large prime factors (Bluestein form): 3 * 11171
                                          (33513, 16368, 16.368e6, 3)])
def test_correlate_code_matches_oracle(gpu, N, P, fs, nbins):
    code, sig = make_case("gps", 7, P, N, fs, 301 % P, 1500.0, 0.8, N + 3)
    pp = oc.phasepoints(fs, N)
    cc, ca = oc.code_spectra(code, P)
    fc = [1500.0 + 500.0 * (i - nbins // 2) for i in range(nbins)]
    ref = oc.correlate_code(sig, pp, cc, ca, fc, N, P, fs)
    got = tga.correlate_code(sig, pp, cc, ca, fc, N, P, fs)
    ties = near_ties(sig, pp, cc, ca, fc, N, P)
    assert_peak_matches(got, ref, P, ties, what=f"N={N}")
    if not any(ties):
        assert got.best_codephase == ref["best_codephase"]  # unfolded: full index in [0, N)


def test_code_spectra_matches_oracle(gpu):
    # 4099 is prime and 33513 = 3 * 11171: both run the Bluestein (chirp-z) form
    for N, P in [(2046, 1023), (32736, 16368), (1500, 1023), (4099, 2046), (33513, 16368)]:
        code = ocodes.code_resampled("gps", 3, float(P), N)
        cc, ca = tga.code_spectra(code, P)
        rc, ra = oc.code_spectra(code, P, mode="fp64")
        for got, ref in ((cc, rc), (ca, ra)):
            err = np.abs(got.astype(np.complex128) - ref)
            assert err.max() <= 1e-5 * np.abs(ref).max(), (N, P)


@pytest.mark.parametrize("delay", [0, 50, 128, 200])
def test_correlate_code_recovers_delay(gpu, delay):
    """correlate.rs:213-240."""
    period = 256
    code = ocodes.code_resampled("gps", 1, float(period), period)
    fs = 1.023e6
    sig = synth.synth_signal(code, period, delay, 0.0, fs, 0.05, 42)
    cc, _ = tga.code_spectra(code, period)
    pk = tga.correlate_code(sig, synth.phasepoints(fs, period), cc, cc, [0.0], period, period,
                            fs)
    assert pk.best_codephase % period == synth.expected_recovered_sample(period, delay)
    assert pk.best_freq_idx == 0 and pk.best_peak > 0


def test_correlate_code_recovers_flipped_second_period(gpu):
    """correlate.rs:242-274."""
    period = 256
    code = ocodes.code_resampled("gps", 1, float(period), 2 * period)
    fs = 1.023e6
    cc, ca = tga.code_spectra(code, period)
    sig = synth.synth_signal_cn0(code, period, 50, 0.0, fs, 60.0, 42, True)
    pk = tga.correlate_code(sig, synth.phasepoints(fs, 2 * period), cc, ca, [0.0], 2 * period,
                            period, fs)
    assert pk.best_codephase % period == synth.expected_recovered_sample(period, 50)
    assert pk.best_peak > 0


def test_correlate_code_noise_only_bounded(gpu):
    """correlate.rs:276-295."""
    period = 256
    code = ocodes.code_resampled("gps", 1, float(period), period)
    fs = 1.023e6
    cc, _ = tga.code_spectra(code, period)
    sig = synth.synth_signal(code, period, 0, 0.0, fs, 0.0, 7)
    pk = tga.correlate_code(sig, synth.phasepoints(fs, period), cc, cc, [0.0], period, period,
                            fs)
    assert pk.best_peak > 0 and math.isfinite(pk.best_peak)
    assert pk.second_peak > 0 and pk.best_codephase < period


ACQ = {"gps": tga.acquire_full, "sbas": tga.acquire_sbas, "qzss": tga.acquire_qzss,
       "galileo": tga.acquire_galileo_single, "beidou": tga.acquire_beidou_single,
       "l1c": tga.acquire_l1c_single}
# (system, prn, fs, period, delay, seed) from the reference's recovery tests
RECOVERY = [("gps", 1, 2.046e6, 2046, 500, 1), ("sbas", 120, 1.023e6, 1023, 400, 41),
            ("qzss", 184, 1.023e6, 1023, 400, 41), ("galileo", 1, 1.023e6, 4092, 900, 11),
            ("beidou", 1, 1.023e6, 10230, 3000, 21), ("l1c", 1, 1.023e6, 10230, 3000, 31)]


@pytest.mark.parametrize("system,prn,fs,period,delay,seed", RECOVERY)
def test_acquire_recovers_delay(gpu, system, prn, fs, period, delay, seed):
    code = ocodes.code_resampled(system, prn, float(period), 2 * period)
    f0 = 1200.0 if system == "gps" else 0.0
    sig = synth.synth_signal(code, period, delay, f0, fs, 0.05, seed)
    r = ACQ[system](sig, synth.phasepoints(fs, len(sig)), fs, 0.0, 3000.0, prn, period)
    assert round(r.codephase_frac * period) == synth.expected_recovered_sample(period, delay)
    assert 0.0 <= r.codephase_frac < 1.0
    if system == "gps":
        assert abs(abs(r.frequency) - 1200.0) <= 250.0
    ref = oc.acquire_single(system, sig, synth.phasepoints(fs, len(sig)), fs, 0.0, 3000.0, prn,
                            period)
    assert abs(r.signal_strength - ref["signal_strength"]) <= 1e-5 * ref["signal_strength"]
    assert r.frequency == ref["frequency"]


# (system, PRN, fs, period, delay, seed, noise-tone step): acquisition.rs, sbas.rs, galileo.rs,
# beidou.rs noise-contrast tests, qzss.rs:473-489 and l1c.rs:377-392
@pytest.mark.parametrize("system,prn,fs,period,delay,seed,w", [
    ("gps", 1, 2.046e6, 2046, 123, 2, 3.3), ("sbas", 121, 1.023e6, 1023, 123, 42, 2.1),
    ("galileo", 1, 1.023e6, 4092, 123, 12, 2.7), ("beidou", 1, 1.023e6, 10230, 123, 22, 1.9),
    ("qzss", 184, 1.023e6, 1023, 300, 52, 3.9), ("l1c", 1, 1.023e6, 10230, 700, 32, 1.3)])
def test_acquire_contrast(gpu, system, prn, fs, period, delay, seed, w):
    code = ocodes.code_resampled(system, prn, float(period), 2 * period)
    n = len(code)
    inj = synth.synth_signal(code, period, delay, 0.0, fs, 0.05, seed)
    noise = (np.sin(np.arange(n) * w) * 0.05).astype(np.float32)
    pp = synth.phasepoints(fs, n)
    a = ACQ[system](inj, pp, fs, 0.0, 3000.0, prn, period)
    b = ACQ[system](noise, pp, fs, 0.0, 3000.0, prn, period)
    assert a.signal_strength > 10.0 * b.signal_strength


@pytest.mark.parametrize("system,prn,fs,period,delay", [("gps", 1, 2.046e6, 2046, 100),
                                                        ("sbas", 120, 1.023e6, 1023, 100),
                                                        ("qzss", 184, 1.023e6, 1023, 100),
                                                        ("galileo", 1, 1.023e6, 4092, 100),
                                                        ("l1c", 1, 1.023e6, 10230, 1000)])
def test_sweep_around_period_boundary(gpu, system, prn, fs, period, delay):
    """acquisition.rs:486-515 and siblings (qzss.rs:491-520, l1c.rs:394-423): lengths
    P-3..P+12 and 2P-3..2P+4, several not whole multiples of the period."""
    lengths = list(range(period - 3, period + 13)) + list(range(2 * period - 3, 2 * period + 5))
    for n in lengths:
        code = ocodes.code_resampled(system, prn, float(period), n)
        assert len(code) == n
        sig = code[(np.arange(n) + delay) % n].astype(np.float32)
        r = ACQ[system](sig, synth.phasepoints(fs, n), fs, 0.0, 3000.0, prn, period)
        assert 0.0 <= r.codephase_frac < 1.0, n
        assert r.signal_strength > 0.0, n


def test_beidou_sweep_long_lengths(gpu):
    """beidou.rs:385-412 at the lengths whose factors include large primes (generic radix)."""
    period, fs = 10230, 1.023e6
    for n in (2 * period - 3, 2 * period + 1, 2 * period + 4, period + 7):
        code = ocodes.code_resampled("beidou", 1, float(period), n)
        sig = code[(np.arange(n) + 1000) % n].astype(np.float32)
        r = tga.acquire_beidou_single(sig, synth.phasepoints(fs, n), fs, 0.0, 3000.0, 1, period)
        assert 0.0 <= r.codephase_frac < 1.0 and r.signal_strength > 0.0, n


def test_error_paths(gpu):
    sig = np.ones(64, np.float32)
    pp = synth.phasepoints(1e6, 64)
    cc = np.ones(64, np.complex64)
    with pytest.raises(tga.TgaError) as e:
        tga.correlate_code(sig, pp, cc, cc, [], 64, 32, 1e6)
    assert e.value.name == "TGA_ERR_EMPTY_BINS"
    bad = sig.copy()
    bad[5] = np.nan
    with pytest.raises(tga.TgaError) as e:
        tga.correlate_code(bad, pp, cc, cc, [0.0], 64, 32, 1e6)
    assert e.value.name == "TGA_ERR_NAN"
    with pytest.raises(tga.TgaError) as e:
        tga.acquire_full(sig, pp, 1e6, 0.0, 3000.0, 39, 32)
    assert e.value.name == "TGA_ERR_PRN_RANGE"
    # the library stays usable after errors
    pk = tga.correlate_code(sig, pp, cc, cc, [0.0], 64, 32, 1e6)
    assert pk.best_peak > 0
