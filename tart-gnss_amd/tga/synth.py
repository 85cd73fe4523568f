"""Seeded synthetic inputs for tests and the benchmark.

Two families:
  * ``synth_signal`` / ``synth_signal_cn0`` / ``phasepoints`` / ``quantize_bipolar_demean``
    follow the reference's test fixtures exactly (testutil.rs:10-81 with its xorshift64 PRNG,
    simulator.rs:237-244), for the ports of the reference's recovery and C/N0 tests.
  * ``synth_observation`` builds multi-antenna packed observations (1- or 2-bit) for the
    benchmark workloads: sum of GNSS signals at given C/N0 (amplitude 2*sqrt(10^(C/N0/10)/fs)
    over unit-variance noise, testutil.rs:60-70), a real carrier at IF + Doppler, per-antenna
    carrier-phase offsets from a plane wave over antenna positions in a disk
    (simulator.rs:101-140 conventions), optional CW tones.  Noise comes from numpy's PCG64
    seeded with the given seed (vectorised; the reference's scalar xorshift would dominate the
    benchmark's setup time).
"""
import math

import numpy as np

from . import api

TAU = 2.0 * math.pi
L1_HZ = 1575.42e6
V_LIGHT = 2.99793e8  # simulator.rs:14
DEFAULT_SEED = 0x5EED_2026  # simulator.rs:18


class XorShift:
    """testutil.rs:84-100 (seed used as is)."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_f64(self):
        x = self.s
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = x
        return (x >> 11) / float(1 << 53)

    def gaussian(self):
        u1 = max(self.next_f64(), 1e-12)
        u2 = self.next_f64()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(TAU * u2)


def phasepoints(fs, n):
    """testutil.rs:10-13."""
    return (TAU / fs) * np.arange(n, dtype=np.float64)


def expected_recovered_sample(period, delay):
    """testutil.rs:45-47."""
    return (period - delay) % period


def synth_signal(code, code_period, delay, f0, fs, noise_amp, seed):
    """testutil.rs:21-40."""
    n = len(code)
    rng = XorShift(seed)
    out = np.empty(n, dtype=np.float32)
    for i in range(n):
        cp = (i + delay) % code_period
        carrier = math.cos(TAU * f0 * i / fs)
        noise = noise_amp * (2.0 * rng.next_f64() - 1.0)
        out[i] = np.float32(code[cp] * carrier + noise)
    return out


def synth_signal_cn0(code, period, delay, f0, fs, cn0_db, seed, flip):
    """testutil.rs:60-81."""
    a = 2.0 * math.sqrt(10.0 ** (cn0_db / 10.0) / fs)
    epoch = period + (period - delay) % period
    rng = XorShift(seed)
    n = len(code)
    out = np.empty(n, dtype=np.float32)
    for i in range(n):
        cp = (i + delay) % period
        sign = -1.0 if (flip and i >= epoch) else 1.0
        carrier = a * math.cos(TAU * f0 * i / fs)
        out[i] = np.float32(sign * code[cp] * carrier + rng.gaussian())
    return out


def quantize_bipolar_demean(signal):
    """simulator.rs:237-244: v >= 0 -> +1 else -1, minus the mean (f64), as f32."""
    b = np.where(np.asarray(signal, dtype=np.float64) >= 0.0, 1.0, -1.0)
    return (b - b.sum() / len(b)).astype(np.float32)


def pack_1bit(v):
    """v >= 0 -> bit 1 (simulator.rs:228-230), MSB-first (simulator.rs:280-292)."""
    bits = (np.asarray(v) >= 0.0).astype(np.uint8)
    return np.packbits(bits, axis=-1)


def pack_2bit(v, threshold=1.0):
    """2-bit sign/magnitude: sign bit (v >= 0), magnitude bit (|v| >= threshold); pairs
    MSB-first, four samples per byte."""
    v = np.asarray(v)
    sign = (v >= 0.0).astype(np.uint8)
    mag = (np.abs(v) >= threshold).astype(np.uint8)
    pair = (sign << 1) | mag
    n = pair.shape[-1]
    pad = (-n) % 4
    if pad:
        pair = np.concatenate([pair, np.zeros(pair.shape[:-1] + (pad,), np.uint8)], axis=-1)
    p = pair.reshape(pair.shape[:-1] + (-1, 4))
    return ((p[..., 0] << 6) | (p[..., 1] << 4) | (p[..., 2] << 2) | p[..., 3]).astype(np.uint8)


def unpack_levels(packed, n, fmt):
    """Inverse of the packers: integer levels (+-1 or +-1/+-3)."""
    packed = np.asarray(packed, dtype=np.uint8)
    if fmt == 1:
        bits = np.unpackbits(packed, axis=-1)[..., :n]
        return bits.astype(np.int8) * 2 - 1
    b = packed[..., :, None] >> np.array([6, 4, 2, 0], np.uint8) & 3
    pair = b.reshape(packed.shape[:-1] + (-1,))[..., :n]
    mag = np.where(pair & 1, 3, 1)
    return np.where(pair & 2, mag, -mag).astype(np.int8)


def random_positions(n, diameter, rng):
    r = diameter / 2.0
    rr = r * np.sqrt(rng.random(n))
    th = TAU * rng.random(n)
    return np.stack([rr * np.cos(th), rr * np.sin(th), np.zeros(n)], axis=1)


# Galileo E1-C secondary code CS25 (Galileo OS SIS ICD, E1-C secondary code "380AD90" hex, first
# 25 bits), one chip per 4 ms primary period; bit 1 -> -1.
E1C_CS25 = "0011100000001010110110010"


def e1c_boc11_code(prn, samples_per_code, num_samples):
    """E1-C chips with the BOC(1,1) subcarrier ([+c, -c] half-chips), resampled like the
    reference's replicas: c[n] = half_chips[floor(n / (spc / 8184)) mod 8184]."""
    chips = api.code_chips("galileo", prn).astype(np.float64)
    half = np.empty(2 * len(chips))
    half[0::2] = chips
    half[1::2] = -chips
    spchip = float(samples_per_code) / len(half)
    idx = np.floor(np.arange(num_samples, dtype=np.float64) / spchip).astype(np.int64) % len(half)
    return half[idx]


def synth_observation(n_ant, L, fs, signals, fmt=1, seed=DEFAULT_SEED, intermediate=4.092e6,
                      tones=None, positions=None, threshold=1.0):
    """Packed observation [n_ant, row_bytes].

    signals: list of dicts {system, prn, cn0, doppler, delay (samples), phase (rad),
             az, el (deg, for the per-antenna plane-wave carrier offset), flip (bool)};
             Galileo only: boc (bool, BOC(1,1) subcarrier on the chips) and secondary (the
             CS25 chip of the first period, 0..24: each later period carries the next chip).
    tones:   optional list per antenna of (frequency Hz, amplitude) CW interferers.
    """
    rng = np.random.default_rng(seed)
    if positions is None:
        positions = random_positions(n_ant, 3.0, rng)
    n = np.arange(L, dtype=np.float64)
    base = []
    for s in signals:
        P = int(fs / 1000.0) if s["system"] in ("gps", "sbas", "qzss") else int(
            fs * {"galileo": 0.004, "beidou": 0.010, "l1c": 0.010}[s["system"]])
        if s.get("boc"):
            code = e1c_boc11_code(s["prn"], float(P), P)
        else:
            code = api.code_resampled(s["system"], s["prn"], float(P), P)
        c = code[(np.arange(L) + int(s["delay"])) % P]
        if s.get("secondary") is not None:  # one CS25 chip per primary period
            k = (np.arange(L) + int(s["delay"])) // P + int(s["secondary"])
            sec = np.array([1.0 - 2.0 * int(b) for b in E1C_CS25])
            c = c * sec[k % 25]
        if s.get("flip"):
            epoch = P + (P - int(s["delay"])) % P
            c = c.copy()
            c[epoch:] *= -1.0
        az, el = math.radians(s.get("az", 0.0)), math.radians(s.get("el", 90.0))
        u = np.array([math.sin(az) * math.cos(el), math.cos(az) * math.cos(el), math.sin(el)])
        base.append((c, 2.0 * math.sqrt(10.0 ** (s["cn0"] / 10.0) / fs),
                     TAU * (intermediate + s["doppler"]) / fs, s.get("phase", 0.0), u))
    rows = []
    for a in range(n_ant):
        x = rng.standard_normal(L)
        for c, amp, w, ph, u in base:
            dt = -float(np.dot(positions[a], u)) / V_LIGHT  # plane-wave arrival offset
            x += amp * c * np.cos(w * n + ph + TAU * L1_HZ * dt)
        if tones is not None:
            for f, amp in tones[a]:
                x += amp * np.cos(TAU * f / fs * n + rng.random() * TAU)
        rows.append(pack_1bit(x) if fmt == 1 else pack_2bit(x, threshold))
    return np.stack(rows)
