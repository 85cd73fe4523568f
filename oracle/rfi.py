"""Narrowband RFI report (oracle restatement of rfi.rs:83-184)."""
import numpy as np

MIN_EXCESS_DB = 15.0   # rfi.rs:36
CLUSTER_HZ = 10_000.0  # rfi.rs:41
MAX_LINES = 3          # rfi.rs:43


def autocorrelation(z, lag):
    """rfi.rs:83-86: sum z[i] z[i+lag] / (n - lag)."""
    n = len(z)
    return float(np.dot(z[:n - lag], z[lag:])) / (n - lag)


def power_spectrum(z):
    """|FFT_f64(z)|^2 for k < n/2 (rfi.rs:96-103)."""
    X = np.fft.fft(np.asarray(z, dtype=np.float64))
    half = len(z) // 2
    return X.real[:half] ** 2 + X.imag[:half] ** 2


def detect_lines(z, fs):
    """rfi.rs:91-138: the line rules applied to |FFT_f64(z)|^2."""
    n = len(z)
    if n < 8:
        return []
    return lines_from_powers(power_spectrum(z), n, fs)


def lines_from_powers(p, n, fs):
    """The tail of detect_lines (rfi.rs:105-137) on the powers p[k], k < n/2, of an n-point
    record: floor = sorted[half/2]; local maxima p[k] > p[k-1] and p[k] >= p[k+1] for
    1 <= k < half-1 more than 15 dB above the floor; strongest first (stable); a line is kept
    only if it lies more than 10 kHz from every kept line; at most 3."""
    half = n // 2
    bin_hz = fs / n
    p = np.asarray(p, dtype=np.float64)
    assert len(p) == half
    floor = np.sort(p)[half // 2]
    lines = []
    k = np.arange(1, half - 1)
    is_max = (p[k] > p[k - 1]) & (p[k] >= p[k + 1])
    for kk in k[is_max]:
        with np.errstate(divide="ignore"):
            excess = 10.0 * np.log10(p[kk] / floor)
        if excess > MIN_EXCESS_DB:
            lines.append({"frequency": kk * bin_hz, "excess_db": float(excess)})
    lines.sort(key=lambda l: -l["excess_db"])  # stable, like Vec::sort_by
    out = []
    for ln in lines:
        if all(abs(c["frequency"] - ln["frequency"]) > CLUSTER_HZ for c in out):
            out.append(ln)
            if len(out) == MAX_LINES:
                break
    return out


def analyze_channel(antenna, bipolar, fs):
    """rfi.rs:142-168."""
    b = np.asarray(bipolar, dtype=np.float64)
    z = b - b.sum() / len(b)
    var = float((z * z).sum()) / len(z)
    if var < 1e-9:
        return {"antenna": antenna, "dead": True, "acf_lag1": None, "acf_lag3": None,
                "acf_lag16": None, "period_samples": None, "lines": []}
    lines = detect_lines(z, fs)
    return {"antenna": antenna, "dead": False,
            "acf_lag1": autocorrelation(z, 1), "acf_lag3": autocorrelation(z, 3),
            "acf_lag16": autocorrelation(z, 16),
            "period_samples": (fs / lines[0]["frequency"]) if lines else None,
            "lines": lines}


def report(channels):
    """RfiReport (rfi.rs:73-77, built by rfi::run at rfi.rs:186-194): antenna indices in order
    and the channels, as dicts in the struct's field order."""
    return {"antenna_numbers": [c["antenna"] for c in channels], "channels": list(channels)}


def report_json(channels):
    """serde_json::to_string_pretty of the report (main.rs:468)."""
    from .jsonfmt import dumps_pretty
    return dumps_pretty(report(channels))
