"""High-level wrappers over the C ABI (see the package docstring for the reference mapping)."""
import ctypes as C
import json
import math
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import check, lib


@dataclass
class CorrelationPeak:
    """correlate.rs:27-38."""
    best_peak: float
    second_peak: float
    best_freq_idx: int
    best_codephase: int


@dataclass
class AcquisitionResult:
    """acquisition.rs:19-30 (and the per-system twins)."""
    prn: int
    signal_strength: float
    second_peak: float
    codephase_frac: float
    frequency: float


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def device_count():
    n = C.c_int(0)
    check(lib().tga_device_count(C.byref(n)))
    return n.value


def _sys(system):
    return L.SYSTEM_IDS[system] if isinstance(system, str) else int(system)


def code_length(system):
    n = C.c_size_t(0)
    check(lib().tga_code_length(_sys(system), C.byref(n)))
    return n.value


def code_chips(system, prn):
    n = code_length(system)
    out = np.empty(n, dtype=np.int8)
    check(lib().tga_code_chips(_sys(system), int(prn), _ptr(out, C.c_int8), n))
    return out


def code_resampled(system, prn, samples_per_code, num_samples):
    out = np.empty(int(num_samples), dtype=np.float64)
    check(lib().tga_code_resampled(_sys(system), int(prn), float(samples_per_code),
                                   int(num_samples), _ptr(out, C.c_double)))
    return out


def gold_code(samples_per_code, prn, num_samples):
    """acquisition.rs:162-165."""
    return code_resampled("gps", prn, samples_per_code, num_samples)


def sbas_gold_code(samples_per_code, prn, num_samples):
    return code_resampled("sbas", prn, samples_per_code, num_samples)


def qzss_gold_code(samples_per_code, prn, num_samples):
    return code_resampled("qzss", prn, samples_per_code, num_samples)


def e1c_code_resampled(samples_per_code, prn, num_samples):
    return code_resampled("galileo", prn, samples_per_code, num_samples)


def b1c_code_resampled(samples_per_code, prn, num_samples):
    return code_resampled("beidou", prn, samples_per_code, num_samples)


def l1c_code_resampled(samples_per_code, prn, num_samples):
    return code_resampled("l1c", prn, samples_per_code, num_samples)


def code_spectra(code, period):
    """correlate.rs:47-58 -> (conj FFT([c,c]), conj FFT([c,-c])) as complex64 arrays."""
    code = _f64(code)
    n = len(code)
    cc = np.empty(n, dtype=np.complex64)
    ca = np.empty(n, dtype=np.complex64)
    check(lib().tga_code_spectra(_ptr(code, C.c_double), n, int(period), _ptr(cc, C.c_float),
                                 _ptr(ca, C.c_float)))
    return cc, ca


def correlate_code(signal_f32, phasepoints, code_complex, code_complex_alt, fc, num_samples,
                   code_period_samples, sampling_freq):
    """correlate.rs:91-183 on the GPU."""
    s = _f32(signal_f32)
    pp = _f64(phasepoints)
    cc = np.ascontiguousarray(code_complex, dtype=np.complex64)
    ca = np.ascontiguousarray(code_complex_alt, dtype=np.complex64)
    fcs = _f64(fc)
    n = int(num_samples)
    if len(s) < n or len(pp) < n or len(cc) < n or len(ca) < n:
        raise ValueError("signal, phasepoints and spectra need num_samples entries")
    pk = L.Peak()
    check(lib().tga_correlate_code(_ptr(s, C.c_float), _ptr(pp, C.c_double),
                                   _ptr(cc, C.c_float), _ptr(ca, C.c_float),
                                   _ptr(fcs, C.c_double) if len(fcs) else None, len(fcs), n,
                                   int(code_period_samples), float(sampling_freq),
                                   C.byref(pk)))
    return CorrelationPeak(pk.best_peak, pk.second_peak, pk.best_freq_idx, pk.best_codephase)


def correlate_code_noalloc(d_signal, d_phasepoints, d_code_complex, d_code_complex_alt, d_fc,
                           n_fc, num_samples, code_period_samples, sampling_freq, d_out,
                           d_nan=None, stream=None):
    """correlate_code on device-resident inputs with no device allocation (integer device
    addresses; the peak is written to d_out, a tga_peak in device memory, the NaN flag to d_nan).
    Only enqueues work on `stream`; the caller synchronises."""
    check(lib().tga_correlate_code_noalloc(
        C.c_void_p(int(d_signal)), C.c_void_p(int(d_phasepoints)),
        C.c_void_p(int(d_code_complex)), C.c_void_p(int(d_code_complex_alt)),
        C.c_void_p(int(d_fc)), int(n_fc), int(num_samples), int(code_period_samples),
        float(sampling_freq), C.c_void_p(int(d_out)),
        C.c_void_p(int(d_nan)) if d_nan else None,
        C.c_void_p(int(stream)) if stream else None))


def search_frequencies(center_freq, search_band, freq_bin_size=300.0):
    """The bin grid of acquire_full and its siblings (acquisition.rs:189-198): round(2 band /
    300 Hz) + 1 bins spread evenly over [center - band, center + band]; a Plan's grid."""
    x = 2.0 * search_band / freq_bin_size
    n = math.floor(x) + (1 if x - math.floor(x) >= 0.5 else 0) + 1  # f64::round: half away
    if n == 1:
        return [float(center_freq)]
    return [center_freq - search_band + i * (2.0 * search_band) / (n - 1.0) for i in range(n)]


def acquire_single(system, signal_f32, phasepoints, sampling_freq, center_freq, search_band,
                   prn, samples_per_period):
    s = _f32(signal_f32)
    pp = _f64(phasepoints)
    if len(pp) < len(s):
        raise ValueError("phasepoints shorter than the signal")
    r = L.AcqResult()
    check(lib().tga_acquire_single(_sys(system), _ptr(s, C.c_float), _ptr(pp, C.c_double),
                                   len(s), float(sampling_freq), float(center_freq),
                                   float(search_band), int(prn), int(samples_per_period),
                                   C.byref(r)))
    return AcquisitionResult(r.prn, r.signal_strength, r.second_peak, r.codephase_frac,
                             r.frequency)


def acquire_full(signal_f32, phasepoints, sampling_freq, center_freq, search_band, prn,
                 samples_per_chunk):
    """acquisition.rs:176-248."""
    return acquire_single("gps", signal_f32, phasepoints, sampling_freq, center_freq,
                          search_band, prn, samples_per_chunk)


def acquire_sbas(signal_f32, phasepoints, sampling_freq, center_freq, search_band, prn,
                 samples_per_chunk):
    return acquire_single("sbas", signal_f32, phasepoints, sampling_freq, center_freq,
                          search_band, prn, samples_per_chunk)


def acquire_qzss(signal_f32, phasepoints, sampling_freq, center_freq, search_band, prn,
                 samples_per_chunk):
    return acquire_single("qzss", signal_f32, phasepoints, sampling_freq, center_freq,
                          search_band, prn, samples_per_chunk)


def acquire_galileo_single(signal_f32, phasepoints, sampling_freq, center_freq, search_band,
                           prn, samples_per_code_period):
    return acquire_single("galileo", signal_f32, phasepoints, sampling_freq, center_freq,
                          search_band, prn, samples_per_code_period)


def acquire_beidou_single(signal_f32, phasepoints, sampling_freq, center_freq, search_band,
                          prn, samples_per_code_period):
    return acquire_single("beidou", signal_f32, phasepoints, sampling_freq, center_freq,
                          search_band, prn, samples_per_code_period)


def acquire_l1c_single(signal_f32, phasepoints, sampling_freq, center_freq, search_band, prn,
                       samples_per_code_period):
    return acquire_single("l1c", signal_f32, phasepoints, sampling_freq, center_freq,
                          search_band, prn, samples_per_code_period)


def estimate_cn0(r_a, table):
    """acr.rs:49-84; ``table`` is a name from ACR_TABLES or its index.  None for no estimate."""
    t = L.ACR_TABLES[table] if isinstance(table, str) else int(table)
    c = C.c_double(0)
    v = C.c_int32(0)
    check(lib().tga_estimate_cn0(float(r_a), t, C.byref(c), C.byref(v)))
    return c.value if v.value else None


@dataclass
class SystemSearch:
    """One system of a batched plan: PRNs (None = full range), absolute bin frequencies,
    coherent length (0 = two code periods, as the reference), and nc_epochs > 0 for the
    non-coherent form (that many one-period epochs, powers summed).  replica "boc11" selects
    the BOC(1,1) E1-C replica (Galileo only, TGA_REPLICA_E1C_BOC11) instead of the reference's
    plain chips."""
    system: str
    fc: list
    prns: list = None
    num_samples: int = 0
    nc_epochs: int = 0
    replica: str = "reference"


PEAK_DTYPE = np.dtype([("best_peak", np.float32), ("second_peak", np.float32),
                       ("best_freq_idx", np.uint64), ("best_codephase", np.uint64)])


class Plan:
    """Batched search plan (tga_plan_*): one observation of n_ant packed rows per search.
    ``precision`` "fp32" is the reference form; "fp16" the f16-storage matrix-core form of the
    fused P = 16368 path (other groups stay fp32)."""

    def __init__(self, sampling_freq, n_ant, record_samples, systems, input_format=1,
                 precision="fp32"):
        self._keep = []
        descs = (L.SystemDesc * len(systems))()
        for i, s in enumerate(systems):
            fc = _f64(s.fc)
            self._keep.append(fc)
            descs[i].system = _sys(s.system)
            descs[i].n_fc = len(fc)
            descs[i].fc = _ptr(fc, C.c_double)
            descs[i].num_samples = int(s.num_samples)
            descs[i].nc_epochs = int(s.nc_epochs)
            descs[i].replica = L.REPLICAS[s.replica]
            if s.prns is None:
                descs[i].n_prn = 0
                descs[i].prns = None
            else:
                pr = np.ascontiguousarray(s.prns, dtype=np.int32)
                self._keep.append(pr)
                descs[i].n_prn = len(pr)
                descs[i].prns = _ptr(pr, C.c_int32)
        d = L.PlanDesc()
        d.sampling_freq = float(sampling_freq)
        d.n_ant = int(n_ant)
        d.input_format = int(input_format)
        d.record_samples = int(record_samples)
        d.n_systems = len(systems)
        d.precision = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        d.systems = descs
        h = C.c_void_p()
        check(lib().tga_plan_create(C.byref(d), C.byref(h)))
        self._h = h
        self.systems = systems
        self.n_ant = int(n_ant)
        rb = C.c_size_t(0)
        check(lib().tga_plan_row_bytes(self._h, C.byref(rb)))
        self.row_bytes = rb.value
        n = C.c_size_t(0)
        check(lib().tga_plan_peak_count(self._h, C.byref(n)))
        self.n_peaks = n.value
        self.prns = []
        for i in range(len(systems)):
            k = C.c_size_t(0)
            check(lib().tga_plan_prns(self._h, i, None, 0, C.byref(k)))
            arr = np.empty(k.value, dtype=np.int32)
            check(lib().tga_plan_prns(self._h, i, _ptr(arr, C.c_int32), k.value, C.byref(k)))
            self.prns.append(arr)

    def close(self):
        if getattr(self, "_h", None):
            check(lib().tga_plan_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _split(self, flat):
        out, off = [], 0
        for pr in self.prns:
            n = self.n_ant * len(pr)
            out.append(flat[off:off + n].reshape(self.n_ant, len(pr)))
            off += n
        return out

    def search(self, packed):
        """packed: uint8 [n_ant, row_bytes] host array -> list (per system) of structured
        arrays [n_ant, n_prn] of PEAK_DTYPE."""
        pk = np.ascontiguousarray(packed, dtype=np.uint8)
        if pk.ndim != 2:
            raise ValueError("packed must be [n_ant, row_bytes]")
        out = np.zeros(self.n_peaks, dtype=PEAK_DTYPE)
        check(lib().tga_search(self._h, _ptr(pk, C.c_uint8), pk.shape[0], pk.shape[1],
                               out.ctypes.data_as(C.POINTER(L.Peak))))
        return self._split(out)

    def search_device(self, dev_ptr, stream=None):
        """Search an observation already in device memory (an integer device address)."""
        out = np.zeros(self.n_peaks, dtype=PEAK_DTYPE)
        check(lib().tga_search_device(self._h, C.c_void_p(dev_ptr),
                                      C.c_void_p(stream) if stream else None,
                                      out.ctypes.data_as(C.POINTER(L.Peak))))
        return self._split(out)

    def carrier(self, sys_index=0):
        """The carrier table of system sys_index's fused group (tga_plan_carrier): complex64
        [bins, samples]."""
        n = C.c_size_t(0)
        check(lib().tga_plan_carrier(self._h, int(sys_index), None, 0, C.byref(n)))
        out = np.empty(n.value, dtype=np.float32)
        check(lib().tga_plan_carrier(self._h, int(sys_index),
                                     out.ctypes.data_as(C.POINTER(C.c_float)), n.value,
                                     C.byref(n)))
        return out.view(np.complex64).reshape(len(self.systems[sys_index].fc), -1)

    def surface_size(self, sys_index=-1):
        """Floats of one system's correlation surface (all systems' for -1)."""
        n = C.c_size_t(0)
        check(lib().tga_plan_surface_size(self._h, int(sys_index), C.byref(n)))
        return n.value

    def surface_device(self, dev_ptr, d_surface, cap, stream=None):
        """Correlation surface of a search of the observation at dev_ptr (tga_plan_surface):
        writes cap >= surface_size() floats at the device address d_surface and returns the
        search's peaks."""
        out = np.zeros(self.n_peaks, dtype=PEAK_DTYPE)
        check(lib().tga_plan_surface(self._h, C.c_void_p(dev_ptr),
                                     C.c_void_p(stream) if stream else None,
                                     C.cast(C.c_void_p(d_surface), C.POINTER(C.c_float)),
                                     int(cap), out.ctypes.data_as(C.POINTER(L.Peak))))
        return self._split(out)

    def surface(self, packed):
        """packed: uint8 [n_ant, row_bytes] host array -> (surfaces, peaks): per system a float32
        array [n_ant, n_prn, n_bin, 2, P] of the rows' values at code phases 0..P-1 on the
        best_peak scale, and the search's peaks.  Device memory comes from torch."""
        import torch
        pk = np.ascontiguousarray(packed, dtype=np.uint8)
        if pk.ndim != 2 or pk.shape != (self.n_ant, self.row_bytes):
            raise ValueError("packed must be [n_ant, row_bytes]")
        dev = torch.from_numpy(pk).cuda()
        n = self.surface_size()
        surf = torch.empty(n, dtype=torch.float32, device=dev.device)
        torch.cuda.synchronize()
        peaks = self.surface_device(dev.data_ptr(), surf.data_ptr(), n)
        flat = surf.cpu().numpy()
        out, off = [], 0
        for i, s in enumerate(self.systems):
            k = self.surface_size(i)
            n_bin = len(s.fc)
            P = k // (self.n_ant * len(self.prns[i]) * n_bin * 2)
            out.append(flat[off:off + k].reshape(self.n_ant, len(self.prns[i]), n_bin, 2, P))
            off += k
        return out, peaks

    def set_timing(self, enable=True):
        check(lib().tga_plan_set_timing(self._h, 1 if enable else 0))

    def set_graph(self, enable=True):
        """hipGraph replay of the whole search from the second search on (default on)."""
        check(lib().tga_plan_set_graph(self._h, 1 if enable else 0))

    def timing(self, kernel=0):
        ms = C.c_double(0)
        n = C.c_int64(0)
        b = C.c_double(0)
        check(lib().tga_plan_get_timing(self._h, int(kernel), C.byref(ms), C.byref(n),
                                        C.byref(b)))
        return ms.value, n.value, b.value

    def rows_matrix_flops(self):
        """(f32, f16) matrix-core FLOP of one search's correlation rows, from the kernels the
        plan's groups run (v_mfma_f32_16x16x4f32 and v_mfma_f32_16x16x32_f16 instructions)."""
        f32 = C.c_double(0)
        f16 = C.c_double(0)
        check(lib().tga_plan_rows_matrix_flops(self._h, C.byref(f32), C.byref(f16)))
        return f32.value, f16.value

    def kernel_name(self, kernel=0):
        return lib().tga_plan_kernel_name(self._h, int(kernel)).decode()

    def describe(self):
        buf = C.create_string_buffer(4096)
        check(lib().tga_plan_describe(self._h, buf, 4096))
        return buf.value.decode()

    def _json(self, fn, first, center_freq, antenna_numbers, tail):
        cf = _f64(center_freq if np.ndim(center_freq) else [center_freq] * len(self.systems))
        an = np.ascontiguousarray(antenna_numbers if antenna_numbers is not None
                                  else range(self.n_ant), dtype=np.int64)
        if len(cf) != len(self.systems) or len(an) != self.n_ant:
            raise ValueError("center_freq per system and antenna_numbers per antenna")
        out = C.c_void_p()
        check(fn(self._h, *first, _ptr(cf, C.c_double), _ptr(an, C.c_int64), *tail,
                 C.byref(out)))
        return _take_string(out)

    @staticmethod
    def _filters(filter_phase_mad, filter_freq_mad):
        return (-1.0 if filter_phase_mad is None else float(filter_phase_mad),
                -1.0 if filter_freq_mad is None else float(filter_freq_mad))

    def _packed_first(self, packed):
        pk = np.ascontiguousarray(packed, dtype=np.uint8)
        return (_ptr(pk, C.c_uint8), pk.shape[0], pk.shape[1]), pk

    @staticmethod
    def _peaks_first(peaks):
        flat = np.ascontiguousarray(np.concatenate([p.reshape(-1) for p in peaks]))
        return (flat.ctypes.data_as(C.POINTER(L.Peak)),), flat

    def acquire_all_json(self, packed, center_freq=4.092e6, antenna_numbers=None, cn0=False,
                         filter_phase_mad=None, filter_freq_mad=None):
        """Search one observation and return the reference's CombinedOutput JSON text."""
        first, _keep = self._packed_first(packed)
        return self._json(lib().tga_acquire_all_json, first, center_freq, antenna_numbers,
                          (1 if cn0 else 0, *self._filters(filter_phase_mad, filter_freq_mad)))

    def peaks_json(self, peaks, center_freq=4.092e6, antenna_numbers=None, cn0=False,
                   filter_phase_mad=None, filter_freq_mad=None):
        """JSON of peaks returned by search() (list per system of [n_ant, n_prn] arrays)."""
        first, _keep = self._peaks_first(peaks)
        return self._json(lib().tga_peaks_json, first, center_freq, antenna_numbers,
                          (1 if cn0 else 0, *self._filters(filter_phase_mad, filter_freq_mad)))

    def cn0(self, peaks):
        """ACR C/N0 (dB-Hz, f64) of search() peaks, per system an [n_ant, n_prn] array with
        NaN where the reference has None (acquisition.rs:341-357); computed natively."""
        first, flat = self._peaks_first(peaks)
        out = np.empty(flat.shape[0], dtype=np.float64)
        check(lib().tga_peaks_cn0(self._h, *first, _ptr(out, C.c_double)))
        res, o = [], 0
        for p in peaks:
            res.append(out[o:o + p.size].reshape(p.shape))
            o += p.size
        return res

    def test_antennas_json(self, packed, center_freq=4.092e6, antenna_numbers=None,
                           min_cn0=44.0, filter_phase_mad=None, filter_freq_mad=None):
        """Search one observation and return the --test-antennas report (main.rs:848-880):
        C/N0 estimates forced on, MAD filters first, then antenna_test::run."""
        first, _keep = self._packed_first(packed)
        return self._json(lib().tga_acquire_antenna_report_json, first, center_freq,
                          antenna_numbers,
                          (*self._filters(filter_phase_mad, filter_freq_mad), float(min_cn0)))

    def peaks_test_antennas_json(self, peaks, center_freq=4.092e6, antenna_numbers=None,
                                 min_cn0=44.0, filter_phase_mad=None, filter_freq_mad=None):
        """The --test-antennas report of peaks returned by search()."""
        first, _keep = self._peaks_first(peaks)
        return self._json(lib().tga_peaks_antenna_report_json, first, center_freq,
                          antenna_numbers,
                          (*self._filters(filter_phase_mad, filter_freq_mad), float(min_cn0)))


def _take_string(out):
    try:
        return C.string_at(out.value).decode()
    finally:
        lib().tga_free_string(out)


def antenna_test_json(combined, min_cn0=44.0):
    """antenna_test::run (antenna_test.rs:105-222) on a CombinedOutput -- the dict (or JSON
    text) that acquire_all_json writes -- returned as the report's JSON text.  Host-only."""
    if isinstance(combined, (str, bytes)):
        combined = json.loads(combined)
    antenna_numbers, svs, rows = [], [], []
    for key in ("gps", "galileo", "beidou", "sbas", "l1c", "qzss"):  # lib.rs:33-44
        out = combined.get(key)
        if out is None:
            continue
        if not antenna_numbers:
            antenna_numbers = list(out["antenna_numbers"])
        for r in out["results"]:
            cn0 = r.get("cn0_acr")
            if cn0 is None:
                continue
            svs.append(r["sv"])
            rows.append([np.nan if c is None else float(c) for c in cn0])
    n_ant = len(antenna_numbers)
    if any(len(r) != n_ant for r in rows):
        raise ValueError("cn0_acr rows must have one entry per antenna")
    an = np.ascontiguousarray(antenna_numbers, dtype=np.int64)
    cn = np.ascontiguousarray(np.array(rows, dtype=np.float64).reshape(len(rows), n_ant))
    labels = (C.c_char_p * max(1, len(svs)))(*[s.encode() for s in svs])
    out = C.c_void_p()
    check(lib().tga_antenna_report_json(len(svs), labels, _ptr(cn, C.c_double), n_ant,
                                        _ptr(an, C.c_int64), float(min_cn0), C.byref(out)))
    return _take_string(out)


def format_f64(x):
    """The JSON text serde_json writes for an f64 (tga_format_f64)."""
    buf = C.create_string_buffer(64)
    check(lib().tga_format_f64(float(x), buf, 64))
    return buf.value.decode()


def _rfi_channels(out):
    def opt(x):
        return None if np.isnan(x) else float(x)

    return [{"antenna": ch.antenna, "dead": bool(ch.dead), "acf_lag1": opt(ch.acf_lag1),
             "acf_lag3": opt(ch.acf_lag3), "acf_lag16": opt(ch.acf_lag16),
             "period_samples": opt(ch.period_samples),
             "lines": [{"frequency": ch.lines[i].frequency, "excess_db": ch.lines[i].excess_db}
                       for i in range(ch.n_lines)]} for ch in out]


def rfi(packed, record_samples, input_format, sampling_freq):
    """Narrowband RFI report per antenna (rfi.rs:142-184) on the GPU; a list of dicts shaped
    like the reference's RfiChannel (None where the reference has None)."""
    pk = np.ascontiguousarray(packed, dtype=np.uint8)
    if pk.ndim != 2:
        raise ValueError("packed must be [n_ant, row_bytes]")
    out = (L.RfiChannel * pk.shape[0])()
    check(lib().tga_rfi(_ptr(pk, C.c_uint8), pk.shape[0], pk.shape[1], int(record_samples),
                        int(input_format), float(sampling_freq), out))
    return _rfi_channels(out)


def rfi_device(dev_ptr, n_ant, row_bytes, record_samples, input_format, sampling_freq):
    """rfi() for an observation already in device memory (an integer device address)."""
    out = (L.RfiChannel * int(n_ant))()
    check(lib().tga_rfi_device(C.c_void_p(int(dev_ptr)), int(n_ant), int(row_bytes),
                               int(record_samples), int(input_format), float(sampling_freq),
                               out))
    return _rfi_channels(out)


def _rfi_structs(channels):
    out = (L.RfiChannel * len(channels))()
    for c, d in zip(out, channels):
        def val(x):
            return float("nan") if x is None else float(x)
        c.antenna = int(d["antenna"])
        c.dead = 1 if d["dead"] else 0
        c.acf_lag1, c.acf_lag3, c.acf_lag16 = (val(d["acf_lag1"]), val(d["acf_lag3"]),
                                               val(d["acf_lag16"]))
        c.period_samples = val(d["period_samples"])
        if len(d["lines"]) > 3:
            raise ValueError("at most 3 lines per channel")
        c.n_lines = len(d["lines"])
        for i, ln in enumerate(d["lines"]):
            c.lines[i].frequency = float(ln["frequency"])
            c.lines[i].excess_db = float(ln["excess_db"])
    return out


def rfi_report_json(channels):
    """The --rfi JSON (serde_json::to_string_pretty of RfiReport, main.rs:468) of channels as
    rfi() returns them; host only."""
    st = _rfi_structs(channels)
    out = C.c_void_p()
    check(lib().tga_rfi_report_json(st, len(channels), C.byref(out)))
    return _take_string(out)


def rfi_json(packed, record_samples, input_format, sampling_freq):
    """rfi() and its JSON report in one call (rfi::run + main.rs:439-478)."""
    pk = np.ascontiguousarray(packed, dtype=np.uint8)
    if pk.ndim != 2:
        raise ValueError("packed must be [n_ant, row_bytes]")
    out = C.c_void_p()
    check(lib().tga_rfi_json(_ptr(pk, C.c_uint8), pk.shape[0], pk.shape[1],
                             int(record_samples), int(input_format), float(sampling_freq),
                             C.byref(out)))
    return _take_string(out)
