"""ctypes binding of libtga.so (the C ABI declared in include/tga.h)."""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# TGA_LIBRARY selects another build of the same library (the sanitizer build of build.py)
LIB_PATH = os.environ.get("TGA_LIBRARY") or os.path.join(_HERE, "libtga.so")

TGA_OK = 0
ERRORS = {
    -1: "TGA_ERR_INVALID_ARGUMENT", -2: "TGA_ERR_EMPTY_BINS", -3: "TGA_ERR_PRN_RANGE",
    -4: "TGA_ERR_ANTENNA_RANGE", -5: "TGA_ERR_NAN", -6: "TGA_ERR_UNSUPPORTED",
    -7: "TGA_ERR_DEVICE", -8: "TGA_ERR_NO_DEVICE", -9: "TGA_ERR_OUT_OF_MEMORY",
}

SYS_GPS, SYS_GALILEO, SYS_BEIDOU, SYS_SBAS, SYS_L1C, SYS_QZSS = range(6)
SYSTEM_IDS = {"gps": SYS_GPS, "galileo": SYS_GALILEO, "beidou": SYS_BEIDOU, "sbas": SYS_SBAS,
              "l1c": SYS_L1C, "qzss": SYS_QZSS}
INPUT_1BIT, INPUT_2BIT = 1, 2
PRECISIONS = {"fp32": 0, "fp16": 1}
ACR_TABLES = {"GPS_L1_CA_ACR_TABLE": 0, "GALILEO_E1C_ACR_TABLE": 1,
              "BEIDOU_B1C_ACR_TABLE": 2, "GPS_L1C_ACR_TABLE": 3}


class TgaError(RuntimeError):
    """A negative status from libtga; ``code`` is the TGA_ERR_* value."""

    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code
        self.name = ERRORS.get(code, str(code))
        self.message = message


class Peak(C.Structure):
    _fields_ = [("best_peak", C.c_float), ("second_peak", C.c_float),
                ("best_freq_idx", C.c_uint64), ("best_codephase", C.c_uint64)]


class AcqResult(C.Structure):
    _fields_ = [("prn", C.c_int32), ("pad", C.c_int32), ("signal_strength", C.c_double),
                ("second_peak", C.c_double), ("codephase_frac", C.c_double),
                ("frequency", C.c_double)]


class SystemDesc(C.Structure):
    _fields_ = [("system", C.c_int32), ("n_prn", C.c_int32), ("prns", C.POINTER(C.c_int32)),
                ("n_fc", C.c_int32), ("fc", C.POINTER(C.c_double)),
                ("num_samples", C.c_int64), ("nc_epochs", C.c_int32), ("replica", C.c_int32)]


REPLICAS = {"reference": 0, "boc11": 1}  # TGA_REPLICA_*


class PlanDesc(C.Structure):
    _fields_ = [("sampling_freq", C.c_double), ("n_ant", C.c_int32),
                ("input_format", C.c_int32), ("record_samples", C.c_int64),
                ("n_systems", C.c_int32), ("precision", C.c_int32),
                ("systems", C.POINTER(SystemDesc))]


class RfiLine(C.Structure):
    _fields_ = [("frequency", C.c_double), ("excess_db", C.c_double)]


class RfiChannel(C.Structure):
    _fields_ = [("antenna", C.c_int32), ("dead", C.c_int32), ("acf_lag1", C.c_double),
                ("acf_lag3", C.c_double), ("acf_lag16", C.c_double),
                ("period_samples", C.c_double), ("n_lines", C.c_int32), ("pad", C.c_int32),
                ("lines", RfiLine * 3)]


_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)

PROTOTYPES = {
    "tga_last_error": (C.c_char_p, []),
    "tga_version": (C.c_char_p, []),
    "tga_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "tga_code_length": (C.c_int, [C.c_int32, C.POINTER(C.c_size_t)]),
    "tga_code_chips": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int8), C.c_size_t]),
    "tga_code_resampled": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_size_t, _dp]),
    "tga_code_spectra": (C.c_int, [_dp, C.c_size_t, C.c_size_t, _fp, _fp]),
    "tga_correlate_code": (C.c_int, [_fp, _dp, _fp, _fp, _dp, C.c_size_t, C.c_size_t,
                                     C.c_size_t, C.c_double, C.POINTER(Peak)]),
    "tga_correlate_code_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t,
                                             C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tga_acquire_single": (C.c_int, [C.c_int32, _fp, _dp, C.c_size_t, C.c_double, C.c_double,
                                     C.c_double, C.c_int32, C.c_size_t, C.POINTER(AcqResult)]),
    "tga_plan_create": (C.c_int, [C.POINTER(PlanDesc), C.POINTER(C.c_void_p)]),
    "tga_plan_destroy": (C.c_int, [C.c_void_p]),
    "tga_plan_row_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "tga_plan_peak_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "tga_plan_prns": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.c_size_t,
                                C.POINTER(C.c_size_t)]),
    "tga_search": (C.c_int, [C.c_void_p, _u8p, C.c_size_t, C.c_size_t, C.POINTER(Peak)]),
    "tga_search_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Peak)]),
    "tga_plan_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "tga_plan_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "tga_plan_get_timing": (C.c_int, [C.c_void_p, C.c_int32, _dp, C.POINTER(C.c_int64), _dp]),
    "tga_plan_kernel_name": (C.c_char_p, [C.c_void_p, C.c_int32]),
    "tga_plan_rows_matrix_flops": (C.c_int, [C.c_void_p, _dp, _dp]),
    "tga_plan_describe": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "tga_plan_surface_size": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_size_t)]),
    "tga_plan_carrier": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_float), C.c_size_t,
                                   C.POINTER(C.c_size_t)]),
    "tga_plan_surface": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float),
                                   C.c_size_t, C.POINTER(Peak)]),
    "tga_estimate_cn0": (C.c_int, [C.c_double, C.c_int32, _dp, C.POINTER(C.c_int32)]),
    "tga_acr_cn0": (C.c_int, [C.c_double, C.c_double, C.c_int32, _dp, C.POINTER(C.c_int32)]),
    "tga_rfi": (C.c_int, [_u8p, C.c_size_t, C.c_size_t, C.c_int64, C.c_int32, C.c_double,
                          C.POINTER(RfiChannel)]),
    "tga_rfi_device": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_int64, C.c_int32,
                                 C.c_double, C.POINTER(RfiChannel)]),
    "tga_acquire_all_json": (C.c_int, [C.c_void_p, _u8p, C.c_size_t, C.c_size_t, _dp,
                                       C.POINTER(C.c_int64), C.c_int32, C.c_double,
                                       C.c_double, C.POINTER(C.c_void_p)]),
    "tga_peaks_json": (C.c_int, [C.c_void_p, C.POINTER(Peak), _dp, C.POINTER(C.c_int64),
                                 C.c_int32, C.c_double, C.c_double, C.POINTER(C.c_void_p)]),
    "tga_peaks_cn0": (C.c_int, [C.c_void_p, C.POINTER(Peak), _dp]),
    "tga_antenna_report_json": (C.c_int, [C.c_size_t, C.POINTER(C.c_char_p), _dp, C.c_size_t,
                                          C.POINTER(C.c_int64), C.c_double,
                                          C.POINTER(C.c_void_p)]),
    "tga_peaks_antenna_report_json": (C.c_int, [C.c_void_p, C.POINTER(Peak), _dp,
                                                C.POINTER(C.c_int64), C.c_double, C.c_double,
                                                C.c_double, C.POINTER(C.c_void_p)]),
    "tga_acquire_antenna_report_json": (C.c_int, [C.c_void_p, _u8p, C.c_size_t, C.c_size_t,
                                                  _dp, C.POINTER(C.c_int64), C.c_double,
                                                  C.c_double, C.c_double,
                                                  C.POINTER(C.c_void_p)]),
    "tga_rfi_report_json": (C.c_int, [C.POINTER(RfiChannel), C.c_size_t,
                                      C.POINTER(C.c_void_p)]),
    "tga_rfi_json": (C.c_int, [_u8p, C.c_size_t, C.c_size_t, C.c_int64, C.c_int32, C.c_double,
                               C.POINTER(C.c_void_p)]),
    "tga_free_string": (None, [C.c_void_p]),
    "tga_format_f64": (C.c_int, [C.c_double, C.c_char_p, C.c_size_t]),
}

_lib = None


def lib():
    """Load libtga.so (built in-tree by tart-gnss_amd/build.py); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python tart-gnss_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != TGA_OK:
        msg = lib().tga_last_error()
        raise TgaError(rc, msg.decode() if msg else "")
    return rc
