"""Python mirror of tart-gnss-acquire's acquisition API over the gfx950 library (libtga.so).

Names and argument meanings follow the reference's Rust functions so tests read like the
reference's own; every call goes through the C ABI of include/tga.h into HIP kernels.  There is
no CPU fallback: if libtga.so or a gfx950 device is missing, calls raise.

  correlate_code       correlate.rs:91-183
  code_spectra         correlate.rs:47-68
  acquire_full / acquire_sbas / acquire_qzss / acquire_galileo_single /
  acquire_beidou_single / acquire_l1c_single
                       acquisition.rs:176-248, sbas.rs:132-206, qzss.rs:135-209,
                       galileo.rs:133-206, beidou.rs:85-158, l1c.rs:94-167
  gold_code / sbas_gold_code / qzss_gold_code / e1c_code_resampled / b1c_code_resampled /
  l1c_code_resampled   replica generators + resampling
  estimate_cn0         acr.rs:49-84
  Plan                 batched search behind acquire_all_* (acquisition.rs:261-396 ...)
  Plan.acquire_all_json / Plan.peaks_json / format_f64
                       CombinedOutput JSON (lib.rs:31-45, main.rs:242-255, 893)
  rfi                  rfi.rs:142-184
  rfi_json / rfi_report_json
                       --rfi JSON report (RfiReport, rfi.rs:73-77, main.rs:439-478)
  antenna_test_json / Plan.test_antennas_json
                       --test-antennas report (antenna_test.rs:105-222, main.rs:848-880)
  read_observation     Observation::from_hdf5 (observation.rs:48-92), host-side HDF5 reader
"""
from ._lib import (ACR_TABLES, INPUT_1BIT, INPUT_2BIT, SYSTEM_IDS, TgaError, lib)  # noqa: F401
from .api import (AcquisitionResult, CorrelationPeak, Plan, SystemSearch,  # noqa: F401
                  acquire_beidou_single, antenna_test_json, acquire_full, acquire_galileo_single,
                  acquire_l1c_single, acquire_qzss, acquire_sbas, acquire_single,
                  b1c_code_resampled, code_chips, code_length, code_resampled, code_spectra,
                  correlate_code, correlate_code_noalloc, device_count, e1c_code_resampled,
                  estimate_cn0, format_f64,
                  gold_code,
                  l1c_code_resampled, qzss_gold_code, rfi, rfi_device, rfi_json,
                  rfi_report_json, sbas_gold_code,
                  search_frequencies)
from .observation import Observation, read_observation  # noqa: F401,E402
