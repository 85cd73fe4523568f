"""Run the CPU tests against the AddressSanitizer + UndefinedBehaviorSanitizer build of libtga
(host code instrumented, device code unchanged; build.py --sanitize).

    python tart-gnss_amd/sanitize.py [--no-build] [pytest args]

The library is loaded through TGA_LIBRARY into a Python started with the clang ASan runtime
preloaded (leak checking off: the interpreter itself keeps allocations to exit); any ASan
report or UBSan finding aborts the test process, so a green run means none occurred.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
# the CPU tests that call into the library (codes, JSON writers, antenna report, ACR, RFI
# report text, plan-descriptor validation, exports)
TESTS = ["test_library_cpu.py", "test_json_cpu.py", "test_antenna_test_cpu.py",
         "test_rfi_json_cpu.py", "test_observation_cpu.py", "test_acr_synth_cpu.py",
         "test_layout_cpu.py"]


def main(argv):
    build_first = "--no-build" not in argv
    argv = [a for a in argv if a != "--no-build"]
    sys.path.insert(0, HERE)
    import build
    lib = build.build(sanitize=True) if build_first else os.path.join(build.SAN_DIR,
                                                                        "libtga.so")
    rt = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not rt:
        raise SystemExit("clang ASan runtime not found under /opt/rocm/llvm")
    env = dict(os.environ, LD_PRELOAD=rt[-1], TGA_LIBRARY=lib,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    cmd = [sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider"]
    cmd += [os.path.join(REPO, "tests", t) for t in TESTS] + argv
    return subprocess.call(cmd, cwd=REPO, env=env)


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
