"""Build libtga.so (HIP, gfx950) in-tree.

    python tart-gnss_amd/build.py [-v]

Steps: generate ``_build/gen/tables_gen.h`` from the data files in ``csrc/tables``, compile
every ``csrc/*.cpp`` / ``csrc/*.hip`` with hipcc for gfx950 (in parallel), link
``tga/libtga.so``.  Host C++ is compiled with ``-ffp-contract=off`` so f64 arithmetic that
must match the reference (grids, ACR interpolation, replica resampling) is never fused.

``build(sanitize=True)`` (``python tart-gnss_amd/build.py --sanitize``) builds the same library
with AddressSanitizer and UndefinedBehaviorSanitizer on the host code only (device code is
unchanged) into ``_build/san/libtga.so``; ``tart-gnss_amd/sanitize.py`` runs the CPU tests
against it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
GEN = os.path.join(BUILD, "gen")
OUT = os.path.join(HERE, "tga", "libtga.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(REPO, "include"), "-I", CSRC, "-I", GEN]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def gen_tables():
    os.makedirs(GEN, exist_ok=True)
    lines = ["// generated by tart-gnss_amd/build.py from csrc/tables/*.txt -- do not edit",
             "#pragma once", ""]
    hexes = []
    with open(os.path.join(CSRC, "tables", "galileo_e1c_hex.txt")) as f:
        for ln in f:
            ln = ln.strip()
            if ln and not ln.startswith("#"):
                hexes.append(ln)
    assert len(hexes) == 50 and all(len(h) == 1023 for h in hexes)
    lines.append("static const char* const kGalileoE1cHex[50] = {")
    lines += ['  "%s",' % h for h in hexes]
    lines += ["};", ""]
    lines.append("struct AcrPoint { double r; double cn0; };")
    tables = []
    cur = None
    with open(os.path.join(CSRC, "tables", "acr_tables.txt")) as f:
        for ln in f:
            ln = ln.strip()
            if not ln or ln.startswith("#"):
                continue
            if ln.startswith("table "):
                _, name, count = ln.split()
                cur = (name, int(count), [])
                tables.append(cur)
                continue
            r, c = ln.split()
            cur[2].append((r, c))
    order = ["GPS_L1_CA_ACR_TABLE", "GALILEO_E1C_ACR_TABLE", "BEIDOU_B1C_ACR_TABLE",
             "GPS_L1C_ACR_TABLE"]
    assert [t[0] for t in tables] == order
    for name, count, pts in tables:
        assert len(pts) == count
        lines.append("static const AcrPoint k_%s[%d] = {" % (name, count))
        lines += ["  {%s, %s}," % p for p in pts]
        lines.append("};")
    lines.append("static const AcrPoint* const kAcrTables[4] = {%s};" %
                 ", ".join("k_" + n for n in order))
    lines.append("static const int kAcrSizes[4] = {%s};" % ", ".join(str(t[1]) for t in tables))
    text = "\n".join(lines) + "\n"
    path = os.path.join(GEN, "tables_gen.h")
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)


# host-only sanitizers: each -fsanitize= applies to the host compilation alone
SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
SAN_DIR = os.path.join(BUILD, "san")


def compile_one(src, verbose, sanitize=False):
    obj = os.path.join(SAN_DIR if sanitize else BUILD, os.path.basename(src) + ".o")
    extra = SANITIZE if sanitize else []
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + extra + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC, "-x", "c++"] + COMMON + extra + ["-ffp-contract=off",
                                                      "-D__HIP_PLATFORM_AMD__",
                                                      "-I", "/opt/rocm/include",
                                                      "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{msgs}")
    if verbose and msgs.strip():
        print(msgs)
    return obj


def build(verbose=False, sanitize=False):
    os.makedirs(SAN_DIR if sanitize else BUILD, exist_ok=True)
    gen_tables()
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".cpp") or f.endswith(".hip"))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose, sanitize), srcs))
    out = os.path.join(SAN_DIR, "libtga.so") if sanitize else OUT
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if sanitize:
        cmd[1:1] = SANITIZE[:6]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, sanitize="--sanitize" in sys.argv))
