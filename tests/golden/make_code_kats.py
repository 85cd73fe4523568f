"""Extract the known-answer chip vectors held in the reference's unit tests.

Reads the octal KAT string arrays from the reference's Rust test modules (as text)
and writes them to ``code_kats.json`` beside this script.  Sources:
  sbas.rs:436-442 (REF_FIRST10, PRN 120..158), qzss.rs:437-441 (REF_FIRST10, PRN 184..206),
  beidou_codes.rs:225-253 (REF_FIRST24 / REF_LAST24, PRN 1..63),
  l1c_codes.rs:256-289 (REF_FIRST24 / REF_LAST24, PRN 1..63).
Run once in a container that has the reference checkout at /root/reference.
"""
import json
import os
import re

REF = "/root/reference/src"


def arr(path, name):
    text = open(os.path.join(REF, path)).read()
    start = text.index("const " + name)
    body = text[start:text.index("];", start)]
    return re.findall(r'"([0-7]+)"', body)


def main():
    out = {
        "sbas_first10_octal": {"prn_base": 120, "values": arr("sbas.rs", "REF_FIRST10")},
        "qzss_first10_octal": {"prn_base": 184, "values": arr("qzss.rs", "REF_FIRST10")},
        "b1cp_first24_octal": {"prn_base": 1, "values": arr("beidou_codes.rs", "REF_FIRST24")},
        "b1cp_last24_octal": {"prn_base": 1, "values": arr("beidou_codes.rs", "REF_LAST24")},
        "l1cp_first24_octal": {"prn_base": 1, "values": arr("l1c_codes.rs", "REF_FIRST24")},
        "l1cp_last24_octal": {"prn_base": 1, "values": arr("l1c_codes.rs", "REF_LAST24")},
    }
    counts = {k: len(v["values"]) for k, v in out.items()}
    assert counts == {"sbas_first10_octal": 39, "qzss_first10_octal": 23,
                      "b1cp_first24_octal": 63, "b1cp_last24_octal": 63,
                      "l1cp_first24_octal": 63, "l1cp_last24_octal": 63}, counts
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "code_kats.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst, counts)


if __name__ == "__main__":
    main()
