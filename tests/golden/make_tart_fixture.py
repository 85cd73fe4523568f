"""Writes tests/golden/tart_2ms.hdf: the reference's TART observation
(data/data_2026-03-09_04_25_10.145497.hdf, 24 antennas, 16.368 MHz, 1-bit) cut to its first
4092 bytes per antenna (32736 samples = 2 ms = two C/A periods).  Everything before the data
block (superblock, group B-tree and heaps, object headers, global heap with the config and
timestamp strings) is copied unchanged; the data dataspace, the data layout size and the
end-of-file address are patched for the shorter rows.  Run from the repository root:

    python tests/golden/make_tart_fixture.py /root/reference/data/data_2026-03-09_04_25_10.145497.hdf
"""
import os
import struct
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "..", "tart-gnss_amd"))
from tga import observation as ob  # noqa: E402

ROW = 4092


def main(src):
    buf = bytearray(open(src, "rb").read())
    h = ob._File(bytes(buf))
    entries = h._group_entries(*h.root[3])
    _, ohdr, _, _ = entries["data"]
    space = layout = None
    for mtype, _, body, _ in h._messages(ohdr):
        if mtype == 0x01:
            space = body
        elif mtype == 0x08:
            layout = body
    ver, rank, flags = buf[space], buf[space + 1], buf[space + 2]
    assert ver == 1 and rank == 2, (ver, rank)
    n_ant, row = struct.unpack_from("<QQ", buf, space + 8)
    struct.pack_into("<Q", buf, space + 16, ROW)
    if flags & 1:  # maximum dimensions follow
        struct.pack_into("<Q", buf, space + 32, ROW)
    assert buf[layout] == 3 and buf[layout + 1] == 1, "contiguous layout expected"
    addr, size = struct.unpack_from("<QQ", buf, layout + 2)
    assert size == n_ant * row
    struct.pack_into("<Q", buf, layout + 10, n_ant * ROW)
    data = b"".join(bytes(buf[addr + a * row: addr + a * row + ROW]) for a in range(n_ant))
    out = bytes(buf[:addr]) + data
    out = bytearray(out)
    struct.pack_into("<Q", out, 24 + 2 * 8, len(out))  # end-of-file address
    dst = os.path.join(os.path.dirname(__file__), "tart_2ms.hdf")
    open(dst, "wb").write(bytes(out))
    print(dst, len(out), "bytes")


if __name__ == "__main__":
    main(sys.argv[1])
