"""TART observation files (HDF5) -> packed antenna rows for Plan.search.

Mirrors Observation::from_hdf5 (observation.rs:48-92): the file holds a scalar string dataset
`config` (JSON, config.rs), a scalar string dataset `timestamp` (ISO-8601) and a 2-D uint8
dataset `data` of [n_ant][row_bytes] bits packed MSB-first (numpy.packbits, observation.rs
:232-240).  The packed rows are exactly the library's TGA_INPUT_1BIT layout, so they go to
the GPU as they are; nothing unpacks them on the host.

The reader understands the subset of HDF5 such files use (the layout h5py writes with its
default, earliest-version settings): superblock version 0, version-1 object headers with
continuation blocks, old-style groups (v1 B-tree of symbol nodes plus a local heap),
contiguous and compact data layouts, fixed-point, fixed-length string and variable-length
string datatypes (global heap).  Anything else raises ValueError naming what was found.
"""
import json
import struct
from dataclasses import dataclass

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"


@dataclass
class Observation:
    config: dict
    timestamp: str
    data: np.ndarray  # uint8 [n_ant][row_bytes], MSB-first packed 1-bit samples

    @property
    def num_antenna(self):
        return int(self.data.shape[0])

    @property
    def num_samples(self):
        return int(self.data.shape[1]) * 8

    @property
    def sampling_frequency(self):
        return float(self.config["sampling_frequency"])

    def antenna(self, ant):
        """observation.rs:140-151: bipolar (+1/-1) samples of one antenna."""
        if not 0 <= ant < self.num_antenna:
            raise IndexError(f"antenna index {ant} out of range (have {self.num_antenna})")
        return np.unpackbits(self.data[ant]).astype(np.float64) * 2.0 - 1.0


class _File:
    def __init__(self, buf):
        self.b = memoryview(buf)
        if bytes(self.b[:8]) != _SIG:
            raise ValueError("not an HDF5 file (signature)")
        ver = self.b[8]
        if ver != 0:
            raise ValueError(f"HDF5 superblock version {ver} (only version 0 is read)")
        self.so, self.sl = self.b[13], self.b[14]
        if self.so != 8 or self.sl != 8:
            raise ValueError("HDF5 offsets/lengths other than 8 bytes")
        # base, free-space, end-of-file, driver addresses, then the root symbol-table entry
        root = 24 + 4 * 8
        self.root = self._entry(root)

    def u(self, off, n):
        return int.from_bytes(self.b[off:off + n], "little")

    def _entry(self, off):
        """Symbol table entry -> (name offset, object header address, cache type, scratch)."""
        name_off, ohdr = self.u(off, 8), self.u(off + 8, 8)
        cache = self.u(off + 16, 4)
        return name_off, ohdr, cache, (self.u(off + 24, 8), self.u(off + 32, 8))

    def _heap_name(self, heap, name_off):
        if bytes(self.b[heap:heap + 4]) != b"HEAP":
            raise ValueError("local heap signature")
        data = self.u(heap + 24, 8)
        start = data + name_off
        end = bytes(self.b[start:start + 4096]).index(b"\0")
        return bytes(self.b[start:start + end]).decode()

    def _group_entries(self, btree, heap):
        out = {}
        stack = [btree]
        while stack:
            node = stack.pop()
            if bytes(self.b[node:node + 4]) != b"TREE":
                raise ValueError("v1 B-tree signature")
            ntype, level, used = self.b[node + 4], self.b[node + 5], self.u(node + 6, 2)
            if ntype != 0:
                raise ValueError("B-tree of a group expected")
            # keys (8 B) and children (8 B) interleave after the two sibling addresses
            p = node + 8 + 16
            for i in range(used):
                child = self.u(p + 8 + 16 * i, 8)
                if level > 0:
                    stack.append(child)
                else:
                    if bytes(self.b[child:child + 4]) != b"SNOD":
                        raise ValueError("symbol node signature")
                    nsym = self.u(child + 6, 2)
                    for k in range(nsym):
                        e = self._entry(child + 8 + 40 * k)
                        out[self._heap_name(heap, e[0])] = e
        return out

    def _messages(self, addr):
        """(type, flags, body offset, size) of every message of a version-1 object header."""
        if self.b[addr] != 1:
            raise ValueError(f"object header version {self.b[addr]} (only version 1 is read)")
        nmsg = self.u(addr + 2, 2)
        blocks = [(addr + 16, self.u(addr + 8, 4))]
        msgs = []
        while blocks and len(msgs) < nmsg:
            start, size = blocks.pop(0)
            p = start
            while p + 8 <= start + size and len(msgs) < nmsg:
                mtype, msize, flags = self.u(p, 2), self.u(p + 2, 2), self.b[p + 4]
                body = p + 8
                if mtype == 0x10:  # continuation
                    blocks.append((self.u(body, 8), self.u(body + 8, 8)))
                msgs.append((mtype, flags, body, msize))
                p = body + msize
        return msgs

    def dataset(self, name):
        name_off, ohdr, cache, scratch = self.root
        if cache == 1:
            btree, heap = scratch
        else:
            btree = heap = None
            for mtype, _, body, _ in self._messages(ohdr):
                if mtype == 0x11:  # symbol table message
                    btree, heap = self.u(body, 8), self.u(body + 8, 8)
            if btree is None:
                raise ValueError("root group without a symbol table")
        entries = self._group_entries(btree, heap)
        if name not in entries:
            raise KeyError(f"dataset {name!r} not found (have {sorted(entries)})")
        return self._read_dataset(entries[name][1])

    def _read_dataset(self, ohdr):
        dims = dtype = layout = None
        for mtype, _, body, size in self._messages(ohdr):
            if mtype == 0x01:
                dims = self._dataspace(body)
            elif mtype == 0x03:
                dtype = self._datatype(body)
            elif mtype == 0x08:
                layout = self._layout(body)
        if dims is None or dtype is None or layout is None:
            raise ValueError("dataset without dataspace, datatype or layout message")
        kind, addr, size = layout
        raw = bytes(self.b[addr:addr + size])
        n = int(np.prod(dims)) if dims else 1
        cls, esize, extra = dtype
        if cls == "vlen-str":
            vals = []
            for i in range(n):
                ln, coll, idx = struct.unpack_from("<IQI", raw, 16 * i)
                vals.append(self._global_object(coll, idx)[:ln].decode(extra))
            return vals[0] if n == 1 else vals  # scalar or single-element (h5py strings)
        if cls == "str":
            vals = [raw[i * esize:(i + 1) * esize].split(b"\0")[0].decode(extra)
                    for i in range(n)]
            return vals[0] if n == 1 else vals
        arr = np.frombuffer(raw[:n * esize], dtype=extra).reshape(dims if dims else ())
        return arr

    def _dataspace(self, body):
        ver, rank, flags = self.b[body], self.b[body + 1], self.b[body + 2]
        off = body + (8 if ver == 1 else 4)
        return tuple(self.u(off + 8 * i, 8) for i in range(rank))

    def _datatype(self, body):
        cls = self.b[body] & 0x0F
        bits0 = self.b[body + 1]
        size = self.u(body + 4, 4)
        if cls == 0:  # fixed point
            signed = bool(bits0 & 0x08)
            big = bool(bits0 & 0x01)
            dt = np.dtype(("i" if signed else "u") + str(size)).newbyteorder(">" if big else "<")
            return "int", size, dt
        if cls == 3:  # fixed-length string
            return "str", size, "ascii" if (self.b[body + 1] >> 4) == 0 else "utf-8"
        if cls == 9:  # variable length
            vtype = bits0 & 0x0F
            cset = (self.b[body + 2] & 0x0F)
            if vtype != 1:
                raise ValueError("variable-length sequence (only strings are read)")
            return "vlen-str", size, "ascii" if cset == 0 else "utf-8"
        raise ValueError(f"datatype class {cls} not supported")

    def _layout(self, body):
        ver = self.b[body]
        if ver != 3:
            raise ValueError(f"data layout message version {ver} (only version 3 is read)")
        cls = self.b[body + 1]
        if cls == 0:  # compact
            size = self.u(body + 2, 2)
            return "compact", body + 4, size
        if cls == 1:  # contiguous
            return "contiguous", self.u(body + 2, 8), self.u(body + 10, 8)
        raise ValueError("chunked datasets are not read")

    def _global_object(self, coll, idx):
        if bytes(self.b[coll:coll + 4]) != b"GCOL":
            raise ValueError("global heap signature")
        size = self.u(coll + 8, 8)
        p, end = coll + 16, coll + size
        while p + 16 <= end:
            oid, osize = self.u(p, 2), self.u(p + 8, 8)
            if oid == 0:
                break
            if oid == idx:
                return bytes(self.b[p + 16:p + 16 + osize])
            p += 16 + ((osize + 7) & ~7)
        raise ValueError(f"global heap object {idx} not found")


def read_observation(path):
    """Observation::from_hdf5 (observation.rs:48-92) for files written by
    Observation.to_hdf5() / the TART telescope software."""
    with open(path, "rb") as f:
        buf = f.read()
    h = _File(buf)
    config = json.loads(h.dataset("config"))
    timestamp = h.dataset("timestamp")
    data = np.ascontiguousarray(h.dataset("data"), dtype=np.uint8)
    if data.ndim != 2:
        raise ValueError(f"data dataset of rank {data.ndim}")
    if "num_antenna" in config and int(config["num_antenna"]) != data.shape[0]:
        raise ValueError(f"data antenna count ({data.shape[0]}) != config num_antenna "
                         f"({config['num_antenna']})")
    return Observation(config=config, timestamp=timestamp, data=data)
