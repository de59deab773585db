"""Minimal HDF5 writer (no h5py, no libhdf5) for the Hamiltonian file layout.

Writes what ``_hdf5.File`` reads and what the HDF5 library's default ("earliest") file format
holds: superblock version 0, version-1 object headers, symbol-table groups (one local heap, one
v1 B-tree of type 0 and symbol-table nodes of 2K = 8 entries each), contiguous datasets,
version-1 attributes, 64-bit little-endian integers and floats, h5py's bool (an 8-bit
enumeration FALSE = 0 / TRUE = 1) and variable-length UTF-8 strings in one global heap
collection.  Format reference: the HDF5 File Format Specification, version 2.0 (sections II.A
superblock, III.A B-trees, III.B local heaps, III.E global heaps, IV.A object headers and the
dataspace, datatype, fill value, layout, attribute and symbol table messages).

A tree is built from ``Group`` nodes (``attrs``: name -> scalar; ``children``: name -> ``Group``
or numpy array / list of str) and written by ``write(path, root)``.
"""

from __future__ import annotations

import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
_LEAF_K = 4  # group leaf node K: symbol-table nodes hold 2K entries
_INTERNAL_K = 16


class Group:
    def __init__(self):
        self.attrs: dict = {}
        self.children: dict = {}


def _pad8(b: bytes) -> bytes:
    return b + b"\x00" * (-len(b) % 8)


# ---------------------------------------------------------------- datatypes


def _dt_int64() -> bytes:
    return struct.pack("<B3sI", 0x10 | 0, bytes([0x08, 0, 0]), 8) + struct.pack("<HH", 0, 64)


def _dt_float64() -> bytes:
    # IEEE little endian, implied leading mantissa bit (normalisation 2), sign at bit 63
    return struct.pack("<B3sI", 0x10 | 1, bytes([0x20, 63, 0]), 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)


def _dt_bool() -> bytes:
    base = struct.pack("<B3sI", 0x10 | 0, bytes([0x08, 0, 0]), 1) + struct.pack("<HH", 0, 8)
    names = _pad8(b"FALSE\x00") + _pad8(b"TRUE\x00")
    return struct.pack("<B3sI", 0x10 | 8, bytes([2, 0, 0]), 1) + base + names + bytes([0, 1])


def _dt_vlen_str() -> bytes:
    base = struct.pack("<B3sI", 0x10 | 0, bytes([0, 0, 0]), 1) + struct.pack("<HH", 0, 8)
    # type 1 (string), null-terminated padding, UTF-8
    return struct.pack("<B3sI", 0x10 | 9, bytes([0x01, 0x01, 0]), 16) + base


def _dataspace(dims) -> bytes:
    return struct.pack("<BBB5x", 1, len(dims), 0) + b"".join(struct.pack("<Q", d) for d in dims)


# ---------------------------------------------------------------- writer


class _Writer:
    def __init__(self):
        self.buf = bytearray(96)  # superblock
        self.strings: list = []   # global heap objects, index = position + 1
        self.gcol = UNDEF

    def alloc(self, data: bytes) -> int:
        addr = len(self.buf)
        self.buf += data
        self.buf += b"\x00" * (-len(self.buf) % 8)
        return addr

    def string_ref(self, s: str) -> bytes:
        raw = s.encode()
        self.strings.append(raw)
        return struct.pack("<IQI", len(raw), self.gcol, len(self.strings))

    # value encoders: (datatype bytes, dims, raw data)
    def encode(self, v):
        if isinstance(v, (bool, np.bool_)):
            return _dt_bool(), [], bytes([1 if v else 0])
        if isinstance(v, (int, np.integer)):
            return _dt_int64(), [], struct.pack("<q", int(v))
        if isinstance(v, (float, np.floating)):
            return _dt_float64(), [], struct.pack("<d", float(v))
        if isinstance(v, str):
            return _dt_vlen_str(), [], self.string_ref(v)
        if isinstance(v, list) and all(isinstance(x, str) for x in v):
            return _dt_vlen_str(), [len(v)], b"".join(self.string_ref(x) for x in v)
        a = np.asarray(v)
        if a.dtype == np.bool_:
            return _dt_bool(), list(a.shape), a.astype(np.uint8).tobytes()
        if a.dtype.kind in "iu":
            return _dt_int64(), list(a.shape), np.ascontiguousarray(a, dtype="<i8").tobytes()
        if a.dtype.kind == "f" or a.size == 0:
            return _dt_float64(), list(a.shape), np.ascontiguousarray(a, dtype="<f8").tobytes()
        raise TypeError(f"cannot store a value of dtype {a.dtype}")

    def attribute(self, name: str, value) -> bytes:
        dt, dims, raw = self.encode(value)
        ds = _dataspace(dims)
        nm = name.encode() + b"\x00"
        return struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(ds)) + _pad8(nm) + _pad8(dt) + _pad8(ds) + raw

    def object_header(self, messages) -> int:
        body = b""
        for mtype, data in messages:
            data = _pad8(data)
            body += struct.pack("<HHB3x", mtype, len(data), 0) + data
        prefix = struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body))
        return self.alloc(prefix + body)

    def dataset(self, value) -> int:
        dt, dims, raw = self.encode(value)
        addr = self.alloc(raw) if raw else UNDEF
        layout = struct.pack("<BBQQ", 3, 1, addr, len(raw))
        fill = struct.pack("<BBBB", 2, 2, 2, 0)  # late allocation, fill never written, no fill value
        return self.object_header([(0x0001, _dataspace(dims)), (0x0003, dt), (0x0005, fill), (0x0008, layout)])

    def group(self, g: Group) -> tuple:
        """Write g's children and g; returns (object header address, B-tree address, heap address)."""
        entries = []
        for name in sorted(g.children):
            child = g.children[name]
            hdr = self.group(child)[0] if isinstance(child, Group) else self.dataset(child)
            entries.append((name, hdr))
        # local heap: offset 0 holds the empty string
        heap = bytearray(8)
        offs = {}
        for name, _ in entries:
            offs[name] = len(heap)
            heap += _pad8(name.encode() + b"\x00")
        heap_size = max(len(heap), 16)
        heap += b"\x00" * (heap_size - len(heap))
        heap_data = self.alloc(bytes(heap))
        heap_addr = self.alloc(b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack("<QQQ", heap_size, UNDEF, heap_data))
        # symbol-table nodes of 2K entries, then one level-0 B-tree node over them
        per = 2 * _LEAF_K
        chunks = [entries[i : i + per] for i in range(0, len(entries), per)] or [[]]
        if len(chunks) > 2 * _INTERNAL_K:
            raise ValueError("more than 256 members in one group")
        snods = []
        for ch in chunks:
            node = b"SNOD" + struct.pack("<BBH", 1, 0, len(ch))
            for name, hdr in ch:
                node += struct.pack("<QQII16x", offs[name], hdr, 0, 0)
            node += b"\x00" * (40 * (per - len(ch)))
            snods.append((self.alloc(node), offs[ch[-1][0]] if ch else 0))
        tree = b"TREE" + struct.pack("<BBHQQ", 0, 0, len(snods), UNDEF, UNDEF) + struct.pack("<Q", 0)
        for addr, last in snods:
            tree += struct.pack("<QQ", addr, last)
        tree += b"\x00" * (16 * (2 * _INTERNAL_K - len(snods)))
        btree_addr = self.alloc(tree)
        messages = [(0x0011, struct.pack("<QQ", btree_addr, heap_addr))]
        messages += [(0x000C, self.attribute(k, v)) for k, v in g.attrs.items()]
        return self.object_header(messages), btree_addr, heap_addr

    def global_heap(self, size: int) -> None:
        self.gcol = len(self.buf)
        self.buf += b"\x00" * size

    def fill_global_heap(self, size: int) -> None:
        body = b""
        for i, raw in enumerate(self.strings):
            body += struct.pack("<HHIQ", i + 1, 1, 0, len(raw)) + _pad8(raw)
        head = b"GCOL" + bytes([1, 0, 0, 0]) + struct.pack("<Q", size)
        free = size - len(head) - len(body)
        if free < 16:
            raise ValueError("global heap collection too small")
        blob = head + body + struct.pack("<HHIQ", 0, 0, 0, free) + b"\x00" * (free - 16)
        self.buf[self.gcol : self.gcol + size] = blob


def _count_strings(g: Group) -> int:
    n = 0
    for v in g.attrs.values():
        n += 1 if isinstance(v, str) else 0
        n += len(v) if isinstance(v, list) and v and all(isinstance(x, str) for x in v) else 0
    for c in g.children.values():
        if isinstance(c, Group):
            n += _count_strings(c)
        elif isinstance(c, list) and all(isinstance(x, str) for x in c):
            n += len(c)
    return n


def _string_bytes(g: Group) -> int:
    total = 0
    for v in list(g.attrs.values()) + list(g.children.values()):
        if isinstance(v, Group):
            total += _string_bytes(v)
        elif isinstance(v, str):
            total += 16 + len(_pad8(v.encode()))
        elif isinstance(v, list) and v and all(isinstance(x, str) for x in v):
            total += sum(16 + len(_pad8(x.encode())) for x in v)
    return total


def write(path: str, root: Group) -> None:
    w = _Writer()
    size = max(4096, 16 + _string_bytes(root) + 16)
    size += -size % 8
    w.global_heap(size)
    root_hdr, btree, heap = w.group(root)
    w.fill_global_heap(size)
    eof = len(w.buf)
    sb = b"\x89HDF\r\n\x1a\n" + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", _LEAF_K, _INTERNAL_K, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sb += struct.pack("<QQII", 0, root_hdr, 1, 0) + struct.pack("<QQ", btree, heap)
    assert len(sb) == 96
    w.buf[0:96] = sb
    with open(path, "wb") as f:
        f.write(bytes(w.buf))
