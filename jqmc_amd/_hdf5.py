"""Minimal read-only HDF5 reader for TREXIO files and Hamiltonian files (no h5py, no libhdf5).

Covers what TREXIO's HDF5 back end and the HDF5 library's default file format write: superblock version 0, version-1 object
headers, symbol-table groups (v1 B-trees of type 0 + local heaps), contiguous and compact
dataset layouts, chunked datasets indexed by v1 B-trees of type 1 (unfiltered), attributes,
fixed-point / floating-point / enumeration (read as integers) / fixed- and variable-length string
datatypes (global heap).
Anything else raises ``NotImplementedError``.  Format reference: the HDF5 File Format
Specification, version 2.0 (sections II.A superblock, III.A B-trees, III.B-D heaps,
IV.A object headers and messages).
"""

from __future__ import annotations

import struct

import numpy as np


class HDF5Error(ValueError):
    pass


class _Datatype:
    def __init__(self, cls, size, signed=True, order="<", strpad=0, vlen_base=None, is_vlen_str=False):
        self.cls = cls
        self.size = size
        self.signed = signed
        self.order = order
        self.strpad = strpad
        self.vlen_base = vlen_base
        self.is_vlen_str = is_vlen_str

    def numpy_dtype(self):
        if self.cls == 0:
            return np.dtype(f"{self.order}{'i' if self.signed else 'u'}{self.size}")
        if self.cls == 1:
            return np.dtype(f"{self.order}f{self.size}")
        raise NotImplementedError(f"datatype class {self.cls} has no fixed numpy dtype")


class Node:
    """A group or dataset of the file."""

    def __init__(self, f, addr, name):
        self.f = f
        self.addr = addr
        self.name = name
        self.messages = f._read_object_header(addr)
        self.attrs = {}
        for mtype, data in self.messages:
            if mtype == 0x000C:
                k, v = f._parse_attribute(data)
                self.attrs[k] = v

    @property
    def is_group(self):
        return any(m == 0x0011 for m, _ in self.messages)

    def children(self):
        for mtype, data in self.messages:
            if mtype == 0x0011:
                btree, heap = struct.unpack_from("<QQ", data, 0)
                return self.f._group_entries(btree, heap)
        return {}

    def __getitem__(self, path):
        node = self
        for part in [p for p in path.split("/") if p]:
            ch = node.children()
            if part not in ch:
                raise KeyError(path)
            node = Node(self.f, ch[part], part)
        return node

    def keys(self):
        return list(self.children().keys())

    def read(self):
        """Dataset contents as a numpy array (strings as a list of str)."""
        dims, dtype, layout = None, None, None
        for mtype, data in self.messages:
            if mtype == 0x0001:
                dims = self.f._parse_dataspace(data)
            elif mtype == 0x0003:
                dtype = self.f._parse_datatype(data, 0)[0]
            elif mtype == 0x0008:
                layout = data
        if dims is None or dtype is None or layout is None:
            raise HDF5Error(f"{self.name} is not a dataset")
        n = int(np.prod(dims)) if dims else 1
        raw = self.f._read_layout(layout, n * dtype.size, dims, dtype.size)
        return self.f._decode(raw, dtype, n, dims)


class File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        if self.buf[:8] != b"\x89HDF\r\n\x1a\n":
            raise HDF5Error("not an HDF5 file")
        if self.buf[8] != 0:
            raise NotImplementedError("only superblock version 0 is supported")
        if self.buf[13] != 8 or self.buf[14] != 8:
            raise NotImplementedError("only 8-byte offsets and lengths are supported")
        root_hdr = struct.unpack_from("<Q", self.buf, 24 + 32 + 8)[0]
        self.root = Node(self, root_hdr, "/")

    def __getitem__(self, path):
        return self.root[path]

    # ------------------------------------------------------------------ object headers

    def _read_object_header(self, addr):
        b = self.buf
        version = b[addr]
        if version != 1:
            raise NotImplementedError(f"object header version {version}")
        nmsg = struct.unpack_from("<H", b, addr + 2)[0]
        size = struct.unpack_from("<I", b, addr + 8)[0]
        blocks = [(addr + 16, size)]
        msgs = []
        while blocks and len(msgs) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(msgs) < nmsg:
                mtype, msize, _flags = struct.unpack_from("<HHB", b, p)
                data = b[p + 8 : p + 8 + msize]
                if mtype == 0x0010:
                    off, ln = struct.unpack_from("<QQ", data, 0)
                    blocks.append((off, ln))
                msgs.append((mtype, data))
                p += 8 + msize
        return msgs

    # ------------------------------------------------------------------ groups

    def _heap_data(self, heap_addr):
        b = self.buf
        if b[heap_addr : heap_addr + 4] != b"HEAP":
            raise HDF5Error("bad local heap")
        size, _free, data_addr = struct.unpack_from("<QQQ", b, heap_addr + 8)
        return b[data_addr : data_addr + size]

    @staticmethod
    def _cstr(data, off):
        end = data.index(b"\x00", off)
        return data[off:end].decode()

    def _group_entries(self, btree_addr, heap_addr):
        heap = self._heap_data(heap_addr)
        out = {}
        stack = [btree_addr]
        while stack:
            a = stack.pop()
            b = self.buf
            if b[a : a + 4] != b"TREE":
                raise HDF5Error("bad group B-tree node")
            ntype, level, used = b[a + 4], b[a + 5], struct.unpack_from("<H", b, a + 6)[0]
            if ntype != 0:
                raise HDF5Error("expected a group B-tree")
            p = a + 24 + 8  # header (24) + key 0
            for _ in range(used):
                child = struct.unpack_from("<Q", b, p)[0]
                p += 16
                if level > 0:
                    stack.append(child)
                else:
                    self._snod(child, heap, out)
        return out

    def _snod(self, addr, heap, out):
        b = self.buf
        if b[addr : addr + 4] != b"SNOD":
            raise HDF5Error("bad symbol table node")
        n = struct.unpack_from("<H", b, addr + 6)[0]
        p = addr + 8
        for _ in range(n):
            name_off, hdr = struct.unpack_from("<QQ", b, p)
            out[self._cstr(heap, name_off)] = hdr
            p += 40

    # ------------------------------------------------------------------ messages

    @staticmethod
    def _parse_dataspace(data):
        version, ndims, flags = data[0], data[1], data[2]
        off = 8 if version == 1 else 4
        if version not in (1, 2):
            raise NotImplementedError(f"dataspace version {version}")
        dims = list(struct.unpack_from(f"<{ndims}Q", data, off)) if ndims else []
        return dims

    def _parse_datatype(self, data, off):
        cv = data[off]
        cls, version = cv & 0x0F, cv >> 4
        bits = data[off + 1] | (data[off + 2] << 8) | (data[off + 3] << 16)
        size = struct.unpack_from("<I", data, off + 4)[0]
        p = off + 8
        if cls == 0:  # fixed point
            order = ">" if bits & 1 else "<"
            signed = bool(bits & 0x8)
            return _Datatype(0, size, signed, order), p + 4
        if cls == 1:  # floating point
            order = ">" if bits & 1 else "<"
            return _Datatype(1, size, True, order), p + 12
        if cls == 3:  # string
            return _Datatype(3, size, strpad=bits & 0x0F), p
        if cls == 9:  # variable length
            vtype = bits & 0x3
            base, p2 = self._parse_datatype(data, p)
            return _Datatype(9, size, vlen_base=base, is_vlen_str=(vtype == 1)), p2
        if cls == 8:  # enumeration (h5py's bool: int8 base, FALSE = 0, TRUE = 1): read as its integers
            n_members = bits & 0xFFFF
            base, p2 = self._parse_datatype(data, p)
            for _ in range(n_members):
                end = data.index(b"\x00", p2)
                p2 += (end - p2 + 1 + 7) // 8 * 8
            p2 += n_members * base.size
            return _Datatype(0, size, base.signed, base.order), p2
        raise NotImplementedError(f"datatype class {cls} (version {version})")

    def _parse_attribute(self, data):
        version = data[0]
        if version != 1:
            raise NotImplementedError(f"attribute message version {version}")
        name_size, dt_size, ds_size = struct.unpack_from("<HHH", data, 2)
        p = 8
        name = data[p : p + name_size].rstrip(b"\x00").decode()
        p += (name_size + 7) // 8 * 8
        dtype = self._parse_datatype(data, p)[0]
        p += (dt_size + 7) // 8 * 8
        dims = self._parse_dataspace(data[p : p + ds_size])
        p += (ds_size + 7) // 8 * 8
        n = int(np.prod(dims)) if dims else 1
        raw = data[p : p + n * dtype.size]
        val = self._decode(raw, dtype, n, dims)
        return name, val

    def _decode(self, raw, dtype, n, dims):
        if dtype.cls in (0, 1):
            arr = np.frombuffer(raw, dtype=dtype.numpy_dtype(), count=n).astype(
                np.float64 if dtype.cls == 1 else np.int64
            )
            return arr.reshape(dims) if dims else arr[0]
        if dtype.cls == 3:
            out = [raw[i * dtype.size : (i + 1) * dtype.size].split(b"\x00")[0].decode().rstrip() for i in range(n)]
            return out if dims else out[0]
        if dtype.cls == 9:
            out = []
            for i in range(n):
                ln, gaddr, gidx = struct.unpack_from("<IQI", raw, i * 16)
                obj = self._global_heap_object(gaddr, gidx)
                out.append(obj.decode() if dtype.is_vlen_str else obj)
            return out if dims else out[0]
        raise NotImplementedError

    def _global_heap_object(self, addr, index):
        b = self.buf
        if b[addr : addr + 4] != b"GCOL":
            raise HDF5Error("bad global heap collection")
        csize = struct.unpack_from("<Q", b, addr + 8)[0]
        p = addr + 16
        end = addr + csize
        while p + 16 <= end:
            idx, _ref, _res, osize = struct.unpack_from("<HHIQ", b, p)
            if idx == 0:
                break
            if idx == index:
                return b[p + 16 : p + 16 + osize]
            p += 16 + (osize + 7) // 8 * 8
        raise HDF5Error("global heap object not found")

    def _read_layout(self, data, nbytes, dims, elsize):
        version = data[0]
        if version == 3:
            cls = data[1]
            if cls == 0:  # compact
                size = struct.unpack_from("<H", data, 2)[0]
                return data[4 : 4 + size]
            if cls == 1:  # contiguous
                addr, size = struct.unpack_from("<QQ", data, 2)
                if addr == 0xFFFFFFFFFFFFFFFF:
                    return b"\x00" * nbytes
                return self.buf[addr : addr + size]
            if cls == 2:  # chunked, v1 B-tree index
                ndim = data[2]
                btree = struct.unpack_from("<Q", data, 3)[0]
                cdims = struct.unpack_from(f"<{ndim}I", data, 11)
                return self._read_chunked(btree, cdims[:-1], dims, elsize)
        raise NotImplementedError(f"data layout version {version}")

    def _read_chunked(self, btree, cdims, dims, elsize):
        out = np.zeros(int(np.prod(dims)) * elsize, dtype=np.uint8)
        full = np.frombuffer(b"", dtype=np.uint8)
        del full
        nd = len(cdims)
        stack = [btree]
        while stack:
            a = stack.pop()
            b = self.buf
            if b[a : a + 4] != b"TREE" or b[a + 4] != 1:
                raise HDF5Error("bad chunk B-tree node")
            level = b[a + 5]
            used = struct.unpack_from("<H", b, a + 6)[0]
            p = a + 24
            ksize = 8 + 8 * (nd + 1)
            for _ in range(used):
                csize, fmask = struct.unpack_from("<II", b, p)
                offs = struct.unpack_from(f"<{nd}Q", b, p + 8)
                child = struct.unpack_from("<Q", b, p + ksize)[0]
                p += ksize + 8
                if level > 0:
                    stack.append(child)
                    continue
                if fmask != 0:
                    raise NotImplementedError("filtered chunks")
                chunk = np.frombuffer(self.buf[child : child + csize], dtype=np.uint8)
                self._scatter_chunk(out, chunk, offs, cdims, dims, elsize)
        return out.tobytes()

    @staticmethod
    def _scatter_chunk(out, chunk, offs, cdims, dims, elsize):
        src = chunk.reshape(*cdims, elsize) if len(cdims) else chunk
        dst = out.reshape(*dims, elsize)
        sl_dst, sl_src = [], []
        for o, c, d in zip(offs, cdims, dims):
            n = min(c, d - o)
            if n <= 0:
                return
            sl_dst.append(slice(o, o + n))
            sl_src.append(slice(0, n))
        dst[tuple(sl_dst)] = src[tuple(sl_src)]
