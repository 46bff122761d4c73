"""A reader for the HDF5 files CMS-Flow reads and writes (XMDF datasets, the
model-parameter file with boundary cell strings and curves, the solution
files its test baselines hold), in pure Python with numpy and zlib, for
machines without h5py.

The reference converts its baselines to text with h5py
(testing/builder/h5_file_to_text.py:21-167: every group's attributes,
datasets and subgroups, recursively); `walk` visits a file in that order and
`read` returns one dataset.  Built from the HDF5 file format specification
(version 0/1 superblocks, version 1 object headers with continuation blocks,
symbol-table groups indexed by version 1 B-trees and local heaps, compact,
contiguous and chunked layouts with deflate and shuffle filters, fixed-point,
floating-point and string datatypes, dataspace and attribute messages of
versions 1-3, variable-length strings from the global heap), which is what
the HDF5 1.8-1.12 libraries write for files created with default properties.

    f = H5File("run_wse.h5")
    for path, kind, obj in f.walk(): ...
    times = f.read("Datasets/Water_Elevation/Times")
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(ValueError):
    pass


class _Dataset:
    def __init__(self, f, path, msgs):
        self.f, self.path, self.msgs = f, path, msgs

    @property
    def shape(self):
        return self.f._dataspace(self.msgs)

    @property
    def attrs(self):
        return self.f._attrs(self.msgs)

    def read(self):
        return self.f._read_data(self.msgs)


class _Group:
    def __init__(self, f, path, msgs):
        self.f, self.path, self.msgs = f, path, msgs

    @property
    def attrs(self):
        return self.f._attrs(self.msgs)

    def members(self):
        return self.f._members(self.msgs)


class H5File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        b = self.buf
        base = b.find(SIGNATURE)
        if base != 0:
            raise H5Error(f"{path}: not an HDF5 file (no signature at offset 0)")
        ver = b[8]
        if ver not in (0, 1):
            raise H5Error(f"{path}: superblock version {ver} not supported")
        self.so, self.sl = b[13], b[14]          # size of offsets, of lengths
        p = 24 if ver == 0 else 28
        p += 4 * self.so                          # base, free space, end of file, driver block
        self.root = self._symbol_entry(p)[1]
        self._heap_cache = {}

    # ------------------------------------------------------------ primitives
    def _u(self, p, n):
        return int.from_bytes(self.buf[p:p + n], "little")

    def _o(self, p):
        return self._u(p, self.so)

    def _l(self, p):
        return self._u(p, self.sl)

    def _symbol_entry(self, p):
        """(link name offset, object header address, cache type, scratch at)"""
        name = self._o(p)
        hdr = self._o(p + self.so)
        cache = self._u(p + 2 * self.so, 4)
        return name, hdr, cache, p + 2 * self.so + 8

    def _messages(self, addr):
        """Header messages of the version 1 object header at addr: [(type, flags, data)]."""
        b = self.buf
        if b[addr:addr + 4] == b"OHDR":
            raise H5Error("version 2 object headers are not supported")
        if b[addr] != 1:
            raise H5Error(f"object header version {b[addr]} at {addr}")
        nmsg = self._u(addr + 2, 2)
        size = self._u(addr + 8, 4)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            p, size = blocks.pop(0)
            end = p + size
            while p + 8 <= end and len(out) < nmsg:
                mtype, msize, flags = self._u(p, 2), self._u(p + 2, 2), b[p + 4]
                data = p + 8
                if mtype == 0x10:                         # continuation
                    blocks.append((self._o(data), self._l(data + self.so)))
                out.append((mtype, flags, data, msize))
                p = data + msize
        return out

    # --------------------------------------------------------------- groups
    def _local_heap(self, addr):
        if addr in self._heap_cache:
            return self._heap_cache[addr]
        if self.buf[addr:addr + 4] != b"HEAP":
            raise H5Error(f"no local heap at {addr}")
        seg = self._o(addr + 8 + 2 * self.sl)
        self._heap_cache[addr] = seg
        return seg

    def _heap_name(self, seg, off):
        p = seg + off
        return self.buf[p:self.buf.index(b"\0", p)].decode()

    def _btree_group(self, addr, seg, out):
        b = self.buf
        if b[addr:addr + 4] != b"TREE":
            raise H5Error(f"no B-tree node at {addr}")
        level, used = b[addr + 5], self._u(addr + 6, 2)
        p = addr + 8 + 2 * self.so + self.sl     # first key skipped
        for _ in range(used):
            child = self._o(p)
            p += self.so + self.sl
            if level > 0:
                self._btree_group(child, seg, out)
                continue
            if b[child:child + 4] != b"SNOD":
                raise H5Error(f"no symbol table node at {child}")
            nsym = self._u(child + 6, 2)
            q = child + 8
            for _ in range(nsym):
                name, hdr, _, _ = self._symbol_entry(q)
                out.append((self._heap_name(seg, name), hdr))
                q += 2 * self.so + 24
        return out

    def _members(self, msgs):
        for mtype, flags, p, _ in msgs:
            if mtype == 0x11:                          # symbol table: B-tree and local heap
                seg = self._local_heap(self._o(p + self.so))
                return self._btree_group(self._o(p), seg, [])
        return []

    def _node(self, path, addr):
        msgs = self._messages(addr)
        kinds = {m[0] for m in msgs}
        if 0x11 in kinds:
            return _Group(self, path, msgs)
        if 0x08 in kinds:
            return _Dataset(self, path, msgs)
        return _Group(self, path, msgs)

    def get(self, path):
        node = self._node("/", self.root)
        for part in [q for q in path.strip("/").split("/") if q]:
            if not isinstance(node, _Group):
                raise KeyError(path)
            found = dict(node.members())
            if part not in found:
                raise KeyError(path)
            node = self._node(node.path.rstrip("/") + "/" + part, found[part])
        return node

    def read(self, path):
        node = self.get(path)
        if not isinstance(node, _Dataset):
            raise KeyError(f"{path} is not a dataset")
        return node.read()

    def walk(self, node=None):
        """Yields (path, "group"|"dataset", node) depth first, a group before its
        members, members in name order (the order h5py lists them)."""
        node = node or self._node("/", self.root)
        yield node.path, "group", node
        for name, addr in sorted(node.members()):
            child = self._node(node.path.rstrip("/") + "/" + name, addr)
            if isinstance(child, _Group):
                yield from self.walk(child)
            else:
                yield child.path, "dataset", child

    # ------------------------------------------------- dataspace and datatype
    def _space(self, p):
        ver, ndims, flags = self.buf[p], self.buf[p + 1], self.buf[p + 2]
        q = p + (8 if ver == 1 else 4)
        if ver == 2 and self.buf[p + 3] == 0:            # scalar
            return ()
        if ver == 2 and self.buf[p + 3] == 2:            # null
            return None
        return tuple(self._l(q + self.sl * i) for i in range(ndims))

    def _dataspace(self, msgs):
        for mtype, _, p, _ in msgs:
            if mtype == 0x01:
                return self._space(p)
        return ()

    def _dtype(self, p):
        """(kind, numpy dtype or None, size): kind 'num', 'str' or 'vstr'."""
        b = self.buf
        cls, bits, size = b[p] & 0x0F, self._u(p + 1, 3), self._u(p + 4, 4)
        order = ">" if bits & 1 else "<"
        if cls == 0:
            return "num", np.dtype(f"{order}{'i' if bits & 8 else 'u'}{size}"), size
        if cls == 1:
            return "num", np.dtype(f"{order}f{size}"), size
        if cls == 3:
            return "str", None, size
        if cls == 9 and (bits & 0x0F) == 1:             # variable-length string
            return "vstr", None, size
        if cls == 9:
            raise H5Error("variable-length sequences are not supported")
        raise H5Error(f"datatype class {cls} not supported")

    def _msg_dtype(self, msgs):
        for mtype, flags, p, _ in msgs:
            if mtype == 0x03:
                if flags & 2:                            # shared: a committed datatype
                    ver = self.buf[p]
                    addr = self._o(p + (8 if ver == 1 else 2))
                    return self._msg_dtype(self._messages(addr))
                return self._dtype(p)
        raise H5Error("no datatype message")

    def _decode(self, raw, kind, dt, size, shape):
        n = int(np.prod(shape)) if shape else 1
        if kind == "num":
            a = np.frombuffer(raw[:n * size], dtype=dt).astype(dt.newbyteorder("="))
            return a.reshape(shape) if shape else a[0]
        if kind == "str":
            out = [raw[i * size:(i + 1) * size].split(b"\0")[0].decode(errors="replace") for i in range(n)]
            return np.array(out, dtype=object).reshape(shape) if shape else out[0]
        out = []                                         # vstr: length, global heap id
        for i in range(n):
            q = i * (4 + self.so + 4)
            length = int.from_bytes(raw[q:q + 4], "little")
            coll = int.from_bytes(raw[q + 4:q + 4 + self.so], "little")
            idx = int.from_bytes(raw[q + 4 + self.so:q + 8 + self.so], "little")
            out.append(self._global_heap(coll, idx)[:length].decode(errors="replace"))
        return np.array(out, dtype=object).reshape(shape) if shape else out[0]

    def _global_heap(self, addr, index):
        b = self.buf
        if b[addr:addr + 4] != b"GCOL":
            raise H5Error(f"no global heap collection at {addr}")
        end = addr + self._l(addr + 8)
        p = addr + 8 + self.sl
        while p + 8 + self.sl <= end:
            idx = self._u(p, 2)
            size = self._l(p + 8)
            if idx == 0:
                break
            if idx == index:
                return b[p + 8 + self.sl:p + 8 + self.sl + size]
            p += 8 + self.sl + ((size + 7) // 8) * 8
        raise H5Error(f"global heap object {index} not found")

    # ------------------------------------------------------------ attributes
    def _attrs(self, msgs):
        out = {}
        b = self.buf
        for mtype, _, p, _ in msgs:
            if mtype != 0x0C:
                continue
            ver = b[p]
            nlen, tlen, slen = self._u(p + 2, 2), self._u(p + 4, 2), self._u(p + 6, 2)
            q = p + (9 if ver == 3 else 8)
            pad = (lambda n: (n + 7) // 8 * 8) if ver == 1 else (lambda n: n)
            name = b[q:q + nlen].split(b"\0")[0].decode()
            q += pad(nlen)
            kind, dt, size = self._dtype(q)
            q += pad(tlen)
            shape = self._space(q)
            q += pad(slen)
            n = int(np.prod(shape)) if shape else 1
            stride = (4 + self.so + 4) if kind == "vstr" else size
            out[name] = self._decode(b[q:q + n * stride], kind, dt, size, shape)
        return out

    # ------------------------------------------------------------------ data
    def _filters(self, msgs):
        b = self.buf
        for mtype, _, p, _ in msgs:
            if mtype != 0x0B:
                continue
            ver, nf = b[p], b[p + 1]
            q = p + (8 if ver == 1 else 2)
            out = []
            for _ in range(nf):
                fid = self._u(q, 2)
                if ver == 1 or fid >= 256:
                    nlen = self._u(q + 2, 2)
                    q += 4
                else:
                    nlen = 0
                    q += 2
                ncd = self._u(q + 2, 2)
                q += 4
                q += ((nlen + 7) // 8 * 8) if ver == 1 else nlen
                cd = [self._u(q + 4 * i, 4) for i in range(ncd)]
                q += 4 * ncd + (4 if ver == 1 and ncd % 2 else 0)
                out.append((fid, cd))
            return out
        return []

    def _unfilter(self, raw, filters, mask, itemsize):
        for i, (fid, _) in reversed(list(enumerate(filters))):
            if mask >> i & 1:
                continue
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:                               # shuffle
                a = np.frombuffer(raw, dtype=np.uint8)
                n = a.size // itemsize
                raw = a[:n * itemsize].reshape(itemsize, n).T.tobytes() + a[n * itemsize:].tobytes()
            elif fid == 3:                               # Fletcher-32: drop the checksum
                raw = raw[:-4]
            else:
                raise H5Error(f"filter {fid} not supported")
        return raw

    def _chunks(self, addr, ndims, out):
        b = self.buf
        if b[addr:addr + 4] != b"TREE" or b[addr + 4] != 1:
            raise H5Error(f"no chunk B-tree node at {addr}")
        level, used = b[addr + 5], self._u(addr + 6, 2)
        ksize = 8 + 8 * (ndims + 1)
        p = addr + 8 + 2 * self.so
        for _ in range(used):
            csize, mask = self._u(p, 4), self._u(p + 4, 4)
            offs = tuple(self._u(p + 8 + 8 * i, 8) for i in range(ndims))
            child = self._o(p + ksize)
            if level > 0:
                self._chunks(child, ndims, out)
            else:
                out.append((offs, child, csize, mask))
            p += ksize + self.so
        return out

    def _read_data(self, msgs):
        b = self.buf
        kind, dt, size = self._msg_dtype(msgs)
        shape = self._dataspace(msgs)
        if shape is None:
            return None
        n = int(np.prod(shape)) if shape else 1
        stride = (4 + self.so + 4) if kind == "vstr" else size
        for mtype, _, p, _ in msgs:
            if mtype != 0x08:
                continue
            ver = b[p]
            if ver == 3:
                cls = b[p + 1]
                if cls == 0:
                    raw = b[p + 4:p + 4 + self._u(p + 2, 2)]
                    return self._decode(raw, kind, dt, size, shape)
                if cls == 1:
                    addr = self._o(p + 2)
                    if addr == UNDEF:
                        return self._decode(bytes(n * stride), kind, dt, size, shape)
                    return self._decode(b[addr:addr + n * stride], kind, dt, size, shape)
                ndims = b[p + 2] - 1
                tree = self._o(p + 3)
                cdims = [self._u(p + 3 + self.so + 4 * i, 4) for i in range(ndims)]
            else:
                ndims = b[p + 1] - 1 if b[p + 2] == 2 else b[p + 1]
                cls = b[p + 2]
                q = p + 8
                if cls == 0:
                    raise H5Error("compact layout of version 1/2 not supported")
                addr = self._o(q)
                q += self.so
                if cls == 1:
                    return self._decode(b[addr:addr + n * stride], kind, dt, size, shape)
                tree = addr
                cdims = [self._u(q + 4 * i, 4) for i in range(ndims)]
            # chunked: assemble the chunks into the array
            filters = self._filters(msgs)
            out = np.zeros(n * stride, dtype=np.uint8)
            full = np.zeros(tuple(shape) + (stride,), dtype=np.uint8) if shape else None
            if tree == UNDEF:
                return self._decode(out.tobytes(), kind, dt, size, shape)
            for offs, addr, csize, mask in self._chunks(tree, ndims, []):
                raw = self._unfilter(b[addr:addr + csize], filters, mask, stride)
                chunk = np.frombuffer(raw, dtype=np.uint8)[:int(np.prod(cdims)) * stride]
                chunk = chunk.reshape(tuple(cdims) + (stride,))
                dst = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cdims, shape))
                src = tuple(slice(0, d.stop - d.start) for d in dst)
                full[dst] = chunk[src]
            return self._decode(full.tobytes(), kind, dt, size, shape)
        raise H5Error("no data layout message")
