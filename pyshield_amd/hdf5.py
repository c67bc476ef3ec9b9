"""A reader for the HDF5 files netCDF-4 writes (NDSL's savepoints), in numpy.

No HDF5 library is importable here, so this module reads the file format
itself, from the published HDF5 File Format Specification (version 3.0).  It
covers what a netCDF-4 savepoint holds: the root group's datasets, their
attributes and their dimension names.

* superblock versions 0-3; object headers version 1 and 2 with continuation
  blocks;
* groups stored as symbol tables (v1 B-tree of symbol nodes over a local
  heap) and as links (link messages; dense storage in a fractal heap indexed
  by a version 2 B-tree);
* datatypes: fixed-point and floating-point numbers of either byte order,
  fixed-length strings, variable-length strings and sequences (global heap),
  object references;
* dataspaces: scalar, null and simple;
* data layouts: compact, contiguous and chunked, with the version 1 B-tree
  (layout message version 3) and the single-chunk, implicit and fixed-array
  chunk indexes (version 4); filters deflate, shuffle and fletcher32;
* attributes stored in the object header (compact storage).

Not read (a clear error when met): extensible-array and version 2 B-tree
chunk indexes (datasets with unlimited dimensions under the newest file
format), densely stored attributes, compound, enumerated, array and bitfield
datatypes, filters other than the three above, external files.

``read(path)`` returns {dataset name: Variable} for the datasets of the root
group; a Variable has ``data`` (numpy array), ``dims`` (netCDF dimension
names from the dimension-scale references, or None) and ``attrs``.
"""

import zlib
from dataclasses import dataclass, field

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = None  # an undefined address


class HDF5Error(ValueError):
    pass


def is_hdf5(path):
    """True when the file carries the HDF5 signature at 0, 512, 1024, ..."""
    with open(path, "rb") as f:
        off = 0
        while True:
            f.seek(off)
            s = f.read(8)
            if len(s) < 8:
                return False
            if s == SIGNATURE:
                return True
            off = 512 if off == 0 else off * 2


@dataclass
class Variable:
    data: np.ndarray
    dims: tuple = None
    attrs: dict = field(default_factory=dict)


@dataclass
class _Datatype:
    cls: int
    size: int
    dtype: object = None     # numpy dtype for numbers and fixed strings
    base: object = None      # base _Datatype of a variable-length type
    vlen_string: bool = False
    ref_type: int = 0


class _File:
    def __init__(self, raw):
        self.raw = raw
        self.b = memoryview(raw)
        self.seen = set()  # the storage forms met (read(..., features=))
        self._superblock()

    # -------------------------------------------------------------- primitives
    def u(self, off, n):
        return int.from_bytes(self.b[off:off + n], "little")

    def addr(self, off):
        v = self.u(off, self.so)
        return UNDEF if v == (1 << (8 * self.so)) - 1 else v + self.base

    def length(self, off):
        return self.u(off, self.sl)

    def sig(self, off, s):
        if self.b[off:off + 4] != s:
            raise HDF5Error(f"expected {s!r} at {off}, found {bytes(self.b[off:off + 4])!r}")

    # ------------------------------------------------------------- superblock
    def _superblock(self):
        off = 0
        while True:
            if off + 8 > len(self.b):
                raise HDF5Error("no HDF5 signature")
            if self.b[off:off + 8] == SIGNATURE:
                break
            off = 512 if off == 0 else off * 2
        ver = self.b[off + 8]
        self.seen.add(f"superblock {ver}")
        if ver in (0, 1):
            self.so, self.sl = self.b[off + 13], self.b[off + 14]
            p = off + 24 + (4 if ver == 1 else 0)
            self.base = 0
            base = self.u(p, self.so)
            self.base = base
            p += 4 * self.so  # base, free-space info, end of file, driver info
            # root group symbol table entry: link name offset, object header address
            self.root = self.addr(p + self.so)
        elif ver in (2, 3):
            self.so, self.sl = self.b[off + 9], self.b[off + 10]
            p = off + 12
            self.base = 0
            self.base = self.u(p, self.so)
            p += 3 * self.so  # base, superblock extension, end of file
            self.root = self.addr(p)
        else:
            raise HDF5Error(f"superblock version {ver} not supported")

    # ---------------------------------------------------------- object headers
    def messages(self, oh):
        """[(type, data offset, size, flags)] of the object header at oh."""
        out = []
        if self.b[oh:oh + 4] == b"OHDR":
            ver, flags = self.b[oh + 4], self.b[oh + 5]
            if ver != 2:
                raise HDF5Error(f"object header version {ver}")
            self.seen.add("object header 2")
            p = oh + 6
            if flags & 0x20:
                p += 16
            if flags & 0x10:
                p += 4
            nsz = 1 << (flags & 3)
            size0 = self.u(p, nsz)
            p += nsz
            blocks = [(p, p + size0)]
            hs = 6 if flags & 0x04 else 4  # message header (with creation order)
            while blocks:
                start, end = blocks.pop(0)
                q = start
                while q + hs <= end:
                    mtype, msize, mflags = self.b[q], self.u(q + 1, 2), self.b[q + 3]
                    q += hs
                    if q + msize > end:
                        break
                    out.append((mtype, q, msize, mflags))
                    if mtype == 0x10:
                        a, n = self.addr(q), self.length(q + self.so)
                        self.sig(a, b"OCHK")
                        blocks.append((a + 4, a + n - 4))
                    q += msize
                # the chunk ends with a 4-byte checksum (and a gap of < 4 bytes)
        else:
            ver = self.b[oh]
            if ver != 1:
                raise HDF5Error(f"object header version {ver} at {oh}")
            self.seen.add("object header 1")
            nmsg, size = self.u(oh + 2, 2), self.u(oh + 8, 4)
            blocks = [(oh + 16, oh + 16 + size)]
            while blocks and len(out) < nmsg + 64:
                start, end = blocks.pop(0)
                q = start
                while q + 8 <= end:
                    mtype, msize, mflags = self.u(q, 2), self.u(q + 2, 2), self.b[q + 4]
                    q += 8
                    out.append((mtype, q, msize, mflags))
                    if mtype == 0x10:
                        a, n = self.addr(q), self.length(q + self.so)
                        blocks.append((a, a + n))
                    q += msize
        return out

    # ----------------------------------------------------------------- groups
    def links(self, oh):
        """{name: object header address} of the group at oh (hard links)."""
        msgs = self.messages(oh)
        out = {}
        for mtype, p, _, _ in msgs:
            if mtype == 0x11:  # symbol table
                self.seen.add("group symbol table")
                self._symbol_table(self.addr(p), self.addr(p + self.so), out)
            elif mtype == 0x06:  # link
                self.seen.add("group compact links")
                name, a = self._link(p)
                if a is not None:
                    out[name] = a
            elif mtype == 0x02:  # link info: dense storage
                flags = self.b[p + 1]
                q = p + 2 + (8 if flags & 1 else 0)
                heap, bt = self.addr(q), self.addr(q + self.so)
                if heap is not None and bt is not None:
                    self.seen.add("group dense links")
                    fh = _FractalHeap(self, heap)
                    for rec in self._btree2_records(bt):
                        name, a = self._link_from(fh.read(rec[4:4 + fh.id_len]), 0)
                        if a is not None:
                            out[name] = a
        return out

    def _link(self, p):
        return self._link_from(self.b, p)

    def _link_from(self, buf, p):
        ver, flags = buf[p], buf[p + 1]
        if ver != 1:
            raise HDF5Error(f"link message version {ver}")
        q = p + 2
        ltype = 0
        if flags & 0x08:
            ltype = buf[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        nlen_size = 1 << (flags & 3)
        nlen = int.from_bytes(buf[q:q + nlen_size], "little")
        q += nlen_size
        name = bytes(buf[q:q + nlen]).decode("utf-8")
        q += nlen
        if ltype != 0:
            return name, None  # soft or external link
        v = int.from_bytes(buf[q:q + self.so], "little")
        return name, v + self.base

    def _symbol_table(self, btree, heap, out):
        self.sig(heap, b"HEAP")
        data = self.addr(heap + 8 + 2 * self.sl)

        def name_at(o):
            s = data + o
            e = self.raw.index(b"\x00", s)
            return bytes(self.b[s:e]).decode("utf-8")

        def walk(node):
            self.sig(node, b"TREE")
            level, used = self.b[node + 5], self.u(node + 6, 2)
            p = node + 8 + 2 * self.so
            for i in range(used):
                child = self.addr(p + self.sl + i * (self.sl + self.so))
                if level > 0:
                    walk(child)
                else:
                    self.sig(child, b"SNOD")
                    n = self.u(child + 6, 2)
                    e = child + 8
                    for _ in range(n):
                        out[name_at(self.length(e))] = self.addr(e + self.so)
                        e += 2 * self.so + 24
        walk(btree)

    def _btree2_records(self, hdr):
        """All records of the version 2 B-tree whose header is at hdr."""
        self.sig(hdr, b"BTHD")
        node_size, rec_size = self.u(hdr + 6, 4), self.u(hdr + 10, 2)
        depth = self.u(hdr + 12, 2)
        root = self.addr(hdr + 16)
        nroot = self.u(hdr + 16 + self.so, 2)
        # records per node at each depth and the widths of the child-pointer
        # counts (H5B2 node_info: max_nrec, cum_max_nrec)
        leaf_max = (node_size - 10) // rec_size
        nrec_size = _enc_size(leaf_max)  # width of every child's record count
        max_nrec, cum_max, cum_size = [leaf_max], [leaf_max], [0]
        for d in range(1, depth + 1):
            ptr = self.so + nrec_size + (cum_size[d - 1] if d > 1 else 0)
            m = (node_size - 10) // (rec_size + ptr)
            max_nrec.append(m)
            cum = (m + 1) * cum_max[d - 1] + m
            cum_max.append(cum)
            cum_size.append(_enc_size(cum))
        recs = []

        def visit(node, nrec, d):
            if d == 0:
                self.sig(node, b"BTLF")
                p = node + 6
                for i in range(nrec):
                    recs.append(bytes(self.b[p + i * rec_size:p + (i + 1) * rec_size]))
                return
            self.sig(node, b"BTIN")
            p = node + 6
            rp = p + nrec * rec_size
            nsz = nrec_size
            csz = cum_size[d - 1] if d > 1 else 0
            for i in range(nrec + 1):
                child = self.addr(rp)
                n = self.u(rp + self.so, nsz)
                rp += self.so + nsz + csz
                visit(child, n, d - 1)
                if i < nrec:
                    recs.append(bytes(self.b[p + i * rec_size:p + (i + 1) * rec_size]))

        if root is not None:
            self.seen.add(f"v2 B-tree depth {depth}")
            visit(root, nroot, depth)
        return recs

    # --------------------------------------------------------------- datasets
    def dataset(self, oh):
        """(data, attrs) of the dataset at oh; data None for a group."""
        msgs = self.messages(oh)
        types = {m[0] for m in msgs}
        attrs = {}
        for mtype, p, size, _ in msgs:
            if mtype == 0x0C:
                name, val = self._attribute(p)
                attrs[name] = val
            elif mtype == 0x15:
                q = p + 2 + (2 if self.b[p + 1] & 1 else 0)
                if self.addr(q) is not None:
                    raise HDF5Error("densely stored attributes are not read")
        if 0x08 not in types:
            return None, attrs
        dt = shape = layout = None
        filters = []
        for mtype, p, size, _ in msgs:
            if mtype == 0x01:
                shape = self._dataspace(p)
            elif mtype == 0x03:
                dt = self._datatype(p)[0]
            elif mtype == 0x08:
                layout = p
            elif mtype == 0x0B:
                filters = self._filters(p)
            elif mtype == 0x07:
                raise HDF5Error("external data files are not read")
        raw, nbytes = self._layout(layout, shape, dt, filters)
        return self._values(raw, dt, shape), attrs

    def _dataspace(self, p):
        ver, ndim, flags = self.b[p], self.b[p + 1], self.b[p + 2]
        if ver == 1:
            q = p + 8
        elif ver == 2:
            if self.b[p + 3] == 2:
                return None  # null dataspace
            q = p + 4
        else:
            raise HDF5Error(f"dataspace version {ver}")
        return tuple(self.length(q + i * self.sl) for i in range(ndim))

    def _datatype(self, p):
        """(_Datatype, encoded size) of the datatype message at p."""
        cv = self.b[p]
        cls, bits = cv & 0x0F, self.u(p + 1, 3)
        size = self.u(p + 4, 4)
        q = p + 8
        if cls == 0:
            order = ">" if bits & 1 else "<"
            kind = "i" if bits & 0x08 else "u"
            return _Datatype(cls, size, np.dtype(f"{order}{kind}{size}")), 12
        if cls == 1:
            order = ">" if bits & 1 else "<"
            if bits & 0x40:
                raise HDF5Error("VAX floating point")
            return _Datatype(cls, size, np.dtype(f"{order}f{size}")), 20
        if cls == 3:
            return _Datatype(cls, size, np.dtype(f"S{size}")), 8
        if cls == 7:
            return _Datatype(cls, size, ref_type=bits & 0x0F), 8
        if cls == 9:
            vtype = bits & 0x0F
            base, n = self._datatype(q)
            return _Datatype(cls, size, base=base, vlen_string=(vtype == 1)), 8 + n
        raise HDF5Error(f"datatype class {cls} is not read")

    def _filters(self, p):
        ver, n = self.b[p], self.b[p + 1]
        q = p + (8 if ver == 1 else 2)
        out = []
        for _ in range(n):
            fid = self.u(q, 2)
            if ver == 1 or fid >= 256:
                nlen = self.u(q + 2, 2)
                q += 4
            else:
                nlen = 0
                q += 2
            flags, nval = self.u(q, 2), self.u(q + 2, 2)
            q += 4
            if ver == 1:
                nlen = (nlen + 7) // 8 * 8
            q += nlen
            vals = [self.u(q + 4 * i, 4) for i in range(nval)]
            q += 4 * nval
            if ver == 1 and nval % 2:
                q += 4
            out.append((fid, flags, vals))
        return out

    def _layout(self, p, shape, dt, filters):
        ver = self.b[p]
        if ver < 3:
            raise HDF5Error(f"data layout version {ver}")
        lclass = self.b[p + 1]
        self.seen.add(f"layout {ver} class {lclass}")
        n = int(np.prod(shape)) if shape else (0 if shape is None else 1)
        nbytes = n * dt.size
        if lclass == 0:
            size = self.u(p + 2, 2)
            return bytes(self.b[p + 4:p + 4 + size]), nbytes
        if lclass == 1:
            a = self.addr(p + 2)
            if a is None:
                return bytes(nbytes), nbytes
            return bytes(self.b[a:a + nbytes]), nbytes
        if lclass != 2:
            raise HDF5Error(f"layout class {lclass}")
        if ver == 3:
            ndim = self.b[p + 2]
            bt = self.addr(p + 3)
            q = p + 3 + self.so
            cdims = [self.u(q + 4 * i, 4) for i in range(ndim)][:-1]
            self.seen.add("chunk index v1 B-tree")
            chunks = self._chunks_btree1(bt, len(cdims)) if bt is not None else []
        else:
            flags, ndim, enc = self.b[p + 2], self.b[p + 3], self.b[p + 4]
            q = p + 5
            cdims = [self.u(q + enc * i, enc) for i in range(ndim)][:-1]
            q += enc * ndim
            itype = self.b[q]
            self.seen.add(f"chunk index {itype}")
            q += 1
            chunks = self._chunks_v4(itype, q, flags, shape, cdims, dt, filters)
        return self._assemble(chunks, shape, cdims, dt, filters), nbytes

    def _chunks_btree1(self, node, nd):
        """[(offsets, address, size, filter mask)] of a version 1 B-tree of chunks."""
        out = []
        ksize = 8 + 8 * (nd + 1)

        def walk(a):
            self.sig(a, b"TREE")
            level, used = self.b[a + 5], self.u(a + 6, 2)
            p = a + 8 + 2 * self.so
            for i in range(used):
                k = p + i * (ksize + self.so)
                size, mask = self.u(k, 4), self.u(k + 4, 4)
                offs = tuple(self.u(k + 8 + 8 * d, 8) for d in range(nd))
                child = self.addr(k + ksize)
                if level > 0:
                    walk(child)
                else:
                    out.append((offs, child, size, mask))
        walk(node)
        return out

    def _chunks_v4(self, itype, q, flags, shape, cdims, dt, filters):
        nchunk = [-(-s // c) for s, c in zip(shape, cdims)]
        csize = int(np.prod(cdims)) * dt.size

        def offsets(i):
            idx = np.unravel_index(i, nchunk)
            return tuple(int(j) * c for j, c in zip(idx, cdims))

        if itype == 1:  # single chunk
            if flags & 2:
                size, mask = self.length(q), self.u(q + self.sl, 4)
                q += self.sl + 4
            else:
                size, mask = csize, 0
            return [((0,) * len(cdims), self.addr(q), size, mask)]
        if itype == 2:  # implicit: chunks in order, unfiltered
            a = self.addr(q)
            return [(offsets(i), a + i * csize, csize, 0) for i in range(int(np.prod(nchunk)))]
        if itype == 3:  # fixed array
            a = self.addr(q + 1)
            return self._fixed_array(a, offsets, csize, bool(filters))
        raise HDF5Error(f"chunk index type {itype} is not read (unlimited dimensions)")

    def _fixed_array(self, hdr, offsets, csize, filtered):
        self.sig(hdr, b"FAHD")
        esize, page_bits = self.b[hdr + 6], self.b[hdr + 7]
        nent = self.length(hdr + 8)
        db = self.addr(hdr + 8 + self.sl)
        if db is None:
            return []
        self.sig(db, b"FADB")
        p = db + 6 + self.so
        page = 1 << page_bits
        paged = nent > page
        if paged:
            npages = -(-nent // page)
            p += (npages + 7) // 8 + 4  # page-init bitmap and the block's checksum
        out = []
        for i in range(nent):
            if paged and i and i % page == 0:
                p += 4  # the previous page's checksum
            a = self.addr(p)
            if filtered:
                size = self.u(p + self.so, esize - self.so - 4)
                mask = self.u(p + esize - 4, 4)
            else:
                size, mask = csize, 0
            if a is not None:
                out.append((offsets(i), a, size, mask))
            p += esize
        return out

    def _assemble(self, chunks, shape, cdims, dt, filters):
        arr = np.zeros(tuple(shape), dtype=np.dtype(f"V{dt.size}"))
        csize = int(np.prod(cdims)) * dt.size
        for offs, a, size, mask in chunks:
            raw = bytes(self.b[a:a + size])
            raw = _unfilter(raw, filters, mask, dt.size, self.seen)
            if len(raw) < csize:
                raise HDF5Error("short chunk")
            c = np.frombuffer(raw[:csize], dtype=arr.dtype).reshape(cdims)
            sl = tuple(slice(o, min(o + cd, s)) for o, cd, s in zip(offs, cdims, shape))
            arr[sl] = c[tuple(slice(0, x.stop - x.start) for x in sl)]
        return arr.tobytes()

    def _values(self, raw, dt, shape):
        if shape is None:
            return None
        n = int(np.prod(shape)) if shape else 1
        if dt.cls in (0, 1, 3):
            a = np.frombuffer(raw[:n * dt.size], dtype=dt.dtype).reshape(shape)
            if dt.cls == 3:
                return np.array([s.split(b"\x00")[0].decode("utf-8", "replace") for s in a.ravel()],
                                dtype=object).reshape(shape)
            return a.astype(dt.dtype.newbyteorder("="))
        if dt.cls == 7:
            return np.array([self.u_bytes(raw, i * dt.size, self.so) + self.base for i in range(n)],
                            dtype=object).reshape(shape)
        if dt.cls == 9:
            out = []
            step = 4 + self.so + 4
            for i in range(n):
                o = i * step
                cnt = self.u_bytes(raw, o, 4)
                coll = self.u_bytes(raw, o + 4, self.so) + self.base
                idx = self.u_bytes(raw, o + 4 + self.so, 4)
                data = self._global_heap(coll, idx) if cnt else b""
                if dt.vlen_string:
                    out.append(data[:cnt].split(b"\x00")[0].decode("utf-8", "replace"))
                elif dt.base.cls == 7:
                    out.append([self.u_bytes(data, j * dt.base.size, self.so) + self.base for j in range(cnt)])
                else:
                    out.append(np.frombuffer(data[:cnt * dt.base.size], dtype=dt.base.dtype).copy())
            return np.array(out + [None], dtype=object)[:-1].reshape(shape)
        raise HDF5Error(f"datatype class {dt.cls}")

    @staticmethod
    def u_bytes(buf, off, n):
        return int.from_bytes(buf[off:off + n], "little")

    def _global_heap(self, coll, idx):
        self.sig(coll, b"GCOL")
        size = self.length(coll + 8)
        p, end = coll + 8 + self.sl, coll + size
        while p + 8 + self.sl <= end:
            i = self.u(p, 2)
            osz = self.length(p + 8)
            if i == 0:
                break
            if i == idx:
                return bytes(self.b[p + 8 + self.sl:p + 8 + self.sl + osz])
            p += 8 + self.sl + (osz + 7) // 8 * 8
        raise HDF5Error(f"global heap object {idx} not found in {coll}")

    def _attribute(self, p):
        ver = self.b[p]
        nlen, tlen, slen = self.u(p + 2, 2), self.u(p + 4, 2), self.u(p + 6, 2)
        q = p + 8 + (1 if ver == 3 else 0)
        pad = (lambda x: (x + 7) // 8 * 8) if ver == 1 else (lambda x: x)
        name = bytes(self.b[q:q + nlen]).split(b"\x00")[0].decode("utf-8")
        q += pad(nlen)
        try:
            dt = self._datatype(q)[0]
        except HDF5Error:
            return name, None  # a datatype this reader does not decode (e.g. REFERENCE_LIST's compound)
        q += pad(tlen)
        shape = self._dataspace(q)
        q += pad(slen)
        if shape is None:
            return name, None
        n = int(np.prod(shape)) if shape else 1
        esize = (4 + self.so + 4) if dt.cls == 9 else dt.size
        val = self._values(bytes(self.b[q:q + n * esize]), dt, shape)
        if shape == () and isinstance(val, np.ndarray):
            val = val[()]
        return name, val


def _enc_size(n):
    """Bytes that hold a count up to n (H5VM_limit_enc_size)."""
    return (max(int(n), 1).bit_length() - 1) // 8 + 1


def _unfilter(raw, filters, mask, typesize, seen=None):
    """Undo the filter pipeline (applied in order on write; undone in reverse)."""
    for i in range(len(filters) - 1, -1, -1):
        if mask & (1 << i):
            continue
        fid, flags, vals = filters[i]
        if seen is not None:
            seen.add(f"filter {fid}")
        if fid == 1:
            raw = zlib.decompress(raw)
        elif fid == 2:
            ts = vals[0] if vals else typesize
            n = len(raw) // ts
            a = np.frombuffer(raw[:n * ts], dtype=np.uint8).reshape(ts, n)
            raw = a.T.tobytes() + raw[n * ts:]
        elif fid == 3:
            raw = raw[:-4]
        else:
            raise HDF5Error(f"filter {fid} is not read")
    return raw


class _FractalHeap:
    """Managed objects of a fractal heap (dense link storage)."""

    def __init__(self, f, hdr):
        self.f = f
        f.sig(hdr, b"FRHP")
        so, sl = f.so, f.sl
        self.id_len = f.u(hdr + 5, 2)
        io_len = f.u(hdr + 7, 2)
        self.flags = f.b[hdr + 9]
        max_man = f.u(hdr + 10, 4)
        p = hdr + 10 + 4 + sl + so + sl + so + sl + sl + sl + sl + sl + sl + sl + sl
        self.width = f.u(p, 2)
        self.start = f.length(p + 2)
        self.max_direct = f.length(p + 2 + sl)
        self.max_heap_bits = f.u(p + 2 + 2 * sl, 2)
        p += 2 + 2 * sl + 2 + 2
        self.root = f.addr(p)
        self.rows = f.u(p + so, 2)
        if io_len:
            raise HDF5Error("filtered fractal heaps are not read")
        # heap ID: offset in ceil(max heap bits / 8) bytes, then the length in
        # the smaller of the bytes of an offset inside the largest direct
        # block and the bytes that hold the largest managed object
        self.off_bytes = (self.max_heap_bits + 7) // 8
        self.len_bytes = min((self.max_direct.bit_length() - 1 + 7) // 8, _enc_size(max_man))
        self.checksummed = bool(self.flags & 2)

    def _direct_blocks(self):
        """[(heap offset, size, address)] of the managed direct blocks."""
        if self.rows == 0:
            self.f.seen.add("fractal heap root direct block")
            return [(0, self.start, self.root)]
        self.f.seen.add("fractal heap root indirect block")
        out = []
        self._indirect(self.root, 0, self.rows, out)
        return out

    def _row_size(self, r):
        return self.start * (1 << max(0, r - 1))

    def _indirect(self, a, base, nrows, out):
        f = self.f
        f.sig(a, b"FHIB")
        p = a + 5 + f.so + self.off_bytes
        max_direct_rows = (self.max_direct // self.start).bit_length() + 1
        off = base
        for r in range(nrows):
            size = self._row_size(r)
            for _ in range(self.width):
                child = f.addr(p)
                p += f.so
                if r < max_direct_rows:
                    if child is not None:
                        out.append((off, size, child))
                elif child is not None:
                    # rows of a child indirect block spanning `size` bytes
                    first_bits = (self.start.bit_length() - 1) + (self.width.bit_length() - 1)
                    self._indirect(child, off, (size.bit_length() - 1) - first_bits + 1, out)
                off += size

    def read(self, hid):
        t = (hid[0] >> 4) & 3
        if t != 0:
            raise HDF5Error("huge or tiny fractal-heap objects are not read")
        off = int.from_bytes(hid[1:1 + self.off_bytes], "little")
        n = int.from_bytes(hid[1 + self.off_bytes:1 + self.off_bytes + self.len_bytes], "little")
        for base, size, a in self._direct_blocks():
            if base <= off < base + size:
                return bytes(self.f.b[a + off - base:a + off - base + n])
        raise HDF5Error(f"fractal heap offset {off} not found")


def read(path, features=None):
    """{name: Variable} of the datasets in the root group of an HDF5 file.

    Dimension names come from the netCDF-4 / HDF5 dimension-scale convention:
    a variable's DIMENSION_LIST attribute references the dimension-scale
    datasets, whose link names are the dimensions.  ``features`` (a set), if
    given, receives the storage forms the file used."""
    with open(path, "rb") as fh:
        f = _File(fh.read())
    links = f.links(f.root)
    by_addr = {a: n for n, a in links.items()}
    out = {}
    for name, a in links.items():
        data, attrs = f.dataset(a)
        if data is None:
            continue
        out[name] = Variable(np.asarray(data), None, attrs)
    for name, v in out.items():
        dl = v.attrs.get("DIMENSION_LIST")
        if dl is not None:
            dims = []
            for refs in np.asarray(dl, dtype=object).ravel():
                refs = list(refs) if refs is not None else []
                dims.append(by_addr.get(refs[0]) if refs else None)
            v.dims = tuple(dims)
        elif v.attrs.get("CLASS") in ("DIMENSION_SCALE", b"DIMENSION_SCALE"):
            v.dims = (name,) * v.data.ndim if v.data.ndim == 1 else None
    if features is not None:
        features |= f.seen
    return out
