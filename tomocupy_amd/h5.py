"""HDF5 files without h5py: a reader for the datasets tomocupy reads and a
writer for the files it writes.

The reference reads DXchange files with h5py (``dataio/reader.py:274-359``:
``/exchange/data``, ``/exchange/data_white``, ``/exchange/data_dark``,
``/exchange/theta``) and writes its ``h5nolinks`` output with h5py
(``dataio/writer.py:172-194``).  h5py is not a dependency here; this module
implements the parts of the HDF5 file format (version 1.x specification)
those files use, on numpy and zlib:

Reading: superblock versions 0-3; object headers versions 1 and 2 (with
continuation blocks); groups as symbol tables (v1 B-tree + local heap) or as
compact link messages; datasets with compact, contiguous or chunked storage
(layout messages versions 1-4; chunk indexes: v1 B-tree, single chunk,
implicit, fixed array with or without paging); filters deflate, shuffle and
fletcher32 (the checksum is not verified); integer and IEEE float element
types of either byte order.  Dense link storage, extensible-array and v2
B-tree chunk indexes, variable-length and compound types raise
``NotImplementedError``.  Datasets are read lazily (``Dataset[...]``
decodes only the chunks a selection touches).

Writing (``File(path, 'w')``): superblock version 0, symbol-table groups,
contiguous little-endian datasets with optional fixed-length string
attributes, written in place by slices (``dset[st:end] = block``).
"""
import os
import struct
import threading
import zlib

import numpy as np

SIGNATURE = b'\x89HDF\r\n\x1a\n'
UNDEF = 0xFFFFFFFFFFFFFFFF


class FormatError(ValueError):
    pass


def _rot(x, k):
    return ((x << k) | (x >> (32 - k))) & 0xFFFFFFFF


def lookup3(data, initval=0):
    """Bob Jenkins' lookup3 ``hashlittle`` (public domain), the checksum of
    HDF5 metadata blocks."""
    M = 0xFFFFFFFF
    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & M
    i = 0
    while n > 12:
        a = (a + int.from_bytes(data[i:i + 4], 'little')) & M
        b = (b + int.from_bytes(data[i + 4:i + 8], 'little')) & M
        c = (c + int.from_bytes(data[i + 8:i + 12], 'little')) & M
        a = (a - c) & M; a ^= _rot(c, 4); c = (c + b) & M      # noqa: E702
        b = (b - a) & M; b ^= _rot(a, 6); a = (a + c) & M      # noqa: E702
        c = (c - b) & M; c ^= _rot(b, 8); b = (b + a) & M      # noqa: E702
        a = (a - c) & M; a ^= _rot(c, 16); c = (c + b) & M     # noqa: E702
        b = (b - a) & M; b ^= _rot(a, 19); a = (a + c) & M     # noqa: E702
        c = (c - b) & M; c ^= _rot(b, 4); b = (b + a) & M      # noqa: E702
        n -= 12
        i += 12
    if n == 0:
        return c
    t = data[i:] + bytes(12 - n)
    a = (a + int.from_bytes(t[0:4], 'little')) & M
    b = (b + int.from_bytes(t[4:8], 'little')) & M
    c = (c + int.from_bytes(t[8:12], 'little')) & M
    c ^= b; c = (c - _rot(b, 14)) & M   # noqa: E702
    a ^= c; a = (a - _rot(c, 11)) & M   # noqa: E702
    b ^= a; b = (b - _rot(a, 25)) & M   # noqa: E702
    c ^= b; c = (c - _rot(b, 16)) & M   # noqa: E702
    a ^= c; a = (a - _rot(c, 4)) & M    # noqa: E702
    b ^= a; b = (b - _rot(a, 14)) & M   # noqa: E702
    c ^= b; c = (c - _rot(b, 24)) & M   # noqa: E702
    return c


# Serialised dataspace selections of virtual-dataset mappings: "all"
# (type 3), or one hyperslab block [start, stop) per axis (type 2, version 1:
# rank, block count, then the inclusive corner coordinates).
def _sel_all():
    return struct.pack('<IIII', 3, 1, 0, 0)


def _sel_block(start, stop):
    rank = len(start)
    body = struct.pack('<II', rank, 1) + b''.join(struct.pack('<I', a) for a in start) + \
        b''.join(struct.pack('<I', b - 1) for b in stop)
    return struct.pack('<IIII', 2, 1, 0, len(body)) + body


def _sel_parse(b, p, shape):
    """(list of (start, stop) blocks or None for "all", next offset)."""
    stype, ver = struct.unpack_from('<II', b, p)
    if stype == 3:  # all
        return None, p + 16
    if stype == 2 and ver == 1:
        _, length = struct.unpack_from('<II', b, p + 8)
        rank, nb = struct.unpack_from('<II', b, p + 16)
        q = p + 24
        blocks = []
        for _ in range(nb):
            lo = struct.unpack_from('<%dI' % rank, b, q)
            hi = struct.unpack_from('<%dI' % rank, b, q + 4 * rank)
            q += 8 * rank
            blocks.append((lo, tuple(h + 1 for h in hi)))
        return blocks, p + 16 + length
    if stype == 2 and ver == 2:
        flags = b[p + 8]
        length, rank = struct.unpack_from('<II', b, p + 9)
        q = p + 17
        if not flags & 1:
            raise NotImplementedError('irregular hyperslab selections')
        dims = [struct.unpack_from('<QQQQ', b, q + 32 * d) for d in range(rank)]
        for st, stride, cnt, blk in dims:
            if cnt != 1 and stride != blk:
                raise NotImplementedError('strided hyperslab selections')
        lo = tuple(st for st, _, _, _ in dims)
        hi = tuple(st + (cnt * blk if cnt != UNDEF else sz - st)
                   for (st, _, cnt, blk), sz in zip(dims, shape))
        return [(lo, hi)], p + 13 + length
    raise NotImplementedError('selection type %d version %d' % (stype, ver))


# ---------------------------------------------------------------------------
# reading


class _Reader:
    """Low-level access to one file: sizes of offsets/lengths and helpers."""

    def __init__(self, f):
        self.f = f
        self.base = 0
        self.so = 8
        self.sl = 8
        self._superblock()

    def read(self, addr, n):
        self.f.seek(self.base + addr)
        b = self.f.read(n)
        if len(b) != n:
            raise FormatError('truncated file at address %d' % addr)
        return b

    def uint(self, b, off, size):
        return int.from_bytes(b[off:off + size], 'little')

    def _superblock(self):
        f = self.f
        pos = 0
        f.seek(0, os.SEEK_END)
        end = f.tell()
        while pos < end:
            f.seek(pos)
            if f.read(8) == SIGNATURE:
                break
            pos = 512 if pos == 0 else 2 * pos
        else:
            raise FormatError('not an HDF5 file')
        f.seek(pos)
        hdr = f.read(128)
        ver = hdr[8]
        if ver in (0, 1):
            self.so, self.sl = hdr[13], hdr[14]
            off = 24 + (4 if ver == 1 else 0)
            so = self.so
            base = self.uint(hdr, off, so)
            off += 4 * so  # base, free-space, end of file, driver info
            # root group symbol table entry
            self.root = self.uint(hdr, off + so, so)
            self.base = base if base != 0 else pos
            if base == 0 and pos != 0:
                self.base = pos
        elif ver in (2, 3):
            self.so, self.sl = hdr[9], hdr[10]
            so = self.so
            off = 12
            base = self.uint(hdr, off, so)
            self.root = self.uint(hdr, off + 3 * so, so)
            self.base = base
        else:
            raise FormatError('superblock version %d not supported' % ver)

    # -- object headers -------------------------------------------------------
    def messages(self, addr):
        """List of (type, data bytes) of the object header at addr."""
        b = self.read(addr, 4)
        if b == b'OHDR':
            return self._messages_v2(addr)
        return self._messages_v1(addr)

    def _messages_v1(self, addr):
        hdr = self.read(addr, 16)
        if hdr[0] != 1:
            raise FormatError('object header version %d at %d' % (hdr[0], addr))
        nmsg = self.uint(hdr, 2, 2)
        size = self.uint(hdr, 8, 4)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            b = self.read(start, length)
            p = 0
            while p + 8 <= length and len(out) < nmsg:
                mtype = self.uint(b, p, 2)
                msize = self.uint(b, p + 2, 2)
                data = b[p + 8:p + 8 + msize]
                p += 8 + msize
                p = (p + 7) & ~7
                if mtype == 0x10:
                    blocks.append((self.uint(data, 0, self.so), self.uint(data, self.so, self.sl)))
                out.append((mtype, data))
        return out

    def _messages_v2(self, addr):
        b = self.read(addr, 6)
        flags = b[5]
        p = 6
        if flags & 0x20:
            p += 16
        if flags & 0x10:
            p += 4
        szlen = 1 << (flags & 3)
        hdr = self.read(addr, p + szlen)
        size = self.uint(hdr, p, szlen)
        p += szlen
        blocks = [(addr + p, size, False)]
        out = []
        track = bool(flags & 0x04)
        while blocks:
            start, length, cont = blocks.pop(0)
            b = self.read(start, length)
            q = 4 if cont else 0
            stop = length - (4 if cont else 0)
            while q + 4 <= stop:
                mtype = b[q]
                msize = self.uint(b, q + 1, 2)
                q += 4 + (2 if track else 0)
                data = b[q:q + msize]
                q += msize
                if mtype == 0x10:
                    blocks.append((self.uint(data, 0, self.so), self.uint(data, self.so, self.sl),
                                   True))
                if mtype != 0:
                    out.append((mtype, data))
        return out

    # -- groups ---------------------------------------------------------------
    def links(self, addr):
        """name -> object header address of the group at addr."""
        out = {}
        for mtype, data in self.messages(addr):
            if mtype == 0x11:  # symbol table
                btree = self.uint(data, 0, self.so)
                heap = self.uint(data, self.so, self.so)
                heap_data = self._local_heap(heap)
                self._group_btree(btree, heap_data, out)
            elif mtype == 0x06:  # link
                name, target = self._link(data)
                if target is not None:
                    out[name] = target
            elif mtype == 0x02:  # link info: dense storage if a fractal heap exists
                flags = data[1]
                p = 2 + (8 if flags & 1 else 0)
                fheap = self.uint(data, p, self.so)
                if fheap != (1 << (8 * self.so)) - 1:
                    raise NotImplementedError('HDF5 groups with dense link storage')
        return out

    def _local_heap(self, addr):
        b = self.read(addr, 8 + 2 * self.sl + self.so)
        if b[:4] != b'HEAP':
            raise FormatError('local heap signature at %d' % addr)
        size = self.uint(b, 8, self.sl)
        data_addr = self.uint(b, 8 + 2 * self.sl, self.so)
        return self.read(data_addr, size)

    def _group_btree(self, addr, heap, out):
        so, sl = self.so, self.sl
        hdr = self.read(addr, 8 + 2 * so)
        if hdr[:4] != b'TREE':
            raise FormatError('B-tree signature at %d' % addr)
        level, used = hdr[5], self.uint(hdr, 6, 2)
        b = self.read(addr + 8 + 2 * so, (used + 1) * sl + used * so)
        for i in range(used):
            child = self.uint(b, sl + i * (sl + so), so)
            if level > 0:
                self._group_btree(child, heap, out)
            else:
                self._snod(child, heap, out)

    def _snod(self, addr, heap, out):
        so = self.so
        hdr = self.read(addr, 8)
        if hdr[:4] != b'SNOD':
            raise FormatError('symbol node signature at %d' % addr)
        n = self.uint(hdr, 6, 2)
        esz = 2 * so + 24
        b = self.read(addr + 8, n * esz)
        for i in range(n):
            e = b[i * esz:(i + 1) * esz]
            name_off = self.uint(e, 0, so)
            obj = self.uint(e, so, so)
            name = heap[name_off:heap.index(b'\0', name_off)].decode('utf-8')
            out[name] = obj

    def _link(self, data):
        flags = data[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = data[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        nlen_size = 1 << (flags & 3)
        nlen = self.uint(data, p, nlen_size)
        p += nlen_size
        name = data[p:p + nlen].decode('utf-8')
        p += nlen
        if ltype == 0:
            return name, self.uint(data, p, self.so)
        return name, None  # soft and external links are not followed

    # -- datasets -------------------------------------------------------------
    def dataset_info(self, addr):
        shape = dtype = layout = None
        filters = []
        for mtype, data in self.messages(addr):
            if mtype == 0x01:
                shape = self._dataspace(data)
            elif mtype == 0x03:
                dtype = self._datatype(data)
            elif mtype == 0x08:
                layout = data
            elif mtype == 0x0B:
                filters = self._filters(data)
        if shape is None or dtype is None or layout is None:
            return None
        return shape, dtype, layout, filters

    def _dataspace(self, data):
        ver, rank, flags = data[0], data[1], data[2]
        p = 8 if ver == 1 else 4
        if ver == 2 and data[3] == 0:
            return ()  # scalar
        return tuple(self.uint(data, p + i * self.sl, self.sl) for i in range(rank))

    def _datatype(self, data):
        cls = data[0] & 0x0F
        bits = data[1]
        size = self.uint(data, 4, 4)
        order = '>' if bits & 1 else '<'
        if cls == 0:
            kind = 'i' if bits & 0x08 else 'u'
            return np.dtype('%s%s%d' % (order, kind, size))
        if cls == 1:
            return np.dtype('%sf%d' % (order, size))
        if cls == 3:
            return np.dtype('S%d' % size)
        raise NotImplementedError('HDF5 datatype class %d' % cls)

    def _filters(self, data):
        ver, nf = data[0], data[1]
        p = 8 if ver == 1 else 2
        out = []
        for _ in range(nf):
            fid = self.uint(data, p, 2)
            p += 2
            nlen = 0
            if ver == 1 or fid >= 256:
                nlen = self.uint(data, p, 2)
                p += 2
            p += 2  # flags
            nval = self.uint(data, p, 2)
            p += 2
            if ver == 1 and nlen:
                p += (nlen + 7) & ~7
            elif nlen:
                p += nlen
            vals = [self.uint(data, p + 4 * i, 4) for i in range(nval)]
            p += 4 * nval
            if ver == 1 and nval % 2:
                p += 4
            out.append((fid, vals))
        return out


class Dataset:
    """A dataset of an open file: ``shape``, ``dtype``, ``ds[sel]`` reads."""

    def __init__(self, r, addr, name):
        info = r.dataset_info(addr)
        if info is None:
            raise FormatError('%s is not a dataset' % name)
        self._r = r
        self.name = name
        self.shape, self.dtype, layout, self._filters = info
        self._parse_layout(layout)

    @property
    def ndim(self):
        return len(self.shape)

    def __len__(self):
        return self.shape[0]

    def __array__(self, dtype=None, copy=None):
        a = self[...]
        return a if dtype is None else a.astype(dtype)

    def _parse_layout(self, d):
        r = self._r
        so, sl = r.so, r.sl
        ver = d[0]
        self._compact = self._contig = None
        self._chunk = None
        self._virtual = None
        if ver in (1, 2):
            rank, cls = d[1], d[2]
            p = 8
            addr = r.uint(d, p, so) if cls != 0 else None
            p += so if cls != 0 else 0
            dims = [r.uint(d, p + 4 * i, 4) for i in range(rank)]
            p += 4 * rank
            if cls == 0:
                size = r.uint(d, p, 4)
                self._compact = d[p + 4:p + 4 + size]
            elif cls == 1:
                self._contig = addr
            else:
                self._chunk = tuple(dims[:-1])
                self._index = ('btree1', addr)
            return
        cls = d[1]
        if cls == 0:
            size = r.uint(d, 2, 2)
            self._compact = d[4:4 + size]
        elif cls == 1:
            self._contig = r.uint(d, 2, so)
        elif cls == 2 and ver == 3:
            rank = d[2]
            addr = r.uint(d, 3, so)
            dims = [r.uint(d, 3 + so + 4 * i, 4) for i in range(rank)]
            self._chunk = tuple(dims[:-1])
            self._index = ('btree1', addr)
        elif cls == 2 and ver == 4:
            flags, rank, enc = d[2], d[3], d[4]
            p = 5
            dims = [r.uint(d, p + enc * i, enc) for i in range(rank)]
            p += enc * rank
            self._chunk = tuple(dims[:-1])
            itype = d[p]
            p += 1
            if itype == 1:  # single chunk
                if flags & 2:
                    fsize = r.uint(d, p, sl)
                    p += sl + 4
                    self._index = ('single', r.uint(d, p, so), fsize)
                else:
                    self._index = ('single', r.uint(d, p, so), None)
            elif itype == 2:  # implicit
                self._index = ('implicit', r.uint(d, p, so))
            elif itype == 3:  # fixed array
                page_bits = d[p]
                self._index = ('farray', r.uint(d, p + 1, so), page_bits)
            else:
                raise NotImplementedError('HDF5 chunk index type %d' % itype)
        elif cls == 3:
            self._virtual = self._parse_vds(r.uint(d, 2, so), r.uint(d, 2 + so, 4))
        else:
            raise FormatError('layout class %d' % cls)

    def _parse_vds(self, gaddr, index):
        r = self._r
        sl = r.sl
        hdr = r.read(gaddr, 8 + sl)
        if hdr[:4] != b'GCOL':
            raise FormatError('global heap signature at %d' % gaddr)
        raw = r.read(gaddr, r.uint(hdr, 8, sl))
        p = 8 + sl
        blob = None
        while p + 8 + sl <= len(raw):
            idx = r.uint(raw, p, 2)
            size = r.uint(raw, p + 8, sl)
            if idx == 0:
                break
            if idx == index:
                blob = raw[p + 8 + sl:p + 8 + sl + size]
                break
            p += 8 + sl + ((size + 7) & ~7)
        if blob is None:
            raise FormatError('virtual dataset mapping %d not found' % index)
        if lookup3(blob[:-4]) != r.uint(blob, len(blob) - 4, 4):
            raise FormatError('virtual dataset mapping checksum mismatch')
        n = r.uint(blob, 1, sl)
        q = 1 + sl
        out = []
        for _ in range(n):
            e = blob.index(b'\0', q)
            fname = blob[q:e].decode('utf-8')
            q = e + 1
            e = blob.index(b'\0', q)
            dname = blob[q:e].decode('utf-8')
            q = e + 1
            src, q = _sel_parse(blob, q, None)
            dst, q = _sel_parse(blob, q, self.shape)
            out.append((fname, dname, src, dst))
        return out

    def _read_virtual(self, idx):
        out = np.zeros(tuple(len(i) for i in idx), self.dtype.newbyteorder('='))
        here = os.path.dirname(os.path.abspath(self._r.f.name))
        for fname, dname, src, dst in self._virtual:
            path = fname if os.path.isabs(fname) or fname == '.' else os.path.join(here, fname)
            if fname == '.':
                path = self._r.f.name
            elif not os.path.exists(path):
                path = fname
            if not os.path.exists(path):
                continue  # missing source: fill value
            dblocks = dst if dst is not None else [((0,) * self.ndim, self.shape)]
            with File(path) as sf:
                sd = sf[dname]
                sblocks = src if src is not None else [((0,) * sd.ndim, sd.shape)]
                if len(sblocks) != 1 or len(dblocks) != 1:
                    raise NotImplementedError('virtual mappings of several blocks')
                (s0, _), (d0, d1) = sblocks[0], dblocks[0]
                sel_dst, sel_src = [], []
                for d, i in enumerate(idx):
                    m = (i >= d0[d]) & (i < d1[d])
                    sel_dst.append(np.nonzero(m)[0])
                    sel_src.append(i[m] - d0[d] + s0[d])
                if all(len(x) for x in sel_dst):
                    out[np.ix_(*sel_dst)] = sd._read(sel_src)
        return out

    # -- chunk lookup ---------------------------------------------------------
    def _chunk_table(self):
        """dict chunk offset tuple -> (address, stored size, filter mask)."""
        if hasattr(self, '_table'):
            return self._table
        kind = self._index[0]
        tab = {}
        if kind == 'btree1':
            self._btree_chunks(self._index[1], tab)
        elif kind == 'single':
            nbytes = self._index[2]
            if nbytes is None:
                nbytes = int(np.prod(self._chunk)) * self.dtype.itemsize
            tab[(0,) * self.ndim] = (self._index[1], nbytes, 0)
        elif kind == 'implicit':
            csz = int(np.prod(self._chunk)) * self.dtype.itemsize
            for i, off in enumerate(self._chunk_offsets()):
                tab[off] = (self._index[1] + i * csz, csz, 0)
        else:
            self._farray_chunks(tab)
        self._table = tab
        return tab

    def _chunk_offsets(self):
        grid = [range(0, s, c) for s, c in zip(self.shape, self._chunk)]
        return [tuple(o) for o in np.stack(np.meshgrid(*grid, indexing='ij'), -1).reshape(
            -1, self.ndim).tolist()] if all(len(g) for g in grid) else []

    def _btree_chunks(self, addr, tab):
        r = self._r
        so = r.so
        if addr == UNDEF or addr == (1 << (8 * so)) - 1:
            return
        hdr = r.read(addr, 8 + 2 * so)
        if hdr[:4] != b'TREE' or hdr[4] != 1:
            raise FormatError('chunk B-tree signature at %d' % addr)
        level, used = hdr[5], r.uint(hdr, 6, 2)
        ksz = 8 + 8 * (self.ndim + 1)
        b = r.read(addr + 8 + 2 * so, (used + 1) * ksz + used * so)
        for i in range(used):
            k = i * (ksz + so)
            size = r.uint(b, k, 4)
            mask = r.uint(b, k + 4, 4)
            off = tuple(r.uint(b, k + 8 + 8 * j, 8) for j in range(self.ndim))
            child = r.uint(b, k + ksz, so)
            if level > 0:
                self._btree_chunks(child, tab)
            else:
                tab[off] = (child, size, mask)

    def _farray_chunks(self, tab):
        r = self._r
        so, sl = r.so, r.sl
        _, haddr, _ = self._index
        h = r.read(haddr, 8 + sl + so + 4)
        if h[:4] != b'FAHD':
            raise FormatError('fixed array header signature at %d' % haddr)
        client, esz, page_bits = h[5], h[6], h[7]
        nent = r.uint(h, 8, sl)
        daddr = r.uint(h, 8 + sl, so)
        if daddr == (1 << (8 * so)) - 1:
            return
        page_n = 1 << page_bits
        prefix = 6 + so
        offs = self._chunk_offsets()
        if nent > page_n:
            npages = (nent + page_n - 1) // page_n
            prefix += (npages + 7) // 8 + 4
            entries = []
            p = daddr + prefix
            for pg in range(npages):
                cnt = min(page_n, nent - pg * page_n)
                entries.append(r.read(p, cnt * esz))
                p += cnt * esz + 4
            raw = b''.join(entries)
        else:
            raw = r.read(daddr + prefix, nent * esz)
        csz = int(np.prod(self._chunk)) * self.dtype.itemsize
        for i in range(min(nent, len(offs))):
            e = raw[i * esz:(i + 1) * esz]
            addr = r.uint(e, 0, so)
            if addr == (1 << (8 * so)) - 1:
                continue
            if client == 1:  # filtered chunks: address, size, mask
                sz_len = esz - so - 4
                size = r.uint(e, so, sz_len)
                mask = r.uint(e, so + sz_len, 4)
                tab[offs[i]] = (addr, size, mask)
            else:
                tab[offs[i]] = (addr, csz, 0)

    def _decode(self, raw, mask):
        for k, (fid, vals) in reversed(list(enumerate(self._filters))):
            if mask & (1 << k):
                continue
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                es = vals[0] if vals else self.dtype.itemsize
                a = np.frombuffer(raw, np.uint8)
                n = a.size // es
                raw = a[:n * es].reshape(es, n).T.tobytes() + a[n * es:].tobytes()
            elif fid == 3:
                raw = raw[:-4]
            else:
                raise NotImplementedError('HDF5 filter %d' % fid)
        return raw

    # -- reads ----------------------------------------------------------------
    def _norm(self, sel):
        if not isinstance(sel, tuple):
            sel = (sel,)
        if any(s is Ellipsis for s in sel):
            i = sel.index(Ellipsis)
            sel = sel[:i] + (slice(None),) * (self.ndim - len(sel) + 1) + sel[i + 1:]
        sel = sel + (slice(None),) * (self.ndim - len(sel))
        out, squeeze = [], []
        for d, (s, n) in enumerate(zip(sel, self.shape)):
            if isinstance(s, (int, np.integer)):
                s = int(s) + (n if s < 0 else 0)
                if not 0 <= s < n:
                    raise IndexError('index %d out of range for axis %d' % (s, d))
                out.append(np.array([s]))
                squeeze.append(d)
            elif isinstance(s, slice):
                out.append(np.arange(n)[s])
            else:
                idx = np.asarray(s)
                if idx.dtype == bool:
                    idx = np.nonzero(idx)[0]
                out.append(idx.astype(np.int64))
        return out, squeeze

    def __getitem__(self, sel):
        idx, squeeze = self._norm(sel)
        res = self._read(idx)
        if squeeze:
            res = res.reshape([len(i) for d, i in enumerate(idx) if d not in squeeze])
        return res

    def _read(self, idx):
        shape = tuple(len(i) for i in idx)
        dt = self.dtype
        if self._virtual is not None:
            return self._read_virtual(idx)
        if self._compact is not None or self._contig is not None:
            if self._compact is not None:
                full = np.frombuffer(self._compact, dt, int(np.prod(self.shape))).reshape(self.shape)
                return full[np.ix_(*idx)].astype(dt.newbyteorder('='))
            if self._contig == UNDEF or self._contig == (1 << (8 * self._r.so)) - 1:
                return np.zeros(shape, dt.newbyteorder('='))
            return self._read_contig(idx)
        out = np.zeros(shape, dt.newbyteorder('='))
        tab = self._chunk_table()
        ch = self._chunk
        # chunks touched by the selection
        per_axis = []
        for d, i in enumerate(idx):
            starts = np.unique((i // ch[d]) * ch[d])
            per_axis.append(starts)
        grids = np.meshgrid(*per_axis, indexing='ij') if per_axis else []
        for off in (zip(*[g.ravel() for g in grids]) if per_axis else []):
            off = tuple(int(o) for o in off)
            if off not in tab:
                continue  # unallocated: fill value 0
            addr, size, mask = tab[off]
            raw = self._decode(self._r.read(addr, size), mask)
            blk = np.frombuffer(raw, dt, int(np.prod(ch))).reshape(ch)
            src, dst = [], []
            for d, i in enumerate(idx):
                m = (i >= off[d]) & (i < off[d] + ch[d])
                dst.append(np.nonzero(m)[0])
                src.append(i[m] - off[d])
            out[np.ix_(*dst)] = blk[np.ix_(*src)]
        return out

    def _read_contig(self, idx):
        dt = self.dtype
        r = self._r
        out = np.empty(tuple(len(i) for i in idx), dt.newbyteorder('='))
        if self.ndim == 0:
            return np.frombuffer(r.read(self._contig, dt.itemsize), dt).astype(out.dtype).reshape(())
        inner = int(np.prod(self.shape[1:])) * dt.itemsize
        for k, i0 in enumerate(idx[0]):
            row = np.frombuffer(r.read(self._contig + int(i0) * inner, inner), dt).reshape(
                self.shape[1:])
            out[k] = row[np.ix_(*idx[1:])] if self.ndim > 1 else row
        return out


class Group:
    def __init__(self, r, addr, name):
        self._r = r
        self._addr = addr
        self.name = name
        self._links = None

    def _items(self):
        if self._links is None:
            self._links = self._r.links(self._addr)
        return self._links

    def keys(self):
        return list(self._items().keys())

    def __contains__(self, path):
        try:
            self[path]
            return True
        except KeyError:
            return False

    def __getitem__(self, path):
        node = self
        for part in [p for p in path.split('/') if p]:
            if not isinstance(node, Group):
                raise KeyError(path)
            items = node._items()
            if part not in items:
                raise KeyError(path)
            addr = items[part]
            name = node.name.rstrip('/') + '/' + part
            info_types = {t for t, _ in node._r.messages(addr)}
            if 0x08 not in info_types:  # no layout message: a group
                node = Group(node._r, addr, name)
            else:
                node = Dataset(node._r, addr, name)
        return node


class File(Group):
    """``File(path)`` reads; ``File(path, 'w')`` creates (see :class:`Writer`)."""

    def __new__(cls, path, mode='r'):
        if mode == 'w':
            return Writer(path)
        return super().__new__(cls)

    def __init__(self, path, mode='r'):
        self._f = open(path, 'rb')
        r = _Reader(self._f)
        super().__init__(r, r.root, '/')

    def close(self):
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# writing


def _pad8(b):
    return b + b'\0' * (-len(b) % 8)


class _WDataset:
    def __init__(self, w, addr, shape, dtype):
        self._w = w
        self.addr = addr
        self.shape = shape
        self.dtype = dtype

    def __setitem__(self, sel, value):
        """Writes whole leading-axis slices ``dset[st:end] = value``."""
        if not isinstance(sel, slice) or sel.step not in (None, 1):
            raise NotImplementedError('writes take a contiguous slice of the first axis')
        st, end, _ = sel.indices(self.shape[0])
        value = np.ascontiguousarray(value, dtype=self.dtype)
        if value.shape != (end - st,) + tuple(self.shape[1:]):
            raise ValueError('shape %s does not match the selection' % (value.shape,))
        inner = int(np.prod(self.shape[1:])) * self.dtype.itemsize
        with self._w.lock:  # writer threads share the file position
            self._w.f.seek(self.addr + st * inner)
            self._w.f.write(value.tobytes())


class Writer:
    """A new HDF5 file (superblock version 0, symbol-table groups, contiguous
    little-endian datasets).  Objects are laid out when the file is closed,
    except dataset storage, which is reserved when the dataset is created so
    that slices can be written as they come."""

    SO = 8

    def __init__(self, path):
        self.lock = threading.Lock()
        self.f = open(path, 'w+b')
        self.tree = {}            # nested dict: name -> dict (group) or _Node
        self.f.write(b'\0' * 4096)  # metadata region, rewritten on close
        self.meta_limit = 4096
        self.eof = 4096

    class _Node:
        def __init__(self, shape, dtype, addr, attrs):
            self.shape, self.dtype, self.addr, self.attrs = shape, dtype, addr, attrs

    def _group_of(self, path, create=True):
        parts = [p for p in path.split('/') if p]
        g = self.tree
        for p in parts[:-1]:
            if p not in g:
                if not create:
                    raise KeyError(path)
                g[p] = {}
            g = g[p]
            if not isinstance(g, dict):
                raise ValueError('%s is not a group' % p)
        return g, parts[-1]

    def create_dataset(self, path, shape=None, dtype=None, data=None, attrs=None):
        if data is not None:
            data = np.asarray(data)
            shape = data.shape if shape is None else tuple(shape)
            dtype = data.dtype if dtype is None else dtype
        dtype = np.dtype(dtype).newbyteorder('<')
        if dtype.kind not in 'uif':
            raise NotImplementedError('dtype %s' % dtype)
        shape = tuple(int(s) for s in shape)
        nbytes = int(np.prod(shape)) * dtype.itemsize
        addr = (self.eof + 7) & ~7
        self.f.seek(addr + nbytes - 1 if nbytes else addr)
        if nbytes:
            self.f.write(b'\0')
        self.eof = addr + nbytes
        g, name = self._group_of(path)
        node = Writer._Node(shape, dtype, addr, dict(attrs or {}))
        g[name] = node
        d = _WDataset(self, addr, shape, dtype)
        d.attrs = node.attrs
        if data is not None and nbytes:
            self.f.seek(addr)
            self.f.write(np.ascontiguousarray(data, dtype).tobytes())
        return d

    def create_virtual_dataset(self, path, shape, dtype, sources, attrs=None):
        """A virtual dataset: ``sources`` is a list of ``(file name, dataset
        name, (start, stop))`` mapping all of each source dataset onto the
        block [start, stop) of this one (relative file names resolve against
        this file's directory)."""
        dtype = np.dtype(dtype).newbyteorder('<')
        shape = tuple(int(s) for s in shape)
        g, name = self._group_of(path)
        node = Writer._Node(shape, dtype, None, dict(attrs or {}))
        node.virtual = [(f, d, tuple(int(a) for a in st), tuple(int(b) for b in sp))
                        for f, d, (st, sp) in sources]
        g[name] = node
        return node

    def __getitem__(self, path):
        g, name = self._group_of(path, create=False)
        node = g[name]
        d = _WDataset(self, node.addr, node.shape, node.dtype)
        d.attrs = node.attrs
        return d

    # -- serialisation --------------------------------------------------------
    def _msg(self, mtype, data):
        data = _pad8(data)
        return struct.pack('<HHB3x', mtype, len(data), 0) + data

    def _ohdr(self, msgs):
        body = b''.join(msgs)
        return struct.pack('<BBHII', 1, 0, len(msgs), 1, len(body)) + b'\0' * 4 + body

    def _dataspace(self, shape):
        return struct.pack('<BBB5x', 1, len(shape), 0) + b''.join(
            struct.pack('<Q', s) for s in shape)

    def _datatype(self, dt):
        if dt.kind in 'ui':
            bits = 0x08 if dt.kind == 'i' else 0
            return struct.pack('<B3BI', 0x10, bits, 0, 0, dt.itemsize) + struct.pack(
                '<HH', 0, 8 * dt.itemsize)
        # IEEE little endian: sign bit location, exponent/mantissa fields
        if dt.itemsize == 4:
            epos, esz, mpos, msz, bias = 23, 8, 0, 23, 127
            sign = 31
        elif dt.itemsize == 8:
            epos, esz, mpos, msz, bias = 52, 11, 0, 52, 1023
            sign = 63
        elif dt.itemsize == 2:
            epos, esz, mpos, msz, bias = 10, 5, 0, 10, 15
            sign = 15
        else:
            raise NotImplementedError('float size %d' % dt.itemsize)
        bits0 = 0x20  # mantissa normalisation: implied msb set
        return struct.pack('<B3BI', 0x11, bits0, sign, 0, dt.itemsize) + struct.pack(
            '<HHBBBBI', 0, 8 * dt.itemsize, epos, esz, mpos, msz, bias)

    def _string_type(self, n):
        # fixed-length string, null-terminated, UTF-8
        return struct.pack('<B3BI', 0x13, 0x10, 0, 0, n)

    def _attr(self, name, value):
        nb = name.encode('utf-8') + b'\0'
        if isinstance(value, (bytes, str)):
            vb = value.encode('utf-8') if isinstance(value, str) else value
            vb = vb + b'\0'
            dt = self._string_type(len(vb))
            ds = struct.pack('<BBB5x', 1, 0, 0)  # scalar
            raw = vb
        else:
            a = np.asarray(value)
            a = a.astype(a.dtype.newbyteorder('<'))
            dt = self._datatype(a.dtype)
            ds = self._dataspace(a.shape)
            raw = a.tobytes()
        data = struct.pack('<BxHHH', 1, len(nb), len(dt), len(ds)) + _pad8(nb) + _pad8(dt) + \
            _pad8(ds) + raw
        return self._msg(0x0C, data)

    def close(self):
        if self.f is None:
            return
        blobs = []          # (address, bytes)
        pos = [(self.eof + 7) & ~7]

        def alloc(n):
            a = pos[0]
            pos[0] = (a + n + 7) & ~7
            return a

        def write_group(g):
            names = sorted(g.keys())
            children = []
            for nm in names:
                v = g[nm]
                children.append(write_group(v) if isinstance(v, dict) else write_dset(v))
            # local heap: "" at offset 0, then the names
            heap = bytearray(b'\0' * 8)
            offs = []
            for nm in names:
                offs.append(len(heap))
                heap += _pad8(nm.encode('utf-8') + b'\0')
            heap += b'\0' * 16  # free block
            free_off = len(heap) - 16
            heap_data = bytes(heap)
            heap_data_addr = alloc(len(heap_data))
            blobs.append((heap_data_addr, heap_data))
            heap_hdr = b'HEAP' + bytes([0, 0, 0, 0]) + struct.pack(
                '<QQQ', len(heap_data), free_off, heap_data_addr)
            heap_addr = alloc(len(heap_hdr))
            # free-list block: next free offset (undefined = 1), size
            heap_data = heap_data[:free_off] + struct.pack('<QQ', 1, 16)
            blobs[-1] = (heap_data_addr, heap_data)
            blobs.append((heap_addr, heap_hdr))
            # symbol node with every entry (leaf K = 4 -> up to 8 entries per node:
            # one node per 8 names, one B-tree level)
            K = 4
            nodes = []
            for i in range(0, max(len(names), 1), 2 * K):
                ent = b''
                part = list(zip(names, offs, children))[i:i + 2 * K]
                for nm, off, ch in part:
                    if ch in stab:
                        ent += struct.pack('<QQII', off, ch, 1, 0) + struct.pack('<QQ', *stab[ch])
                    else:
                        ent += struct.pack('<QQII16x', off, ch, 0, 0)
                ent += b'\0' * ((2 * K - len(part)) * 40)
                snod = b'SNOD' + bytes([1, 0]) + struct.pack('<H', len(part)) + ent
                a = alloc(len(snod))
                blobs.append((a, snod))
                nodes.append((a, part[-1][1] if part else 0))
            if not names:
                nodes = []
            # B-tree (type 0, level 0); internal node K = 16 -> up to 32 children
            if len(nodes) > 32:
                raise NotImplementedError('more than 256 names in one group')
            keys = struct.pack('<Q', 0)
            for a, last in nodes:
                keys += struct.pack('<QQ', a, last)
            tree = b'TREE' + bytes([0, 0]) + struct.pack('<H', len(nodes)) + struct.pack(
                '<QQ', UNDEF, UNDEF) + keys
            tree += b'\0' * ((32 - len(nodes)) * 16)
            tree_addr = alloc(len(tree))
            blobs.append((tree_addr, tree))
            oh = self._ohdr([self._msg(0x11, struct.pack('<QQ', tree_addr, heap_addr))])
            a = alloc(len(oh))
            blobs.append((a, oh))
            stab[a] = (tree_addr, heap_addr)
            return a

        def write_dset(node):
            nbytes = int(np.prod(node.shape)) * node.dtype.itemsize
            if getattr(node, 'virtual', None) is not None:
                # mapping list in a global heap collection (checksummed blob),
                # layout message version 4, class 3
                blob = bytes([0]) + struct.pack('<Q', len(node.virtual))
                for fname, dname, st, sp in node.virtual:
                    blob += fname.encode('utf-8') + b'\0' + dname.encode('utf-8') + b'\0'
                    blob += _sel_all() + _sel_block(st, sp)
                blob += struct.pack('<I', lookup3(blob))
                obj = struct.pack('<HHIQ', 1, 0, 0, len(blob)) + _pad8(blob)
                size = max(4096, 16 + len(obj) + 16)
                free = size - 16 - len(obj)
                gcol = b'GCOL' + bytes([1, 0, 0, 0]) + struct.pack('<Q', size) + obj + \
                    struct.pack('<HHIQ', 0, 0, 0, free)
                gcol += b'\0' * (size - len(gcol))
                gaddr = alloc(size)
                blobs.append((gaddr, gcol))
                layout = struct.pack('<BBQI', 4, 3, gaddr, 1)
                fill = struct.pack('<BBBB', 2, 2, 2, 0)
            else:
                layout = struct.pack('<BBQQ', 3, 1, node.addr, nbytes)
                fill = struct.pack('<BBBB', 2, 1, 2, 0)
            msgs = [self._msg(0x01, self._dataspace(node.shape)),
                    self._msg(0x03, self._datatype(node.dtype)),
                    # fill value message version 2: no fill value defined
                    self._msg(0x05, fill),
                    self._msg(0x08, layout)]
            for k, v in node.attrs.items():
                msgs.append(self._attr(k, v))
            oh = self._ohdr(msgs)
            a = alloc(len(oh))
            blobs.append((a, oh))
            return a

        stab = {}  # group header address -> (B-tree, heap): cached in the entries
        root = write_group(self.tree)
        end = pos[0]
        for a, b in blobs:
            self.f.seek(a)
            self.f.write(b)
        sb = SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack('<HHI', 4, 16, 0)
        sb += struct.pack('<QQQQ', 0, UNDEF, end, UNDEF)
        sb += struct.pack('<QQII', 0, root, 1, 0) + struct.pack('<QQ', *stab[root])
        self.f.seek(0)
        self.f.write(sb)
        self.f.truncate(end)
        self.f.close()
        self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
