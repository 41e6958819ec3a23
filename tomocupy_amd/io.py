"""File formats on either side of the reconstruction path.

Input (reference: ``dataio/reader.py:274-359``, DXchange layout
``/exchange/data``, ``data_white``, ``data_dark``, ``theta`` in degrees):
* HDF5 files, read by :mod:`tomocupy_amd.h5` (no h5py needed; contiguous,
  chunked, compressed and virtual datasets);
* ``.npz`` archives with the same four keys (``exchange/`` prefix optional).
When ``theta`` is absent the angles are ``linspace(0, pi, nproj)`` including
the endpoint, as ``reader.py:317``.

Output (reference: ``dataio/writer.py``): one uncompressed single-strip
little-endian TIFF per slice (float32, or float16 for half plans, sample
format IEEE float), written by :func:`write_tiff`; the reference's ``h5``
(a virtual dataset over one part file per chunk) and ``h5nolinks`` (one
dataset) files, written by :class:`H5RecWriter`; or one ``.npy`` stack.
"""
import os
import struct

import numpy as np

_KEYS = ('data', 'data_white', 'data_dark', 'theta')


class Dataset:
    """Lazy access to the projections ``(nproj, nz, ni)`` uint16 and the
    flat/dark fields of a DXchange-like file."""

    def __init__(self, path):
        self.path = path
        ext = os.path.splitext(path)[1].lower()
        if ext in ('.h5', '.hdf5', '.hdf', '.nxs'):
            from . import h5
            self._f = h5.File(path)
            self._get = lambda k: self._f['/exchange/' + k] if ('/exchange/' + k) in self._f else None
        elif ext == '.npz':
            self._f = np.load(path, allow_pickle=False, mmap_mode='r')
            files = set(self._f.files)

            def get(k):
                for name in (k, 'exchange/' + k, 'exchange_' + k):
                    if name in files:
                        return self._f[name]
                return None
            self._get = get
        else:
            raise ValueError('unsupported input %s (expected .h5 or .npz)' % path)
        self.data = self._get('data')
        self.flat = self._get('data_white')
        self.dark = self._get('data_dark')
        if self.data is None or self.flat is None or self.dark is None:
            raise ValueError('%s lacks data, data_white or data_dark' % path)
        self.nproj, self.nz, self.ni = (int(v) for v in self.data.shape)
        th = self._get('theta')
        if th is not None and len(th) == self.nproj:
            self.theta = (np.asarray(th, dtype=np.float64) / 180.0 * np.pi).astype(np.float32)
        else:
            self.theta = np.linspace(0, np.pi, self.nproj).astype(np.float32)

    def rows(self, z0, z1, cols=None, proj=None):
        """``(data, dark, flat)`` uint16 for detector rows [z0, z1), columns
        ``cols = (x0, x1)`` (all by default) and projections ``proj`` (an index
        array or a slice; all by default, dark and flat frames always all)."""
        x0, x1 = (0, self.ni) if cols is None else (int(cols[0]), int(cols[1]))
        if proj is not None:
            d = np.asarray(self.data[:, z0:z1, x0:x1])[proj]
            return (np.ascontiguousarray(d, dtype=np.uint16),
                    np.ascontiguousarray(self.dark[:, z0:z1, x0:x1], dtype=np.uint16),
                    np.ascontiguousarray(self.flat[:, z0:z1, x0:x1], dtype=np.uint16))
        return (np.ascontiguousarray(self.data[:, z0:z1, x0:x1], dtype=np.uint16),
                np.ascontiguousarray(self.dark[:, z0:z1, x0:x1], dtype=np.uint16),
                np.ascontiguousarray(self.flat[:, z0:z1, x0:x1], dtype=np.uint16))


def write_tiff(path, img):
    """Write a 2-D float32 or float16 array as an uncompressed TIFF."""
    img = np.ascontiguousarray(img)
    if img.dtype not in (np.float32, np.float16) or img.ndim != 2:
        raise ValueError('write_tiff takes a 2-D float32/float16 array')
    h, w = img.shape
    bits = img.dtype.itemsize * 8
    data = img.astype(img.dtype.newbyteorder('<'), copy=False).tobytes()
    entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, bits), (259, 3, 1, 1),
               (262, 3, 1, 1), (273, 4, 1, 0), (277, 3, 1, 1), (278, 4, 1, h),
               (279, 4, 1, len(data)), (339, 3, 1, 3)]
    ifd_off = 8
    data_off = ifd_off + 2 + 12 * len(entries) + 4
    out = bytearray(b'II*\x00' + struct.pack('<I', ifd_off))
    out += struct.pack('<H', len(entries))
    for tag, typ, cnt, val in entries:
        if tag == 273:
            val = data_off
        if typ == 3:
            out += struct.pack('<HHIHH', tag, typ, cnt, val, 0)
        else:
            out += struct.pack('<HHII', tag, typ, cnt, val)
    out += struct.pack('<I', 0)
    out += data
    with open(path, 'wb') as f:
        f.write(out)


def read_tiff(path):
    """Read a TIFF written by :func:`write_tiff`."""
    b = open(path, 'rb').read()
    if b[:4] != b'II*\x00':
        raise ValueError('not a little-endian TIFF')
    off = struct.unpack_from('<I', b, 4)[0]
    n = struct.unpack_from('<H', b, off)[0]
    tags = {}
    for i in range(n):
        tag, typ, cnt = struct.unpack_from('<HHI', b, off + 2 + 12 * i)
        fmt = '<H' if typ == 3 else '<I'
        tags[tag] = struct.unpack_from(fmt, b, off + 2 + 12 * i + 8)[0]
    w, h, bits = tags[256], tags[257], tags[258]
    dt = np.float32 if bits == 32 else np.float16
    return np.frombuffer(b, dtype=np.dtype(dt).newbyteorder('<'), count=w * h,
                         offset=tags[273]).reshape(h, w).astype(dt)


class H5RecWriter:
    """The reference's HDF5 outputs of a full reconstruction
    (``dataio/writer.py:127-194, 277-285``).

    ``h5``: ``<base>.h5`` holds ``/exchange/data`` ``(nz, n, n)`` as a virtual
    dataset over ``<base>_parts/p%04d.h5`` (one ``/exchange/data`` per chunk,
    written as the chunk arrives); ``h5nolinks``: ``<base>.h5`` holds
    ``/exchange/data`` itself, filled chunk by chunk.  Both carry the
    attributes ``command``, ``axes`` = ``z:y:x``, ``description`` =
    ``ReconData`` and ``units`` = ``counts`` (writer.py:146-155).
    """

    def __init__(self, base, fmt, nz, n, dtype, chunk_rows, command):
        from . import h5
        self.base, self.fmt, self.dtype = base, fmt, np.dtype(dtype)
        attrs = {'command': command, 'axes': 'z:y:x', 'description': 'ReconData',
                 'units': 'counts'}
        self.fname = base + '.h5'
        if fmt == 'h5':
            parts = base + '_parts'
            os.makedirs(parts, exist_ok=True)
            rel = os.path.basename(parts)
            srcs = []
            for k, (st, lz) in enumerate(chunk_rows):
                srcs.append(('%s/p%04d.h5' % (rel, k), '/exchange/data',
                             ((st, 0, 0), (st + lz, n, n))))
            with h5.File(self.fname, 'w') as f:
                f.create_virtual_dataset('/exchange/data', (nz, n, n), self.dtype, srcs, attrs)
            self._f = None
        elif fmt == 'h5nolinks':
            self._f = h5.File(self.fname, 'w')
            self._d = self._f.create_dataset('/exchange/data', shape=(nz, n, n),
                                             dtype=self.dtype, attrs=attrs)
        else:
            raise ValueError('unknown HDF5 output format %r' % fmt)

    def write_chunk(self, k, st, rec):
        """Chunk ``k`` (``rec`` of shape ``(lz, n, n)``) at output slice ``st``."""
        from . import h5
        if self.fmt == 'h5':
            with h5.File('%s_parts/p%04d.h5' % (self.base, k), 'w') as f:
                f.create_dataset('/exchange/data', data=np.asarray(rec, self.dtype))
        else:
            self._d[st:st + len(rec)] = np.asarray(rec, self.dtype)

    def close(self):
        if self._f is not None:
            self._f.close()
            self._f = None
