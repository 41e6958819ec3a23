"""Writes the HDF5 fixtures of tests/test_h5.py with h5py (a third-party
library, not the reference): run with an interpreter that has h5py, e.g.
``/opt/conda/bin/python3.9 tests/golden/make_h5_fixtures.py`` (h5py 3.3.0,
HDF5 1.10.6).  The arrays come from numpy's seeded generator and are saved
beside the files in ``expected.npz``, so the tests compare what the reader
decodes with what h5py was given, not with what h5py reads back.

Files (DXchange layout, reader.py:286-318: /exchange/data, data_white,
data_dark, theta in degrees):
* v0_contig.h5      default file format (superblock 0, symbol-table groups),
                    contiguous storage, theta big-endian float32
* v0_gzip.h5        chunked (1, nz, n) with shuffle + deflate, plus a
                    fletcher32 dataset and a big-endian uint16 dataset
* latest_farray.h5  libver='latest' (superblock 3, version-2 object headers,
                    link messages, layout version 4): fixed-array chunk
                    index, single-chunk index, a paged fixed array (1100
                    chunks), a filtered fixed array
* vds.h5 + vds_parts/p0000.h5, p0001.h5: a virtual dataset over two part
                    files, as the reference's h5 output (writer.py:127-160)
"""
import os

import h5py
import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h5')


def main():
    os.makedirs(os.path.join(HERE, 'vds_parts'), exist_ok=True)
    rng = np.random.default_rng(2024)
    exp = {}
    data = rng.integers(0, 65535, (6, 4, 10), dtype=np.uint16)
    flat = rng.integers(20000, 40000, (3, 4, 10), dtype=np.uint16)
    dark = rng.integers(90, 110, (2, 4, 10), dtype=np.uint16)
    theta = np.linspace(0, 180, 6).astype(np.float32)
    exp.update(data=data, flat=flat, dark=dark, theta=theta)
    with h5py.File(os.path.join(HERE, 'v0_contig.h5'), 'w') as f:
        f.create_dataset('/exchange/data', data=data)
        f.create_dataset('/exchange/data_white', data=flat)
        f.create_dataset('/exchange/data_dark', data=dark)
        f.create_dataset('/exchange/theta', data=theta.astype('>f4'))
        f.create_dataset('/measurement/instrument/name', data=np.arange(3, dtype=np.int32))
    with h5py.File(os.path.join(HERE, 'v0_gzip.h5'), 'w') as f:
        f.create_dataset('/exchange/data', data=data, chunks=(1, 4, 10), compression='gzip',
                         shuffle=True)
        f.create_dataset('/exchange/data_white', data=flat, chunks=(2, 3, 7), fletcher32=True)
        f.create_dataset('/exchange/data_dark', data=dark.astype('>u2'))
        f.create_dataset('/exchange/theta', data=theta.astype(np.float64))
    paged = rng.standard_normal((1100, 2, 3)).astype(np.float32)
    exp['paged'] = paged
    with h5py.File(os.path.join(HERE, 'latest_farray.h5'), 'w', libver='latest') as f:
        f.create_dataset('/exchange/data', data=data, chunks=(2, 3, 4))
        f.create_dataset('/exchange/data_white', data=flat, chunks=(3, 4, 10))
        f.create_dataset('/exchange/data_dark', data=dark, chunks=(1, 4, 10), compression='gzip')
        f.create_dataset('/exchange/theta', data=theta)
        f.create_dataset('/extra/paged', data=paged, chunks=(1, 2, 3))
    vol = rng.standard_normal((5, 3, 4)).astype(np.float32)
    exp['vol'] = vol
    layout = h5py.VirtualLayout(shape=vol.shape, dtype='float32')
    for k, (st, en) in enumerate([(0, 3), (3, 5)]):
        fn = 'vds_parts/p%04d.h5' % k
        with h5py.File(os.path.join(HERE, fn), 'w') as f:
            f.create_dataset('/exchange/data', data=vol[st:en], chunks=(1, 3, 4))
        layout[st:en] = h5py.VirtualSource(fn, '/exchange/data', shape=(en - st, 3, 4))
    with h5py.File(os.path.join(HERE, 'vds.h5'), 'w') as f:
        f.create_virtual_dataset('/exchange/data', layout)
    np.savez(os.path.join(HERE, 'expected.npz'), **exp)


if __name__ == '__main__':
    main()
