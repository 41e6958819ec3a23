"""``python -m tomocupy_amd recon`` end to end on a small synthetic data set.

* full: every slice's TIFF against the FP64 oracle chain (per-slice rel-L2
  <= 1e-5; the chunking pairs slices within a chunk of nsino-per-chunk, as the
  conveyor does);
* try: the TIFF of each trial centre against the oracle reconstructed at that
  centre (<= 1e-5);
* a width with a prime factor 17 through fw stripe removal and the fused
  filter + backprojection (<= 1e-5).
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from oracle import preproc_fp64 as pp
from tomocupy_amd import io, synth
from tomocupy_amd.cli import main

pytestmark = pytest.mark.gpu


def _dataset(tmp_path, n=128, nproj=90, nz=6):
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(7, n, theta, nz, ndark=3, nflat=4)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    return str(p), data, dark, flat, theta


def _rel(got, ref):
    return np.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum())


def test_cli_full(tmp_path):
    path, data, dark, flat, theta = _dataset(tmp_path)
    out = str(tmp_path / 'rec')
    main(['recon', '--file-name', path, '--reconstruction-type', 'full', '--nsino-per-chunk', '4',
          '--out-path-name', out, '--fbp-filter', 'shepp', '--rotation-axis', '63.5'])
    n, nz = data.shape[2], data.shape[1]
    ne = 4 * n
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    for c0 in range(0, nz, 4):
        s = slice(c0, min(nz, c0 + 4))
        sino = pp.proc_chain(data[:, s], dark[:, s], flat[:, s])
        lz = sino.shape[0]
        sino = np.concatenate([sino, np.zeros((4 - lz,) + sino.shape[1:])])
        ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'shepp'), ne, 63.5),
                                th)[:lz]
        for k in range(lz):
            got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % (c0 + k)))
            err = _rel(got.astype(np.float64), ref[k])
            print('slice', c0 + k, err)
            assert err <= 1e-5


def test_cli_try(tmp_path):
    path, data, dark, flat, theta = _dataset(tmp_path, nz=4)
    out = str(tmp_path / 'try')
    main(['recon', '--file-name', path, '--reconstruction-type', 'try', '--out-path-name', out,
          '--center-search-width', '3', '--center-search-step', '1.5', '--nsino', '0.5'])
    n = data.shape[2]
    ne = 4 * n
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    idz = int(0.5 * (data.shape[1] - 1))
    sino = pp.proc_chain(data[:, idz:idz + 1], dark[:, idz:idz + 1], flat[:, idz:idz + 1])
    files = sorted(f for f in os.listdir(out) if f.endswith('.tiff'))
    shifts = np.arange(-3, 3, 1.5)
    assert len(files) == len(shifts)
    # the GPU batch: nsino-per-chunk (8) copies of the sinogram, shifts on the
    # first len(shifts) and zero on the rest, paired c with c + 4
    ncz = 8
    shv = np.zeros(ncz)
    shv[:len(shifts)] = shifts
    batch = np.concatenate([sino] * ncz)
    ref = fr.backprojection(ff.fbp_filter_center(batch, ff.calc_filter(ne, 'parzen'), ne, n / 2,
                                                 shv), th)
    for k, sh in enumerate(shifts):
        centre = n / 2 - sh
        got = io.read_tiff(os.path.join(out, 'recon_slice%04d_center%05.2f.tiff' % (idz, centre)))
        err = _rel(got.astype(np.float64), ref[k])
        print('centre', centre, err)
        assert err <= 1e-5


def test_cli_double_fov_full_and_try(tmp_path):
    # 360-degree scan with the axis near the left border: the CLI flips,
    # tapers and pads to 2 ni (reader.py:143-148, proc_functions.py:109-124)
    ni, nproj, nz = 64, 90, 2
    theta = (np.arange(nproj) * (2 * np.pi / nproj)).astype(np.float32)
    data, dark, flat, _ = synth.raw_counts(7, ni, theta, nz, ndark=3, nflat=4)
    p = tmp_path / 'scan360.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    centeri = 10.0
    n = 2 * ni
    ne = 4 * n
    sino = pp.proc_chain(data, dark, flat)
    padded, center, flip = pp.pad360(sino, centeri)
    assert flip
    out = str(tmp_path / 'rec360')
    main(['recon', '--file-name', str(p), '--reconstruction-type', 'full', '--nsino-per-chunk',
          '2', '--out-path-name', out, '--file-type', 'double_fov', '--rotation-axis',
          str(centeri)])
    ref = fr.backprojection(ff.fbp_filter_center(padded, ff.calc_filter(ne, 'parzen'), ne,
                                                 center), th)
    for k in range(nz):
        got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % k))
        assert got.shape == (n, n)
        err = _rel(got.astype(np.float64), ref[k])
        print('double-fov full slice', k, err)
        assert err <= 1e-5
    # try mode: the shifts are mirrored with the data (reader.py:233-234);
    # file names carry the centres of the original detector frame
    outt = str(tmp_path / 'try360')
    main(['recon', '--file-name', str(p), '--reconstruction-type', 'try', '--out-path-name',
          outt, '--file-type', 'double_fov', '--rotation-axis', str(centeri),
          '--center-search-width', '2', '--center-search-step', '1', '--nsino', '0.0',
          '--nsino-per-chunk', '4'])
    shifts = np.arange(-2, 2, 1.0)
    ncz = 4
    shv = np.zeros(ncz)
    shv[:len(shifts)] = -shifts
    batch = np.concatenate([padded[:1]] * ncz)
    reft = fr.backprojection(ff.fbp_filter_center(batch, ff.calc_filter(ne, 'parzen'), ne,
                                                  center, shv), th)
    for k, sh in enumerate(shifts):
        got = io.read_tiff(os.path.join(outt, 'recon_slice%04d_center%05.2f.tiff'
                                        % (0, centeri - sh)))
        err = _rel(got.astype(np.float64), reft[k])
        print('double-fov try centre', centeri - sh, err)
        assert err <= 1e-5


def test_cli_full_float16_crops_to_power_of_two(tmp_path):
    # reader.py:155-165: float16 crops the width to 2^int(log2 ni) (centred)
    # and moves the centre by (2^k - ni)/2; here 96 -> 64 columns from 16
    ni, nproj, nz = 96, 60, 2
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(3, ni, theta, nz, ndark=3, nflat=4)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    out = str(tmp_path / 'rec16')
    main(['recon', '--file-name', str(p), '--reconstruction-type', 'full', '--nsino-per-chunk',
          '2', '--out-path-name', out, '--dtype', 'float16', '--rotation-axis', '50'])
    n, st, centre = 64, 16, 50 + (64 - 96) / 2
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    sino = pp.proc_chain(data[:, :, st:st + n], dark[:, :, st:st + n], flat[:, :, st:st + n])
    ne = 256  # 4n, a power of two already
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, centre),
                            th)
    for k in range(nz):
        got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % k))
        assert got.shape == (n, n) and got.dtype == np.float16
        err = _rel(got.astype(np.float64), ref[k])
        print('slice', k, err)
        assert err <= 2e-3


def test_cli_full_projection_and_column_selection(tmp_path):
    # reader.py:124-181: --start-proj/--end-proj, --blocked-views (angles of
    # [0.5, 1.2] modulo pi dropped), --start-column/--end-column (width 70 ->
    # 70 - 70 % 2 = 70, centre moved by the start column)
    n0, nproj, nz = 100, 90, 2
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(4, n0, theta, nz, ndark=3, nflat=4)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    out = str(tmp_path / 'rec')
    main(['recon', '--file-name', str(p), '--reconstruction-type', 'full', '--nsino-per-chunk',
          '2', '--out-path-name', out, '--start-proj', '5', '--end-proj', '80',
          '--blocked-views', '[0.5,1.2]', '--start-column', '20', '--end-column', '90',
          '--rotation-axis', '55'])
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)[5:80]
    keep = ((th % np.pi) < 0.5) | (((th - 0.5) % np.pi) > 0.7)
    ids = np.arange(5, 80)[keep]
    assert 0 < len(ids) < 75
    n = 70
    sino = pp.proc_chain(data[ids][:, :, 20:90], dark[:, :, 20:90], flat[:, :, 20:90])
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, 35.0),
                            th[keep])
    for k in range(nz):
        got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % k)).astype(np.float64)
        assert got.shape == (n, n)
        err = _rel(got, ref[k])
        print('slice', k, err)
        assert err <= 1e-5


def test_cli_full_width_with_prime_17(tmp_path):
    # a detector width with a prime factor 17 (as 2448 = 2^4 3^2 17) through
    # recon with fw stripe removal: raw counts -> fw -> fused filter +
    # backprojection, every slice against the oracle chain
    path, data, dark, flat, theta = _dataset(tmp_path, n=136, nproj=60, nz=4)
    out = str(tmp_path / 'rec')
    main(['recon', '--file-name', path, '--reconstruction-type', 'full', '--nsino-per-chunk', '4',
          '--out-path-name', out, '--remove-stripe-method', 'fw', '--fw-level', '3',
          '--rotation-axis', '67.5'])
    n, nz = data.shape[2], data.shape[1]
    ne = 4 * n
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    sino = pp.proc_chain(data, dark, flat, stripe='fw', fw_level=3)
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, 67.5), th)
    for k in range(nz):
        got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % k))
        err = _rel(got.astype(np.float64), ref[k])
        print('slice', k, err)
        assert err <= 1e-5


@pytest.mark.parametrize("fmt", ['h5', 'h5nolinks'])
def test_cli_full_hdf5_in_and_out(tmp_path, fmt):
    """DXchange HDF5 input (written by tomocupy_amd.h5) and the reference's
    h5 (virtual dataset over part files) and h5nolinks outputs, rows 2..6 of 8:
    the output covers the whole height with the rows at their absolute
    positions (writer.py:130-141, 172-175); each slice against the oracle."""
    from tomocupy_amd import h5
    n, nproj, nz = 128, 90, 8
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(9, n, theta, nz, ndark=3, nflat=4)
    p = str(tmp_path / 'scan.h5')
    with h5.File(p, 'w') as f:
        f.create_dataset('/exchange/data', data=data)
        f.create_dataset('/exchange/data_white', data=flat)
        f.create_dataset('/exchange/data_dark', data=dark)
        f.create_dataset('/exchange/theta', data=theta.astype(np.float64) * 180 / np.pi)
    out = str(tmp_path / 'rec')
    main(['recon', '--file-name', p, '--reconstruction-type', 'full', '--nsino-per-chunk', '2',
          '--start-row', '2', '--end-row', '6', '--out-path-name', out, '--fbp-filter', 'parzen',
          '--save-format', fmt])
    ne = 4 * n
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    with h5.File(out + '.h5') as f:
        d = f['/exchange/data']
        assert d.shape == (nz, n, n)
        vol = d[...]
    assert not vol[:2].any() and not vol[6:].any()
    for c0 in (2, 4):
        s = slice(c0, c0 + 2)
        sino = pp.proc_chain(data[:, s], dark[:, s], flat[:, s])
        ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, n / 2),
                                th)
        for k in range(2):
            err = _rel(vol[c0 + k].astype(np.float64), ref[k])
            print(fmt, 'slice', c0 + k, err)
            assert err <= 1e-5
