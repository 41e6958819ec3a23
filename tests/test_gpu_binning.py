"""Binning (``utils.downsample``, reference ``src/tomocupy/utils.py:157-167``,
applied by ``dataio/reader.py:350-389``) on the GPU:

* ``pre_bin`` against the oracle's block sums: bit-exact (float32 sums of at
  most 64 uint16 counts are exact);
* zinger removal on the binned float counts: bit-exact against the oracle;
* the whole chunk preprocessing on binned frames (none / fw / ti, dezinger)
  against the FP64 oracle chain: per-slice rel-L2 <= 1e-5;
* ``recon --binning 1`` (full and try) against the oracle: width trimmed to a
  multiple of 4, centre and sizes halved, raw slice ids in the try names.
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

RTOL32 = 1e-5


def _rel_slices(got, ref):
    nz = ref.shape[0]
    return np.sqrt(((got - ref) ** 2).reshape(nz, -1).sum(1) / (ref ** 2).reshape(nz, -1).sum(1))


@pytest.mark.parametrize("b", [0, 1, 2, 3])
def test_bin_bit_exact(b):
    from tomocupy_amd import ProcFunctions
    rows, nz, n = 5, 3, 37
    print('binning', b, flush=True)
    rng = np.random.default_rng(b)
    raw = rng.integers(0, 65536, size=(rows, nz << b, n << b)).astype(np.uint16)
    raw[0, 0, :] = 65535  # the largest sums: 65535 * 4^b
    pf = ProcFunctions(rows, nz, n, 1, 1, binning=b)
    got = pf.downsample(torch.from_numpy(raw).cuda())
    torch.cuda.synchronize()
    ref = pp.downsample(raw, b)
    assert got.dtype == torch.float32 and tuple(got.shape) == (rows, nz, n)
    assert np.array_equal(got.cpu().numpy().astype(np.float64), ref)
    pf.free()


@pytest.mark.parametrize("w", [3, 5])
def test_outliers_on_binned_counts_bit_exact(w):
    from tomocupy_amd import ProcFunctions
    rows, nz, n = 21, 2, 70
    rng = np.random.default_rng(30 + w)
    d = rng.integers(80000, 120000, size=(rows, nz, n)).astype(np.float32)
    hits = rng.random(d.shape) < 0.03
    d[hits] = rng.integers(160000, 260000, size=int(hits.sum())).astype(np.float32)
    pf = ProcFunctions(rows, nz, n, 1, 1, dezinger=w, dezinger_threshold=20000, binning=1)
    g = torch.from_numpy(d.copy()).cuda()
    pf.remove_outliers(g)
    torch.cuda.synchronize()
    got = g.cpu().numpy()
    ref = pp.remove_outliers(d, w, 20000)
    assert np.array_equal(got, ref), int((got != ref).sum())
    assert (got != d).any()
    pf.free()


@pytest.mark.parametrize("method,dez", [('none', 0), ('none', 3), ('fw', 0), ('ti', 0)])
def test_binned_chain_matches_oracle(method, dez):
    from tomocupy_amd import ProcFunctions
    b = 1
    n, nproj, nz = 96, 60, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(2, n << b, theta, nz << b, ndark=4, nflat=6,
                                           stripes=True)
    if dez:
        rng = np.random.default_rng(9)
        for a in (data, dark, flat):
            a[rng.random(a.shape) < 0.005] = 65000
    print('binned chain', method, 'dezinger', dez, flush=True)
    kw = dict(stripe=method, dezinger=dez, dezinger_threshold=30000, fw_level=3, binning=b)
    ref = pp.proc_chain(data, dark, flat, **kw)
    pf = ProcFunctions(nproj, nz, n, 4, 6, remove_stripe_method=method, fw_level=3,
                       dezinger=dez, dezinger_threshold=30000, binning=b)
    cu = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (data, dark, flat)]
    out = pf.proc_chunk(*cu)
    torch.cuda.synchronize()
    assert tuple(out.shape) == (nz, nproj, n)
    err = _rel_slices(out.cpu().numpy().astype(np.float64), ref)
    print('rel-L2', err)
    assert np.all(err <= RTOL32), err
    # the raw frames are left as they were
    assert np.array_equal(cu[0].cpu().numpy(), data)
    pf.free()


def _scan(tmp_path, ni, nproj, nz):
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(5, ni, theta, nz, ndark=3, nflat=4)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    return str(p), data, dark, flat, th


def test_cli_full_binning(tmp_path):
    # raw width 130 -> trimmed to 128 (a multiple of 2^(b+1)) -> 64 binned;
    # 14 raw rows -> 7 slices in chunks of 4 and 3; centre 130/2 / 2
    ni_raw, nproj, nz_raw = 130, 80, 14
    path, data, dark, flat, th = _scan(tmp_path, ni_raw, nproj, nz_raw)
    out = str(tmp_path / 'rec')
    main(['recon', '--file-name', path, '--reconstruction-type', 'full', '--nsino-per-chunk',
          '4', '--out-path-name', out, '--fbp-filter', 'shepp', '--binning', '1'])
    n, ne, centre, ncz = 64, 256, ni_raw / 2 / 2, 4
    for c0 in range(0, 7, ncz):
        lz = min(ncz, 7 - c0)
        rows = slice(2 * c0, 2 * (c0 + lz))
        sino = pp.proc_chain(data[:, rows, :128], dark[:, rows, :128], flat[:, rows, :128],
                             binning=1)
        sino = np.concatenate([sino, np.zeros((ncz - lz,) + sino.shape[1:])])
        ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'shepp'), ne,
                                                     centre), th)[:lz]
        for k in range(lz):
            got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % (c0 + k)))
            assert got.shape == (n, n)
            err = np.sqrt(((got - ref[k]) ** 2).sum() / (ref[k] ** 2).sum())
            print('slice', c0 + k, err)
            assert err <= RTOL32


def test_cli_try_binning(tmp_path):
    ni_raw, nproj, nz_raw = 128, 80, 8
    path, data, dark, flat, th = _scan(tmp_path, ni_raw, nproj, nz_raw)
    out = str(tmp_path / 'try')
    main(['recon', '--file-name', path, '--reconstruction-type', 'try', '--out-path-name', out,
          '--center-search-width', '4', '--center-search-step', '1', '--nsino', '0.5',
          '--binning', '1'])
    n, ne, centre = 64, 256, 32.0
    # reader.py:196-197: raw row int(0.5 (4*2 - 1) / 2) * 2 = 2
    idz = int(0.5 * (4 * 2 - 1) / 2) * 2
    # reader.py:229-232: shifts step 1*2 raw pixels / 2, centres saved in raw pixels
    shifts = np.arange(-4, 4, 2).astype(np.float32) / 2
    files = sorted(f for f in os.listdir(out) if f.endswith('.tiff'))
    assert len(files) == len(shifts)
    sino = pp.proc_chain(data[:, idz:idz + 2], dark[:, idz:idz + 2], flat[:, idz:idz + 2],
                         binning=1)
    ncz = 8
    shv = np.zeros(ncz)
    shv[:len(shifts)] = shifts
    ref = fr.backprojection(ff.fbp_filter_center(np.concatenate([sino] * ncz),
                                                 ff.calc_filter(ne, 'parzen'), ne, centre, shv),
                            th)
    for k, sh in enumerate(shifts):
        name = 'recon_slice%04d_center%05.2f.tiff' % (idz, (centre - sh) * 2)
        got = io.read_tiff(os.path.join(out, name))
        assert got.shape == (n, n)
        err = np.sqrt(((got - ref[k]) ** 2).sum() / (ref[k] ** 2).sum())
        print(name, err)
        assert err <= RTOL32
