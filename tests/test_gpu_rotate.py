"""Projection rotation (``ProcFunctions.rotate_proj``, reference
``processing/proc_functions.py:126-138``) on the GPU against the oracle
(``ndimage.rotate``, pinned in ``test_oracle_preproc.py``), and in
recon_steps (``--rotate-proj-angle``, the reference's
``tests/test_steps_full.py:19``)."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from oracle import preproc_fp64 as pp
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("angle,order", [(1.5, 1), (-7.25, 1), (90.0, 1), (200.0, 1),
                                         (1.5, 0), (33.0, 0), (1.5, 3), (-7.25, 3), (90.0, 3),
                                         (200.0, 2), (4.0, 2), (-3.0, 4), (130.0, 4),
                                         (2.0, 5), (-45.0, 5)])
def test_rotate_proj_matches_oracle(angle, order):
    from tomocupy_amd import ProcFunctions
    rng = np.random.default_rng(int(abs(angle) * 4) + order)
    y, x = np.mgrid[0:40, 0:52]
    a = np.stack([np.exp(-((x - 20 - 3 * p) ** 2 + (y - 22) ** 2) / 80.0) +
                  0.01 * rng.standard_normal((40, 52)) for p in range(3)]).astype(np.float32)
    ref = pp.rotate_proj(a.astype(np.float64), angle, order)
    g = torch.from_numpy(a.copy()).cuda()
    ProcFunctions.rotate_proj(g, angle, order)
    torch.cuda.synchronize()
    got = g.cpu().numpy().astype(np.float64)
    if order == 0:
        # nearest sample: identical except where a coordinate sits within
        # rounding of a half-sample (none for these angles)
        assert np.array_equal(got, ref.astype(np.float32).astype(np.float64))
    else:
        err = np.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum())
        print('angle', angle, 'rel-L2', err)
        assert err <= 1e-6


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_rotate_proj_spline_large(order):
    # padded planes wider and taller than one 256-thread workgroup
    from tomocupy_amd import ProcFunctions
    rng = np.random.default_rng(order)
    a = rng.uniform(0.2, 1.0, (2, 300, 420)).astype(np.float32)
    ref = pp.rotate_proj(a.astype(np.float64), 2.5, order)
    g = torch.from_numpy(a.copy()).cuda()
    ProcFunctions.rotate_proj(g, 2.5, order)
    torch.cuda.synchronize()
    got = g.cpu().numpy().astype(np.float64)
    err = np.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum())
    print('order', order, 'rel-L2', err, 'max', np.abs(got - ref).max())
    assert err <= 1e-6


def test_rotate_proj_rejects_orders_above_5():
    # as ndimage ('spline order not supported')
    from tomocupy_amd import ProcFunctions, _lib
    g = torch.ones((1, 8, 8), dtype=torch.float32, device='cuda')
    with pytest.raises(_lib.TomoError):
        ProcFunctions.rotate_proj(g, 3.0, 6)


@pytest.mark.parametrize("order", [1, 3])
def test_cli_recon_steps_rotate(tmp_path, order):
    from tomocupy_amd import io
    from tomocupy_amd.cli import main
    n, nproj, nz = 64, 40, 8
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(12, n, theta, nz, ndark=3, nflat=4)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    out = str(tmp_path / 'rec')
    main(['recon_steps', '--file-name', str(p), '--reconstruction-type', 'full',
          '--nsino-per-chunk', '4', '--out-path-name', out, '--rotate-proj-angle', '1.5',
          '--rotate-proj-order', str(order)])
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    res = pp.rotate_proj(pp.darkflat_correction(data, dark, flat), 1.5, order)
    sino = np.ascontiguousarray(pp.minus_log(res).swapaxes(0, 1))
    ne = 4 * n
    for c0 in range(0, nz, 4):
        ref = fr.backprojection(ff.fbp_filter_center(sino[c0:c0 + 4], ff.calc_filter(ne, 'parzen'),
                                                     ne, n / 2), th)
        for k in range(4):
            got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % (c0 + k))).astype(np.float64)
            err = np.sqrt(((got - ref[k]) ** 2).sum() / (ref[k] ** 2).sum())
            print('slice', c0 + k, err)
            assert err <= 1e-5
