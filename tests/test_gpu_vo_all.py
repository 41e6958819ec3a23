"""``vo-all`` ring removal (``remove_all_stripe``, reference
``processing/remove_stripe.py:311-497``) on the GPU against the FP64 oracle
(pinned in ``test_oracle_vo_all.py``).

The GPU sorts, medians and interpolation are selections and FP64-rounded
forms of the same float32 inputs, so located columns match the oracle's
whenever no ratio lies within rounding of a threshold (the planted stripes are
far from it); per-slice relative L2 <= 1e-5.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import preproc_fp64 as pp
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu

RTOL32 = 1e-5


def _rel_slices(got, ref, axis):
    g = np.moveaxis(got, axis, 0).reshape(got.shape[axis], -1)
    r = np.moveaxis(ref, axis, 0).reshape(ref.shape[axis], -1)
    return np.sqrt(((g - r) ** 2).sum(1) / (r ** 2).sum(1))


def _chunk(n, nproj, nz, seed):
    """Transmission chunk (nproj, nz, n) float32 with random, dead and gain
    stripes, from the GPU dark/flat correction of synthetic counts."""
    from tomocupy_amd import ProcFunctions
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(seed, n, theta, nz, ndark=3, nflat=4, stripes=True)
    data[:, 0, n // 3] = 15000                       # dead column
    data[:, nz - 1, (2 * n) // 3] = (data[:, nz - 1, (2 * n) // 3] * 1.3).astype(np.uint16)
    pf = ProcFunctions(nproj, nz, n, 3, 4)
    cu = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (data, dark, flat)]
    res = pf.darkflat_correction(*cu)
    torch.cuda.synchronize()
    pf.free()
    return res, (data, dark, flat)


@pytest.mark.parametrize("n,nproj,nz,snr,la,sm", [
    (256, 180, 4, 3.0, 61, 21),
    (200, 97, 2, 3.0, 10, 4),      # even windows, odd nproj
    (128, 64, 1, 2.0, 31, 9),      # one slice
    (160, 60, 2, 3.0, 101, 65),    # windows above 64: the LDS median kernel
])
def test_remove_all_stripe_matches_oracle(n, nproj, nz, snr, la, sm):
    from tomocupy_amd import ProcFunctions
    print('vo-all', n, nproj, nz, snr, la, sm, flush=True)
    res, _ = _chunk(n, nproj, nz, 11)
    ref = pp.remove_all_stripe(res.cpu().numpy().astype(np.float64), snr, la, sm, 1)
    pf = ProcFunctions(nproj, nz, n, 3, 4, remove_stripe_method='vo-all', vo_all_snr=snr,
                       vo_all_la_size=la, vo_all_sm_size=sm)
    inp = res.cpu().numpy()
    pf.remove_all_stripe(res)
    torch.cuda.synchronize()
    got = res.cpu().numpy().astype(np.float64)
    err = _rel_slices(got, ref, 1)
    print('rel-L2', err, 'changed', np.abs(ref - inp).max())
    assert np.all(err <= RTOL32), err
    assert np.abs(ref - inp).max() > 1e-3   # the stripes were acted on
    # a second call on the same plan reuses the scratch
    res2 = torch.from_numpy(inp).cuda()
    pf.remove_all_stripe(res2)
    torch.cuda.synchronize()
    assert np.array_equal(res2.cpu().numpy().astype(np.float64), got)
    pf.free()


@pytest.mark.parametrize("n,nproj,nz", [(64, 1500, 2), (32, 3000, 1), (16, 8000, 1)])
def test_remove_all_stripe_long_columns(n, nproj, nz):
    # column sorts of 8, 16 and 32 keys per thread (1500, 3000, 8000 angles)
    from tomocupy_amd import ProcFunctions
    print('vo-all long columns', n, nproj, nz, flush=True)
    res, _ = _chunk(n, nproj, nz, 12)
    ref = pp.remove_all_stripe(res.cpu().numpy().astype(np.float64), 3.0, 9, 5, 1)
    pf = ProcFunctions(nproj, nz, n, 3, 4, remove_stripe_method='vo-all', vo_all_la_size=9,
                       vo_all_sm_size=5)
    pf.remove_all_stripe(res)
    torch.cuda.synchronize()
    err = _rel_slices(res.cpu().numpy().astype(np.float64), ref, 1)
    print('rel-L2', err)
    assert np.all(err <= RTOL32), err
    pf.free()


def test_vo_all_chain_matches_oracle():
    from tomocupy_amd import ProcFunctions
    n, nproj, nz = 192, 120, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(3, n, theta, nz, ndark=3, nflat=4, stripes=True)
    ref = pp.proc_chain(data, dark, flat, stripe='vo-all')
    pf = ProcFunctions(nproj, nz, n, 3, 4, remove_stripe_method='vo-all')
    cu = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (data, dark, flat)]
    out = pf.proc_chunk(*cu)
    torch.cuda.synchronize()
    err = _rel_slices(out.cpu().numpy().astype(np.float64), ref, 0)
    print('chain rel-L2', err)
    assert np.all(err <= RTOL32), err
    pf.free()


@pytest.mark.parametrize("n,nproj,nz,la,sm", [
    (128, 90, 2, 31, 9),     # 2-D window 9 x 9
    (96, 30, 1, 9, 21),      # reflected at both rank ends (21 of 30 ranks)
    (64, 200, 1, 11, 64),    # the largest 2-D window
])
def test_remove_all_stripe_dim2_matches_oracle(n, nproj, nz, la, sm):
    # _rs_sort3 with dim 2: median over a sm x sm window of detector and rank
    from tomocupy_amd import ProcFunctions
    print('vo-all dim 2', n, nproj, nz, la, sm, flush=True)
    res, _ = _chunk(n, nproj, nz, 13)
    ref = pp.remove_all_stripe(res.cpu().numpy().astype(np.float64), 3.0, la, sm, 2)
    pf = ProcFunctions(nproj, nz, n, 3, 4, remove_stripe_method='vo-all', vo_all_la_size=la,
                       vo_all_sm_size=sm, vo_all_dim=2)
    pf.remove_all_stripe(res)
    torch.cuda.synchronize()
    err = _rel_slices(res.cpu().numpy().astype(np.float64), ref, 1)
    print('rel-L2', err)
    assert np.all(err <= RTOL32), err
    pf.free()


def test_vo_all_rejects_unsupported():
    from tomocupy_amd import ProcFunctions, _lib
    pf = ProcFunctions(16, 2, 32, 1, 1, remove_stripe_method='vo-all', vo_all_dim=2,
                       vo_all_sm_size=17)
    res = torch.ones((16, 2, 32), dtype=torch.float32, device='cuda')
    with pytest.raises(_lib.TomoError):   # 2-D window taller than the 16 ranks
        pf.remove_all_stripe(res)
    pf.vo_all_dim, pf.vo_all_la_size = 3, 5
    with pytest.raises(_lib.TomoError):
        pf.remove_all_stripe(res)
    pf.vo_all_dim, pf.vo_all_la_size = 1, 256
    with pytest.raises(_lib.TomoError):
        pf.remove_all_stripe(res)
    pf.free()
    pf = ProcFunctions(8193, 1, 32, 1, 1, remove_stripe_method='vo-all')
    res = torch.ones((8193, 1, 32), dtype=torch.float32, device='cuda')
    with pytest.raises(_lib.TomoError):
        pf.remove_all_stripe(res)
    pf.free()


def test_cli_full_vo_all(tmp_path):
    # recon --remove-stripe-method vo-all with non-default windows: every
    # slice against the oracle chain + backprojection
    import os
    from oracle import filter_fp64 as ff
    from oracle import fourierrec_fp64 as fr
    from tomocupy_amd import io
    from tomocupy_amd.cli import main
    n, nproj, nz = 128, 90, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(8, n, theta, nz, ndark=3, nflat=4, stripes=True)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    out = str(tmp_path / 'rec')
    main(['recon', '--file-name', str(p), '--reconstruction-type', 'full', '--nsino-per-chunk',
          '4', '--out-path-name', out, '--remove-stripe-method', 'vo-all', '--vo-all-la-size',
          '31', '--vo-all-sm-size', '11', '--vo-all-snr', '2.5'])
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    sino = pp.proc_chain(data, dark, flat, stripe='vo-all', vo_snr=2.5, vo_la_size=31,
                         vo_sm_size=11)
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(4 * n, 'parzen'), 4 * n,
                                                 n / 2), th)
    for k in range(nz):
        got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % k)).astype(np.float64)
        err = np.sqrt(((got - ref[k]) ** 2).sum() / (ref[k] ** 2).sum())
        print('slice', k, err)
        assert err <= RTOL32
