"""Phase retrieval (reference ``processing/retrieve_phase.py``) and the
recon_steps passes (``rec_steps.py:114-243``) on the GPU against the FP64
oracle (pinned in ``test_oracle_phase.py``): per-projection and per-slice
relative L2 <= 1e-5 (float32 FFTs of the padded projections)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from oracle import phase_fp64 as ph
from oracle import preproc_fp64 as pp
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu

RTOL32 = 1e-5


def _rel(got, ref):
    g = got.reshape(got.shape[0], -1)
    r = ref.reshape(ref.shape[0], -1)
    return np.sqrt(((g - r) ** 2).sum(1) / (r ** 2).sum(1))


def _projections(nproj, dy, dz, seed):
    """Transmission-like projections: smooth phantom attenuation plus noise."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:dy, 0:dz]
    out = np.empty((nproj, dy, dz), np.float32)
    for p in range(nproj):
        cx, cy = rng.uniform(0.3, 0.7) * dz, rng.uniform(0.3, 0.7) * dy
        r2 = ((x - cx) / (0.25 * dz)) ** 2 + ((y - cy) / (0.3 * dy)) ** 2
        out[p] = np.exp(-0.8 * np.clip(1 - r2, 0, None)) + 0.01 * rng.standard_normal((dy, dz))
    return out


@pytest.mark.parametrize("method,dy,dz,kw", [
    ('paganin', 64, 96, dict(pixel_size=1e-4, dist=5.0, energy=20.0, alpha=1e-3)),
    ('paganin', 128, 128, dict(pixel_size=0.65e-4, dist=20.0, energy=25.0, alpha=0.0)),
    ('Gpaganin', 64, 96, dict(pixel_size=1e-4, dist=5.0, energy=20.0, db=1500.0, W=2e-4)),
    ('farago', 96, 64, dict(pixel_size=1e-4, dist=0.5, energy=20.0, db=1500.0)),
])
def test_retrieve_phase_matches_oracle(method, dy, dz, kw):
    from tomocupy_amd import retrieve_phase as rp
    print('phase', method, dy, dz, kw, flush=True)
    data = _projections(5, dy, dz, 3)
    ref = ph.retrieve_phase(data.astype(np.float64), method, **kw)
    g = torch.from_numpy(data.copy()).cuda()
    if method == 'farago':
        rp.farago_filter(g, kw['pixel_size'], kw['dist'], kw['energy'], kw['db'])
    else:
        rp.paganin_filter(g, kw['pixel_size'], kw['dist'], kw['energy'], kw.get('alpha', 1e-3),
                          method, kw.get('db', 1000.0), kw.get('W', 2e-4))
    torch.cuda.synchronize()
    err = _rel(g.cpu().numpy().astype(np.float64), ref)
    print('rel-L2', err, 'changed', np.abs(ref - data).max())
    assert np.all(err <= RTOL32), err
    assert np.abs(ref - data).max() > 1e-3
    rp.clear_plans()


def test_phase_plan_rejects_odd_and_bad_parameters():
    from tomocupy_amd import _lib
    from tomocupy_amd.retrieve_phase import PhasePlan
    with pytest.raises(_lib.TomoError):
        PhasePlan(63, 64, 'paganin', 1e-4, 5.0, 20.0)
    with pytest.raises(_lib.TomoError):
        PhasePlan(64, 64, 'paganin', 1e-4, 0.0, 20.0)   # no propagation distance


@pytest.mark.parametrize("method", ['paganin', 'none'])
def test_recon_steps_matches_oracle(method):
    # pass 1 by row chunks (4 + 2), pass 2 by projection chunks, pass 3 with
    # the fused -log + swap, against the oracle's whole-volume chain
    from tomocupy_amd.rec_steps import GPURecSteps
    n, nproj, nz, ncz = 96, 60, 6, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(4, n, theta, nz, ndark=3, nflat=4, stripes=True)
    kw = dict(energy=20.0, propagation_distance=50.0, pixel_size=1.0, retrieve_phase_alpha=1e-3)
    print('recon_steps', method, flush=True)
    rs = GPURecSteps(n, nproj, nz, ncz, torch.from_numpy(theta).cuda(), n / 2, 3, 4,
                     fbp_filter='parzen', nproj_per_chunk=7, retrieve_phase_method=method,
                     remove_stripe_method='fw', fw_level=3, **kw)
    out = {}
    rs.recon_steps_all(data, dark, flat, lambda k, r: out.__setitem__(k, r.copy()))
    rs.free()
    got = np.concatenate([out[k] for k in sorted(out)]).astype(np.float64)
    res = pp.remove_stripe_fw(pp.darkflat_correction(data, dark, flat), 1.0, 'sym16', 3)
    if method != 'none':
        res = ph.retrieve_phase(res, method, pixel_size=1e-4, dist=5.0, energy=20.0, alpha=1e-3)
    sino = np.ascontiguousarray(pp.minus_log(res).swapaxes(0, 1))
    ne = 4 * n
    ref = []
    for c0 in range(0, nz, ncz):
        s = sino[c0:c0 + ncz]
        lz = s.shape[0]
        s = np.concatenate([s, np.zeros((ncz - lz,) + s.shape[1:])])
        ref.append(fr.backprojection(ff.fbp_filter_center(s, ff.calc_filter(ne, 'parzen'), ne,
                                                          n / 2), theta)[:lz])
    ref = np.concatenate(ref)
    err = _rel(got, ref)
    print('rel-L2', err)
    assert np.all(err <= RTOL32), err


def test_cli_recon_steps_paganin(tmp_path):
    # recon_steps --retrieve-phase-method paganin with the reference test's
    # settings (tests/test_steps_full.py:12) on a small scan
    import os
    from tomocupy_amd import io
    from tomocupy_amd.cli import main
    n, nproj, nz = 64, 48, 6
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(6, n, theta, nz, ndark=3, nflat=4)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    out = str(tmp_path / 'rec')
    main(['recon_steps', '--file-name', str(p), '--reconstruction-type', 'full',
          '--nsino-per-chunk', '4', '--out-path-name', out, '--rotation-axis', '32',
          '--retrieve-phase-method', 'paganin', '--retrieve-phase-alpha', '0.0001',
          '--propagation-distance', '60', '--energy', '20', '--pixel-size', '1.17'])
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    res = ph.retrieve_phase(pp.darkflat_correction(data, dark, flat), 'paganin',
                            pixel_size=1.17e-4, dist=6.0, energy=20.0, alpha=1e-4)
    sino = np.ascontiguousarray(pp.minus_log(res).swapaxes(0, 1))
    ne = 4 * n
    for c0 in range(0, nz, 4):
        s = sino[c0:c0 + 4]
        lz = s.shape[0]
        s = np.concatenate([s, np.zeros((4 - lz,) + s.shape[1:])])
        ref = fr.backprojection(ff.fbp_filter_center(s, ff.calc_filter(ne, 'parzen'), ne, 32.0),
                                th)
        for k in range(lz):
            got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % (c0 + k))).astype(np.float64)
            err = np.sqrt(((got - ref[k]) ** 2).sum() / (ref[k] ** 2).sum())
            print('slice', c0 + k, err)
            assert err <= RTOL32


def test_cli_recon_steps_try_paganin(tmp_path):
    # recon_steps --reconstruction-type try: the whole scan phase-retrieved,
    # then one slice at several centres
    import os
    from tomocupy_amd import io
    from tomocupy_amd.cli import main
    n, nproj, nz = 64, 48, 6
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(9, n, theta, nz, ndark=3, nflat=4)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    out = str(tmp_path / 'try')
    main(['recon_steps', '--file-name', str(p), '--reconstruction-type', 'try',
          '--out-path-name', out, '--rotation-axis', '32', '--center-search-width', '2',
          '--center-search-step', '1', '--nsino', '0.5', '--retrieve-phase-method', 'paganin',
          '--retrieve-phase-alpha', '0.0001', '--propagation-distance', '60', '--energy', '20',
          '--pixel-size', '1.17'])
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    res = ph.retrieve_phase(pp.darkflat_correction(data, dark, flat), 'paganin',
                            pixel_size=1.17e-4, dist=6.0, energy=20.0, alpha=1e-4)
    idz = int(0.5 * (nz - 1))
    sino = np.ascontiguousarray(pp.minus_log(res).swapaxes(0, 1))[idz:idz + 1]
    shifts = np.arange(-2, 2, 1.0)
    ncz = 8
    shv = np.zeros(ncz)
    shv[:len(shifts)] = shifts
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(np.concatenate([sino] * ncz),
                                                 ff.calc_filter(ne, 'parzen'), ne, 32.0, shv), th)
    files = sorted(f for f in os.listdir(out) if f.endswith('.tiff'))
    assert len(files) == len(shifts)
    for k, sh in enumerate(shifts):
        name = 'recon_slice%04d_center%05.2f.tiff' % (idz, 32.0 - sh)
        got = io.read_tiff(os.path.join(out, name)).astype(np.float64)
        err = np.sqrt(((got - ref[k]) ** 2).sum() / (ref[k] ** 2).sum())
        print(name, err)
        assert err <= RTOL32


def test_cli_recon_steps_paganin_float16(tmp_path):
    # the reference's test_steps_full.py:13 case: Paganin by steps in float16
    # (width 80 cropped to 64 columns, centre moved by -8)
    import os
    from tomocupy_amd import io
    from tomocupy_amd.cli import main
    ni, nproj, nz = 80, 48, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(10, ni, theta, nz, ndark=3, nflat=4)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    out = str(tmp_path / 'rec')
    main(['recon_steps', '--file-name', str(p), '--reconstruction-type', 'full',
          '--nsino-per-chunk', '4', '--out-path-name', out, '--rotation-axis', '40',
          '--dtype', 'float16', '--retrieve-phase-method', 'paganin',
          '--retrieve-phase-alpha', '0.0001', '--propagation-distance', '60', '--energy', '20',
          '--pixel-size', '1.17'])
    n, st, centre = 64, 8, 40 - 8
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    sl = slice(st, st + n)
    res = ph.retrieve_phase(pp.darkflat_correction(data[:, :, sl], dark[:, :, sl], flat[:, :, sl]),
                            'paganin', pixel_size=1.17e-4, dist=6.0, energy=20.0, alpha=1e-4)
    sino = np.ascontiguousarray(pp.minus_log(res).swapaxes(0, 1))
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, centre),
                            th)
    for k in range(nz):
        got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % k))
        assert got.dtype == np.float16 and got.shape == (n, n)
        err = np.sqrt(((got.astype(np.float64) - ref[k]) ** 2).sum() / (ref[k] ** 2).sum())
        print('slice', k, err)
        assert err <= 2e-3


def test_cli_recon_steps_binning_paganin(tmp_path):
    # binning 1 then Paganin by steps: 8 raw rows x 128 raw columns -> 4 x 64
    import os
    from tomocupy_amd import io
    from tomocupy_amd.cli import main
    ni_raw, nproj, nz_raw = 128, 40, 8
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(11, ni_raw, theta, nz_raw, ndark=3, nflat=4)
    p = tmp_path / 'scan.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark,
             theta=theta.astype(np.float64) * 180 / np.pi)
    out = str(tmp_path / 'rec')
    main(['recon_steps', '--file-name', str(p), '--reconstruction-type', 'full',
          '--nsino-per-chunk', '4', '--out-path-name', out, '--binning', '1',
          '--retrieve-phase-method', 'paganin', '--retrieve-phase-alpha', '0.0001',
          '--propagation-distance', '60', '--energy', '20', '--pixel-size', '1.17'])
    n, nz = 64, 4
    th = (theta.astype(np.float64) * 180 / np.pi / 180 * np.pi).astype(np.float32)
    b = [pp.downsample(a, 1) for a in (data, dark, flat)]
    res = ph.retrieve_phase(pp.darkflat_correction(*b), 'paganin', pixel_size=1.17e-4, dist=6.0,
                            energy=20.0, alpha=1e-4)
    sino = np.ascontiguousarray(pp.minus_log(res).swapaxes(0, 1))
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne,
                                                 ni_raw / 2 / 2), th)
    for k in range(nz):
        got = io.read_tiff(os.path.join(out, 'recon_%05d.tiff' % k)).astype(np.float64)
        err = np.sqrt(((got - ref[k]) ** 2).sum() / (ref[k] ** 2).sum())
        print('slice', k, err)
        assert err <= RTOL32
