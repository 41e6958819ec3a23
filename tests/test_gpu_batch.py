"""Batching form and plan autotuner.

* A batch of independent requests of different sizes (n = 64, 96, 128, 192,
  nproj = round(1.5 n), random start angle, odd and even nz), run eagerly,
  matches the FP64 oracle per request (per-slice rel-L2 <= 1e-5).
* The captured HIP graph replayed three times, with other GPU work in
  between, reproduces the eager outputs bitwise.
* The batched gridding (one gridding launch for all requests) gives the
  per-request launches' outputs bitwise, eagerly and over graph replays.
* Autotuner options change only rounding: outputs within 1e-6 rel-L2 of the
  default plan.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu


def _requests(seed=5):
    rng = np.random.default_rng(seed)
    reqs = []
    for k, n in enumerate([64, 96, 128, 192, 128, 64]):
        nproj = int(round(1.5 * n))
        th0 = rng.uniform(0, np.pi)
        theta = (th0 + np.arange(nproj) * np.pi / nproj).astype(np.float32)
        nz = int(rng.integers(1, 6))
        reqs.append({'n': n, 'nz': nz, 'theta': theta, 'center': n / 2 + rng.uniform(-2, 2),
                     'sino': synth.sinograms(20 + k, n, theta, nz)})
    return reqs


def test_batch_eager_matches_oracle():
    from tomocupy_amd.batch import BatchRec
    reqs = _requests()
    print('batch of', [(r['n'], r['nz']) for r in reqs], flush=True)
    b = BatchRec(reqs, fbp_filter='parzen', streams=3)
    for r, d in zip(reqs, b.inputs()):
        d.copy_(torch.from_numpy(r['sino'].astype(np.float32)))
    b.run_eager()
    torch.cuda.synchronize()
    for r, o in zip(reqs, b.outputs()):
        n, nz = r['n'], r['nz']
        ne = 4 * n
        ref = fr.backprojection(ff.fbp_filter_center(r['sino'], ff.calc_filter(ne, 'parzen'), ne,
                                                     r['center']), r['theta'])
        got = o.double().cpu().numpy()
        err = np.sqrt(((got - ref) ** 2).reshape(nz, -1).sum(1) / (ref ** 2).reshape(nz, -1).sum(1))
        print('request n=%d nz=%d rel-L2 %s' % (n, nz, err))
        assert np.all(err <= 1e-5), err
    b.free()


def test_batch_graph_replays_match_eager():
    from tomocupy_amd.batch import BatchRec
    reqs = _requests(9)
    b = BatchRec(reqs, fbp_filter='ramp', streams=4)
    for r, d in zip(reqs, b.inputs()):
        d.copy_(torch.from_numpy(r['sino'].astype(np.float32)))
    b.run_eager()
    torch.cuda.synchronize()
    eager = [o.clone() for o in b.outputs()]
    b.capture()
    for rep in range(3):
        for o in b.outputs():
            o.fill_(float('nan'))
        scratch = torch.randn(1 << 20, device='cuda')
        scratch.mul_(2.0)
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        for e, o in zip(eager, b.outputs()):
            assert torch.equal(e, o), 'replay %d differs from the eager run' % rep
    b.free()


def test_batched_gridding_equals_per_request():
    from tomocupy_amd.batch import BatchRec
    reqs = _requests(13)
    print('batch of', [(r['n'], r['nz']) for r in reqs], flush=True)
    outs = {}
    for mode in ('streams', 'batched'):
        b = BatchRec(reqs, fbp_filter='parzen', streams=2, grid=mode)
        for r, d in zip(reqs, b.inputs()):
            d.copy_(torch.from_numpy(r['sino'].astype(np.float32)))
        b.run_eager()
        torch.cuda.synchronize()
        outs[mode] = [o.clone() for o in b.outputs()]
        if mode == 'batched':
            b.capture()
            for rep in range(3):
                for o in b.outputs():
                    o.fill_(float('nan'))
                torch.randn(1 << 20, device='cuda').mul_(2.0)
                torch.cuda.synchronize()
                b.run()
                torch.cuda.synchronize()
                bad = [i for i, (e, o) in enumerate(zip(outs[mode], b.outputs()))
                       if not torch.equal(e, o)]
                assert not bad, 'replay %d: requests %s differ' % (rep, bad)
        b.free()
    for i, (e, o) in enumerate(zip(outs['streams'], outs['batched'])):
        assert torch.isfinite(o).all()
        assert torch.equal(e, o), 'request %d differs' % i


def test_batched_gridding_rejects_16bit():
    from tomocupy_amd.batch import BatchRec
    with pytest.raises(ValueError):
        BatchRec(_requests()[:1], dtype='float16', grid='batched')


def test_autotune_options_change_only_rounding(tmp_path, monkeypatch):
    from tomocupy_amd import BackprojFunctions
    from tomocupy_amd.autotune import tune
    monkeypatch.setenv('TOMOHIP_TUNE_CACHE', str(tmp_path / 'tune.json'))
    n, nproj, nz = 256, 180, 4
    best = tune(n, nproj, nz, grid_chunks=(128, 512), reps=2, verbose=True)
    assert best['grid_chunk'] in (128, 512) and best['fft'] in ('static', 'generic')
    assert tune(n, nproj, nz) == best  # cached
    theta = torch.from_numpy(synth.angles(nproj)).cuda()
    sino = torch.from_numpy(synth.sinograms(30, n, synth.angles(nproj), nz).astype(np.float32)).cuda()
    outs = []
    for opts in (None, {'grid_chunk': 128, 'fft': 'generic'}, best):
        bf = BackprojFunctions(n, nproj, nz, theta, n / 2, 'float32', 'parzen', rec_options=opts)
        o = torch.zeros((nz, n, n), device='cuda')
        bf.fbp_filter_backprojection(o, sino)
        outs.append(o.double())
        bf.free()
    torch.cuda.synchronize()
    for o in outs[1:]:
        rel = (torch.linalg.norm(o - outs[0]) / torch.linalg.norm(outs[0])).item()
        print('tuned vs default rel-L2 %.3g' % rel)
        assert rel < 1e-6
