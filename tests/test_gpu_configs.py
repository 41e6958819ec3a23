"""The benchmark configurations at (or near) their own sizes, against the FP64
oracle.

* Full pipeline (n=2048, 1500 angles, one slice pair of the uint16 chunk with
  column-gain stripes, Fourier-wavelet sym16 level 7 sigma 1, ramp filter,
  centre 1024.5), through the fused kernels the benchmark runs
  (``proc_chunk`` + ``fbp_filter_backprojection``), stage by stage:
  the sinograms against the oracle chain where the rays are not starved, and
  the reconstruction of those same sinograms against the oracle filter and
  backprojection; per-slice rel-L2 <= 1e-5.
* Precision form (n=4096, 3000 angles, one slice pair of uint16 counts,
  dark/flat + -log, Parzen, centre n/2) end to end against the FP64 oracle
  chain on the same counts: float32 <= 1e-5, float16 <= 2e-3,
  bfloat16 <= 1e-2 per slice (rel-L2); max-abs printed.
* Batching form: the first request of each size (256 ... 2048, n=768 and 1536
  with 3*2^k transforms) of the 96-request mix (seed 2024) through
  ``BatchRec`` at its own nz, slices 0 and ceil(nz/2) (one complex pair)
  against the oracle (<= 1e-5); and the whole 96-request batch replayed from
  one HIP graph equals its eager run bitwise.
* Differentiation form (n=1024, 180 angles, nz=32): the dot-product test,
  the gradient of the Tikhonov least-squares loss on one slice pair against
  FP64, and 5 CG iterations against the same iteration with the FP64 oracle
  operators.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from oracle import preproc_fp64 as pp
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu


def _rel(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    num = np.sqrt(((got - ref) ** 2).reshape(len(ref), -1).sum(1))
    den = np.sqrt((ref ** 2).reshape(len(ref), -1).sum(1))
    return num / np.maximum(den, 1e-300)


def _maxabs(got, ref):
    return np.abs(np.asarray(got, np.float64) - ref).max() / np.abs(ref).max()


# ---- full pipeline at n=2048, 1500 angles ---------------------------------

def test_fullsize_pipeline_pair_stagewise():
    from tomocupy_amd import BackprojFunctions, ProcFunctions
    n, nproj, nz, center = 2048, 1500, 2, 1024.5
    print('full pipeline pair n=%d nproj=%d fw sym16/7 ramp centre %g' % (n, nproj, center),
          flush=True)
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(2, n, theta, nz, stripes=True)
    pf = ProcFunctions(nproj, nz, n, dark.shape[0], flat.shape[0], 'float32',
                       remove_stripe_method='fw', fw_sigma=1.0, fw_filter='sym16', fw_level=7)
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), center, 'float32',
                           'ramp')
    g = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (data, dark, flat)]
    sino = pf.proc_chunk(*g)
    obj = torch.empty((nz, n, n), device='cuda')
    bf.fbp_filter_backprojection(obj, sino)
    torch.cuda.synchronize()
    sino_np = sino.double().cpu().numpy()
    # stage 1: sinograms, where the oracle transmission is well above zero
    res = pp.remove_stripe_fw(pp.darkflat_correction(data, dark, flat), 1.0, 'sym16', 7)
    ref_sino = np.ascontiguousarray(pp.minus_log(res.copy()).swapaxes(0, 1))
    ok = np.ascontiguousarray(res.swapaxes(0, 1)) > 1e-3
    err = np.array([np.linalg.norm((sino_np[z] - ref_sino[z])[ok[z]]) /
                    np.linalg.norm(ref_sino[z][ok[z]]) for z in range(nz)])
    print('sinogram rel-L2 (%.4f of the rays not starved) %s' % (ok.mean(), err))
    assert np.all(err <= 1e-5), err
    # stage 2: filter + backprojection of the same sinograms
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(sino_np, ff.calc_filter(ne, 'ramp'), ne, center),
                            theta)
    got = obj.double().cpu().numpy()
    err = _rel(got, ref)
    print('reconstruction rel-L2 %s, max-abs/max %.3g' % (err, _maxabs(got, ref)))
    assert np.all(err <= 1e-5), err
    pf.free()
    bf.free()


# ---- precision form at n=4096, 3000 angles --------------------------------

P3 = {'n': 4096, 'nproj': 3000, 'nz': 2}
TOL3 = {'float32': 1e-5, 'float16': 2e-3, 'bfloat16': 1e-2}


@pytest.fixture(scope='module')
def config3():
    n, nproj, nz = P3['n'], P3['nproj'], P3['nz']
    print('precision form data + oracle n=%d nproj=%d' % (n, nproj), flush=True)
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(3, n, theta, nz)
    sino = pp.proc_chain(data, dark, flat)
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, n / 2),
                            theta)
    return theta, data, dark, flat, ref


@pytest.mark.parametrize('dtype', ['float32', 'float16', 'bfloat16'])
def test_precision_form_n4096(config3, dtype):
    from tomocupy_amd import BackprojFunctions, ProcFunctions, _lib
    theta, data, dark, flat, ref = config3
    n, nproj, nz = P3['n'], P3['nproj'], P3['nz']
    print('precision form', dtype, flush=True)
    pf = ProcFunctions(nproj, nz, n, dark.shape[0], flat.shape[0], dtype)
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), n / 2, dtype, 'parzen')
    g = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (data, dark, flat)]
    sino = pf.proc_chunk(*g)
    obj = torch.empty((nz, n, n), dtype=_lib.torch_dtype(dtype), device='cuda')
    bf.fbp_filter_backprojection(obj, sino)
    torch.cuda.synchronize()
    got = obj.double().cpu().numpy()
    err = _rel(got, ref)
    print('%s n=%d nproj=%d rel-L2 per slice %s (max %.3g), max-abs/max %.3g'
          % (dtype, n, nproj, err, err.max(), _maxabs(got, ref)))
    assert np.all(err <= TOL3[dtype]), err
    pf.free()
    bf.free()


# ---- batching form ---------------------------------------------------------

def _first_of_each_size():
    seen = {}
    for r in synth.batch_requests(96, 2024):
        seen.setdefault(r['n'], r)
    return [seen[n] for n in sorted(seen)]


def test_batch_mix_sizes_match_oracle():
    from tomocupy_amd.batch import BatchRec
    reqs = []
    for k, r in enumerate(_first_of_each_size()):
        n, nz = r['n'], r['nz']
        a_b = synth.sinograms(40 + k, n, r['theta'], 2)
        nzc = (nz + 1) // 2
        # every pair (c, c + nzc) holds the same two slices A, B
        sino = np.stack([a_b[0] if z < nzc else a_b[1] for z in range(nz)])
        reqs.append(dict(r, sino=sino, pair=a_b if nz >= 2 else a_b[:1]))
    print('batch sizes', [(r['n'], len(r['theta']), r['nz']) for r in reqs], flush=True)
    b = BatchRec(reqs, fbp_filter='parzen', streams=4)
    for r, d in zip(reqs, b.inputs()):
        d.copy_(torch.from_numpy(r['sino'].astype(np.float32)))
    b.run_eager()
    torch.cuda.synchronize()
    for r, o in zip(reqs, b.outputs()):
        n, nz = r['n'], r['nz']
        nzc = (nz + 1) // 2
        ne = 4 * n
        ref = fr.backprojection(ff.fbp_filter_center(r['pair'], ff.calc_filter(ne, 'parzen'), ne,
                                                     r['center']), r['theta'])
        idx = [0, nzc] if nz >= 2 else [0]
        got = o[idx].double().cpu().numpy()
        err = _rel(got, ref)
        print('n=%d nproj=%d nz=%d slices %s rel-L2 %s' % (n, len(r['theta']), nz, idx, err),
              flush=True)
        assert np.all(err <= 1e-5), err
    b.free()


def test_batch_96_requests_graph_equals_eager():
    from tomocupy_amd.batch import BatchRec
    reqs = synth.batch_requests(96, 2024)
    cache = {}
    for r in reqs:
        n = r['n']
        if n not in cache:
            cache[n] = synth.sinograms(4, n, r['theta'], 1)[0].astype(np.float32)
    print('96-request batch, %d slices' % sum(r['nz'] for r in reqs), flush=True)
    b = BatchRec(reqs, fbp_filter='parzen', streams=4)
    for r, d in zip(reqs, b.inputs()):
        d.copy_(torch.from_numpy(cache[r['n']])[None].expand(r['nz'], -1, -1))
    b.run_eager()
    torch.cuda.synchronize()
    eager = [o.clone() for o in b.outputs()]
    b.capture()
    for rep in range(2):
        for o in b.outputs():
            o.fill_(float('nan'))
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        bad = [i for i, (e, o) in enumerate(zip(eager, b.outputs())) if not torch.equal(e, o)]
        print('replay %d: %d of %d requests differ' % (rep, len(bad), len(eager)), flush=True)
        assert not bad
    b.free()
    # the same batch with one gridding launch for all 96 requests
    b = BatchRec(reqs, fbp_filter='parzen', grid='batched')
    for r, d in zip(reqs, b.inputs()):
        d.copy_(torch.from_numpy(cache[r['n']])[None].expand(r['nz'], -1, -1))
    b.run()
    torch.cuda.synchronize()
    bad = [i for i, (e, o) in enumerate(zip(eager, b.outputs())) if not torch.equal(e, o)]
    print('batched gridding: %d of %d requests differ' % (len(bad), len(eager)), flush=True)
    assert not bad
    b.free()


# ---- differentiation form --------------------------------------------------

def test_differentiation_form_n1024():
    from tomocupy_amd import FourierRec
    from tomocupy_amd.autograd import backproject, cgls, project
    n, nproj, nz, lam = 1024, 180, 32, 1e-3
    print('differentiation form n=%d nproj=%d nz=%d' % (n, nproj, nz), flush=True)
    theta = synth.angles(nproj)
    rec = FourierRec(n, nproj, nz, torch.from_numpy(theta).cuda(), 'float32')
    gen = torch.Generator(device='cuda').manual_seed(21)
    # dot-product test over the whole chunk
    x = torch.randn((nz, n, n), device='cuda', generator=gen)
    y = torch.randn((nz, nproj, n), device='cuda', generator=gen)
    ax, aty = project(x, rec), backproject(y, rec)
    lhs = torch.sum(ax.double() * y.double()).item()
    rhs = torch.sum(x.double() * aty.double()).item()
    dot = abs(lhs - rhs) / (torch.linalg.norm(ax.double()) * torch.linalg.norm(y.double())).item()
    print('dot test %.3g' % dot, flush=True)
    assert dot < 1e-5
    # gradient of |A x - y|^2 + lam |x|^2 against FP64 on the pair (0, nz/2)
    s4 = synth.sinograms(5, n, theta, 4)
    y32 = np.concatenate([s4] * (nz // 4))
    yt = torch.from_numpy(y32.astype(np.float32)).cuda()
    xg = (0.01 * x).requires_grad_(True)
    r = project(xg, rec).double() - yt.double()
    (torch.sum(r * r) + lam * torch.sum(xg.double() ** 2)).backward()
    pair = [0, nz // 2]
    x64 = xg.detach()[pair].double().cpu().numpy()
    ref = 2 * fr.backprojection(fr.forward(x64, theta) - y32[pair], theta) + 2 * lam * x64
    gerr = np.linalg.norm(xg.grad[pair].double().cpu().numpy() - ref) / np.linalg.norm(ref)
    print('gradient rel-L2 vs FP64 %.3g' % gerr, flush=True)
    assert gerr < 1e-5
    # 5 CG iterations: slices z and z + 16 of y32 are the same slice s4[z % 4],
    # so every inner product over the 32 slices is 4x the one over the 4
    # distinct pairs (s_j, s_j), and FP64 CG on y8 = [s0..s3, s0..s3] has the
    # same step sizes and iterates
    y8 = np.concatenate([s4, s4])

    def normal(v):
        return fr.backprojection(fr.forward(v, theta), theta) + lam * v

    x8 = np.zeros((8, n, n))
    rr8 = fr.backprojection(y8, theta)
    p8 = rr8.copy()
    rho = np.sum(rr8 * rr8)
    for _ in range(5):
        q = normal(p8)
        a = rho / np.sum(p8 * q)
        x8 += a * p8
        rr8 -= a * q
        rn = np.sum(rr8 * rr8)
        p8 = rr8 + (rn / rho) * p8
        rho = rn
    xg32 = cgls(rec, yt, lam=lam, niter=5).double().cpu().numpy()
    map8 = [(z % 4) + 4 * (z >= nz // 2) for z in range(nz)]
    err = np.linalg.norm(xg32 - x8[map8]) / np.linalg.norm(x8[map8])
    print('CG (5 iterations) iterate rel-L2 vs FP64 %.3g' % err, flush=True)
    assert err < 1e-4
