"""Full-size checks at the benchmark shapes (n = 2048, 1500 angles, chunks of 8)
and at the largest supported size (n = 4096), where the oracle is too slow
for whole chunks.

* n = 2048, 1500 angles, 8 slices, the fused chunk pipeline (fw, ramp,
  centre 1024.5): bitwise reproducible, and linear in the sinograms
  (BP(a x + b y) = a BP(x) + b BP(y), rel-L2 <= 1e-6).
* n = 2048 with 300 angles and n = 4096 with 96 angles (2n = 8192 compiled
  transforms), one slice pair, against the FP64 oracle (<= 1e-5).
* The adjoint dot test at n = 2048, 1500 angles (<= 1e-5).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return (torch.linalg.norm((a - b).double()) / torch.linalg.norm(b.double())).item()


def test_fullsize_pipeline_deterministic_and_linear():
    from tomocupy_amd import BackprojFunctions, ProcFunctions
    n, nproj, nz = 2048, 1500, 8
    print('full-size chunk', n, nproj, nz, flush=True)
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(2, n, theta, 2, stripes=True)
    reps = nz // 2
    g = [torch.from_numpy(np.ascontiguousarray(np.concatenate([a] * reps, axis=1))).cuda()
         for a in (data, dark, flat)]
    pf = ProcFunctions(nproj, nz, n, dark.shape[0], flat.shape[0], remove_stripe_method='fw')
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), n / 2 + 0.5, 'float32',
                           'ramp')
    sino = pf.proc_chunk(*g)
    o1 = torch.empty((nz, n, n), device='cuda')
    o2 = torch.empty((nz, n, n), device='cuda')
    bf.fbp_filter_backprojection(o1, sino)
    sino2 = pf.proc_chunk(*g)
    bf.fbp_filter_backprojection(o2, sino2)
    torch.cuda.synchronize()
    assert torch.equal(sino, sino2) and torch.equal(o1, o2)
    y = torch.randn_like(sino)
    oy = torch.empty_like(o1)
    bf.fbp_filter_backprojection(oy, y)
    mix = torch.empty_like(o1)
    bf.fbp_filter_backprojection(mix, 0.75 * sino - 1.5 * y)
    torch.cuda.synchronize()
    rel = _rel(mix, 0.75 * o1 - 1.5 * oy)
    print('linearity rel-L2 %.3g' % rel)
    assert rel <= 1e-6
    assert torch.isfinite(o1).all()


@pytest.mark.parametrize("n,nproj", [(2048, 300), (4096, 96)])
def test_large_n_matches_oracle(n, nproj):
    from tomocupy_amd import BackprojFunctions
    print('large n', n, nproj, flush=True)
    theta = synth.angles(nproj)
    sino = synth.sinograms(9, n, theta, 2)
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, n / 2),
                            theta)
    bf = BackprojFunctions(n, nproj, 2, torch.from_numpy(theta).cuda(), n / 2, 'float32',
                           'parzen')
    obj = torch.zeros((2, n, n), device='cuda')
    bf.fbp_filter_backprojection(obj, torch.from_numpy(sino.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    got = obj.double().cpu().numpy()
    err = np.sqrt(((got - ref) ** 2).reshape(2, -1).sum(1) / (ref ** 2).reshape(2, -1).sum(1))
    print('n=%d rel-L2 %s' % (n, err))
    assert np.all(err <= 1e-5), err


def test_fullsize_adjoint_dot():
    from tomocupy_amd import FourierRec
    n, nproj, nz = 2048, 1500, 4
    rec = FourierRec(n, nproj, nz, torch.from_numpy(synth.angles(nproj)).cuda())
    gen = torch.Generator(device='cuda').manual_seed(4)
    x = torch.randn((nz, n, n), device='cuda', generator=gen)
    y = torch.randn((nz, nproj, n), device='cuda', generator=gen)
    ax = torch.empty_like(y)
    aty = torch.empty_like(x)
    rec.forward(ax, x)
    rec.backprojection(aty, y)
    torch.cuda.synchronize()
    lhs = torch.sum(ax.double() * y.double()).item()
    rhs = torch.sum(x.double() * aty.double()).item()
    rel = abs(lhs - rhs) / (torch.linalg.norm(ax.double()) * torch.linalg.norm(y.double())).item()
    print('full-size dot test %.3g' % rel)
    assert rel <= 1e-5
