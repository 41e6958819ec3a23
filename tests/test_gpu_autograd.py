"""Differentiation form: the USFFT projector behind torch.autograd.

Checks, on the GPU against the FP64 oracle:
* dot-product test |<Ax,y> - <x,A^T y>| / (|Ax| |y|) < 1e-5 through the
  autograd functions;
* central finite differences of the Tikhonov least-squares loss on 64 seeded
  directions (the loss is quadratic, so the central difference is exact up to
  rounding): relative deviation < 1e-3 per direction;
* the autograd gradient against the FP64 oracle gradient
  2 A^T (A x - y) + 2 lam x: rel-L2 < 1e-5;
* CG least squares (20 iterations, 32 angles) against the same iteration run
  with the FP64 oracle operators: iterate rel-L2 < 1e-4 after 5 iterations;
  objective within 2% at every iteration (float32 rounding of the operator
  slowly de-orthogonalises the search directions; the FP64 run has the same
  start and data);
* the fused CG vector updates against the same updates in torch.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import fourierrec_fp64 as fr
from tomocupy_amd import FourierRec, synth

pytestmark = pytest.mark.gpu


def _plan(n, nproj, nz, seed=0):
    theta = synth.angles(nproj)
    return FourierRec(n, nproj, nz, torch.from_numpy(theta).cuda(), 'float32'), theta


def test_autograd_dot_product():
    from tomocupy_amd.autograd import backproject, project
    n, nproj, nz = 128, 90, 4
    rec, _ = _plan(n, nproj, nz)
    g = torch.Generator(device='cuda').manual_seed(3)
    x = torch.randn((nz, n, n), device='cuda', generator=g)
    y = torch.randn((nz, nproj, n), device='cuda', generator=g)
    ax = project(x, rec)
    aty = backproject(y, rec)
    lhs = torch.sum(ax.double() * y.double()).item()
    rhs = torch.sum(x.double() * aty.double()).item()
    rel = abs(lhs - rhs) / (torch.linalg.norm(ax.double()).item() * torch.linalg.norm(y.double()).item())
    print('dot test', lhs, rhs, rel)
    assert rel < 1e-5


def _loss(x, y, rec, lam):
    from tomocupy_amd.autograd import project
    r = project(x, rec).double() - y.double()
    return torch.sum(r * r) + lam * torch.sum(x.double() * x.double())


def test_autograd_finite_differences():
    n, nproj, nz, lam = 128, 60, 2, 0.1
    rec, theta = _plan(n, nproj, nz)
    y = torch.from_numpy(synth.sinograms(5, n, theta, nz).astype(np.float32)).cuda()
    g = torch.Generator(device='cuda').manual_seed(11)
    x = (0.01 * torch.randn((nz, n, n), device='cuda', generator=g)).requires_grad_(True)
    L = _loss(x, y, rec, lam)
    L.backward()
    grad = x.grad.detach().double()
    from tomocupy_amd.autograd import project
    res = (project(x.detach(), rec) - y).double()
    worst = 0.0
    for k in range(64):
        d = torch.randn((nz, n, n), device='cuda', generator=g)
        ad = project(d, rec).double()
        h = (torch.linalg.norm(res) / torch.linalg.norm(ad)).item()
        with torch.no_grad():
            lp = _loss(x.detach() + h * d, y, rec, lam).item()
            lm = _loss(x.detach() - h * d, y, rec, lam).item()
        fd = (lp - lm) / (2 * h)
        an = torch.sum(grad * d.double()).item()
        rel = abs(fd - an) / max(abs(an), 1e-30)
        worst = max(worst, rel)
    print('finite differences: worst relative deviation over 64 directions %.3g' % worst)
    assert worst < 1e-3


def test_autograd_gradient_matches_fp64():
    n, nproj, nz, lam = 96, 48, 2, 0.05
    rec, theta = _plan(n, nproj, nz)
    rng = np.random.default_rng(7)
    x64 = rng.standard_normal((nz, n, n)) * fr.circ_mask(n)
    y64 = synth.sinograms(6, n, theta, nz)
    ref = 2 * fr.backprojection(fr.forward(x64, theta) - y64, theta) + 2 * lam * x64
    x = torch.from_numpy(x64.astype(np.float32)).cuda().requires_grad_(True)
    y = torch.from_numpy(y64.astype(np.float32)).cuda()
    _loss(x, y, rec, lam).backward()
    got = x.grad.double().cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print('gradient rel-L2 vs FP64 %.3g' % err)
    assert err < 1e-5


def test_cgls_matches_fp64_iteration():
    """CG iterates in float32 follow the FP64 iteration closely at first; later
    they drift along the poorly conditioned directions (sparse angles), which
    barely move the objective, so the objective is compared at every step."""
    from tomocupy_amd.autograd import cgls, project
    n, nproj, nz, lam, niter = 96, 32, 2, 1e-3, 20
    rec, theta = _plan(n, nproj, nz)
    y64 = synth.sinograms(8, n, theta, nz)
    # FP64 CG on the normal equations with the oracle operators
    def normal(v):
        return fr.backprojection(fr.forward(v, theta), theta) + lam * v
    x = np.zeros((nz, n, n))
    r = fr.backprojection(y64, theta)
    p = r.copy()
    rr = np.sum(r * r)
    objective, xs = [], []
    for _ in range(niter):
        q = normal(p)
        a = rr / np.sum(p * q)
        x += a * p
        r -= a * q
        rn = np.sum(r * r)
        p = r + (rn / rr) * p
        rr = rn
        res = fr.forward(x, theta) - y64
        objective.append(np.sum(res * res) + lam * np.sum(x * x))
        xs.append(x.copy())
    assert all(b <= a * (1 + 1e-12) for a, b in zip(objective, objective[1:]))
    y = torch.from_numpy(y64.astype(np.float32)).cuda()
    got_obj, got_x = [], []

    def cb(k, xk, rr):
        res = (project(xk, rec) - y).double()
        got_obj.append((torch.sum(res * res) + lam * torch.sum(xk.double() ** 2)).item())
        got_x.append(xk.double().cpu().numpy())

    cgls(rec, y, lam=lam, niter=niter, callback=cb)
    rel_obj = np.abs(np.array(got_obj) - np.array(objective)) / np.array(objective)
    err5 = np.linalg.norm(got_x[4] - xs[4]) / np.linalg.norm(xs[4])
    errN = np.linalg.norm(got_x[-1] - xs[-1]) / np.linalg.norm(xs[-1])
    print('CG objective rel. deviation max %.3g; iterate rel-L2 at 5: %.3g, at %d: %.3g'
          % (rel_obj.max(), err5, niter, errN))
    print('objective fp32', np.array(got_obj))
    print('objective fp64', np.array(objective))
    assert err5 < 1e-4
    assert rel_obj.max() < 2e-2
    assert got_obj[-1] < 0.05 * got_obj[0]


@pytest.mark.parametrize("n", [1, 1000, 3000001])
def test_dot_f64_matches_numpy(n):
    """The CG inner product (tomo_dot_f64): FP64 accumulation of float32
    products, against numpy in FP64, and bitwise equal on repetition."""
    from tomocupy_amd.autograd import _dot
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    got = _dot(ta, tb)
    again = _dot(ta, tb)
    ref = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
    scale = float(np.abs(a.astype(np.float64) * b.astype(np.float64)).sum())
    assert got.dtype == torch.float64
    assert abs(float(got) - ref) <= 1e-13 * scale
    assert float(got) == float(again)


@pytest.mark.parametrize("n", [1, 5000, 3000001])
def test_cg_updates_match_torch(n):
    """The fused CG vector updates (tomo_cg_step, tomo_cg_direction) against
    the same updates in torch with the float32 step sizes: x and r within one
    rounding (one fused multiply-add against a multiply and an add), the
    returned r.r bitwise equal to tomo_dot_f64 of the updated r, and the
    zero-denominator rule (alpha = 0 when p.q <= 0, beta = 0 when r.r <= 0)."""
    from tomocupy_amd import _lib
    from tomocupy_amd.autograd import _dot
    L = _lib.lib()
    st = _lib.stream_ptr(None)
    rng = np.random.default_rng(n + 11)
    x, r, p, q = (torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
                  for _ in range(4))
    rr = torch.zeros(1025, dtype=torch.float64, device='cuda')
    pq = torch.zeros(1025, dtype=torch.float64, device='cuda')
    out = torch.zeros(1025, dtype=torch.float64, device='cuda')
    rr[0], pq[0] = 3.5, 1.75
    for case in ('step', 'zero'):
        if case == 'zero':
            pq[0] = -1.0
        alpha = torch.where(pq[0] > 0, rr[0] / torch.where(pq[0] > 0, pq[0], 1.0),
                            torch.zeros((), dtype=torch.float64, device='cuda')).float()
        xe = x + alpha * p
        re = r - alpha * q
        _lib.check(L.tomo_cg_step(x.data_ptr(), r.data_ptr(), p.data_ptr(), q.data_ptr(), n,
                                  rr.data_ptr(), pq.data_ptr(), out.data_ptr(), st),
                   'tomo_cg_step')
        torch.cuda.synchronize()
        tol = 2 * torch.finfo(torch.float32).eps
        assert torch.allclose(x, xe, rtol=tol, atol=tol * float(alpha.abs()) * 8)
        assert torch.allclose(r, re, rtol=tol, atol=tol * float(alpha.abs()) * 8)
        assert float(out[0]) == float(_dot(r, r))
        if case == 'zero':
            assert torch.equal(x, xe) and torch.equal(r, re)
    # direction: beta = rr_new / rr
    rnew = torch.zeros(1025, dtype=torch.float64, device='cuda')
    for rr0, rn0 in ((2.0, 0.5), (0.0, 0.5)):
        rr[0], rnew[0] = rr0, rn0
        beta = np.float32(rn0 / rr0) if rr0 > 0 else np.float32(0.0)
        pe = r + float(beta) * p
        _lib.check(L.tomo_cg_direction(p.data_ptr(), r.data_ptr(), n, rr.data_ptr(),
                                       rnew.data_ptr(), st), 'tomo_cg_direction')
        torch.cuda.synchronize()
        eps = torch.finfo(torch.float32).eps
        assert torch.allclose(p, pe, rtol=2 * eps,
                              atol=4 * eps * float(beta) * float(p.abs().max()))
        if rr0 == 0.0:
            assert torch.equal(p, r)
