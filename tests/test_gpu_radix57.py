"""Detector widths with prime factors above 3 (the reference's cuFFT takes any
length; here the LDS Stockham transforms have butterflies for the primes
5..17 and direct-DFT passes for the primes 19..127): backprojection, the fused
filter + backprojection, the unfused filter and the forward projection
(autograd) against the FP64 oracle (per-slice rel-L2 <= 1e-5)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu

RTOL32 = 1e-5


def _rel(got, ref):
    g = np.asarray(got, np.float64).reshape(len(ref), -1)
    r = np.asarray(ref, np.float64).reshape(len(ref), -1)
    return np.sqrt(((g - r) ** 2).sum(1) / (r ** 2).sum(1))


def _sinos(rng, nz, nproj, n):
    d = rng.standard_normal((nz, nproj, n))
    f = np.fft.rfft(d, axis=-1)
    f *= np.clip(1 - 2 * np.arange(f.shape[-1]) / n, 0, None) ** 3
    return np.fft.irfft(f, n, axis=-1)


@pytest.mark.parametrize("n,nproj,nz", [(80, 60, 4), (70, 48, 2), (100, 75, 3), (140, 90, 2),
                                        (350, 64, 2), (68, 40, 2), (88, 50, 3), (286, 40, 2),
                                        (254, 30, 2), (228, 40, 2), (184, 36, 2),
                                        (2448, 16, 1)])
def test_backprojection_5_7(n, nproj, nz):
    from tomocupy_amd import FourierRec
    print('n', n, flush=True)
    rng = np.random.default_rng(n)
    theta = synth.angles(nproj)
    data = _sinos(rng, nz, nproj, n)
    ref = fr.backprojection(data, theta)
    rec = FourierRec(n, nproj, nz, torch.from_numpy(theta).cuda(), 'float32')
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    rec.backprojection(obj, torch.from_numpy(data.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    err = _rel(obj.cpu().numpy(), ref)
    print('rel-L2', err)
    assert np.all(err <= RTOL32), err


@pytest.mark.parametrize("n,filt", [(70, 'shepp'), (100, 'parzen'), (68, 'parzen'),
                                    (286, 'hann'), (184, 'cosine')])
def test_filter_backprojection_5_7(n, filt):
    from tomocupy_amd import BackprojFunctions
    nproj, nz = 60, 4
    rng = np.random.default_rng(3 * n)
    theta = synth.angles(nproj)
    data = rng.uniform(0, 1, (nz, nproj, n))
    sht = np.array([0.0, 0.5, -1.25, 2.0])
    centre = n / 2 + 0.3
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(data, ff.calc_filter(ne, filt), ne, centre, sht),
                            theta)
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), centre, 'float32', filt)
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    d = torch.from_numpy(data.astype(np.float32)).cuda()
    bf.fbp_filter_backprojection(obj, d, torch.from_numpy(sht.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    err = _rel(obj.cpu().numpy(), ref)
    print('fused rel-L2', err)
    assert np.all(err <= RTOL32), err
    # the unfused filter (k_flt_rows, ne = 4n) then the backprojection
    d = torch.from_numpy(data.astype(np.float32)).cuda()
    bf.fbp_filter_center(d, torch.from_numpy(sht.astype(np.float32)).cuda())
    obj2 = torch.zeros_like(obj)
    bf.cl_rec.backprojection(obj2, d)
    torch.cuda.synchronize()
    err2 = _rel(obj2.cpu().numpy(), ref)
    print('unfused rel-L2', err2)
    assert np.all(err2 <= RTOL32), err2
    bf.free()


@pytest.mark.parametrize("n,nproj,nz", [(70, 30, 2), (68, 24, 3), (228, 20, 2), (254, 16, 2)])
def test_forward_general_widths(n, nproj, nz):
    # the forward projection (adjoint) at widths with 5/7, 17, 19 and 127:
    # against the FP64 oracle adjoint and the dot-product test
    from tomocupy_amd import FourierRec
    print('n', n, flush=True)
    rng = np.random.default_rng(n + 1)
    theta = synth.angles(nproj, 0.1)
    rec = FourierRec(n, nproj, nz, torch.from_numpy(theta).cuda(), 'float32')
    x = rng.standard_normal((nz, n, n))
    y = rng.standard_normal((nz, nproj, n))
    xg = torch.from_numpy(x.astype(np.float32)).cuda()
    yg = torch.from_numpy(y.astype(np.float32)).cuda()
    Ax = torch.zeros((nz, nproj, n), dtype=torch.float32, device='cuda')
    Aty = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    rec.forward(Ax, xg)
    rec.backprojection(Aty, yg)
    torch.cuda.synchronize()
    lhs = float((Ax.double() * yg.double()).sum())
    rhs = float((xg.double() * Aty.double()).sum())
    rel = abs(lhs - rhs) / (float(Ax.double().norm()) * float(yg.double().norm()))
    ref = fr.forward(xg.cpu().numpy().astype(np.float64), theta)
    err = _rel(Ax.cpu().numpy(), ref)
    print('dot test', rel, 'rel-L2', err)
    assert rel < 1e-5
    assert np.all(err <= RTOL32), err
