"""Compile-time transforms of the sizes 2^k 3^a 5^b (2n = 768, 1536, 3072,
6144; 1280, 2304, 2560, 3840, 4608, 5120, with the n-point spectrum kernels
compiled where n is such a size too and run-time otherwise):
the backprojection, the fused filter + backprojection and the forward
projection against the FP64 oracle (per-slice rel-L2 <= 1e-5), and the
compiled kernels and the run-time radix plans of the same plan shape
(``fft='generic'``) both against the oracle (float32 <= 1e-5, float16
storage <= 2e-3)."""
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


@pytest.mark.parametrize("n,nproj,nz", [(384, 40, 2), (768, 30, 3), (1536, 16, 2),
                                        (3072, 8, 1), (640, 30, 2), (1152, 20, 3),
                                        (1280, 16, 2), (1920, 12, 2), (2304, 8, 2),
                                        (2560, 8, 1)])
def test_backprojection_static3(n, nproj, nz):
    from tomocupy_amd import FourierRec
    print('n', n, flush=True)
    rng = np.random.default_rng(n)
    theta = synth.angles(nproj, 0.2)
    data = synth.sinograms(int(rng.integers(1000)), n, theta, nz)
    ref = fr.backprojection(data, theta)
    rec = FourierRec(n, nproj, nz, torch.from_numpy(theta).cuda(), 'float32')
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    rec.backprojection(obj, torch.from_numpy(data.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    err = _rel(obj.cpu().numpy(), ref)
    print('rel-L2', err)
    assert np.all(err <= RTOL32), err
    rec.free()


@pytest.mark.parametrize("n", [384, 768, 1536, 640, 1280, 2304])
def test_filter_backprojection_static3(n):
    from tomocupy_amd import BackprojFunctions
    nproj, nz = 24, 3
    theta = synth.angles(nproj)
    data = synth.sinograms(7, n, theta, nz)
    sht = np.array([0.0, 0.75, -1.5])
    centre = n / 2 - 0.4
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(data, ff.calc_filter(ne, 'parzen'), ne, centre,
                                                 sht), theta)
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), centre, 'float32',
                           'parzen')
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    bf.fbp_filter_backprojection(obj, torch.from_numpy(data.astype(np.float32)).cuda(),
                                 torch.from_numpy(sht.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    err = _rel(obj.cpu().numpy(), ref)
    print('n', n, 'fused rel-L2', err)
    assert np.all(err <= RTOL32), err
    bf.free()


@pytest.mark.parametrize("n", [384, 768, 640, 1152])
def test_forward_static3(n):
    from tomocupy_amd import FourierRec
    nproj, nz = 20, 2
    rng = np.random.default_rng(n + 7)
    theta = synth.angles(nproj, 0.1)
    rec = FourierRec(n, nproj, nz, torch.from_numpy(theta).cuda(), 'float32')
    x = rng.standard_normal((nz, n, n))
    xg = torch.from_numpy(x.astype(np.float32)).cuda()
    Ax = torch.zeros((nz, nproj, n), dtype=torch.float32, device='cuda')
    rec.forward(Ax, xg)
    torch.cuda.synchronize()
    ref = fr.forward(xg.cpu().numpy().astype(np.float64), theta)
    err = _rel(Ax.cpu().numpy(), ref)
    print('n', n, 'forward rel-L2', err)
    assert np.all(err <= RTOL32), err
    rec.free()


@pytest.mark.parametrize("n,dtype", [(768, 'float32'), (1024, 'float32'), (1536, 'float32'),
                                     (3072, 'float32'), (768, 'float16'), (1536, 'float16'),
                                     (1280, 'float32'), (2304, 'float32'), (1920, 'float32'),
                                     (1280, 'float16')])
def test_compiled_equals_runtime_plans(n, dtype):
    from tomocupy_amd import BackprojFunctions
    from tomocupy_amd.backproj_functions import padded_length
    # parzen: with a high-pass filter (ramp, shepp) on smooth data and few
    # angles, the float32 FFT filter's rounding grows with n; float16 storage
    # at the usual nproj = 1.5 n
    nproj, nz = (48 if dtype == 'float32' else 3 * n // 2), 2
    th = synth.angles(nproj)
    theta = torch.from_numpy(th).cuda()
    td = torch.float32 if dtype == 'float32' else torch.float16
    sino = synth.sinograms(3, n, th, nz)
    data = torch.from_numpy(sino.astype(np.float32)).to(td).cuda()
    ne = padded_length(n, dtype)  # 16-bit plans pad to a power of two
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, n / 2),
                            th)
    outs = []
    for fft in ('static', 'generic'):
        bf = BackprojFunctions(n, nproj, nz, theta, n / 2, dtype, 'parzen',
                               rec_options={'fft': fft})
        o = torch.zeros((nz, n, n), dtype=td, device='cuda')
        bf.fbp_filter_backprojection(o, data)
        torch.cuda.synchronize()
        err = _rel(o.double().cpu().numpy(), ref)
        print('n', n, dtype, fft, 'rel-L2 vs FP64', err, flush=True)
        assert np.all(err <= (RTOL32 if dtype == 'float32' else 2e-3)), err
        outs.append(o.double())
        bf.free()
    rel = (torch.linalg.norm(outs[0] - outs[1]) / torch.linalg.norm(outs[1])).item()
    print('n', n, dtype, 'compiled vs run-time rel-L2 %.3g' % rel)
    assert rel <= (RTOL32 if dtype == 'float32' else 2e-3)
