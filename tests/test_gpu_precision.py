"""Precision form: 16-bit storage (IEEE half, bfloat16) with float32 arithmetic.

Tolerances, per slice, relative L2 against the FP64 oracle:
* float16: 2e-3 (backprojection, fused filter + backprojection, full chain);
* bfloat16: 1e-2 (8-bit significand: each rounding of a stored stage costs
  about 2e-3 rms, and the path stores four stages).
The maximum absolute error relative to the largest oracle value is printed.

Where the input sinograms are stored in 16 bits, the oracle runs on the same
16-bit-rounded input: the ramp-type filters amplify the input rounding (white
noise) relative to the filtered signal, an error of the storage format that
the reference's float16 path has as well, not of the kernels.  The config-1
chain compares end to end against the unrounded oracle.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from oracle import preproc_fp64 as pp
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu

TOL = {'float16': 2e-3, 'bfloat16': 1e-2}


def _rel(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    num = np.sqrt(((got - ref) ** 2).reshape(len(ref), -1).sum(1))
    den = np.sqrt((ref ** 2).reshape(len(ref), -1).sum(1))
    mx = np.abs(got - ref).max() / np.abs(ref).max()
    return num / np.maximum(den, 1e-300), mx


def _filtered(n, nproj, nz, cfg, name, center):
    theta = synth.angles(nproj)
    sino = synth.sinograms(cfg, n, theta, nz)
    ne = 4 * n
    return theta, sino, ff.fbp_filter_center(sino, ff.calc_filter(ne, name), ne, center)


@pytest.mark.parametrize("dtype", ['float16', 'bfloat16'])
@pytest.mark.parametrize("n,nproj,nz", [(256, 180, 4), (384, 200, 3)])
def test_backprojection_16bit(dtype, n, nproj, nz):
    from tomocupy_amd import FourierRec
    td = {'float16': torch.float16, 'bfloat16': torch.bfloat16}[dtype]
    print('16-bit backprojection', dtype, n, nproj, nz, flush=True)
    theta, _, filt = _filtered(n, nproj, nz, 4, 'parzen', n / 2)
    ref = fr.backprojection(filt, theta)
    rec = FourierRec(n, nproj, nz, torch.from_numpy(theta).cuda(), dtype)
    obj = torch.zeros((nz, n, n), dtype=td, device='cuda')
    rec.backprojection(obj, torch.from_numpy(filt).to(td).cuda())
    torch.cuda.synchronize()
    err, mx = _rel(obj.float().cpu().numpy(), ref)
    print('%s rel-L2 %s max-abs/max %.3g' % (dtype, err, mx))
    assert np.all(err <= TOL[dtype]), err


@pytest.mark.parametrize("dtype", ['float16', 'bfloat16'])
def test_fused_filter_backprojection_16bit(dtype):
    from tomocupy_amd import BackprojFunctions
    td = {'float16': torch.float16, 'bfloat16': torch.bfloat16}[dtype]
    n, nproj, nz, center = 256, 240, 4, 129.5
    print('16-bit fused', dtype, flush=True)
    theta, sino, _ = _filtered(n, nproj, nz, 5, 'ramp', center)
    ne = 4 * n
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), center, dtype, 'ramp')
    obj = torch.zeros((nz, n, n), dtype=td, device='cuda')
    d = torch.from_numpy(sino).to(td).cuda()
    sino_r = d.double().cpu().numpy()
    ref = fr.backprojection(ff.fbp_filter_center(sino_r, ff.calc_filter(ne, 'ramp'), ne, center),
                            theta)
    bf.fbp_filter_backprojection(obj, d)
    torch.cuda.synchronize()
    err, mx = _rel(obj.float().cpu().numpy(), ref)
    print('%s fused rel-L2 %s max-abs/max %.3g' % (dtype, err, mx))
    assert np.all(err <= TOL[dtype]), err


@pytest.mark.parametrize("dtype", ['float16', 'bfloat16'])
def test_filter_center_16bit(dtype):
    from tomocupy_amd import BackprojFunctions
    td = {'float16': torch.float16, 'bfloat16': torch.bfloat16}[dtype]
    n, nproj, nz, center = 128, 60, 3, 66.25
    theta, sino, _ = _filtered(n, nproj, nz, 6, 'shepp', center)
    ne = 4 * n if dtype != 'float16' else 512
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), center, dtype, 'shepp')
    d = torch.from_numpy(sino).to(td).cuda()
    ref = ff.fbp_filter_center(d.double().cpu().numpy(), ff.calc_filter(ne, 'shepp'), ne, center)
    bf.fbp_filter_center(d)
    torch.cuda.synchronize()
    err, mx = _rel(d.float().cpu().numpy(), ref)
    print('%s filter rel-L2 %s' % (dtype, err))
    assert np.all(err <= TOL[dtype]), err


@pytest.mark.parametrize("dtype", ['float16', 'bfloat16'])
def test_config1_chain_16bit(dtype):
    """Config-1 shape through preprocessing (float32 arithmetic, 16-bit
    sinogram storage), the fused Parzen filter and the backprojection."""
    from tomocupy_amd import BackprojFunctions, ProcFunctions
    td = {'float16': torch.float16, 'bfloat16': torch.bfloat16}[dtype]
    n, nproj, nz = 512, 720, 16
    print('config-1 16-bit chain', dtype, flush=True)
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(1, n, theta, nz)
    sino = pp.proc_chain(data, dark, flat)
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, n / 2),
                            theta)
    pf = ProcFunctions(nproj, nz, n, dark.shape[0], flat.shape[0], dtype)
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), n / 2, dtype, 'parzen')
    g = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (data, dark, flat)]
    out = pf.proc_chunk(*g)
    assert out.dtype == td
    obj = torch.zeros((nz, n, n), dtype=td, device='cuda')
    bf.fbp_filter_backprojection(obj, out)
    torch.cuda.synchronize()
    err, mx = _rel(obj.float().cpu().numpy(), ref)
    print('config-1 %s rel-L2 max %.3g median %.3g, max-abs/max %.3g'
          % (dtype, err.max(), np.median(err), mx))
    assert np.all(err <= TOL[dtype]), err


def test_forward_rejects_16bit():
    from tomocupy_amd import FourierRec, _lib
    theta = synth.angles(30)
    rec = FourierRec(64, 30, 2, torch.from_numpy(theta).cuda(), 'float16')
    with pytest.raises(_lib.TomoError):
        rec.forward(torch.zeros((2, 30, 64), device='cuda'), torch.zeros((2, 64, 64), device='cuda'))
