"""Double field of view (reference ``proc_functions.py:109-124`` with the
centre rule of ``reader.py:143-148``) on the GPU against the FP64 oracle:
the padding kernel alone, and a 360-degree reconstruction at n = 2 ni."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from oracle import preproc_fp64 as pp
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("centeri,dtype", [(50.0, 'float32'), (11.5, 'float32'),
                                            (47.25, 'float16')])
def test_pad360_matches_oracle(centeri, dtype):
    from tomocupy_amd import ProcFunctions, _lib
    ni, nproj, nz = 64, 30, 3
    print('pad360', centeri, dtype, flush=True)
    rng = np.random.default_rng(int(centeri * 4))
    sino = rng.random((nz, nproj, ni)).astype(np.float32)
    pf = ProcFunctions(nproj, nz, ni, 1, 1, dtype, file_type='double_fov', center=centeri)
    td = _lib.torch_dtype(dtype)
    g = torch.from_numpy(sino).cuda().to(td)
    out = pf.pad360(g)
    torch.cuda.synchronize()
    got = out.float().cpu().numpy()
    ref, center, flip = pp.pad360(g.float().cpu().numpy(), centeri)
    assert flip == (centeri < ni // 2) and pf.center360 == center
    tol = 1e-7 if dtype == 'float32' else 1e-3
    assert np.allclose(got, ref, rtol=tol, atol=tol), np.max(np.abs(got - ref))
    assert np.all(got[..., ni:] == 0)
    pf.free()


def test_double_fov_reconstruction_matches_oracle():
    from tomocupy_amd import BackprojFunctions, ProcFunctions
    ni, nproj, nz = 64, 90, 2
    centeri = 12.0                                     # axis near the left border
    theta = (np.arange(nproj) * (2 * np.pi / nproj)).astype(np.float32)
    data, dark, flat, _ = synth.raw_counts(1, ni, theta, nz, ndark=3, nflat=4)
    sino_ref = pp.proc_chain(data, dark, flat)         # (nz, nproj, ni)
    padded, center, _ = pp.pad360(sino_ref, centeri)
    n = 2 * ni
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(padded, ff.calc_filter(ne, 'parzen'), ne,
                                                 center), theta)
    pf = ProcFunctions(nproj, nz, ni, 3, 4, file_type='double_fov', center=centeri)
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), pf.center360,
                           'float32', 'parzen')
    cu = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (data, dark, flat)]
    sino = pf.proc_chunk(*cu)
    assert tuple(sino.shape) == (nz, nproj, n)
    obj = torch.empty((nz, n, n), device='cuda')
    bf.fbp_filter_backprojection(obj, sino)
    torch.cuda.synchronize()
    got = obj.double().cpu().numpy()
    err = np.sqrt(((got - ref) ** 2).reshape(nz, -1).sum(1) / (ref ** 2).reshape(nz, -1).sum(1))
    print('double-fov rel-L2', err)
    assert np.all(err <= 1e-5), err
    pf.free()
    bf.free()
