"""Float32 accuracy of the fused filter + backprojection on smooth data with
few angles, where a high-pass filter discards most of a row's energy: the
fused filter transforms the differenced row (d[j] = z[j] - z[j-1], spectra
H2 / (1 - W^k), DC from the row sum), so its rounding follows the row's
differences.  Before that form the error grew with n (5.5e-6 at n = 1024 and
1.1e-5 at n = 2048 for 24 angles with the shepp filter) while the
backprojection of the FP64-filtered rows stayed at 9e-7.

Bounds: per-slice rel-L2 against the FP64 oracle <= 5e-6 for the ramp
family, <= 1e-5 for 'none' (which keeps the direct form)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu


def _rel(got, ref):
    g = np.asarray(got, np.float64).reshape(len(ref), -1)
    r = np.asarray(ref, np.float64).reshape(len(ref), -1)
    return np.sqrt(((g - r) ** 2).sum(1) / (r ** 2).sum(1))


@pytest.mark.parametrize("n,filt,bound", [(1024, 'shepp', 5e-6), (1024, 'ramp', 5e-6),
                                          (2048, 'shepp', 5e-6), (1024, 'parzen', 5e-6),
                                          (512, 'none', 1e-5)])
def test_fused_filter_few_angles(n, filt, bound):
    from tomocupy_amd import BackprojFunctions
    nproj, nz = 24, 2
    theta = synth.angles(nproj)
    sino = synth.sinograms(3, n, theta, nz)
    ne = 4 * n
    centre = n / 2 + 0.25
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, filt), ne, centre),
                            theta)
    bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), centre, 'float32', filt)
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    bf.fbp_filter_backprojection(obj, torch.from_numpy(sino.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    err = _rel(obj.cpu().numpy(), ref)
    print('n', n, filt, 'fused rel-L2', err, flush=True)
    assert np.all(err <= bound), err
    bf.free()
