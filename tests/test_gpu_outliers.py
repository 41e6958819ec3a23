"""Zinger removal (``ProcFunctions.remove_outliers``, reference
``processing/proc_functions.py:96-107``) on the GPU against the oracle
(``scipy.ndimage.median_filter`` + the replacement rule, pinned by a brute-force
window median in ``test_oracle_preproc.py``): bit-exact on uint16 counts."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import preproc_fp64 as pp
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu


def _zingered(rows, nz, n, seed):
    rng = np.random.default_rng(seed)
    d = rng.integers(20000, 30000, size=(rows, nz, n)).astype(np.uint16)
    hits = rng.random(d.shape) < 0.03
    d[hits] = rng.integers(40000, 65535, size=int(hits.sum())).astype(np.uint16)
    return d


@pytest.mark.parametrize("w", [2, 3, 4, 5, 6, 7, 8, 9])
def test_remove_outliers_bit_exact(w):
    from tomocupy_amd import ProcFunctions
    rows, nz, n = 37, 3, 70
    print('dezinger width', w, flush=True)
    data = _zingered(rows, nz, n, 100 + w)
    pf = ProcFunctions(rows, nz, n, 2, 3, dezinger=w, dezinger_threshold=3000)
    g = torch.from_numpy(data.copy()).cuda()
    pf.remove_outliers(g)
    torch.cuda.synchronize()
    got = g.cpu().numpy()
    ref = pp.remove_outliers(data, w, 3000)
    assert np.array_equal(got, ref), int((got != ref).sum())
    assert (got != data).any()
    pf.free()


@pytest.mark.parametrize("rows,n", [(1, 5), (2, 3), (3, 130)])
def test_remove_outliers_small_and_reflected(rows, n):
    # windows wider than the array: the reflection wraps more than once
    from tomocupy_amd import ProcFunctions
    data = _zingered(rows, 2, n, 7 * rows + n)
    pf = ProcFunctions(rows, 2, n, 1, 1, dezinger=7, dezinger_threshold=100)
    g = torch.from_numpy(data.copy()).cuda()
    pf.remove_outliers(g)
    torch.cuda.synchronize()
    assert np.array_equal(g.cpu().numpy(), pp.remove_outliers(data, 7, 100))
    pf.free()


def test_remove_outliers_off_and_too_wide():
    from tomocupy_amd import ProcFunctions, _lib
    data = _zingered(4, 1, 8, 3)
    pf = ProcFunctions(4, 1, 8, 1, 1, dezinger=0)
    g = torch.from_numpy(data.copy()).cuda()
    pf.remove_outliers(g)
    torch.cuda.synchronize()
    assert np.array_equal(g.cpu().numpy(), data)
    pf.dezinger = 10
    with pytest.raises(_lib.TomoError):
        pf.remove_outliers(g)
    pf.free()


def test_chain_with_dezinger_matches_oracle():
    from tomocupy_amd import ProcFunctions
    n, nproj, nz = 96, 50, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(1, n, theta, nz, ndark=4, nflat=6)
    rng = np.random.default_rng(5)
    for a in (data, dark, flat):
        hits = rng.random(a.shape) < 0.01
        a[hits] = 65000
    ref = pp.proc_chain(data, dark, flat, dezinger=3, dezinger_threshold=5000)
    pf = ProcFunctions(nproj, nz, n, 4, 6, dezinger=3, dezinger_threshold=5000)
    cu = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (data, dark, flat)]
    out = pf.proc_chunk(*cu)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    err = np.sqrt(((got - ref) ** 2).reshape(nz, -1).sum(1) / (ref ** 2).reshape(nz, -1).sum(1))
    print('dezinger chain rel-L2', err)
    assert np.all(err <= 1e-5), err
    # the inputs were cleaned in place, as the reference's proc_sino does
    assert np.array_equal(cu[0].cpu().numpy(), pp.remove_outliers(data, 3, 5000))
    pf.free()
