"""Titarenko ring removal (reference ``processing/remove_stripe.py:296-309``)
on the GPU against the FP64 oracle (pinned by a direct circular convolution in
``test_oracle_preproc.py``), alone and inside the preprocessing chain."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import preproc_fp64 as pp
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("beta,mask,n", [(0.022, 1.0, 96), (0.1, 0.5, 128), (0.3, 0.25, 64)])
def test_ti_matches_oracle(beta, mask, n):
    from tomocupy_amd import ProcFunctions
    nproj, nz = 40, 3
    print('ti', beta, mask, n, flush=True)
    rng = np.random.default_rng(21)
    res = (0.5 + 0.4 * rng.random((nproj, nz, n))
           * (1 + 0.05 * (rng.random(n) < 0.1))[None, None, :]).astype(np.float32)
    ref = pp.remove_stripe_ti(res.astype(np.float64), beta, mask)
    pf = ProcFunctions(nproj, nz, n, 1, 1, remove_stripe_method='ti', ti_beta=beta,
                       ti_mask=mask)
    g = torch.from_numpy(res.copy()).cuda()
    pf.remove_stripe_ti(g)
    torch.cuda.synchronize()
    got = g.cpu().numpy().astype(np.float64)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print('ti max rel err', err)
    assert err <= 1e-6, err
    assert not np.array_equal(got, res)
    pf.free()


def test_ti_chain_matches_oracle():
    from tomocupy_amd import ProcFunctions
    n, nproj, nz = 96, 50, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(1, n, theta, nz, ndark=4, nflat=6, stripes=True)
    ref = pp.proc_chain(data, dark, flat, stripe='ti', ti_beta=0.05, ti_mask=1.0)
    pf = ProcFunctions(nproj, nz, n, 4, 6, remove_stripe_method='ti', ti_beta=0.05)
    cu = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (data, dark, flat)]
    out = pf.proc_chunk(*cu)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    err = np.sqrt(((got - ref) ** 2).reshape(nz, -1).sum(1) / (ref ** 2).reshape(nz, -1).sum(1))
    print('ti chain rel-L2', err)
    assert np.all(err <= 1e-5), err
    pf.free()


def test_ti_rejects_odd_width():
    from tomocupy_amd import ProcFunctions, _lib
    pf = ProcFunctions(4, 1, 7, 1, 1, remove_stripe_method='ti')
    with pytest.raises(_lib.TomoError):
        pf.remove_stripe_ti(torch.ones((4, 1, 7), device='cuda'))
    pf.free()
