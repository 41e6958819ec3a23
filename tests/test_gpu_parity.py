"""GPU parity of the HIP fourierrec path against the FP64 oracle.

Tolerance (float32 path): per-slice relative L2 error <= 1e-5 against the FP64
oracle, for the backprojection, the filter and the full chain.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import filter_fp64 as ff
from oracle import fourierrec_fp64 as fr
from oracle import preproc_fp64 as pp
from tomocupy_amd import synth

pytestmark = pytest.mark.gpu

RTOL32 = 1e-5


def _cuda(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def _rel_l2_per_slice(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    num = np.sqrt(((got - ref) ** 2).reshape(len(ref), -1).sum(1))
    den = np.sqrt((ref ** 2).reshape(len(ref), -1).sum(1))
    return num / np.maximum(den, 1e-300)


def _smooth_sinograms(rng, nz, nproj, n):
    """Random sinograms with a Parzen-like roll-off (as filtered data has)."""
    d = rng.standard_normal((nz, nproj, n))
    f = np.fft.rfft(d, axis=-1)
    k = np.arange(f.shape[-1]) / n
    f *= np.clip(1 - 2 * k, 0, None) ** 3
    return np.fft.irfft(f, n, axis=-1)


@pytest.mark.parametrize("n,nproj,nz,start", [(64, 48, 4, 0.0), (96, 60, 3, 0.2),
                                              (128, 90, 2, 0.0), (48, 30, 5, 0.7)])
def test_backprojection_matches_oracle(n, nproj, nz, start):
    from tomocupy_amd import FourierRec
    print('case n=%d nproj=%d nz=%d' % (n, nproj, nz), flush=True)
    rng = np.random.default_rng(n * 7 + nz)
    theta = synth.angles(nproj, start)
    data = _smooth_sinograms(rng, nz, nproj, n)
    ref = fr.backprojection(data, theta)
    rec = FourierRec(n, nproj, nz, _cuda(theta), 'float32')
    d = _cuda(data)
    d0 = d.clone()
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    rec.backprojection(obj, d)
    torch.cuda.synchronize()
    err = _rel_l2_per_slice(obj.cpu().numpy(), ref)
    print('rel-L2 per slice', err)
    assert np.all(err <= RTOL32), err
    assert torch.equal(d, d0), 'input data was modified'
    rec.free()


def test_backprojection_deterministic():
    from tomocupy_amd import FourierRec
    n, nproj, nz = 128, 120, 8
    rng = np.random.default_rng(5)
    theta = synth.angles(nproj)
    d = _cuda(_smooth_sinograms(rng, nz, nproj, n))
    rec = FourierRec(n, nproj, nz, _cuda(theta))
    a = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    b = torch.zeros_like(a)
    rec.backprojection(a, d)
    rec.backprojection(b, d)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("n,nproj,nz", [(64, 40, 4), (96, 30, 3)])
def test_forward_is_adjoint(n, nproj, nz):
    from tomocupy_amd import FourierRec
    rng = np.random.default_rng(11)
    theta = synth.angles(nproj, 0.1)
    rec = FourierRec(n, nproj, nz, _cuda(theta))
    x = _cuda(rng.standard_normal((nz, n, n)))
    y = _cuda(rng.standard_normal((nz, nproj, n)))
    Ax = torch.zeros((nz, nproj, n), dtype=torch.float32, device='cuda')
    Aty = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    rec.forward(Ax, x)
    rec.backprojection(Aty, y)
    torch.cuda.synchronize()
    lhs = float((Ax.double() * y.double()).sum())
    rhs = float((x.double() * Aty.double()).sum())
    rel = abs(lhs - rhs) / (float(Ax.double().norm()) * float(y.double().norm()))
    print('dot test', lhs, rhs, rel)
    assert rel < 1e-5
    # and against the FP64 oracle adjoint
    ref = fr.forward(x.cpu().numpy().astype(np.float64), theta)
    err = _rel_l2_per_slice(Ax.cpu().numpy(), ref)
    assert np.all(err <= RTOL32), err


@pytest.mark.parametrize("name,center,sht", [('parzen', 64.0, None), ('ramp', 63.5, [0.0, 1.5, -2.25]),
                                             ('shepp', 70.25, None), ('hann', 64.0, [0.5, 0.0, 3.0])])
def test_filter_center_matches_oracle(name, center, sht):
    from tomocupy_amd import FBPFilter
    n, nproj, nz = 128, 37, 3
    ne = 4 * n
    rng = np.random.default_rng(3)
    data = rng.standard_normal((nz, nproj, n)) + 2.0
    w64 = ff.calc_filter(ne, name)
    shv = np.zeros(nz) if sht is None else np.array(sht)
    ref = ff.fbp_filter_center(data, w64, ne, center, shv)
    flt = FBPFilter(ne, nproj, nz, 'float32', n_unpadded=n)
    wf = flt.calc_filter(name)
    d = _cuda(data)
    flt.filter_center(d, wf, center, None if sht is None else _cuda(shv))
    torch.cuda.synchronize()
    err = _rel_l2_per_slice(d.cpu().numpy(), ref)
    print('filter rel-L2', err)
    assert np.all(err <= RTOL32), err


def test_filter_padded_matches_oracle():
    from tomocupy_amd import FBPFilter
    ne, nproj, nz = 256, 9, 2
    rng = np.random.default_rng(4)
    tmp = rng.standard_normal((nz, nproj, ne))
    w = rng.standard_normal((nz, ne // 2 + 1)) + 1j * rng.standard_normal((nz, ne // 2 + 1))
    ref = ff.filter_padded(tmp, w)
    flt = FBPFilter(ne, nproj, nz, 'float32')
    d = _cuda(tmp)
    flt.filter(d, torch.as_tensor(w.astype(np.complex64)).cuda())
    torch.cuda.synchronize()
    err = _rel_l2_per_slice(d.cpu().numpy(), ref)
    assert np.all(err <= RTOL32), err


@pytest.mark.parametrize("flat_linear,bright", [(False, 1.0), (True, 1.0), (False, 10.0)])
def test_preprocess_matches_oracle(flat_linear, bright):
    from tomocupy_amd import ProcFunctions
    n, nproj, nz = 96, 50, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(1, n, theta, nz, ndark=4, nflat=6)
    ref = pp.proc_chain(data, dark, flat, bright_ratio=bright, flat_linear=flat_linear)
    pf = ProcFunctions(nproj, nz, n, 4, 6, bright_ratio=bright, flat_linear=flat_linear)
    out = pf.proc_chunk(_cuda(data, torch.uint16), _cuda(dark, torch.uint16), _cuda(flat, torch.uint16))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    err = _rel_l2_per_slice(got, ref)
    print('preprocess rel-L2', err)
    assert np.all(err <= RTOL32), err
    # dark/flat alone, projection order
    res = pf.darkflat_correction(_cuda(data, torch.uint16), _cuda(dark, torch.uint16),
                                 _cuda(flat, torch.uint16))
    ref2 = pp.darkflat_correction(data, dark, flat, bright, flat_linear)
    torch.cuda.synchronize()
    assert np.allclose(res.cpu().numpy(), ref2, rtol=1e-5, atol=1e-6)


def test_minus_log_rules_on_gpu():
    from tomocupy_amd import ProcFunctions
    pf = ProcFunctions(1, 1, 6, 1, 1)
    x = _cuda(np.array([[[-1.0, 0.0, 1.0, np.e, np.nan, np.inf]]]))
    out = pf.proc_proj_swap(x)
    torch.cuda.synchronize()
    assert np.allclose(out.cpu().numpy().ravel(), [0, 0, 0, -1, 6.0, 0.0], atol=1e-6)


def test_config1_full_chain():
    """Config-1 shape: n=512, nproj=720, nz=16, Parzen, dark/flat + -log, centre n/2."""
    from tomocupy_amd import BackprojFunctions, ProcFunctions
    n, nproj, nz = 512, 720, 16
    print('config-1 chain', flush=True)
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(1, n, theta, nz)
    sino = pp.proc_chain(data, dark, flat)
    ne = 4 * n
    filt = ff.fbp_filter_center(sino, ff.calc_filter(ne, 'parzen'), ne, n / 2)
    ref = fr.backprojection(filt, theta)
    pf = ProcFunctions(nproj, nz, n, dark.shape[0], flat.shape[0])
    bf = BackprojFunctions(n, nproj, nz, _cuda(theta), n / 2, 'float32', 'parzen')
    out = pf.proc_chunk(_cuda(data, torch.uint16), _cuda(dark, torch.uint16), _cuda(flat, torch.uint16))
    bf.fbp_filter_center(out)
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    bf.cl_rec.backprojection(obj, out)
    torch.cuda.synchronize()
    err = _rel_l2_per_slice(obj.cpu().numpy(), ref)
    print('config-1 rel-L2 max %.3g median %.3g' % (err.max(), np.median(err)))
    assert np.all(err <= RTOL32), err


@pytest.mark.parametrize("wname,level,sigma,shape", [('sym16', 7, 1.0, (180, 3, 256)),
                                                     ('db5', 4, 2.0, (97, 2, 130)),
                                                     ('haar', 3, 1.0, (64, 2, 64)),
                                                     ('sym5', 5, 1.5, (120, 1, 100)),
                                                     # 27 and 52 kept notch frequencies
                                                     ('db5', 5, 4.0, (300, 2, 128)),
                                                     ('sym16', 3, 8.0, (400, 2, 256))])
def test_remove_stripe_fw_matches_oracle(wname, level, sigma, shape):
    from tomocupy_amd import ProcFunctions
    nproj, nz, n = shape
    print('fw case', wname, level, sigma, shape, flush=True)
    rng = np.random.default_rng(level)
    yy = np.arange(nproj)[:, None, None]
    xx = np.arange(n)[None, None, :]
    data = 1 + 0.2 * np.sin(yy / 7.0) * np.cos(xx / 5.0) + 0.05 * rng.standard_normal((nproj, nz, n))
    data += 0.1 * (rng.random((1, nz, n)) < 0.05)  # column stripes
    ref = pp.remove_stripe_fw(data, sigma, wname, level)
    pf = ProcFunctions(nproj, nz, n, 1, 1, remove_stripe_method='fw', fw_sigma=sigma,
                       fw_filter=wname, fw_level=level)
    d = _cuda(data)
    pf.remove_stripe_fw(d)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    err = _rel_l2_per_slice(got.swapaxes(0, 1), ref.swapaxes(0, 1))
    print('fw rel-L2', err)
    assert np.all(err <= RTOL32), err


def test_config1_chain_with_fw():
    """Config-1 shape with Fourier-wavelet stripe removal (sym16, level 7,
    sigma 1), column stripes and the ramp filter at centre n/2 + 0.5, on a
    well-exposed phantom (attenuation x0.25, no photon starvation)."""
    from tomocupy_amd import BackprojFunctions, ProcFunctions
    n, nproj, nz = 512, 720, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(2, n, theta, nz, stripes=True, atten=0.25)
    sino = pp.proc_chain(data, dark, flat, stripe='fw')
    ne = 4 * n
    filt = ff.fbp_filter_center(sino, ff.calc_filter(ne, 'ramp'), ne, n / 2 + 0.5)
    ref = fr.backprojection(filt, theta)
    pf = ProcFunctions(nproj, nz, n, dark.shape[0], flat.shape[0], remove_stripe_method='fw')
    bf = BackprojFunctions(n, nproj, nz, _cuda(theta), n / 2 + 0.5, 'float32', 'ramp')
    out = pf.proc_chunk(_cuda(data, torch.uint16), _cuda(dark, torch.uint16), _cuda(flat, torch.uint16))
    bf.fbp_filter_center(out)
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    bf.cl_rec.backprojection(obj, out)
    torch.cuda.synchronize()
    err = _rel_l2_per_slice(obj.cpu().numpy(), ref)
    print('config-1+fw rel-L2 max %.3g median %.3g' % (err.max(), np.median(err)))
    assert np.all(err <= RTOL32), err


def test_starved_chain_with_fw_stagewise():
    """Photon-starved counts (the default attenuation): after fw, pixels with
    near-zero transmission make minus_log's ``x <= 0 -> 1`` clamp flip on
    rounding-level differences, so the chain is ill-conditioned end to end.
    Parity is checked stage by stage instead: the fw output before the log
    against the oracle, then -log + filter + backprojection of that same
    float32 input against the oracle."""
    from tomocupy_amd import BackprojFunctions, ProcFunctions
    n, nproj, nz = 512, 720, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(2, n, theta, nz, stripes=True)
    pf = ProcFunctions(nproj, nz, n, dark.shape[0], flat.shape[0], remove_stripe_method='fw')
    res = pf.proc_sino(_cuda(data, torch.uint16), _cuda(dark, torch.uint16), _cuda(flat, torch.uint16))
    torch.cuda.synchronize()
    res_np = res.cpu().numpy()
    ref_res = pp.remove_stripe_fw(pp.darkflat_correction(data, dark, flat), 1.0, 'sym16', 7)
    err = _rel_l2_per_slice(res_np.swapaxes(0, 1), ref_res.swapaxes(0, 1))
    print('starved fw pre-log rel-L2', err)
    assert np.all(err <= RTOL32), err
    out = pf.proc_proj_swap(res)
    bf = BackprojFunctions(n, nproj, nz, _cuda(theta), n / 2 + 0.5, 'float32', 'ramp')
    bf.fbp_filter_center(out)
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    bf.cl_rec.backprojection(obj, out)
    torch.cuda.synchronize()
    sino = np.ascontiguousarray(pp.minus_log(res_np.astype(np.float64)).swapaxes(0, 1))
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(sino, ff.calc_filter(ne, 'ramp'), ne, n / 2 + 0.5),
                            theta)
    err = _rel_l2_per_slice(obj.cpu().numpy(), ref)
    print('starved post-fw chain rel-L2', err)
    assert np.all(err <= RTOL32), err


def test_conveyor_recon_all_matches_oracle():
    """GPURec.recon_all (3 streams, pinned double buffers, writer threads) over
    3 chunks of 4 slices with a partial last chunk, against the oracle chain
    applied chunk by chunk (the pairing is within a chunk)."""
    from tomocupy_amd.rec import GPURec
    n, nproj, nz, ncz = 128, 90, 10, 4
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(3, n, theta, nz, ndark=3, nflat=5)
    rec = GPURec(n, nproj, nz, ncz, _cuda(theta), n / 2, 3, 5, fbp_filter='parzen',
                 max_write_threads=2)
    out = np.full((nz, n, n), np.nan, np.float32)

    def read(k):
        s = slice(k * ncz, min(nz, (k + 1) * ncz))
        return data[:, s], dark[:, s], flat[:, s]

    def write(k, r):
        out[k * ncz:k * ncz + len(r)] = r

    rec.recon_all(read, write)
    ne = 4 * n
    w = ff.calc_filter(ne, 'parzen')
    for k in range(3):
        s = slice(k * ncz, min(nz, (k + 1) * ncz))
        sino = pp.proc_chain(data[:, s], dark[:, s], flat[:, s])
        lz = sino.shape[0]
        sino = np.concatenate([sino, np.zeros((ncz - lz, nproj, n))])
        ref = fr.backprojection(ff.fbp_filter_center(sino, w, ne, n / 2), theta)[:lz]
        err = _rel_l2_per_slice(out[s], ref)
        print('chunk', k, err)
        assert np.all(err <= RTOL32), err


@pytest.mark.parametrize("n,nproj,nz,name,center,sht", [
    (128, 90, 4, 'parzen', 64.0, None), (96, 60, 3, 'ramp', 47.5, [0.0, 1.25, -2.0]),
    (256, 100, 2, 'shepp', 130.25, None), (64, 33, 5, 'hann', 30.0, [0.5, 0.0, 2.0, -1.0, 0.25])])
def test_fused_filter_backprojection_matches_oracle(n, nproj, nz, name, center, sht):
    from tomocupy_amd import BackprojFunctions
    print('fused case', n, nproj, nz, name, center, flush=True)
    rng = np.random.default_rng(n + nz)
    theta = synth.angles(nproj, 0.05)
    data = rng.standard_normal((nz, nproj, n)) + np.linspace(0, 2, n)
    shv = np.zeros(nz) if sht is None else np.array(sht, dtype=np.float64)
    ne = 4 * n
    ref = fr.backprojection(ff.fbp_filter_center(data, ff.calc_filter(ne, name), ne, center, shv),
                            theta)
    bf = BackprojFunctions(n, nproj, nz, _cuda(theta), center, 'float32', name)
    d = _cuda(data)
    d0 = d.clone()
    obj = torch.zeros((nz, n, n), dtype=torch.float32, device='cuda')
    bf.fbp_filter_backprojection(obj, d, 0 if sht is None else _cuda(shv))
    torch.cuda.synchronize()
    err = _rel_l2_per_slice(obj.cpu().numpy(), ref)
    print('fused rel-L2', err)
    assert np.all(err <= RTOL32), err
    assert torch.equal(d, d0)


@pytest.mark.parametrize("flat_linear,dtype", [(False, 'float32'), (True, 'float32'),
                                                (False, 'float16')])
def test_fused_fw_preprocessing_matches_unfused(flat_linear, dtype):
    """pre_process_fw (dark/flat in the first wavelet pass, -log + swap in the
    last) against darkflat -> remove_stripe_fw -> minus_log_swap."""
    from tomocupy_amd import ProcFunctions
    nproj, nz, n = 180, 4, 256
    theta = synth.angles(nproj)
    data, dark, flat, _ = synth.raw_counts(3, n, theta, nz, ndark=3, nflat=6, stripes=True,
                                           atten=0.25)
    g = [_cuda(a, torch.uint16) for a in (data, dark, flat)]
    pf = ProcFunctions(nproj, nz, n, dark.shape[0], flat.shape[0], dtype, flat_linear=flat_linear,
                       remove_stripe_method='fw', fw_level=5)
    fused = pf.proc_chunk(*g).float()
    res = pf.proc_sino(*g)
    unfused = pf.proc_proj_swap(res).float()
    torch.cuda.synchronize()
    rel = (torch.linalg.norm(fused - unfused) / torch.linalg.norm(unfused)).item()
    print('fused vs unfused rel-L2 %.3g' % rel)
    assert rel <= 1e-6
    ref = pp.proc_chain(data, dark, flat, stripe='fw', fw_level=5, flat_linear=flat_linear)
    err = _rel_l2_per_slice(fused.cpu().numpy(), ref)
    print('fused vs oracle', err)
    assert np.all(err <= (1e-5 if dtype == 'float32' else 1e-3)), err
