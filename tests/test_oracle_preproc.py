"""Pins the FP64 preprocessing oracle (dark/flat, -log, fw stripe removal)."""
import os

import numpy as np
import pytest

from oracle import preproc_fp64 as pp

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'dwt2_sym16.npz')


def test_darkflat_closed_form():
    data = np.array([[[1000, 2100]]], np.uint16)
    dark = np.array([[[100, 100]], [[110, 90]]], np.uint16)
    flat = np.array([[[2000, 3000]], [[2200, 3000]]], np.uint16)
    r = pp.darkflat_correction(data, dark, flat)
    f0 = np.array([2100, 3000]) * np.float64(np.float32(1 + 1e-5)) - np.array([105, 95])
    assert np.allclose(r[0, 0], (np.array([1000, 2100]) - np.array([105, 95])) / f0)
    r2 = pp.darkflat_correction(data, dark, flat, bright_ratio=2.0)
    f0 = np.array([1050, 1500]) * np.float64(np.float32(1 + 1e-5)) - np.array([105, 95])
    assert np.allclose(r2[0, 0], (np.array([1000, 2100]) - np.array([105, 95])) / f0)


def test_flat_linear_interpolates():
    nproj = 5
    data = np.full((nproj, 1, 1), 500.0)
    dark = np.zeros((1, 1, 1))
    flat = np.array([100.0, 100.0, 999.0, 300.0, 300.0]).reshape(5, 1, 1)
    r = pp.darkflat_correction(data, dark, flat, flat_linear=True)
    f = (100 + (300 - 100) * np.linspace(0, 1, nproj)) * np.float64(np.float32(1 + 1e-5))
    assert np.allclose(r[:, 0, 0], 500 / f)


def test_minus_log_rules():
    x = np.array([-1.0, 0.0, 1.0, np.e, np.nan, np.inf])
    r = pp.minus_log(x)
    assert np.allclose(r, [0, 0, 0, -1, 6.0, 0.0])


@pytest.mark.parametrize("i", [0, 1, 2])
def test_dwt2_matches_pywt(i):
    g = np.load(GOLD)
    ll, (lh, hl, hh) = pp.dwt2(g['x%d' % i], 'sym16')
    for a, k in ((ll, 'cA'), (lh, 'cH'), (hl, 'cV'), (hh, 'cD')):
        assert np.allclose(a, g['%s%d' % (k, i)], atol=1e-12)
    rec = pp.idwt2(g['in_cA%d' % i], (g['in_cH%d' % i], g['in_cV%d' % i], g['in_cD%d' % i]), 'sym16')
    r = g['rec%d' % i]
    assert np.allclose(rec[:r.shape[0], :r.shape[1]], r, atol=1e-12)


def test_coeff_len():
    g = np.load(GOLD)
    for N, L in g['coeff_len_sym16']:
        assert pp.dwt_coeff_len(N, 32) == L


@pytest.mark.parametrize("wname", ['haar', 'db5', 'sym5', 'sym16'])
def test_perfect_reconstruction(wname):
    x = np.random.default_rng(2).standard_normal((3, 70, 53))
    ll, bands = pp.dwt2(x, wname)
    y = pp.idwt2(ll, bands, wname)
    assert np.allclose(y[:, :70, :53], x, atol=1e-10)


def test_damping_closed_form():
    # even band height: (k+0.5) for k < my/2, and the Nyquist entry repeats my/2-0.5
    yh = np.array([0.5, 1.5, 2.5, 2.5])
    assert np.allclose(pp.fw_damping(6, 1.0), 1 - np.exp(-yh ** 2 / 2))
    assert np.allclose(pp.fw_damping(7, 2.0), 1 - np.exp(-np.arange(4) ** 2 / 8))


def test_fw_removes_constant_stripe():
    """A column-constant (angle-independent) additive stripe is strongly
    attenuated while a smooth field passes almost unchanged."""
    nproj, n = 96, 64
    yy, xx = np.mgrid[0:nproj, 0:n]
    smooth = 1 + 0.1 * np.sin(2 * np.pi * yy / nproj) * np.cos(2 * np.pi * xx / n)
    stripe = np.zeros(n)
    stripe[[10, 31, 50]] = 0.05
    data = (smooth + stripe[None, :])[:, None, :]
    out = pp.remove_stripe_fw(data, 1.0, 'sym16', 3)[:, 0]
    assert np.abs(out - smooth)[:, [10, 31, 50]].mean() < 0.3 * 0.05


def _brute_outliers(data, w, thr):
    """Pure-Python window median with explicit half-sample reflection, the
    window offsets -(w//2) .. w-1-w//2 and rank (w*w)//2 (tiny arrays only)."""
    rows, nz, n = data.shape

    def refl(m, N):
        P = 2 * N
        m %= P
        return P - 1 - m if m >= N else m
    out = data.copy()
    for p in range(rows):
        for z in range(nz):
            for x in range(n):
                win = sorted(data[refl(p + a - w // 2, rows), z, refl(x + b - w // 2, n)].item()
                             for a in range(w) for b in range(w))
                f = win[(w * w) // 2]
                v = data[p, z, x].item()
                if v > f and v - f > thr:
                    out[p, z, x] = f
    return out


@pytest.mark.parametrize("w", [2, 3, 4, 5])
def test_remove_outliers_matches_brute_force(w):
    rng = np.random.default_rng(40 + w)
    data = rng.integers(900, 1100, size=(7, 2, 9)).astype(np.uint16)
    data[rng.random(data.shape) < 0.1] = 60000  # zingers
    got = pp.remove_outliers(data, w, 500)
    ref = _brute_outliers(data, w, 500)
    assert np.array_equal(got, ref)
    assert (got == 60000).sum() < (data == 60000).sum()


def test_remove_outliers_rank_and_origin_matter():
    # a planted off-by-one in the rank or the window origin changes the result
    rng = np.random.default_rng(7)
    data = rng.integers(0, 65535, size=(6, 1, 8)).astype(np.uint16)
    ref = _brute_outliers(data, 4, 0)
    got = pp.remove_outliers(data, 4, 0)
    assert np.array_equal(got, ref)
    shifted = np.roll(_brute_outliers(np.roll(data, 1, axis=2), 4, 0), -1, axis=2)
    assert not np.array_equal(shifted, ref)


@pytest.mark.parametrize("w", [2, 3])
def test_remove_outliers_float_counts_match_brute_force(w):
    # binned frames are float sums of counts; same median and rule
    rng = np.random.default_rng(90 + w)
    data = rng.integers(3600, 4400, size=(6, 2, 8)).astype(np.float32)
    data[rng.random(data.shape) < 0.1] = 250000.0
    got = pp.remove_outliers(data, w, 2000)
    assert got.dtype == np.float32
    assert np.array_equal(got, _brute_outliers(data, w, 2000))


@pytest.mark.parametrize("b", [1, 2, 3])
def test_downsample_is_block_sum(b):
    # utils.downsample: pairwise column sums then pairwise row sums, b times
    # each = the sum over each 2^b x 2^b block (reshape form, brute force on
    # one block, total conserved)
    s = 2 ** b
    rng = np.random.default_rng(b)
    x = rng.integers(0, 65536, size=(3, 2 * s, 3 * s)).astype(np.uint16)
    got = pp.downsample(x, b)
    ref = x.astype(np.float64).reshape(3, 2, s, 3, s).sum(axis=(2, 4))
    assert got.shape == (3, 2, 3) and np.array_equal(got, ref)
    assert got[1, 1, 2] == sum(int(x[1, s + a, 2 * s + c]) for a in range(s) for c in range(s))
    assert got.sum() == x.astype(np.float64).sum()
    assert np.array_equal(pp.downsample(x, 0), x.astype(np.float64))


def test_remove_outliers_width_zero_is_identity():
    data = np.arange(60, dtype=np.uint16).reshape(3, 2, 10)
    assert np.array_equal(pp.remove_outliers(data, 0, 5000), data)


def _fw_correction_form(data, sigma, wname, level):
    """The order of operations of the GPU kernels: only the notch correction
    D_k of the stripe band goes through the synthesis bank,
    d_k = IDWT(d_{k+1}, 0, D_k, 0), and the output is x_0 + d_0."""
    nproj, nz, ni = data.shape
    npad = nproj + nproj // 8
    st = (npad - nproj) // 2
    x0 = np.zeros((nz, npad, ni))
    x0[:, st:st + nproj] = np.asarray(data, np.float64).swapaxes(0, 1)
    sli, D, shapes = x0, [], []
    for _ in range(level):
        sli, (lh, hl, hh) = pp.dwt2(sli, wname)
        f = np.fft.rfft(hl, axis=-2) * pp.fw_damping(hl.shape[-2], sigma)[:, None]
        D.append(np.fft.irfft(f, hl.shape[-2], axis=-2) - hl)
        shapes.append(hl.shape[-2:])
    d = None
    for k in range(level)[::-1]:
        h, w = shapes[k]
        dd = np.zeros((nz, h, w)) if d is None else d[..., :h, :w]
        zero = np.zeros_like(D[k])
        d = pp.idwt2(dd, (zero, D[k], zero), wname)
    out = x0 + d[..., :npad, :ni]
    return out[:, st:st + nproj].swapaxes(0, 1)


@pytest.mark.parametrize("wname,level,sigma,shape", [
    ('sym16', 3, 1.0, (90, 2, 64)), ('db5', 4, 2.0, (181, 1, 129)),
    ('haar', 5, 1.5, (60, 2, 47)), ('sym5', 2, 1.0, (100, 1, 100))])
def test_fw_correction_form_equals_reference_order(wname, level, sigma, shape):
    rng = np.random.default_rng(3)
    data = 1 + 0.3 * rng.random(shape) + 0.2 * (np.arange(shape[2]) % 7 == 0)
    a = pp.remove_stripe_fw(data, sigma, wname, level)
    b = _fw_correction_form(data, sigma, wname, level)
    assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(a))


def test_ti_equals_direct_circular_convolution():
    rng = np.random.default_rng(11)
    nproj, nz, ni, beta = 9, 2, 16, 0.05
    data = 1 + 0.1 * rng.random((nproj, nz, ni))
    got = pp.remove_stripe_ti(data, beta, 1.0)
    r = (1 - beta) / (1 + beta)
    gam = np.array([beta * r ** min(k, ni - k) for k in range(ni)])
    gam[0] -= 1
    v = data.mean(axis=0)
    v = v - v[:, :1]
    c = np.array([[sum(v[z, j] * gam[(x - j) % ni] for j in range(ni)) for x in range(ni)]
                  for z in range(nz)])
    assert np.allclose(got, data + c[None], rtol=0, atol=1e-13)


def test_ti_mask_and_constant_rows():
    rng = np.random.default_rng(12)
    data = 1 + 0.1 * rng.random((5, 1, 20))
    # a zero-width mask changes nothing; half the width changes only the centre
    assert np.array_equal(pp.remove_stripe_ti(data, 0.1, 0.0), data)
    half = pp.remove_stripe_ti(data, 0.1, 0.5)
    changed = np.any(half != data, axis=(0, 1))
    assert changed[5:15].all() and not changed[:5].any() and not changed[15:].any()
    # rows constant along the detector carry no stripe: unchanged
    flat = np.broadcast_to(rng.random((5, 1, 1)), (5, 1, 20)).copy()
    assert np.allclose(pp.remove_stripe_ti(flat, 0.1, 1.0), flat, atol=1e-15)


def test_pad360_taper_closed_forms():
    # the taper t^5 (126 - 420 t + 540 t^2 - 315 t^3 + 70 t^4) is the order-9
    # smoothstep: 1 at t = 1, 1/2 at t = 1/2, 0 at t = 0, monotone in between
    ni = 40
    ones = np.ones((1, 1, ni))
    out, center, flip = pp.pad360(ones, 30.0)       # w = 2 (40 - 30) = 20
    assert not flip and center == 30.0 and out.shape == (1, 1, 2 * ni)
    taper = out[0, 0, ni - 20:ni]
    assert taper[0] == 1.0 and abs(taper[10] - 0.5) < 1e-15
    assert np.all(np.diff(taper) < 0) and np.all(out[0, 0, :ni - 20] == 1.0)
    assert np.all(out[0, 0, ni:] == 0.0)


def test_pad360_flip_rule():
    ni = 32
    x = np.arange(ni, dtype=np.float64)[None, None, :]
    out, center, flip = pp.pad360(x, 5.0)            # axis in the left half
    assert flip and center == ni - 5.0
    w = int(2 * (ni - center))
    assert np.array_equal(out[0, 0, :ni - w], x[0, 0, ::-1][:ni - w])


def test_rotate_proj_conventions():
    # ndimage.rotate(..., axes=(2, 1), reshape=False): identity at 0, a
    # quarter turn of a square projection is np.rot90 (k = 1 in the (y, x)
    # plane), and at any angle the bilinear value at (y, x) is the one at
    # R (y, x) + c - R c with R = [[cos, sin], [-sin, cos]], c the centre,
    # neighbours clamped (brute force on a small projection)
    rng = np.random.default_rng(12)
    a = rng.uniform(0, 1, (2, 5, 7))
    assert np.allclose(pp.rotate_proj(a, 0.0), a, atol=1e-15)
    sq = rng.uniform(0, 1, (1, 6, 6))
    assert np.allclose(pp.rotate_proj(sq, 90.0), np.rot90(sq, 1, axes=(1, 2)), atol=1e-12)
    ang = 17.0
    c, s = np.cos(np.deg2rad(ang)), np.sin(np.deg2rad(ang))
    ny, nx = a.shape[1:]
    cy, cx = (ny - 1) / 2, (nx - 1) / 2
    got = pp.rotate_proj(a, ang, 1)
    for p in range(2):
        for y in range(ny):
            for x in range(nx):
                iy = c * y + s * x + cy - (c * cy + s * cx)
                ix = -s * y + c * x + cx - (-s * cy + c * cx)
                y0, x0 = int(np.floor(iy)), int(np.floor(ix))
                wy, wx = iy - y0, ix - x0

                def at(yy, xx):
                    return a[p, min(max(yy, 0), ny - 1), min(max(xx, 0), nx - 1)]
                v = ((1 - wy) * ((1 - wx) * at(y0, x0) + wx * at(y0, x0 + 1)) +
                     wy * ((1 - wx) * at(y0 + 1, x0) + wx * at(y0 + 1, x0 + 1)))
                assert abs(got[p, y, x] - v) < 1e-12, (p, y, x)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_rotate_proj_spline_reproduces_polynomials(order):
    # B-splines of order k reproduce polynomials of degree <= k: away from the
    # edges (where the 'nearest' extension bends the image), the rotated
    # value at (y, x) is the polynomial at R (y, x) + c - R c
    ny, nx = 48, 56
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)

    def f(yy, xx):
        return 0.3 + 0.02 * yy - 0.015 * xx + 1e-3 * yy * xx - 4e-4 * xx * xx

    ang = 11.0
    got = pp.rotate_proj(f(y, x)[None], ang, order)[0]
    c, s = np.cos(np.deg2rad(ang)), np.sin(np.deg2rad(ang))
    cy, cx = (ny - 1) / 2, (nx - 1) / 2
    iy = c * y + s * x + cy - (c * cy + s * cx)
    ix = -s * y + c * x + cx - (-s * cy + c * cx)
    inner = (iy > 14) & (iy < ny - 15) & (ix > 14) & (ix < nx - 15)
    assert inner.sum() > 200
    assert np.max(np.abs(got - f(iy, ix))[inner]) < 1e-6
    # and a spline interpolates: the identity at angle 0
    rng = np.random.default_rng(4)
    a = rng.uniform(0, 1, (1, 9, 11))
    assert np.allclose(pp.rotate_proj(a, 0.0, order), a, atol=1e-12)
