"""Pins of the FP64 phase-retrieval oracle (reference
``processing/retrieve_phase.py``)."""
import numpy as np
import pytest

from oracle import phase_fp64 as ph


def _brute(img, method, ps, dist, energy, alpha=0.0, db=1000.0, W=2e-4):
    """Explicit loops: edge padding, the 2-D DFT, the filter from its
    formulas at each shifted frequency, the inverse DFT, the crop."""
    dy, dz = img.shape
    wl = 2 * ph.PI * ph.PLANCK_CONSTANT * ph.SPEED_OF_LIGHT / energy
    pad_pix = np.ceil(ph.PI * wl * dist / ps ** 2)
    py = int((2 ** np.ceil(np.log2(dy + pad_pix)) - dy) / 2)
    pz = int((2 ** np.ceil(np.log2(dz + pad_pix)) - dz) / 2)
    ny, nx = dy + 2 * py, dz + 2 * pz
    prj = np.empty((ny, nx))
    for i in range(ny):
        for j in range(nx):
            prj[i, j] = img[min(max(i - py, 0), dy - 1), min(max(j - pz, 0), dz - 1)]

    def coord(k, n):  # reciprocal coordinate at fftshifted index k
        j = (k - n // 2) % n
        return (2 * j - (n - 1)) * 0.5 / ((n - 1) * ps)

    def factor(u, v):
        if method == 'paganin':
            return 1 / (wl * dist * (u * u + v * v) / (4 * ph.PI) + alpha)
        if method == 'farago':
            t = ph.PI * wl * dist * (u * u + v * v)
            return 1 / (np.cos(t) + db * np.sin(t))
        aph = db * dist * wl / (4 * ph.PI)
        kf = np.cos(u * 2 * ph.PI * ps) + np.cos(v * 2 * ph.PI * ps)
        return 1 / (1 - 2 * aph / W ** 2 * (kf - 2))
    H = np.array([[factor(coord(a, ny), coord(b, nx)) for b in range(nx)] for a in range(ny)])
    H /= H.max()
    yy, xx = np.arange(ny), np.arange(nx)
    Ey = np.exp(-2j * np.pi * np.outer(yy, yy) / ny)
    Ex = np.exp(-2j * np.pi * np.outer(xx, xx) / nx)
    F = Ey @ prj @ Ex.T
    G = F * H
    back = (np.conj(Ey) @ G @ np.conj(Ex).T) / (ny * nx)
    return np.real(back)[py:py + dy, pz:pz + dz]


@pytest.mark.parametrize("method", ['paganin', 'Gpaganin', 'farago'])
def test_retrieve_phase_matches_brute_force(method):
    rng = np.random.default_rng(1)
    img = rng.uniform(0.5, 1.0, (6, 10))
    kw = dict(pixel_size=1e-4, dist=4.0, energy=20.0, alpha=1e-3, db=50.0, W=2e-4)
    got = ph.retrieve_phase(img[None], method, **kw)[0]
    ref = _brute(img, method, kw['pixel_size'], kw['dist'], kw['energy'], kw['alpha'], kw['db'],
                 kw['W'])
    assert np.allclose(got, ref, rtol=0, atol=1e-12)
    assert not np.allclose(got, img, atol=1e-3)


def test_constant_projection_is_unchanged():
    # the normalised filter is 1 at the zero bin, the grid's smallest w2, when
    # the factor decreases with w2: always for paganin and Gpaganin, for
    # farago while pi lambda dist w2 stays below the first pole
    img = np.full((1, 8, 12), 0.7)
    for method, dist in (('paganin', 20.0), ('Gpaganin', 20.0), ('farago', 1.0)):
        got = ph.retrieve_phase(img, method, pixel_size=1e-4, dist=dist, energy=30.0, db=200.0)
        assert np.allclose(got, 0.7, atol=1e-13), method


def test_pad_width_closed_form():
    wl = ph.wavelength(20.0)
    pad_pix = np.ceil(ph.PI * wl * 50.0 / 1e-8)        # ceil(pi lambda dist / ps^2)
    assert ph.pad_width(2048, 1e-4, wl, 50.0) == int((2 ** np.ceil(np.log2(2048 + pad_pix))
                                                      - 2048) / 2)
    assert ph.pad_width(2048, 1e-4, wl, 50.0) == 1024    # 2048 + a few -> 4096
    assert ph.pad_width(6, 1e-4, wl, 1e-9) == 1          # pad_pix = 1: 6 -> 8


def test_large_alpha_is_identity():
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 1, (1, 6, 8))
    got = ph.retrieve_phase(img, 'paganin', pixel_size=1e-4, dist=1.0, energy=30.0, alpha=1e12)
    assert np.allclose(got, img, atol=1e-9)
