"""Pins the FP64 filter oracle (calc_filter/_wint, fbp_filter_center)."""
import numpy as np
import pytest

from oracle import filter_fp64 as ff


def test_wint_interior_closed_form():
    ne = 1024
    t = np.arange(0, ne / 2 + 1) / ne
    w = ff.wint(12, t)
    r = ne * w[1:] / t[1:]
    # away from the first bins and the last-40 linear tail the weights are t/ne
    assert np.allclose(r[30:-45], 1.0, atol=1e-8)
    # last 40 replaced by a linear ramp through w[-40]
    N = len(t)
    assert np.allclose(w[-40:], w[-40] / (N - 40) * np.arange(N - 40, N))


@pytest.mark.parametrize("name", ff.FILTERS)
def test_calc_filter_windows(name):
    ne = 512
    t = np.arange(0, ne / 2 + 1) / ne
    w = ff.calc_filter(ne, name)
    base = ff.wint(12, t) * ne * 0.5
    d = 0.5
    win = {'none': None, 'ramp': np.ones_like(t), 'shepp': np.sinc(t / (2 * d)),
           'cosine': np.cos(np.pi * t / (2 * d)), 'cosine2': np.cos(np.pi * t / (2 * d)) ** 2,
           'hamming': .54 + .46 * np.cos(np.pi * t / d), 'hann': (1 + np.cos(np.pi * t / d)) / 2,
           'parzen': (1 - t / d) ** 3}[name]
    if name == 'none':
        exp = np.full_like(t, ne * 0.5)
        exp[0] *= 2
    else:
        exp = base * win
    exp = 2 * exp * (exp >= 0)
    exp[0] *= 2
    assert np.allclose(w, exp / ne, rtol=1e-12, atol=1e-15)


def test_filter_integer_shift_is_roll():
    rng = np.random.default_rng(0)
    n, ne = 64, 256
    x = np.zeros((1, 3, n))
    x[0, :, 20:40] = rng.standard_normal((3, 20))
    w = ff.calc_filter(ne, 'none')
    a = ff.fbp_filter_center(x, w, ne, n / 2, sht=0.0)
    b = ff.fbp_filter_center(x, w, ne, n / 2, sht=3.0)
    # 'none' keeps w=1 (so the unnormalised C2R scales by ne) except the DC
    # bin, which calc_filter scales by 4
    dc = 3 * x.sum(axis=-1, keepdims=True)
    assert np.allclose(a, ne * x + dc, atol=1e-9)
    # a shift of +3 delays the row by 3 samples
    assert np.allclose(b[0, :, 10:50], np.roll(a, 3, axis=-1)[0, :, 10:50], atol=1e-9)


def test_filter_matches_spatial_convolution():
    """The padded R2C*w*C2R equals a circular convolution of the edge-padded row
    with the real impulse response irfft(w)."""
    rng = np.random.default_rng(1)
    n, ne = 32, 128
    x = rng.standard_normal((1, 2, n))
    w = ff.calc_filter(ne, 'parzen')
    out = ff.fbp_filter_center(x, w, ne, n / 2)
    h = np.fft.irfft(w, ne) * ne
    pad = ne // 2 - n // 2
    xp = np.concatenate([np.repeat(x[..., :1], pad, -1), x, np.repeat(x[..., -1:], ne - pad - n, -1)], -1)
    conv = np.array([[sum(h[(j - i) % ne] * xp[0, r, i] for i in range(ne)) for j in range(ne)]
                     for r in range(2)])
    assert np.allclose(out[0], conv[:, pad:pad + n], atol=1e-10)
