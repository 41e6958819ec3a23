"""FP64 restatement of tomocupy's FBP filter design and filtering.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Citations: ``src/tomocupy/reconstruction/fbp_filter.py`` (``F``),
``src/tomocupy/reconstruction/backproj_functions.py`` (``B``),
``src/cuda/cfunc_filter.cu`` (``C``), ``src/include/kernels_filter.cuh`` (``K``).

Parity status: ``wint`` is pinned by its interior closed form
(``ne*w_j/t_j == 1`` away from both ends) and the last-40 linear rule;
``calc_filter`` by the window closed forms and the ``wfa[0]`` doubling;
``fbp_filter_center`` by the integer-shift == roll identity, the near-identity
of the ``none`` filter and a direct spatial-domain convolution
(``tests/test_oracle_filter.py``).
"""

import numpy as np

FILTERS = ('none', 'ramp', 'shepp', 'cosine', 'cosine2', 'hamming', 'hann', 'parzen')


def wint(n, t):
    """High-order quadrature weights (F:90-118), FP64.

    ``n`` is the local polynomial order (12 in the reference), ``t`` the
    sample points ``j/ne``.  The last 40 weights are replaced by a linear ramp.
    """
    t = np.asarray(t, dtype=np.float64)
    N = len(t)
    s = np.linspace(1e-40, 1, n)
    tmp1 = np.arange(n)
    tmp2 = np.arange(1, n + 2)
    iv = np.linalg.inv(np.exp(np.outer(tmp1, np.log(s))))
    u = np.diff(np.exp(np.outer(tmp2, np.log(s))) * np.tile(1.0 / tmp2[..., None], [1, n]))
    W1 = iv @ u[1:n + 1, :]
    W2 = iv @ u[0:n, :]
    p = 1.0 / np.concatenate((np.arange(1, n), (n - 1) * np.ones(N - 2 * (n - 1) - 1),
                              np.arange(n - 1, 0, -1)))
    w = np.zeros(N)
    for j in range(N - n + 1):
        W = ((t[j + n - 1] - t[j]) ** 2) * W1 + (t[j + n - 1] - t[j]) * t[j] * W2
        for k in range(n - 1):
            w[j:j + n] += p[j + k] * W[:, k]
    w[-40:] = (w[-40]) / (N - 40) * np.arange(N - 40, N)
    return w


def calc_filter(ne, name):
    """``FBPFilter.calc_filter`` (F:60-88) in FP64 (the reference's final
    float32 cast is not applied).  Returns ``(ne//2+1,)`` real weights with the
    C2R ``1/ne`` folded in."""
    d = 0.5
    t = np.arange(0, ne / 2 + 1) / ne
    if name == 'none':
        wfa = ne * 0.5 + t * 0
        wfa[0] *= 2
    elif name == 'ramp':
        wfa = ne * 0.5 * wint(12, t)
    elif name == 'shepp':
        wfa = ne * 0.5 * wint(12, t) * np.sinc(t / (2 * d)) * (t / d <= 2)
    elif name == 'cosine':
        wfa = ne * 0.5 * wint(12, t) * np.cos(np.pi * t / (2 * d)) * (t / d <= 1)
    elif name == 'cosine2':
        wfa = ne * 0.5 * wint(12, t) * (np.cos(np.pi * t / (2 * d))) ** 2 * (t / d <= 1)
    elif name == 'hamming':
        wfa = ne * 0.5 * wint(12, t) * (.54 + .46 * np.cos(np.pi * t / d)) * (t / d <= 1)
    elif name == 'hann':
        wfa = ne * 0.5 * wint(12, t) * (1 + np.cos(np.pi * t / d)) / 2.0 * (t / d <= 1)
    elif name == 'parzen':
        wfa = ne * 0.5 * wint(12, t) * pow(1 - t / d, 3) * (t / d <= 1)
    else:
        raise ValueError(name)
    wfa = 2 * wfa * (wfa >= 0)
    wfa[0] *= 2
    return wfa / ne


def filter_weights(wfilter, ne, n, center, sht):
    """Per-slice complex filter ``w[z,k] = W_k exp(-2 pi i (-center + sht_z + n/2) t_k)``
    with ``t = rfftfreq(ne)`` (B:84, B:103-104)."""
    t = np.fft.rfftfreq(ne)
    sht = np.asarray(sht, dtype=np.float64).reshape(-1)
    return wfilter[None, :] * np.exp(-2j * np.pi * (-center + sht[:, None] + n / 2) * t[None, :])


def fbp_filter_center(data, wfilter, ne, center, sht=0.0):
    """``BackprojFunctions.fbp_filter_center`` (B:93-109) with
    ``cfunc_filter::filter`` (C:49-61, K:4-18), FP64.

    ``data``: ``(nz, nproj, n)``.  Edge-replicate pad to ``ne`` (B:97-100),
    R2C, multiply by ``w[z]``, unnormalised C2R (the Nyquist bin contributes its
    real part only, as ``numpy.fft.irfft``), crop.  Returns a new array.
    """
    data = np.asarray(data, dtype=np.float64)
    nz, nproj, n = data.shape
    pad = ne // 2 - n // 2
    tmp = np.empty((nz, nproj, ne))
    tmp[:, :, pad:pad + n] = data
    tmp[:, :, :pad] = data[:, :, :1]
    tmp[:, :, pad + n:] = data[:, :, -1:]
    if np.isscalar(sht):
        sht = np.full(nz, sht)
    w = filter_weights(wfilter, ne, n, center, sht)
    out = np.fft.irfft(np.fft.rfft(tmp, axis=-1) * w[:, None, :], n=ne, axis=-1) * ne
    return out[:, :, pad:pad + n]


def filter_padded(tmp, w):
    """``cfunc_filter::filter`` alone (C:49-61): in-place R2C, ``mulw``, C2R on
    an already padded ``(nz, nproj, ne)`` array with complex ``w (nz, ne//2+1)``."""
    ne = tmp.shape[-1]
    return np.fft.irfft(np.fft.rfft(tmp, axis=-1) * w[:, None, :], n=ne, axis=-1) * ne
