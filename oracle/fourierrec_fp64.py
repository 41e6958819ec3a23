"""FP64 restatement of tomocupy's USFFT backprojector (``cfunc_fourierrec``).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the checker, never the
thing measured or shipped.

Every step follows the reference kernel it names, in FP64 throughout: the
reference's float32 ``__expf``/``__cosf``/``__sinf`` rounding, float atomics
and cuFFT algorithm choice are deliberately not imitated.  Citations are to
``src/include/kernels_fourierrec.cuh`` (``K``), ``src/cuda/cfunc_fourierrec.cu``
(``C``) and ``src/tomocupy/reconstruction/fourierrec.py`` (``P``).

Parity status: pinned by the dot-product test against :func:`forward`, by a
brute-force non-uniform DFT at small n, by geometric special cases (zero,
detector-constant and angle-independent sinograms, an off-centre point source)
and by the swap symmetry of the pair packing (``tests/test_oracle_fourierrec.py``).
"""

import ctypes
import os

import numpy as np

EPS = 1e-3

_HERE = os.path.dirname(os.path.abspath(__file__))
_CLIB = None


def _clib():
    """The C gridding of ``grid_fp64.c`` (``libgrid_fp64.so``, built by
    ``oracle/Makefile``), or None when it is not built: the NumPy loops below
    then do the same FP64 arithmetic, more slowly."""
    global _CLIB
    if _CLIB is None:
        path = os.path.join(_HERE, 'libgrid_fp64.so')
        _CLIB = False
        if os.environ.get('ORACLE_NUMPY_GRID') != '1' and os.path.exists(path):
            lib = ctypes.CDLL(path)
            dp = ctypes.POINTER(ctypes.c_double)
            i64 = ctypes.c_int64
            for f in (lib.grid_scatter, lib.grid_interp):
                f.argtypes = [dp, dp, i64, i64, i64, ctypes.c_int, ctypes.c_double, dp]
                f.restype = None
            _CLIB = lib
    return _CLIB or None


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def params(n):
    """Oversampling constants of the ctor (C:6-8).

    ``mu = -ln(eps)/(2 n^2)``; ``m = ceil(2n/pi * sqrt(-mu ln eps + (mu n)^2/4))``,
    which is 4 for every n.  Returns ``(m, mu)``.
    """
    mu = -np.log(EPS) / (2.0 * n * n)
    m = int(np.ceil(2 * n / np.pi * np.sqrt(-mu * np.log(EPS) + (mu * n) ** 2 / 4)))
    return m, mu


def takexy(theta, n):
    """Polar sample coordinates (K:23-36).

    ``x[p,j] = (j - n//2)/n * cos(theta_p)``, ``y[p,j] = -(j - n//2)/n * sin(theta_p)``;
    values ``>= 0.5`` are clamped to ``0.5 - 1e-5`` (the lower side is not clamped).
    """
    th = np.asarray(theta, dtype=np.float64)
    j = (np.arange(n) - n // 2) / n
    x = j[None, :] * np.cos(th)[:, None]
    y = -j[None, :] * np.sin(th)[:, None]
    x = np.where(x >= 0.5, 0.5 - 1e-5, x)
    y = np.where(y >= 0.5, 0.5 - 1e-5, y)
    return x, y


def _shift_signs(n):
    """``1 - 2*((t+1) % 2)`` used by ifftshiftc/fftshiftc (K:121-132, K:149-159)."""
    t = np.arange(n)
    return (1 - 2 * ((t + 1) % 2)).astype(np.float64)


def pack(data):
    """Real-pair packing (P:63-64): ``g[c] = data[c] + 1j*data[c+nz/2]``.

    Odd nz (a documented extension; the reference requires even nz) pairs the
    last slice with an implicit zero slice.
    """
    nz = data.shape[0]
    nzc = (nz + 1) // 2
    g = np.zeros((nzc,) + data.shape[1:], dtype=np.complex128)
    g.real = data[:nzc]
    g.imag[: nz - nzc] = data[nzc:]
    return g


def unpack(f, nz):
    """Unpacking (P:66-68): ``obj[c] = Re f[c]``, ``obj[c+nz/2] = Im f[c]``."""
    nzc = f.shape[0]
    obj = np.empty((nz,) + f.shape[1:], dtype=np.float64)
    obj[:nzc] = f.real
    obj[nzc:] = f.imag[: nz - nzc]
    return obj


def gather_weights(x, y, n, m, mu):
    """Separable Gaussian weights of the gather kernel (K:76-98).

    Returns ``b0, b1`` (int, lower footprint corner in centred cell units) and
    ``e0[...,i0]``, ``e1[...,i1]`` so that the tap weight is ``c0*e0[i0]*e1[i1]``
    with ``c0 = pi/(mu*4n^2)``.
    """
    c1 = -np.pi ** 2 / mu
    b0 = np.floor(2 * n * x).astype(np.int64) - m
    b1 = np.floor(2 * n * y).astype(np.int64) - m
    i = np.arange(2 * m + 1)
    w0 = (b0[..., None] + i) / (2.0 * n) - x[..., None]
    w1 = (b1[..., None] + i) / (2.0 * n) - y[..., None]
    return b0, b1, np.exp(c1 * w0 * w0), np.exp(c1 * w1 * w1)


def backprojection(data, theta, chunk=64):
    """``FourierRec.backprojection`` (P:61-68 -> C:67-93), FP64.

    ``data``: filtered sinograms ``(nz, nproj, n)`` (real).  Returns ``obj``
    ``(nz, n, n)``.  ``data`` is not modified.
    """
    data = np.asarray(data, dtype=np.float64)
    nz, nproj, n = data.shape
    m, mu = params(n)
    S = 2 * n + 2 * m
    g = pack(data)
    nzc = g.shape[0]

    # ifftshiftc -> 1D forward FFT -> ifftshiftc (C:78-80)
    s = _shift_signs(n)
    G = s * np.fft.fft(s * g, axis=-1)

    # gather (K:60-119): scatter-add of each sample onto its 9x9 footprint
    fde = np.zeros((nzc, S * S), dtype=np.complex128)
    lib = _clib()
    if lib is not None:
        Gc = np.ascontiguousarray(G, dtype=np.complex128)
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64))
        lib.grid_scatter(_dptr(Gc.view(np.float64)), _dptr(th), nzc, nproj, n, m, mu,
                         _dptr(fde.view(np.float64)))
        chunk = nproj  # the loop below then has nothing left to do
        p_start = nproj
    else:
        x, y = takexy(theta, n)
        b0, b1, e0, e1 = gather_weights(x, y, n, m, mu)
        c0 = np.pi / (mu * (4.0 * n * n))
        v = G * (4.0 / n)
        p_start = 0
    for p0 in range(p_start, nproj, chunk):
        p1 = min(nproj, p0 + chunk)
        rows = (n + m + b1[p0:p1])[..., None, None] + np.arange(2 * m + 1)[:, None]
        cols = (n + m + b0[p0:p1])[..., None, None] + np.arange(2 * m + 1)[None, :]
        idx = (rows * S + cols).reshape(-1)
        w = (c0 * e1[p0:p1][..., :, None] * e0[p0:p1][..., None, :])
        for c in range(nzc):
            val = (w * v[c, p0:p1][..., None, None]).reshape(-1)
            fde[c] += np.bincount(idx, weights=val.real, minlength=S * S)
            fde[c] += 1j * np.bincount(idx, weights=val.imag, minlength=S * S)
    fde = fde.reshape(nzc, S, S)

    # wrap (K:38-58): fold the m-wide border into the periodic interior
    r = np.arange(S)
    border = (r < m) | (r >= 2 * n + m)
    bmask = border[:, None] | border[None, :]
    rr, qq = np.nonzero(bmask)
    r0 = (rr - m + 2 * n) % (2 * n) + m
    q0 = (qq - m + 2 * n) % (2 * n) + m
    for c in range(nzc):
        np.add.at(fde[c], (r0, q0), fde[c][rr, qq])

    # fftshiftc -> 2D inverse FFT on the interior (unnormalised) -> fftshiftc (C:86-89)
    sS = _shift_signs(S)
    sh = sS[:, None] * sS[None, :]
    fde *= sh
    inner = fde[:, m:m + 2 * n, m:m + 2 * n]
    fde[:, m:m + 2 * n, m:m + 2 * n] = np.fft.ifft2(inner, axes=(-2, -1)) * (2 * n) ** 2
    fde *= sh

    # divphi (K:4-21): crop with the "ty+1" row shift and deapodise
    t = np.arange(n)
    d = t / n - 0.5
    phi = np.exp(mu * n * n * (d[:, None] ** 2 + d[None, :] ** 2)) / nproj * (1 - n % 4)
    f = fde[:, n // 2 + m + 1:n // 2 + m + 1 + n, n // 2 + m:n // 2 + m + n] * phi

    # circ (K:161-174)
    f *= circ_mask(n)
    return unpack(f, nz)


def circ_mask(n):
    """``4x^2 + 4y^2 < 1`` with ``x = (tx - n//2)/n`` (K:161-174); rows are y."""
    t = (np.arange(n) - n // 2) / n
    return ((4 * t[None, :] ** 2 + 4 * t[:, None] ** 2) < 1).astype(np.float64)


def forward(obj, theta, nproj=None, chunk=64):
    """Exact real adjoint of :func:`backprojection` (USFFT forward projector).

    Not in the reference for standard tomography; it mirrors the structure of
    ``cfunc_usfft2d::fwd`` (``src/cuda/cfunc_usfft2d.cu:53-70``): every step of
    :func:`backprojection` is replaced by its adjoint, in reverse order.  With
    ``<u,v> = sum(u*v)`` over real arrays,
    ``<backprojection(d), o> == <d, forward(o)>`` to rounding.
    """
    obj = np.asarray(obj, dtype=np.float64)
    nz, n, _ = obj.shape
    theta = np.asarray(theta)
    nproj = len(theta)
    m, mu = params(n)
    S = 2 * n + 2 * m
    nzc = (nz + 1) // 2

    # unpack^T = pack
    f = np.zeros((nzc, n, n), dtype=np.complex128)
    f.real = obj[:nzc]
    f.imag[: nz - nzc] = obj[nzc:]
    # circ^T, divphi^T (zero-embed at the same offset)
    f *= circ_mask(n)
    t = np.arange(n)
    d = t / n - 0.5
    phi = np.exp(mu * n * n * (d[:, None] ** 2 + d[None, :] ** 2)) / nproj * (1 - n % 4)
    fde = np.zeros((nzc, S, S), dtype=np.complex128)
    fde[:, n // 2 + m + 1:n // 2 + m + 1 + n, n // 2 + m:n // 2 + m + n] = f * phi
    # fftshiftc^T = fftshiftc; (unnormalised inverse FFT)^T = unnormalised forward FFT
    sS = _shift_signs(S)
    sh = sS[:, None] * sS[None, :]
    fde *= sh
    inner = fde[:, m:m + 2 * n, m:m + 2 * n].copy()
    fde[:] = 0
    fde[:, m:m + 2 * n, m:m + 2 * n] = np.fft.fft2(inner, axes=(-2, -1))
    fde *= sh
    # wrap^T: every border cell takes the value of its periodic interior image
    r = np.arange(S)
    border = (r < m) | (r >= 2 * n + m)
    bmask = border[:, None] | border[None, :]
    rr, qq = np.nonzero(bmask)
    r0 = (rr - m + 2 * n) % (2 * n) + m
    q0 = (qq - m + 2 * n) % (2 * n) + m
    fde[:, rr, qq] = fde[:, r0, q0]
    fde = fde.reshape(nzc, S * S)
    # gather^T: interpolate the grid at the polar samples with the same weights
    G = np.zeros((nzc, nproj, n), dtype=np.complex128)
    lib = _clib()
    if lib is not None:
        Fc = np.ascontiguousarray(fde)
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64))
        lib.grid_interp(_dptr(Fc.view(np.float64)), _dptr(th), nzc, nproj, n, m, mu,
                        _dptr(G.view(np.float64)))
        p_start = nproj
    else:
        x, y = takexy(theta, n)
        b0, b1, e0, e1 = gather_weights(x, y, n, m, mu)
        c0 = np.pi / (mu * (4.0 * n * n))
        p_start = 0
    for p0 in range(p_start, nproj, chunk):
        p1 = min(nproj, p0 + chunk)
        rows = (n + m + b1[p0:p1])[..., None, None] + np.arange(2 * m + 1)[:, None]
        cols = (n + m + b0[p0:p1])[..., None, None] + np.arange(2 * m + 1)[None, :]
        idx = rows * S + cols
        w = c0 * e1[p0:p1][..., :, None] * e0[p0:p1][..., None, :]
        for c in range(nzc):
            G[c, p0:p1] = (w * fde[c][idx]).sum(axis=(-2, -1))
    G *= 4.0 / n
    # (ifftshiftc . FFT . ifftshiftc)^T = ifftshiftc . n*IFFT . ifftshiftc
    s = _shift_signs(n)
    g = s * (np.fft.ifft(s * G, axis=-1) * n)
    # pack^T = unpack
    out = np.empty((nz, nproj, n), dtype=np.float64)
    out[:nzc] = g.real
    out[nzc:] = g.imag[: nz - nzc]
    return out


def direct_nudft_backprojection(data, theta):
    """Brute-force counterpart of :func:`backprojection` for tiny sizes.

    Replaces gridding + oversampled FFT + deapodisation by the exact
    non-uniform sum the USFFT approximates:
    ``f[ty,tx] = sum_{p,k} (4/n) G[p,k]/nproj * exp(2 pi i (x_k X + y_k Y))`` with
    ``X = tx - n/2``, ``Y = ty + 1 - n/2`` and ``G`` the centred 1D DFT of the
    packed sinogram.  (The Gaussian kernel ``c0*exp(c1 t^2)`` summed over the
    2n x 2n grid has Fourier transform ``exp(-mu |X|^2)``, which ``divphi``
    cancels up to the one-row offset between the crop and the deapodisation
    rows, which is reproduced here.)  Agrees with the USFFT to its ~1e-3 accuracy (m=4, eps=1e-3):
    a check of the mathematics and geometry, not of the constants.
    """
    data = np.asarray(data, dtype=np.float64)
    nz, nproj, n = data.shape
    g = pack(data)
    s = _shift_signs(n)
    G = s * np.fft.fft(s * g, axis=-1) * (1 - n % 4) * (4.0 / n)
    x, y = takexy(theta, n)
    X = np.arange(n) - n // 2
    Y = np.arange(n) + 1 - n // 2
    ex = np.exp(2j * np.pi * x.reshape(-1)[:, None] * X[None, :])  # (P*n, n)
    ey = np.exp(2j * np.pi * y.reshape(-1)[:, None] * Y[None, :])
    f = np.einsum('cs,sy,sx->cyx', G.reshape(G.shape[0], -1), ey, ex) / nproj
    # divphi deapodises at the unshifted row ty while the crop reads row ty+1
    # (K:14-17), leaving a factor exp(mu*((ty-n/2)^2 - (ty+1-n/2)^2)) per row.
    m, mu = params(n)
    ty = np.arange(n)
    f *= np.exp(mu * ((ty - n / 2) ** 2 - (ty + 1 - n // 2) ** 2))[:, None]
    f *= circ_mask(n)
    return unpack(f, nz)
