"""Seeded synthetic tomography data: ellipse phantoms, analytic sinograms and
uint16 detector counts with dark/flat frames.

Geometry (the convention of the fourierrec backprojector): image pixel
``(row, col)`` has coordinates ``X = col - n/2``, ``Y = row - n/2``; the
projection at angle ``theta`` and detector bin ``j`` is the line integral over
``X cos(theta) - Y sin(theta) = j - n/2``.

Generator (all FP64, numpy ``default_rng`` seeds fixed per slice):

* phantom: 10 ellipses, semi-axes ``U(0.05,0.4)*n/2``, centres uniform in the
  disc of radius ``n/2``, orientation ``U(0, pi)``, attenuation
  ``U(0.1,1.0)*1e-2`` per pixel; seed ``1000*cfg + slice``;
* sinogram: analytic chord lengths ``2ab sqrt(h^2 - s'^2)/h^2``;
* counts (seed ``+10**6``): flat gain ``30000*U(0.98,1.02)`` per pixel, dark
  ``U(95,105)``; projections ``Poisson(gain*exp(-p) + dark)``, flats
  ``Poisson(gain + dark)``, darks ``Poisson(dark)``; optional column-gain
  stripes on the projections; clipped to uint16.
"""

import numpy as np


def ellipses(n, seed, count=10):
    """Ellipse parameters ``(cx, cy, a, b, phi, rho)`` as an ``(count, 6)`` array."""
    rng = np.random.default_rng(seed)
    r = n / 2
    a = rng.uniform(0.05, 0.4, count) * r
    b = rng.uniform(0.05, 0.4, count) * r
    rad = r * np.sqrt(rng.uniform(0, 1, count))
    ang = rng.uniform(0, 2 * np.pi, count)
    phi = rng.uniform(0, np.pi, count)
    rho = rng.uniform(0.1, 1.0, count) * 1e-2
    return np.stack([rad * np.cos(ang), rad * np.sin(ang), a, b, phi, rho], axis=1)


def sinogram(ell, theta, n):
    """Analytic projections ``(nproj, n)`` of the ellipses at detector bin centres."""
    theta = np.asarray(theta, dtype=np.float64)
    s = np.arange(n) - n / 2
    c, sn = np.cos(theta)[:, None], np.sin(theta)[:, None]
    out = np.zeros((len(theta), n))
    for cx, cy, a, b, phi, rho in ell:
        # projection direction normal n_theta = (cos, -sin) in (X, Y)
        nx, ny = c, -sn
        e1 = nx * np.cos(phi) + ny * np.sin(phi)
        e2 = -nx * np.sin(phi) + ny * np.cos(phi)
        h2 = (a * e1) ** 2 + (b * e2) ** 2
        s0 = cx * c - cy * sn
        d2 = h2 - (s[None, :] - s0) ** 2
        out += np.where(d2 > 0, 2 * a * b * np.sqrt(np.maximum(d2, 0)) / h2 * rho, 0.0)
    return out


def phantom_image(ell, n):
    """Rasterised phantom ``(n, n)`` (pixel centres), for visual/loose checks."""
    X = np.arange(n)[None, :] - n / 2
    Y = np.arange(n)[:, None] - n / 2
    img = np.zeros((n, n))
    for cx, cy, a, b, phi, rho in ell:
        u = (X - cx) * np.cos(phi) + (Y - cy) * np.sin(phi)
        v = -(X - cx) * np.sin(phi) + (Y - cy) * np.cos(phi)
        img += rho * ((u / a) ** 2 + (v / b) ** 2 <= 1)
    return img


def angles(nproj, start=0.0):
    """``theta_p = start + p*pi/nproj`` (endpoint excluded), float32."""
    return (start + np.arange(nproj) * np.pi / nproj).astype(np.float32)


def sinograms(cfg, n, theta, nz, z0=0, atten=1.0):
    """Analytic sinograms ``(nz, nproj, n)`` FP64 for slices ``z0..z0+nz-1``;
    ``atten`` scales every attenuation coefficient."""
    return atten * np.stack([sinogram(ellipses(n, 1000 * cfg + z), theta, n)
                             for z in range(z0, z0 + nz)])


def raw_counts(cfg, n, theta, nz, ndark=10, nflat=20, stripes=False, z0=0, atten=1.0):
    """Detector data in tomocupy's chunk layout.

    Returns ``data (nproj, nz, n)``, ``dark (ndark, nz, n)``, ``flat (nflat, nz, n)``
    as uint16, plus the analytic sinograms ``(nz, nproj, n)`` they were made from.
    """
    sino = sinograms(cfg, n, theta, nz, z0, atten)
    nproj = len(theta)
    data = np.empty((nproj, nz, n), np.uint16)
    dark = np.empty((ndark, nz, n), np.uint16)
    flat = np.empty((nflat, nz, n), np.uint16)
    for z in range(nz):
        rng = np.random.default_rng(1000 * cfg + z0 + z + 10 ** 6)
        gain = 30000 * rng.uniform(0.98, 1.02, n)
        dk = rng.uniform(95, 105, n)
        lam = gain[None, :] * np.exp(-sino[z]) + dk[None, :]
        if stripes:
            cols = rng.random(n) < 0.05
            sg = np.where(cols, rng.uniform(0.97, 1.03, n), 1.0)
            lam = lam * sg[None, :]
        data[:, z] = np.clip(rng.poisson(lam), 0, 65535)
        flat[:, z] = np.clip(rng.poisson(np.broadcast_to(gain + dk, (nflat, n))), 0, 65535)
        dark[:, z] = np.clip(rng.poisson(np.broadcast_to(dk, (ndark, n))), 0, 65535)
    return data, dark, flat, sino


BATCH_SIZES = (256, 512, 768, 1024, 1536, 2048)


def batch_requests(count=96, seed=2024, sizes=BATCH_SIZES):
    """The batching form's mixed request list: per request ``n`` drawn from
    ``sizes``, ``nproj = round(1.5 n)`` angles ``theta0 + p pi/nproj`` from a
    random start ``theta0`` in [0, pi), ``nz`` in [1, 32], centre ``n/2``.
    Returns dicts with ``n``, ``nz``, ``theta`` (float32) and ``center``."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        nproj = int(round(1.5 * n))
        th0 = rng.uniform(0, np.pi)
        theta = (th0 + np.arange(nproj) * np.pi / nproj).astype(np.float32)
        nz = int(rng.integers(1, 33))
        reqs.append({'n': n, 'nz': nz, 'theta': theta, 'center': n / 2})
    return reqs
