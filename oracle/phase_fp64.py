"""FP64 restatement of tomocupy's single-step phase retrieval.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``src/tomocupy/processing/retrieve_phase.py`` (``R``): ``paganin_filter``
(R:59-104, methods ``paganin`` and ``Gpaganin``), ``farago_filter``
(R:106-141), ``_retrieve_phase`` (R:143-161), ``_calc_pad`` (R:164-200),
the filter factors (R:203-216) and the reciprocal grids (R:219-304), with
NumPy FFTs in FP64.  The reference's edge replication writes
``prj[-px:] = prj[-px-1]``, which for a zero pad width would overwrite the
whole image; a zero width cannot occur for a propagation distance > 0 and
even sizes (the pad is at least one sample), so the padding here is plain
edge replication (``np.pad(mode='edge')``).

Parity status: pinned in ``tests/test_oracle_phase.py`` against a brute-force
DFT with explicit padding on tiny images, a constant image (unchanged: the
normalised filter is 1 at the zero bin), the large-alpha limit (identity) and
the pad-width closed form.
"""
import numpy as np

BOLTZMANN_CONSTANT = 1.3806488e-16  # R:46
SPEED_OF_LIGHT = 299792458e+2       # [cm/s] R:47
PI = 3.14159265359                  # R:48
PLANCK_CONSTANT = 6.58211928e-19    # [keV*s] R:49


def wavelength(energy):
    """R:52-53 [cm]"""
    return 2 * PI * PLANCK_CONSTANT * SPEED_OF_LIGHT / energy


def pad_width(dim, pixel_size, wl, dist):
    """``_calc_pad_width`` (R:214-216)."""
    pad_pix = np.ceil(PI * wl * dist / pixel_size ** 2)
    return int((pow(2, np.ceil(np.log2(dim + pad_pix))) - dim) * 0.5)


def reciprocal_coord(pixel_size, num_grid):
    """``_reciprocal_coord`` (R:281-304)."""
    n = num_grid - 1
    rc = np.arange(-n, num_grid, 2, dtype=np.float64)
    return rc * (0.5 / (n * pixel_size))


def filter_grid(method, shape, pixel_size, dist, energy, alpha=0.0, db=1000.0, W=2e-4):
    """The unshifted factor on the (rows, cols) reciprocal grid (R:83-95,
    131-133, 203-212, 219-278)."""
    ny, nx = shape
    wl = wavelength(energy)
    ry, rx = reciprocal_coord(pixel_size, ny), reciprocal_coord(pixel_size, nx)
    if method == 'Gpaganin':
        ky = np.cos(ry * 2 * PI * pixel_size)
        kx = np.cos(rx * 2 * PI * pixel_size)
        kf = ky[:, None] + kx[None, :]
        aph = db * (dist * wl) / (4 * PI)
        return 1 / (1.0 - (2 * aph / (W ** 2)) * (kf - 2))
    w2 = ry[:, None] ** 2 + rx[None, :] ** 2
    if method == 'paganin':
        return 1 / (wl * dist * w2 / (4 * PI) + alpha)
    if method == 'farago':
        return 1 / (np.cos(PI * wl * dist * w2) + db * np.sin(PI * wl * dist * w2))
    raise ValueError(method)


def retrieve_phase(data, method='paganin', pixel_size=1e-4, dist=50, energy=20, alpha=1e-3,
                   db=1000, W=2e-4):
    """``paganin_filter`` / ``farago_filter`` on ``data (nproj, dy, dz)``:
    returns the filtered projections (FP64)."""
    data = np.asarray(data, dtype=np.float64)
    _, dy, dz = data.shape
    wl = wavelength(energy)
    py = pad_width(dy, pixel_size, wl, dist)
    pz = pad_width(dz, pixel_size, wl, dist)
    h = np.fft.fftshift(filter_grid(method, (dy + 2 * py, dz + 2 * pz), pixel_size, dist,
                                    energy, alpha, db, W))
    h = h / h.max()
    out = np.empty_like(data)
    for m in range(data.shape[0]):
        prj = np.pad(data[m], ((py, py), (pz, pz)), mode='edge')
        proj = np.real(np.fft.ifft2(np.fft.fft2(prj) * h))
        out[m] = proj[py:dy + py, pz:dz + pz]
    return out
