"""``tomocupy recon`` for the fourierrec path.

``python -m tomocupy_amd recon --file-name F [options]`` mirrors the
reference's command (``src/tomocupy/__main__.py``, flags and defaults of
``config.py``) for ``--reconstruction-algorithm fourierrec``:

* ``--reconstruction-type full``: every row chunk through the conveyor
  (:class:`tomocupy_amd.rec.GPURec`, ``rec.py:113-208``), one TIFF per slice
  ``<out>/recon_%05d.tiff`` (``writer.py:258-264``);
* ``recon_steps --reconstruction-type full|try``: the whole scan resident on the
  GPU through :class:`tomocupy_amd.rec_steps.GPURecSteps` (rec_steps.py:114-125)
  with phase retrieval (``--retrieve-phase-method paganin|Gpaganin|farago``,
  ``--energy`` keV, ``--propagation-distance`` mm, ``--pixel-size`` um) and
  projection rotation (``--rotate-proj-angle``, ``--rotate-proj-order`` 0 to 5);
* ``--reconstruction-type try``: the slices at ``--nsino`` (fractions of the
  height, ``reader.py:192-197``) reconstructed at the centres
  ``centre - arange(-w, w, step)`` (``reader.py:224-246``), each a per-slice
  shift of the same sinogram (``rec.py:210-275``), one TIFF per centre
  ``<out>/recon_slice%04d_center%05.2f.tiff`` (``writer.py:309-313``).

Stripe removal: ``fw``, ``ti`` and ``vo-all``; zinger removal: ``--dezinger``; 360-degree
scans with the axis near the border: ``--file-type double_fov``; binning
``--binning b`` (2^b x 2^b sums of counts on the GPU, ``reader.py:96-140``).
As the reference, the width is trimmed to a multiple of ``2^(b+1)``
(``reader.py:127-133``); ``--start-column/--end-column``,
``--start-proj/--end-proj`` and ``--blocked-views`` select the data
(``reader.py:124-181``); float16 crops the width to a power of two
(``reader.py:155-165``).  Settings outside this path (other algorithms,
laminography, automatic centre search, HDF5 output) are not offered.  Input
formats: see :mod:`tomocupy_amd.io`.
"""
import argparse
import ast
import os
import sys
import time

import numpy as np


def _bool(s):
    return str(s).lower() in ('true', '1', 'yes')


def parser():
    prog = os.path.basename(sys.argv[0]) if sys.argv and sys.argv[0] else 'tomocupy'
    if prog in ('__main__.py', '-c', '-m', ''):
        prog = 'python -m tomocupy_amd'
    ap = argparse.ArgumentParser(prog=prog, description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest='cmd', required=True)
    r = sub.add_parser('recon', help='reconstruct a DXchange-like data set')
    _recon_args(r)
    s = sub.add_parser('recon_steps', help='reconstruct by steps (whole scan resident; phase '
                                           'retrieval)')
    _recon_args(s)
    s.add_argument('--retrieve-phase-method', default='none',
                   choices=['none', 'paganin', 'Gpaganin', 'farago'])
    s.add_argument('--energy', type=float, default=20.0)
    s.add_argument('--propagation-distance', type=float, default=0.0)
    s.add_argument('--pixel-size', type=float, default=1.0)
    s.add_argument('--retrieve-phase-alpha', type=float, default=0.0)
    s.add_argument('--retrieve-phase-delta-beta', type=float, default=1500.0)
    s.add_argument('--retrieve-phase-W', type=float, default=2e-4)
    s.add_argument('--nproj-per-chunk', type=int, default=8)
    s.add_argument('--rotate-proj-angle', type=float, default=0.0)
    s.add_argument('--rotate-proj-order', type=int, default=1)
    return ap


def _recon_args(r):
    r.add_argument('--file-name', required=True)
    r.add_argument('--out-path-name', default=None)
    r.add_argument('--reconstruction-type', default='try', choices=['full', 'try'])
    r.add_argument('--reconstruction-algorithm', default='fourierrec', choices=['fourierrec'])
    r.add_argument('--file-type', default='standard', choices=['standard', 'double_fov'])
    r.add_argument('--rotation-axis', type=float, default=-1.0)
    r.add_argument('--center-search-width', type=float, default=50.0)
    r.add_argument('--center-search-step', type=float, default=0.5)
    r.add_argument('--nsino', default='0.5')
    r.add_argument('--nsino-per-chunk', type=int, default=8)
    r.add_argument('--start-row', type=int, default=0)
    r.add_argument('--end-row', type=int, default=-1)
    r.add_argument('--start-column', type=int, default=0)
    r.add_argument('--end-column', type=int, default=-1)
    r.add_argument('--start-proj', type=int, default=0)
    r.add_argument('--end-proj', type=int, default=-1)
    r.add_argument('--blocked-views', default='none')
    r.add_argument('--fbp-filter', default='parzen',
                   choices=['none', 'ramp', 'shepp', 'hann', 'hamming', 'parzen', 'cosine',
                            'cosine2'])
    r.add_argument('--dtype', default='float32', choices=['float32', 'float16', 'bfloat16'])
    r.add_argument('--remove-stripe-method', default='None',
                   choices=['None', 'none', 'fw', 'ti', 'vo-all'])
    r.add_argument('--fw-sigma', type=float, default=1.0)
    r.add_argument('--fw-filter', default='sym16', choices=['haar', 'db5', 'sym5', 'sym16'])
    r.add_argument('--fw-level', type=int, default=7)
    r.add_argument('--ti-beta', type=float, default=0.022)
    r.add_argument('--ti-mask', type=float, default=1.0)
    r.add_argument('--vo-all-snr', type=float, default=3.0)
    r.add_argument('--vo-all-la-size', type=int, default=61)
    r.add_argument('--vo-all-sm-size', type=int, default=21)
    r.add_argument('--vo-all-dim', type=int, default=1, choices=[1, 2])
    r.add_argument('--minus-log', default='True')
    r.add_argument('--flat-linear', default='False')
    r.add_argument('--bright-ratio', type=float, default=1.0)
    r.add_argument('--dezinger', type=int, default=0)
    r.add_argument('--dezinger-threshold', type=int, default=5000)
    r.add_argument('--binning', type=int, default=0, choices=[0, 1, 2, 3])
    r.add_argument('--max-write-threads', type=int, default=8)
    r.add_argument('--save-format', default='tiff', choices=['tiff', 'h5', 'h5nolinks', 'npy'])


def _out_dir(a, kind):
    if a.out_path_name:
        return a.out_path_name
    base = os.path.splitext(os.path.basename(a.file_name))[0]
    root = os.path.dirname(os.path.abspath(a.file_name)) + '_rec'
    return os.path.join(root, 'try_center', base) if kind == 'try' else \
        os.path.join(root, base + '_rec')


def _vo(a):
    return dict(vo_all_snr=a.vo_all_snr, vo_all_la_size=a.vo_all_la_size,
                vo_all_sm_size=a.vo_all_sm_size, vo_all_dim=a.vo_all_dim)


def _common(a, ds):
    """Row range, raw column range, binned width and centre, stripe method
    (reader.py:82-165): the raw width is trimmed to a multiple of 2^(b+1),
    then width, centre and height are divided by 2^b; for float16 the width
    is cropped to its largest power of two, centred, and the centre moves by
    (2^k - ni)/2 (reader.py:155-165; the crop is applied to the binned
    columns, which equals the reference's raw crop without binning)."""
    z0 = a.start_row
    z1 = ds.nz if a.end_row == -1 else a.end_row
    if not (0 <= z0 < z1 <= ds.nz):
        raise SystemExit('invalid --start-row/--end-row for %d rows' % ds.nz)
    b = a.binning
    centre = ds.ni / 2 if a.rotation_axis == -1 else a.rotation_axis
    x0 = a.start_column
    x1 = ds.ni if a.end_column == -1 else a.end_column
    if not (0 <= x0 < x1 <= ds.ni):
        raise SystemExit('invalid --start-column/--end-column for %d columns' % ds.ni)
    nsh = (x1 - x0) % (2 ** (b + 1))
    if nsh:
        x1 -= nsh
        print('decreasing projection width by %d pixel to operate with multiple of 2 sizes; '
              'new width %d' % (nsh, x1 - x0))
    if (z1 - z0) >> b < 1:
        raise SystemExit('fewer than 2^%d rows to bin' % b)
    stripe = a.remove_stripe_method if a.remove_stripe_method in ('fw', 'ti', 'vo-all') else 'none'
    ni, centre = (x1 - x0) >> b, (centre - x0) / 2 ** b
    if a.dtype == 'float16':
        p2 = 2 ** int(np.log2(ni))
        st = (ni - p2) // 2
        centre += (p2 - ni) / 2
        if p2 != ni:
            print('crop data to the power of 2 sizes to work with 16bit precision, output size '
                  'in x dimension %d' % p2)
        x0, x1, ni = x0 + (st << b), x0 + ((st + p2) << b), p2
    return z0, z1, (x0, x1), ni, centre, stripe


def _projections(a, ds):
    """Projection selection (reader.py:167-181): [start-proj, end-proj), then
    the angles outside every blocked range [st, end] (modulo pi).  Returns
    (index array into the data, angles)."""
    p1 = ds.nproj if a.end_proj == -1 else a.end_proj
    if not (0 <= a.start_proj < p1 <= ds.nproj):
        raise SystemExit('invalid --start-proj/--end-proj for %d projections' % ds.nproj)
    theta = ds.theta[a.start_proj:p1]
    ids = np.arange(len(theta))
    if a.blocked_views != 'none':
        tmp = ast.literal_eval(a.blocked_views)
        if not isinstance(tmp[0], (list, tuple)):
            tmp = [tmp]
        for st, end in tmp:
            keep = np.where(((theta % np.pi) < st) | (((theta - st) % np.pi) > end - st))[0]
            ids = np.intersect1d(ids, keep)
    return np.arange(a.start_proj, p1)[ids], np.ascontiguousarray(theta[ids])


def recon_full(a, ds):
    from .io import write_tiff
    from .rec import GPURec
    z0, z1, cols, ni, centre, stripe = _common(a, ds)
    pids, theta = _projections(a, ds)
    b = a.binning
    nz = (z1 - z0) >> b
    out = _out_dir(a, 'full')
    os.makedirs(out, exist_ok=True)
    rec = GPURec(ni, len(pids), nz, a.nsino_per_chunk, theta, centre,
                 ds.dark.shape[0], ds.flat.shape[0], dtype=a.dtype, fbp_filter=a.fbp_filter,
                 remove_stripe_method=stripe, fw_sigma=a.fw_sigma, fw_filter=a.fw_filter,
                 fw_level=a.fw_level, minus_log=_bool(a.minus_log), bright_ratio=a.bright_ratio,
                 flat_linear=_bool(a.flat_linear), max_write_threads=a.max_write_threads,
                 dezinger=a.dezinger, dezinger_threshold=a.dezinger_threshold,
                 ti_beta=a.ti_beta, ti_mask=a.ti_mask, file_type=a.file_type, binning=b,
                 **_vo(a))
    ncz, nrec = rec.ncz, rec.n
    zb = z0 >> b  # first output slice id (rec.py:170)
    odt = np.float16 if a.dtype == 'float16' else np.float32
    stack = np.lib.format.open_memmap(os.path.join(out, 'recon.npy'), mode='w+', dtype=odt,
                                      shape=(nz, nrec, nrec)) if a.save_format == 'npy' else None
    h5w = None
    if a.save_format in ('h5', 'h5nolinks'):
        from .io import H5RecWriter
        # the whole height of the file, chunks at their absolute rows
        # (writer.py:130-141, 172-175)
        rows = [(zb + k * ncz, rec.lzchunk[k]) for k in range(rec.nzchunk)]
        h5w = H5RecWriter(out, a.save_format, ds.nz >> b, nrec, odt, rows,
                          ' '.join([os.path.basename(sys.argv[0])] + sys.argv[1:]))

    def read(k):
        s0 = z0 + ((k * ncz) << b)
        return ds.rows(s0, s0 + (rec.lzchunk[k] << b), cols, pids)

    def write(k, r):
        if h5w is not None:
            h5w.write_chunk(k, zb + k * ncz, r)
            return
        for kk in range(len(r)):
            fid = zb + k * ncz + kk
            if stack is not None:
                stack[fid - zb] = r[kk]
            else:
                write_tiff(os.path.join(out, 'recon_%05d.tiff' % fid), r[kk])

    t = time.perf_counter()
    rec.recon_all(read, write)
    rec.free()
    if stack is not None:
        stack.flush()
    if h5w is not None:
        h5w.close()
        out = h5w.fname
    dt = time.perf_counter() - t
    print('reconstructed %d slices of %dx%d in %.2f s (%.1f slices/s) -> %s'
          % (nz, nrec, nrec, dt, nz / dt, out))
    return out


def recon_try(a, ds, steps=False):
    """Try mode; ``steps`` (``recon_steps --reconstruction-type try``,
    rec_steps.py:114-125 then backproj_parallel.py:360-413): the whole scan
    is processed first (phase retrieval mixes rows), then the slices are
    taken from the processed volume."""
    import torch
    from . import BackprojFunctions, ProcFunctions, _lib
    from .io import write_tiff
    z0, z1, cols, ni, centre, stripe = _common(a, ds)
    b = a.binning
    nz = (z1 - z0) >> b
    out = _out_dir(a, 'try')
    os.makedirs(out, exist_ok=True)
    fractions = ast.literal_eval(a.nsino)
    if not isinstance(fractions, (list, tuple)):
        fractions = [fractions]
    # raw row of each slice, a multiple of 2^b (reader.py:196-197)
    ids = [(int(f * (nz * 2 ** b - 1) / 2 ** b)) << b for f in fractions]
    # shifts in binned pixels, centres saved in raw pixels (reader.py:229-232)
    shifts = (np.arange(-a.center_search_width, a.center_search_width,
                        a.center_search_step * 2 ** b).astype(np.float32) / 2 ** b)
    # saved centres (reader.py:231-232): (centre before the float16 crop -
    # shift) 2^b + st_n, st_n the start column, or for float16 the crop start
    # (binned; the reference replaces st_n with it, reader.py:157)
    st16 = (cols[0] - a.start_column) >> b if a.dtype == 'float16' else 0
    st_n = st16 if a.dtype == 'float16' else a.start_column
    centres = (centre + st16 - shifts) * 2 ** b + st_n
    ncz = max(2, a.nsino_per_chunk)
    td = _lib.torch_dtype(a.dtype)
    pids, th = _projections(a, ds)
    theta = torch.from_numpy(th).cuda()
    pf = ProcFunctions(len(pids), 1, ni, ds.dark.shape[0], ds.flat.shape[0], a.dtype,
                       minus_log=_bool(a.minus_log), bright_ratio=a.bright_ratio,
                       flat_linear=_bool(a.flat_linear), remove_stripe_method=stripe,
                       fw_sigma=a.fw_sigma, fw_filter=a.fw_filter, fw_level=a.fw_level,
                       dezinger=a.dezinger, dezinger_threshold=a.dezinger_threshold,
                       ti_beta=a.ti_beta, ti_mask=a.ti_mask, file_type=a.file_type,
                       center=centre, binning=b, **_vo(a))
    nrec, rcentre, sgn = ni, centre, 1.0
    if a.file_type == 'double_fov':
        # reader.py:143-148 and 233-234: the axis moves to the right half of a
        # 2 ni wide sinogram; mirrored data take mirrored shifts
        nrec = 2 * ni
        if centre < ni // 2:
            rcentre, sgn = ni - centre, -1.0
    bf = BackprojFunctions(nrec, len(pids), ncz, theta, rcentre, a.dtype, a.fbp_filter)
    written = []
    if steps:
        rs = _steps(a, ds, ni, nz, centre, stripe, pids, th)
        vol = rs.proc_volume(*(torch.from_numpy(x).cuda()
                               for x in ds.rows(z0, z0 + (nz << b), cols, pids)))
    for idz in ids:
        if steps:
            sino = rs.sino_row(vol, idz >> b)
        else:
            data, dark, flat = (torch.from_numpy(x).cuda()
                                for x in ds.rows(z0 + idz, z0 + idz + 2 ** b, cols, pids))
            sino = pf.proc_chunk(data, dark, flat)  # (1, nproj, nrec)
        batch = sino.expand(ncz, -1, -1).contiguous()
        obj = torch.empty((ncz, nrec, nrec), dtype=td, device='cuda')
        for c0 in range(0, len(shifts), ncz):
            lz = min(ncz, len(shifts) - c0)
            sht = torch.zeros(ncz, dtype=torch.float32, device='cuda')
            sht[:lz] = torch.from_numpy(sgn * shifts[c0:c0 + lz]).cuda()
            bf.fbp_filter_backprojection(obj, batch, sht)
            host = obj[:lz].float().cpu().numpy()
            for kk in range(lz):
                img = host[kk].astype(np.float16) if a.dtype == 'float16' else host[kk]
                path = os.path.join(out, 'recon_slice%04d_center%05.2f.tiff'
                                    % (z0 + idz, centres[c0 + kk]))
                write_tiff(path, img)
                written.append(path)
    pf.free()
    bf.free()
    if steps:
        rs.free()
    print('wrote %d try images (%d slices x %d centres) -> %s'
          % (len(written), len(ids), len(shifts), out))
    return written


def _steps(a, ds, ni, nz, centre, stripe, pids, theta):
    import torch
    from .rec_steps import GPURecSteps
    b = a.binning
    return GPURecSteps(ni, len(pids), nz, a.nsino_per_chunk, torch.from_numpy(theta).cuda(),
                     centre, ds.dark.shape[0], ds.flat.shape[0], dtype=a.dtype,
                     fbp_filter=a.fbp_filter, nproj_per_chunk=a.nproj_per_chunk,
                     retrieve_phase_method=a.retrieve_phase_method, energy=a.energy,
                     propagation_distance=a.propagation_distance, pixel_size=a.pixel_size,
                     retrieve_phase_alpha=a.retrieve_phase_alpha,
                     retrieve_phase_delta_beta=a.retrieve_phase_delta_beta,
                     retrieve_phase_W=a.retrieve_phase_W, file_type=a.file_type,
                     rotate_proj_angle=a.rotate_proj_angle, rotate_proj_order=a.rotate_proj_order,
                     remove_stripe_method=stripe, fw_sigma=a.fw_sigma, fw_filter=a.fw_filter,
                     fw_level=a.fw_level, minus_log=_bool(a.minus_log),
                     bright_ratio=a.bright_ratio, flat_linear=_bool(a.flat_linear),
                     dezinger=a.dezinger, dezinger_threshold=a.dezinger_threshold,
                     ti_beta=a.ti_beta, ti_mask=a.ti_mask, binning=b, **_vo(a))


def recon_steps_full(a, ds):
    """``tomocupy recon_steps --reconstruction-type full`` (rec_steps.py:114-125):
    the whole scan through GPURecSteps, one TIFF per slice."""
    from .io import write_tiff
    z0, z1, cols, ni, centre, stripe = _common(a, ds)
    b = a.binning
    nz = (z1 - z0) >> b
    out = _out_dir(a, 'full')
    os.makedirs(out, exist_ok=True)
    pids, theta = _projections(a, ds)
    rs = _steps(a, ds, ni, nz, centre, stripe, pids, theta)
    zb, ncz = z0 >> b, rs.ncz
    data, dark, flat = ds.rows(z0, z0 + (nz << b), cols, pids)
    t = time.perf_counter()

    def write(k, r):
        for kk in range(len(r)):
            write_tiff(os.path.join(out, 'recon_%05d.tiff' % (zb + k * ncz + kk)), r[kk])

    rs.recon_steps_all(data, dark, flat, write)
    rs.free()
    dt = time.perf_counter() - t
    print('reconstructed %d slices of %dx%d by steps in %.2f s -> %s'
          % (nz, rs.n, rs.n, dt, out))
    return out


# Largest --fw-sigma: the notch keeps the frequencies whose damping
# exp(-k^2/(2 sigma^2)) is >= 1e-9, at most 64 per level (about 6.44 sigma + 1).
FW_SIGMA_MAX = 9.9


def main(argv=None):
    ap = parser()
    a = ap.parse_args(argv)
    if getattr(a, 'remove_stripe_method', None) == 'fw' and not 0 < a.fw_sigma <= FW_SIGMA_MAX:
        ap.error('--fw-sigma must be in (0, %g] (the notch keeps at most 64 frequencies per '
                 'wavelet level)' % FW_SIGMA_MAX)
    from .io import Dataset
    ds = Dataset(a.file_name)
    if a.cmd == 'recon_steps':
        if a.reconstruction_type == 'full':
            recon_steps_full(a, ds)
        else:
            recon_try(a, ds, steps=True)
        return 0
    if a.reconstruction_type == 'full':
        recon_full(a, ds)
    else:
        recon_try(a, ds)
    return 0


if __name__ == '__main__':
    sys.exit(main())
