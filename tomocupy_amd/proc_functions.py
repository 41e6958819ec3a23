"""Chunk preprocessing with the reference's interface.

Mirrors ``ProcFunctions`` (``src/tomocupy/processing/proc_functions.py:46-198``)
for the steps on the fourierrec path: ``remove_outliers`` (dezinger),
stripe removal (``fw``, ``ti``, ``vo-all``),
``darkflat_correction``, ``minus_log``, ``proc_sino`` (outliers, dark/flat,
stripe removal), ``proc_proj`` (minus log) and ``pad360`` (double field of
view: ``file_type='double_fov'`` with the rotation axis ``center`` in detector
columns doubles the sinogram width), and the reader's binning
(``binning=b``: ``downsample``, ``utils.py:157-167``, sums of 2^b x 2^b blocks
of counts as float32; ``n`` and ``ncz`` are then the binned sizes and the raw
frames are ``(rows, ncz 2^b, n 2^b)``).
Settings the reference reads from the global ``args`` are constructor
arguments.  ``proc_chunk`` is the fused form used by the conveyor: raw uint16
chunk -> dark/flat -> [stripe removal] -> minus log -> sinogram order.
Arithmetic is float32; ``dtype`` sets the storage of the sinogram-order
output (the reference computes in its dtype, ``proc_functions.py:67-74``).
"""
import ctypes

from . import _lib


class ProcFunctions:
    def __init__(self, nproj, ncz, n, ndark, nflat, dtype='float32', minus_log=True,
                 bright_ratio=1.0, flat_linear=False, remove_stripe_method='none',
                 fw_sigma=1.0, fw_filter='sym16', fw_level=7, dezinger=0,
                 dezinger_threshold=5000, ti_beta=0.022, ti_mask=1.0, file_type='standard',
                 center=None, binning=0, vo_all_snr=3.0, vo_all_la_size=61,
                 vo_all_sm_size=21, vo_all_dim=1):
        if remove_stripe_method not in ('none', 'fw', 'ti', 'vo-all'):
            raise NotImplementedError('remove_stripe_method %r is not implemented'
                                      % remove_stripe_method)
        self.nproj, self.ncz, self.n = int(nproj), int(ncz), int(n)
        self.ndark, self.nflat = int(ndark), int(nflat)
        if dtype not in _lib.DTYPES:
            raise ValueError('dtype must be one of %s' % sorted(_lib.DTYPES))
        self.dtype = dtype
        self.take_minus_log = bool(minus_log)
        self.bright_ratio = float(bright_ratio)
        self.flat_linear = bool(flat_linear)
        self.remove_stripe_method = remove_stripe_method
        self.fw_sigma, self.fw_filter, self.fw_level = float(fw_sigma), fw_filter, int(fw_level)
        self.dezinger, self.dezinger_threshold = int(dezinger), int(dezinger_threshold)
        self.ti_beta, self.ti_mask = float(ti_beta), float(ti_mask)
        self.vo_all_snr, self.vo_all_la_size = float(vo_all_snr), int(vo_all_la_size)
        self.vo_all_sm_size, self.vo_all_dim = int(vo_all_sm_size), int(vo_all_dim)
        self._res = None  # float32 transmission chunk of the unfused ti/fw paths
        if int(binning) not in (0, 1, 2, 3):
            raise ValueError('binning must be 0, 1, 2 or 3 (config.py:154-158)')
        self.binning = int(binning)
        self._binned = {}  # float32 binned data/dark/flat of the chunk
        if file_type not in ('standard', 'double_fov'):
            raise ValueError('file_type must be standard or double_fov')
        self.file_type = file_type
        if file_type == 'double_fov':
            # reader.py:143-148: the axis is moved to the right half (flip)
            if center is None:
                raise ValueError('double_fov needs the rotation axis position (center)')
            self.flip = float(center) < self.n // 2
            self.center360 = self.n - float(center) if self.flip else float(center)
            self._sino_ni = None
        self.nout = 2 * self.n if file_type == 'double_fov' else self.n
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().pre_create(ctypes.byref(h), self.nproj, self.ncz, self.n,
                                         self.ndark, self.nflat), 'pre_create')
        self._h = h

    def _shapes(self):
        import torch
        return ((self.nproj, self.ncz, self.n), (self.ndark, self.ncz, self.n),
                (self.nflat, self.ncz, self.n), torch.uint16)

    def downsample(self, data, out=None):
        """``utils.downsample(data, binning)`` (``utils.py:157-167``) on the
        GPU: uint16 ``(rows, ncz 2^b, n 2^b)`` -> float32 ``(rows, ncz, n)``
        block sums (exact)."""
        import torch
        b = self.binning
        if data.dim() != 3:
            raise ValueError('downsample: expected (rows, %d, %d) uint16 frames, got %s'
                             % (self.ncz << b, self.n << b, tuple(data.shape)))
        rows = int(data.shape[0])
        shape = (rows, self.ncz << b, self.n << b)
        if out is None:
            out = torch.empty((rows, self.ncz, self.n), dtype=torch.float32, device=data.device)
        _lib.check(_lib.lib().pre_bin(
            self._h, _lib.dev_ptr(data, shape, torch.uint16, 'data'),
            _lib.dev_ptr(out, (rows, self.ncz, self.n), torch.float32, 'out'), rows, b,
            _lib.stream_ptr(None)), 'pre_bin')
        return out

    def _downsample3(self, data, dark, flat):
        import torch
        res = []
        for key, x in (('data', data), ('dark', dark), ('flat', flat)):
            buf = self._binned.get(key)
            rows = int(x.shape[0])
            if buf is None or buf.device != x.device or buf.shape[0] != rows:
                buf = torch.empty((rows, self.ncz, self.n), dtype=torch.float32, device=x.device)
                self._binned[key] = buf
            res.append(self.downsample(x, buf))
        return res

    def remove_outliers(self, data):
        """Zinger removal in place on a ``(rows, ncz, n)`` array of counts,
        uint16 or float32 (binned) (``proc_functions.py:96-107``): pixels more
        than ``dezinger_threshold`` above the median of their ``dezinger`` x 1
        x ``dezinger`` window (angle, slice, detector) take the median value."""
        import torch
        if self.dezinger > 0:
            if data.dim() != 3 or tuple(data.shape[1:]) != (self.ncz, self.n):
                raise ValueError('remove_outliers: expected (rows, %d, %d), got %s'
                                 % (self.ncz, self.n, tuple(data.shape)))
            f32 = data.dtype == torch.float32
            fn = 'pre_remove_outliers_f32' if f32 else 'pre_remove_outliers'
            _lib.check(getattr(_lib.lib(), fn)(
                self._h, _lib.dev_ptr(data, tuple(data.shape),
                                      torch.float32 if f32 else torch.uint16, 'data'),
                int(data.shape[0]), self.dezinger, self.dezinger_threshold,
                _lib.stream_ptr(None)), fn)
        return data

    def _outliers(self, data, dark, flat):
        # proc_sino order (proc_functions.py:146-148): data, dark, flat
        if self.dezinger > 0:
            self.remove_outliers(data)
            self.remove_outliers(dark)
            self.remove_outliers(flat)

    def darkflat_correction(self, data, dark, flat, res=None):
        """``(data - dark0) / flat0`` (``proc_functions.py:55-74``) on uint16
        or float32 (binned) counts of the plan's ``(rows, ncz, n)``."""
        import torch
        sd, sk, sf, u16 = self._shapes()
        if res is None:
            res = torch.empty(sd, dtype=torch.float32, device=data.device)
        tin = torch.float32 if data.dtype == torch.float32 else u16
        fn = 'pre_darkflat_f32' if tin == torch.float32 else 'pre_darkflat'
        _lib.check(getattr(_lib.lib(), fn)(
            self._h, _lib.dev_ptr(data, sd, tin, 'data'), _lib.dev_ptr(dark, sk, tin, 'dark'),
            _lib.dev_ptr(flat, sf, tin, 'flat'), _lib.dev_ptr(res, sd, torch.float32, 'res'),
            self.bright_ratio, int(self.flat_linear), _lib.stream_ptr(None)), fn)
        return res

    def remove_stripe_fw(self, res):
        import torch
        sd = (self.nproj, self.ncz, self.n)
        _lib.check(_lib.lib().pre_remove_stripe_fw(
            self._h, _lib.dev_ptr(res, sd, torch.float32, 'res'), self.fw_sigma,
            self.fw_filter.encode(), self.fw_level, _lib.stream_ptr(None)), 'pre_remove_stripe_fw')
        return res

    def remove_stripe_ti(self, res):
        """Titarenko ring removal in place on the float32 chunk
        (``remove_stripe.py:296-309``)."""
        import torch
        sd = (self.nproj, self.ncz, self.n)
        _lib.check(_lib.lib().pre_remove_stripe_ti(
            self._h, _lib.dev_ptr(res, sd, torch.float32, 'res'), self.ti_beta, self.ti_mask,
            _lib.stream_ptr(None)), 'pre_remove_stripe_ti')
        return res

    @staticmethod
    def rotate_proj(data, angle, order=1, out=None):
        """``rotate_proj`` (``proc_functions.py:126-138``): every projection of
        the float32 ``data (nproj, ny, nx)`` rotated by ``angle`` degrees
        (``ndimage.rotate(..., reshape=False, axes=(2, 1), mode='nearest')``),
        interpolation order 0 to 5; in place (through a temporary) unless
        ``out`` is given."""
        import torch
        if abs(float(angle)) == 0.0:
            if out is not None:
                out.copy_(data)
                return out
            return data
        dst = torch.empty_like(data) if out is None else out
        shape = tuple(data.shape)
        work = None
        if int(order) >= 2:  # spline coefficients of the 12-padded planes
            work = torch.empty((shape[0], shape[1] + 24, shape[2] + 24), dtype=torch.float32,
                               device=data.device)
        _lib.check(_lib.lib().pre_rotate_proj(
            _lib.dev_ptr(data, shape, torch.float32, 'data'),
            _lib.dev_ptr(dst, shape, torch.float32, 'out'), shape[0], shape[1], shape[2],
            float(angle), int(order), None if work is None else work.data_ptr(),
            _lib.stream_ptr(None)), 'pre_rotate_proj')
        if out is None:
            data.copy_(dst)
            return data
        return dst

    def remove_all_stripe(self, res):
        """Vo's combined ring removal (``remove_stripe.py:311-497``, the
        ``vo-all`` method) in place on the float32 chunk: dead, large and
        small stripes."""
        import torch
        sd = (self.nproj, self.ncz, self.n)
        _lib.check(_lib.lib().pre_remove_all_stripe(
            self._h, _lib.dev_ptr(res, sd, torch.float32, 'res'), self.vo_all_snr,
            self.vo_all_la_size, self.vo_all_sm_size, self.vo_all_dim, _lib.stream_ptr(None)),
            'pre_remove_all_stripe')
        return res

    def _remove_stripe_res(self, res):
        # stripe removal on the float32 transmission chunk (proc_functions.py:150-159)
        if self.remove_stripe_method == 'fw':
            self.remove_stripe_fw(res)
        elif self.remove_stripe_method == 'ti':
            self.remove_stripe_ti(res)
        elif self.remove_stripe_method == 'vo-all':
            self.remove_all_stripe(res)
        return res

    def pad360(self, sino, out=None):
        """Double field of view (``proc_functions.py:109-124``) on sinograms
        ``(ncz, nproj, n)``: flip when the axis is in the left half, taper the
        last ``max(1, int(2 (n - center)))`` columns, zero-pad to ``2n``."""
        import torch
        td = _lib.torch_dtype(self.dtype)
        if out is None:
            out = torch.empty((self.ncz, self.nproj, 2 * self.n), dtype=td, device=sino.device)
        _lib.check(_lib.lib().pre_pad360(
            self._h, _lib.dev_ptr(sino, (self.ncz, self.nproj, self.n), td, 'sino'),
            _lib.dev_ptr(out, (self.ncz, self.nproj, 2 * self.n), td, 'out'),
            _lib.DTYPES[self.dtype], self.center360, int(self.flip), _lib.stream_ptr(None)),
            'pre_pad360')
        return out

    def proc_sino(self, data, dark, flat, res=None):
        """outliers, dark/flat correction, then stripe removal
        (``proc_functions.py:140-161``); the float32 transmission chunk
        ``(nproj, ncz, n)``.  With binning the raw frames are binned first."""
        if self.binning > 0:
            data, dark, flat = self._downsample3(data, dark, flat)
        self._outliers(data, dark, flat)
        res = self.darkflat_correction(data, dark, flat, res)
        return self._remove_stripe_res(res)

    def log_swap_rows(self, vol, z0, lz, out=None):
        """Sinograms ``(ncz, nproj, n)`` of rows ``z0 .. z0+lz-1`` of a
        resident float32 volume ``(nproj, nz_total, n)``: minus log
        (``proc_functions.py:76-83``, as ``proc_proj``) and the swap to
        sinogram order in one pass; rows from ``lz`` on are zero."""
        import torch
        td = _lib.torch_dtype(self.dtype)
        if out is None:
            out = torch.empty((self.ncz, self.nproj, self.n), dtype=td, device=vol.device)
        ld = int(vol.shape[1])
        _lib.check(_lib.lib().pre_minus_log_swap_rows(
            self._h, _lib.dev_ptr(vol, (self.nproj, ld, self.n), torch.float32, 'vol'), ld,
            int(z0), int(lz), _lib.dev_ptr(out, (self.ncz, self.nproj, self.n), td, 'out'),
            _lib.DTYPES[self.dtype], int(self.take_minus_log), _lib.stream_ptr(None)),
            'pre_minus_log_swap_rows')
        return out

    def proc_proj_swap(self, data, out=None):
        """minus log (``proc_functions.py:163-198``) fused with the swap to
        sinogram order (``rec.py:174``): ``(nproj, ncz, n) -> (ncz, nproj, n)``."""
        import torch
        sd = (self.nproj, self.ncz, self.n)
        td = _lib.torch_dtype(self.dtype)
        if out is None:
            out = torch.empty((self.ncz, self.nproj, self.n), dtype=td, device=data.device)
        _lib.check(_lib.lib().pre_minus_log_swap(
            self._h, _lib.dev_ptr(data, sd, torch.float32, 'data'),
            _lib.dev_ptr(out, (self.ncz, self.nproj, self.n), td, 'out'),
            _lib.DTYPES[self.dtype], int(self.take_minus_log), _lib.stream_ptr(None)),
            'pre_minus_log_swap')
        return out

    def proc_chunk(self, data, dark, flat, out=None, res=None):
        """Raw uint16 chunk -> filtered-ready sinograms ``(ncz, nproj, nout)``
        (``nout = 2n`` for double field of view, else ``n``).  With
        ``dezinger > 0`` the outliers of ``data``, ``dark`` and ``flat`` are
        removed in place first, as ``proc_sino`` does; with ``binning > 0``
        the raw frames ``(rows, ncz 2^b, n 2^b)`` are binned first (into
        float32 buffers of the object; the raw frames are not modified)."""
        import torch
        if self.file_type == 'double_fov':
            if self._sino_ni is None or self._sino_ni.device != data.device:
                self._sino_ni = torch.empty((self.ncz, self.nproj, self.n),
                                            dtype=_lib.torch_dtype(self.dtype), device=data.device)
            self._proc_chunk(data, dark, flat, self._sino_ni, res)
            return self.pad360(self._sino_ni, out)
        return self._proc_chunk(data, dark, flat, out, res)

    def _proc_chunk(self, data, dark, flat, out, res):
        import torch
        if self.binning > 0:
            return self._proc_chunk_f32(*self._downsample3(data, dark, flat), out)
        sd, sk, sf, u16 = self._shapes()
        self._outliers(data, dark, flat)
        td = _lib.torch_dtype(self.dtype)
        if out is None:
            out = torch.empty((self.ncz, self.nproj, self.n), dtype=td, device=data.device)
        if self.remove_stripe_method == 'fw':
            # dark/flat fused into the first wavelet pass, minus_log + swap
            # into the last (the transmission chunk is never written)
            _lib.check(_lib.lib().pre_process_fw(
                self._h, _lib.dev_ptr(data, sd, u16, 'data'), _lib.dev_ptr(dark, sk, u16, 'dark'),
                _lib.dev_ptr(flat, sf, u16, 'flat'),
                _lib.dev_ptr(out, (self.ncz, self.nproj, self.n), td, 'out'),
                _lib.DTYPES[self.dtype], self.bright_ratio, int(self.flat_linear),
                int(self.take_minus_log), self.fw_sigma, self.fw_filter.encode(), self.fw_level,
                _lib.stream_ptr(None)), 'pre_process_fw')
            return out
        if self.remove_stripe_method in ('ti', 'vo-all'):
            # needs the whole transmission chunk (means, sorts over angles):
            # dark/flat into a float32 chunk, ring removal in place, then
            # -log + swap
            if self._res is None or self._res.device != data.device:
                self._res = torch.empty(sd, dtype=torch.float32, device=data.device)
            self.darkflat_correction(data, dark, flat, self._res)
            self._remove_stripe_res(self._res)
            return self.proc_proj_swap(self._res, out)
        _lib.check(_lib.lib().pre_process(
            self._h, _lib.dev_ptr(data, sd, u16, 'data'), _lib.dev_ptr(dark, sk, u16, 'dark'),
            _lib.dev_ptr(flat, sf, u16, 'flat'),
            _lib.dev_ptr(out, (self.ncz, self.nproj, self.n), td, 'out'), _lib.DTYPES[self.dtype],
            self.bright_ratio, int(self.flat_linear), int(self.take_minus_log),
            _lib.stream_ptr(None)), 'pre_process')
        return out

    def _proc_chunk_f32(self, data, dark, flat, out):
        # binned float32 counts: dark/flat (+ -log + swap) in one pass, or via
        # the float32 transmission chunk for stripe removal
        import torch
        sd, sk, sf, _ = self._shapes()
        f32 = torch.float32
        self._outliers(data, dark, flat)
        td = _lib.torch_dtype(self.dtype)
        if out is None:
            out = torch.empty((self.ncz, self.nproj, self.n), dtype=td, device=data.device)
        if self.remove_stripe_method != 'none':
            if self._res is None or self._res.device != data.device:
                self._res = torch.empty(sd, dtype=f32, device=data.device)
            self.darkflat_correction(data, dark, flat, self._res)
            self._remove_stripe_res(self._res)
            return self.proc_proj_swap(self._res, out)
        _lib.check(_lib.lib().pre_process_f32(
            self._h, _lib.dev_ptr(data, sd, f32, 'data'), _lib.dev_ptr(dark, sk, f32, 'dark'),
            _lib.dev_ptr(flat, sf, f32, 'flat'),
            _lib.dev_ptr(out, (self.ncz, self.nproj, self.n), td, 'out'), _lib.DTYPES[self.dtype],
            self.bright_ratio, int(self.flat_linear), int(self.take_minus_log),
            _lib.stream_ptr(None)), 'pre_process_f32')
        return out

    def free(self):
        h, self._h = getattr(self, '_h', None), None
        if h is not None and h.value:
            _lib.lib().pre_destroy(h)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
