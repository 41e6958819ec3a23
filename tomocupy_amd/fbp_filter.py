"""FBP filter with the reference's interface.

Mirrors ``FBPFilter`` (``src/tomocupy/reconstruction/fbp_filter.py:46-118``):
``FBPFilter(ne, ntheta, nz, dtype)`` with ``filter(data, w, stream)`` on padded
rows and ``calc_filter(name)``.  ``calc_filter`` runs the FP64 quadrature
design natively on the host (``flt_calc_filter``) and returns a float32 GPU
tensor, as the reference returns a float32 CuPy array.
"""
import ctypes

from . import _lib


class FBPFilter:
    def __init__(self, n, ntheta, nz, dtype='float32', n_unpadded=None):
        """``n`` is the padded FFT length ``ne`` as in the reference;
        ``n_unpadded`` is the detector width used by :meth:`filter_center`
        (defaults to ``n``)."""
        if dtype not in _lib.DTYPES:
            raise ValueError('dtype must be one of %s' % sorted(_lib.DTYPES))
        self.n = int(n)
        self.ntheta, self.nz, self.dtype = int(ntheta), int(nz), dtype
        self.n_unpadded = int(n_unpadded) if n_unpadded is not None else self.n
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().flt_create(ctypes.byref(h), self.n_unpadded, self.n, self.ntheta,
                                         self.nz, _lib.DTYPES[dtype]), 'flt_create')
        self._h = h

    def filter(self, data, w, stream=None):
        """In place on padded ``data (nz, ntheta, ne)`` of the plan's dtype with
        complex64 ``w (nz, ne/2+1)`` (``cfunc_filter::filter``)."""
        import torch
        pd = _lib.dev_ptr(data, (self.nz, self.ntheta, self.n), _lib.torch_dtype(self.dtype),
                          'data')
        if w.dtype == torch.complex64:
            w = torch.view_as_real(w.contiguous())
        pw = _lib.dev_ptr(w.contiguous(), (self.nz, self.n // 2 + 1, 2), torch.float32, 'w')
        _lib.check(_lib.lib().flt_filter(self._h, pd, pw, _lib.stream_ptr(stream)), 'flt_filter')
        return data

    def filter_center(self, data, wfilter, center, sht=None, stream=None):
        """Fused pad + centre phase ramp + filter + crop on ``data (nz, ntheta, n_unpadded)``."""
        import torch
        pd = _lib.dev_ptr(data, (self.nz, self.ntheta, self.n_unpadded),
                          _lib.torch_dtype(self.dtype), 'data')
        pw = _lib.dev_ptr(wfilter, (self.n // 2 + 1,), torch.float32, 'wfilter')
        ps = None
        if sht is not None:
            ps = _lib.dev_ptr(sht, (self.nz,), torch.float32, 'sht')
        _lib.check(_lib.lib().flt_filter_center(self._h, pd, pw, float(center), ps,
                                                _lib.stream_ptr(stream)), 'flt_filter_center')
        return data

    def calc_filter(self, filter):
        import torch
        w = _lib.calc_filter(self.n, filter)
        return torch.as_tensor(w.astype('float32')).cuda()

    def free(self):
        h, self._h = getattr(self, '_h', None), None
        if h is not None and h.value:
            _lib.lib().flt_destroy(h)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
