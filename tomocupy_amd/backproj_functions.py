"""Filter + backprojection driver with the reference's interface.

Mirrors ``BackprojFunctions`` (``src/tomocupy/reconstruction/backproj_functions.py:46-109``)
for the fourierrec algorithm.  The reference reads the process-global
``args``/``params`` namespaces; this class takes the same values as explicit
arguments (``n``, ``nproj``, ``ncz``, ``theta``, ``center``, ``dtype``,
``fbp_filter``).
"""
import math

from .fbp_filter import FBPFilter
from .fourierrec import FourierRec


def padded_length(n, dtype='float32'):
    """``ne = 4n``, rounded up to a power of two for float16
    (``backproj_functions.py:50-54``)."""
    ne = 4 * n
    if dtype == 'float16':
        ne = 2 ** int(math.ceil(math.log2(ne)))
    return ne


class BackprojFunctions:
    def __init__(self, n, nproj, ncz, theta, center, dtype='float32', fbp_filter='parzen',
                 reconstruction_algorithm='fourierrec', rec_options=None):
        """``rec_options``: FourierRec tuning options (``grid_chunk``, ``fft``),
        e.g. from :func:`tomocupy_amd.autotune.tune`."""
        if reconstruction_algorithm != 'fourierrec':
            raise NotImplementedError('only the fourierrec algorithm is implemented')
        self.n, self.nproj, self.ncz = int(n), int(nproj), int(ncz)
        self.center = float(center)
        self.dtype = dtype
        self.ne = padded_length(self.n, dtype)
        self.cl_rec = FourierRec(self.n, self.nproj, self.ncz, theta, dtype, **(rec_options or {}))
        self.cl_filter = FBPFilter(self.ne, self.nproj, self.ncz, dtype, n_unpadded=self.n)
        self.wfilter = self.cl_filter.calc_filter(fbp_filter)
        self.pad = self.ne // 2 - self.n // 2

    def fbp_filter_center(self, data, sht=0):
        """FBP filtering of ``data (ncz, nproj, n)`` in place, with the rotation
        centre shift; ``sht`` is a scalar or a ``(ncz,)`` float32 GPU tensor."""
        import torch
        if isinstance(sht, torch.Tensor):
            shv = sht.to(torch.float32).contiguous()
        elif sht == 0:
            shv = None
        else:
            shv = torch.full((self.ncz,), float(sht), dtype=torch.float32, device=data.device)
        return self.cl_filter.filter_center(data, self.wfilter, self.center, shv)

    def fbp_filter_backprojection(self, obj, data, sht=0, stream=None):
        """``fbp_filter_center`` then ``backprojection`` in one pass
        (``rec.py:175-176``): ``obj (ncz, n, n)`` <- filtered backprojection of
        ``data (ncz, nproj, n)``, both of the plan's dtype.  ``data`` is not
        modified."""
        import torch
        from . import _lib
        td = _lib.torch_dtype(self.dtype)
        po = _lib.dev_ptr(obj, (self.ncz, self.n, self.n), td, 'obj')
        pd = _lib.dev_ptr(data, (self.ncz, self.nproj, self.n), td, 'data')
        L = _lib.lib()
        sp = _lib.stream_ptr(stream)
        if isinstance(sht, torch.Tensor):
            # per-slice shifts: kernels for this call
            shv = sht.to(torch.float32).contiguous()
            ps = _lib.dev_ptr(shv, (self.ncz,), torch.float32, 'sht')
            self._prepared = None
            _lib.check(L.fr_filter_backprojection(
                self.cl_rec._handle(), self.cl_filter._h, po, pd, self.wfilter.data_ptr(),
                self.center, ps, sp), 'fr_filter_backprojection')
            return obj
        self.prepare(sht, stream)
        _lib.check(L.fr_filter_backprojection_prepared(
            self.cl_rec._handle(), self.cl_filter._h, po, pd, sp),
            'fr_filter_backprojection_prepared')
        return obj

    def prepare(self, sht=0, stream=None):
        """Builds the filter kernels for the centre and the scalar shift
        ``sht`` unless the last call built them (``flt_prepare``)."""
        import torch
        from . import _lib
        key = (self.center, float(sht))
        if getattr(self, '_prepared', None) == key:
            return
        ps = None
        if sht != 0:
            self._shv = torch.full((self.ncz,), float(sht), dtype=torch.float32, device='cuda')
            ps = self._shv.data_ptr()
        _lib.check(_lib.lib().flt_prepare(self.cl_filter._h, self.wfilter.data_ptr(), self.center,
                                          ps, _lib.stream_ptr(stream)), 'flt_prepare')
        self._prepared = key

    def free(self):
        self.cl_rec.free()
        self.cl_filter.free()
