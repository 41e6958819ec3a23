"""USFFT backprojector with the reference's interface.

Mirrors ``FourierRec`` (``src/tomocupy/reconstruction/fourierrec.py:46-68``):
``FourierRec(n, nproj, nz, theta, dtype).backprojection(obj, data, stream)``.
Differences, all documented deviations:

* ``stream`` is optional and defaults to torch's current stream;
* ``data`` is never modified (the reference protects it with a packed copy;
  here the packing happens inside the kernels);
* odd ``nz`` is accepted (the last slice pairs with an implicit zero slice);
* ``theta`` is copied when the plan is created (the reference keeps a borrowed
  device pointer alive, ``fourierrec.py:50``);
* ``forward`` is the exact adjoint (USFFT forward projection), new;
* ``dtype`` ``'float16'`` / ``'bfloat16'`` store sinograms, spectra, grids and
  images in 16 bits with float32 arithmetic (the reference's float16 module
  computes in half precision throughout, ``cfunc_fourierrecfp16.i``).
"""
import ctypes

import numpy as np

from . import _lib


class FourierRec:
    """Backprojection by the Fourier-based method."""

    def __init__(self, n, nproj, nz, theta, dtype='float32', grid_chunk=0, fft='static'):
        """``grid_chunk`` and ``fft`` are tuning options (see
        :mod:`tomocupy_amd.autotune`): the maximum sample-list entries per
        gridding work item (0 = default) and ``'static'`` (size-specialised
        transforms where available) or ``'generic'`` radix plans."""
        import torch
        if dtype not in _lib.DTYPES:
            raise ValueError('dtype must be one of %s' % sorted(_lib.DTYPES))
        self.n, self.nproj, self.nz, self.dtype = int(n), int(nproj), int(nz), dtype
        if not isinstance(theta, torch.Tensor):
            theta = torch.as_tensor(np.asarray(theta, dtype=np.float32))
        if not theta.is_cuda:
            theta = theta.cuda()
        theta = theta.to(torch.float32).contiguous()
        if theta.numel() != self.nproj:
            raise ValueError('theta has %d angles, expected nproj=%d' % (theta.numel(), self.nproj))
        self.theta = theta
        L = _lib.lib()
        h = ctypes.c_void_p()
        if fft not in ('static', 'generic'):
            raise ValueError("fft must be 'static' or 'generic'")
        self.grid_chunk, self.fft = int(grid_chunk), fft
        _lib.check(L.fr_create_opts(ctypes.byref(h), self.n, self.nproj, self.nz, theta.data_ptr(),
                                    _lib.DTYPES[dtype], self.grid_chunk, int(fft == 'generic'),
                                    _lib.stream_ptr(None)), 'fr_create')
        self._h = h

    def backprojection(self, obj, data, stream=None):
        """``obj (nz, n, n)`` <- backprojection of filtered sinograms ``data (nz, nproj, n)``,
        both of the plan's dtype."""
        td = _lib.torch_dtype(self.dtype)
        po = _lib.dev_ptr(obj, (self.nz, self.n, self.n), td, 'obj')
        pd = _lib.dev_ptr(data, (self.nz, self.nproj, self.n), td, 'data')
        _lib.check(_lib.lib().fr_backprojection(self._handle(), po, pd, _lib.stream_ptr(stream)),
                   'fr_backprojection')
        return obj

    def forward(self, data, obj, stream=None):
        """``data (nz, nproj, n)`` <- adjoint of :meth:`backprojection` applied to ``obj``
        (float32 plans)."""
        import torch
        pd = _lib.dev_ptr(data, (self.nz, self.nproj, self.n), torch.float32, 'data')
        po = _lib.dev_ptr(obj, (self.nz, self.n, self.n), torch.float32, 'obj')
        _lib.check(_lib.lib().fr_forward(self._handle(), pd, po, _lib.stream_ptr(stream)),
                   'fr_forward')
        return data

    def _handle(self):
        if self._h is None:
            raise RuntimeError('FourierRec plan has been freed')
        return self._h

    def free(self):
        """Release device memory; idempotent (``cfunc_fourierrec::free``)."""
        h, self._h = getattr(self, '_h', None), None
        if h is not None and h.value:
            _lib.lib().fr_destroy(h)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
