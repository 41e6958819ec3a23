"""Differentiable USFFT projector and iterative least-squares reconstruction.

``FourierRec.forward`` (the USFFT forward projection A, ``fr_forward``) and
``FourierRec.backprojection`` (the reference's fourierrec backprojection,
``fourierrec.py:61-68``) are exact adjoints of each other over a pair-packed
chunk: <A x, y> = <x, A^T y> with the real inner product summed over all
slices.  ``Project`` exposes A as a ``torch.autograd.Function`` whose backward
is the hand-written adjoint, ``Backproject`` the transpose pair.

``cgls`` solves min ||A x - y||^2 + lam ||x||^2 with conjugate gradients on
the normal equations (A^T A + lam I) x = A^T y, every operator application a
pair of HIP launches sequences on the plan's stream.
"""
import torch

from . import _lib
from .fourierrec import FourierRec


def _apply_forward(rec, obj):
    data = torch.empty((rec.nz, rec.nproj, rec.n), dtype=torch.float32, device=obj.device)
    rec.forward(data, obj.contiguous())
    return data


def _apply_backprojection(rec, data):
    obj = torch.empty((rec.nz, rec.n, rec.n), dtype=torch.float32, device=data.device)
    rec.backprojection(obj, data.contiguous())
    return obj


class Project(torch.autograd.Function):
    """``y = A x``: images ``(nz, n, n)`` -> sinograms ``(nz, nproj, n)``."""

    @staticmethod
    def forward(ctx, obj, rec):
        ctx.rec = rec
        return _apply_forward(rec, obj.detach())

    @staticmethod
    def backward(ctx, grad):
        return _apply_backprojection(ctx.rec, grad), None


class Backproject(torch.autograd.Function):
    """``x = A^T y``: sinograms ``(nz, nproj, n)`` -> images ``(nz, n, n)``."""

    @staticmethod
    def forward(ctx, data, rec):
        ctx.rec = rec
        return _apply_backprojection(rec, data.detach())

    @staticmethod
    def backward(ctx, grad):
        return _apply_forward(ctx.rec, grad), None


def project(obj, rec):
    return Project.apply(obj, rec)


def backproject(data, rec):
    return Backproject.apply(data, rec)


def _dot(a, b, out=None):
    """sum(a * b) of two float32 tensors, accumulated in FP64 in a fixed
    order on the GPU (``tomo_dot_f64``); a 0-d float64 tensor."""
    if out is None:
        out = torch.empty(1025, dtype=torch.float64, device=a.device)
    a = a.contiguous()
    b = b.contiguous()
    _lib.check(_lib.lib().tomo_dot_f64(a.data_ptr(), b.data_ptr(), a.numel(), out.data_ptr(),
                                       _lib.stream_ptr(None)), 'tomo_dot_f64')
    return out[0].clone()


def cgls(rec, data, lam=0.0, niter=50, x0=None, callback=None):
    """Tikhonov-regularised least squares by CG on the normal equations.

    ``rec`` is a :class:`FourierRec` plan, ``data`` the measured sinograms
    ``(nz, nproj, n)``.  Returns the images ``(nz, n, n)``.  Inner products are
    accumulated in float64 on the GPU and the step sizes stay on the GPU, so an
    iteration does not wait for the host; the vector updates are two fused
    passes (``tomo_cg_step``: x, r and the new r.r; ``tomo_cg_direction``).  ``callback(k, x, rr)`` is called
    after every iteration with the normal-equation residual norm squared ``rr``
    (a 0-d float64 tensor).
    """
    if not isinstance(rec, FourierRec):
        raise TypeError('rec must be a FourierRec plan')
    lam = float(lam)

    def normal(v):
        q = _apply_backprojection(rec, _apply_forward(rec, v))
        if lam:
            q.add_(v, alpha=lam)
        return q

    if x0 is None:
        x = torch.zeros((rec.nz, rec.n, rec.n), dtype=torch.float32, device=data.device)
        r = _apply_backprojection(rec, data)
    else:
        x = x0.detach().clone().to(torch.float32)
        r = _apply_backprojection(rec, data) - normal(x)
    p = r.clone()
    x = x.contiguous()
    r = r.contiguous()
    nel = r.numel()
    # FP64 scalars and partial sums on the device (tomo_dot_f64 layout):
    # pq, and rr / rr_new alternating between two buffers
    wpq = torch.empty(1025, dtype=torch.float64, device=data.device)
    wrr = [torch.empty(1025, dtype=torch.float64, device=data.device) for _ in range(2)]
    L = _lib.lib()
    st = _lib.stream_ptr(None)
    _lib.check(L.tomo_dot_f64(r.data_ptr(), r.data_ptr(), nel, wrr[0].data_ptr(), st),
               'tomo_dot_f64')
    cur = 0
    for k in range(niter):
        q = normal(p).contiguous()
        _lib.check(L.tomo_dot_f64(p.data_ptr(), q.data_ptr(), nel, wpq.data_ptr(), st),
                   'tomo_dot_f64')
        # alpha = rr / pq; x += alpha p; r -= alpha q; rr_new = r.r
        _lib.check(L.tomo_cg_step(x.data_ptr(), r.data_ptr(), p.data_ptr(), q.data_ptr(), nel,
                                  wrr[cur].data_ptr(), wpq.data_ptr(), wrr[1 - cur].data_ptr(),
                                  st), 'tomo_cg_step')
        if callback is not None:
            callback(k, x, wrr[1 - cur][0].clone())
        # beta = rr_new / rr; p = r + beta p
        _lib.check(L.tomo_cg_direction(p.data_ptr(), r.data_ptr(), nel, wrr[cur].data_ptr(),
                                       wrr[1 - cur].data_ptr(), st), 'tomo_cg_direction')
        cur = 1 - cur
    return x
