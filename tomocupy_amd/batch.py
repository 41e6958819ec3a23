"""Batching form: many independent reconstructions as one HIP graph.

Each request (its own n, nproj, angles, centre and slice count) gets its
filter and backprojection plans and device buffers at construction.  A
``BatchRec`` then runs the fused filter + backprojection of every request
(``fr_filter_backprojection``, the per-chunk step of the reference's
``rec.py:175-176``) either as eager launches or as a captured HIP graph that
fans the requests out over ``streams`` streams (longest first, each to the
least loaded stream) and joins them, so one graph launch replaces the host
launch sequence of the whole batch.  Plans own their scratch, so requests on
different streams never share memory.

With ``grid='batched'`` (float32) the requests share ONE gridding launch
instead: ``fr_batch_filter_backprojection`` runs every request's filter +
spectrum stage (fanned out over 4 streams the batch owns), then grids all
requests together from one work list (the largest requests' items first, each
request's items together; the small requests' short items fill the tail
instead of each small request launching a grid that leaves most CUs idle),
then every request's FFT and output stages (fanned out again).

This is the request-level scheduler the reference does not have (it runs one
volume per process); the per-request arithmetic is exactly the chunk step,
and both modes give the same result bit for bit.
"""
import numpy as np

from . import _lib
from .backproj_functions import BackprojFunctions


def request_bytes(n, nproj, nz, dtype='float32'):
    """Algorithmic bytes of one request's filtered backprojection (the cost
    model used to balance streams)."""
    r = 4 if dtype == 'float32' else 2
    nzc = (nz + 1) // 2
    return (nz * nproj * n * r + 4 * nzc * nproj * n * r
            + nzc * (2 * (2 * n) ** 2 * 2 * r + 2 * 2 * n * n * 2 * r) + nz * n * n * r)


class BatchRec:
    def __init__(self, requests, dtype='float32', fbp_filter='parzen', streams=4,
                 rec_options=None, grid='streams'):
        """``requests``: list of dicts with ``n``, ``nz``, ``theta`` (angles in
        radians, array-like) and optionally ``center`` (default n/2) and
        ``rec_options`` (FourierRec tuning options, e.g. from the autotuner).
        ``grid``: 'streams' (per-request launches fanned out over ``streams``
        streams) or 'batched' (one gridding launch for the whole batch,
        float32 only)."""
        import torch
        if not requests:
            raise ValueError('empty batch')
        if grid not in ('streams', 'batched'):
            raise ValueError("grid must be 'streams' or 'batched'")
        if grid == 'batched' and dtype != 'float32':
            raise ValueError("grid='batched' needs float32 storage")
        self.dtype = dtype
        self.grid = grid
        td = _lib.torch_dtype(dtype)
        self.items = []
        for r in requests:
            n, nz = int(r['n']), int(r['nz'])
            theta = np.asarray(r['theta'], dtype=np.float32)
            nproj = len(theta)
            center = float(r.get('center', n / 2))
            opts = r.get('rec_options', rec_options)
            bf = BackprojFunctions(n, nproj, nz, torch.from_numpy(theta).cuda(), center, dtype,
                                   fbp_filter, rec_options=opts)
            self.items.append({
                'n': n, 'nproj': nproj, 'nz': nz, 'bf': bf,
                'data': torch.zeros((nz, nproj, n), dtype=td, device='cuda'),
                'obj': torch.zeros((nz, n, n), dtype=td, device='cuda'),
                'cost': request_bytes(n, nproj, nz, dtype)})
        self.streams = [torch.cuda.Stream() for _ in range(max(1, int(streams)))]
        # longest processing time first onto the least loaded stream
        load = [0] * len(self.streams)
        self.assign = [[] for _ in self.streams]
        for i in sorted(range(len(self.items)), key=lambda i: -self.items[i]['cost']):
            s = int(np.argmin(load))
            self.assign[s].append(i)
            load[s] += self.items[i]['cost']
        self.graph = None
        self._batch = None
        if grid == 'batched':
            import ctypes
            k = len(self.items)
            plans = (ctypes.c_void_p * k)(*[it['bf'].cl_rec._handle() for it in self.items])
            h = ctypes.c_void_p()
            _lib.check(_lib.lib().fr_batch_create(ctypes.byref(h), plans, k), 'fr_batch_create')
            self._batch = h
            self._flts = (ctypes.c_void_p * k)(*[it['bf'].cl_filter._h for it in self.items])
            self._objs = (ctypes.c_void_p * k)(*[it['obj'].data_ptr() for it in self.items])
            self._datas = (ctypes.c_void_p * k)(*[it['data'].data_ptr() for it in self.items])

    @property
    def slices(self):
        return sum(it['nz'] for it in self.items)

    def inputs(self):
        """Per-request input sinogram tensors ``(nz, nproj, n)`` to fill."""
        return [it['data'] for it in self.items]

    def outputs(self):
        """Per-request output tensors ``(nz, n, n)``."""
        return [it['obj'] for it in self.items]

    def _launch(self):
        import torch
        cur = torch.cuda.current_stream()
        if self._batch is not None:
            for it in self.items:
                it['bf'].prepare(0, cur)
            _lib.check(_lib.lib().fr_batch_filter_backprojection(
                self._batch, self._flts, self._objs, self._datas, _lib.stream_ptr(cur)),
                'fr_batch_filter_backprojection')
            return
        for s, idx in zip(self.streams, self.assign):
            if not idx:
                continue
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                for i in idx:
                    it = self.items[i]
                    it['bf'].fbp_filter_backprojection(it['obj'], it['data'])
        for s, idx in zip(self.streams, self.assign):
            if idx:
                cur.wait_stream(s)

    def run_eager(self, sequential=False):
        """Launch every request from the host (no graph); ``sequential``: all
        on the current stream, one request after the other."""
        if sequential:
            for it in self.items:
                it['bf'].fbp_filter_backprojection(it['obj'], it['data'])
        else:
            self._launch()

    def capture(self):
        """Capture the whole batch into one HIP graph (after one eager run
        that finishes any lazy set-up)."""
        import torch
        self._launch()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch()
        torch.cuda.synchronize()
        self.graph = g
        return g

    def run(self):
        """Replay the captured graph (capturing it first if needed)."""
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def free(self):
        self.graph = None
        if self._batch is not None:
            _lib.lib().fr_batch_destroy(self._batch)
            self._batch = None
        for it in self.items:
            it['bf'].free()
        self.items = []
