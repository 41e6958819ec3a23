"""Chunk conveyor with the reference's structure (``src/tomocupy/rec.py:62-208``).

``GPURec.recon_all(read_chunk, write_chunk)`` reconstructs a volume chunk by
chunk with three HIP streams: stream 1 uploads chunk k from pinned host
memory, stream 2 processes chunk k-1 (fused preprocessing, FBP filter,
backprojection), stream 3 downloads chunk k-2 into a pinned buffer that a
writer thread hands to ``write_chunk`` -- the conveyor of ``rec.py:161-205``.
Settings the reference reads from its global ``args``/``params`` are
constructor arguments.  ``recon_resident`` runs the same per-chunk work on a
volume that is already in GPU memory.

Deviation: the reference leaves stale slices from an earlier chunk in the
tail of a partial last chunk (``rec.py:188-190``), where they are paired with
valid slices in the complex packing; here the tail is reset to an empty
sinogram (data 0, dark 0, flat 1), so valid slices never pair with stale data.
"""
import threading

import numpy as np

from .backproj_functions import BackprojFunctions
from .proc_functions import ProcFunctions


class _WRThread:
    """Writer thread slot (``utils.py:123-145``)."""

    def __init__(self):
        self.thread = None
        self.done = threading.Event()
        self.done.set()
        self.error = None

    def run(self, fun, args):
        self.done.clear()

        def wrapper():
            try:
                fun(*args)
            except BaseException as e:  # surfaced by join()
                self.error = e
            finally:
                self.done.set()

        self.thread = threading.Thread(target=wrapper, daemon=True)
        self.thread.start()

    def busy(self):
        return not self.done.is_set()

    def join(self):
        if self.thread is not None:
            self.thread.join()
        if self.error is not None:
            err, self.error = self.error, None
            raise err


def _free_slot(slots):
    """Index of an idle writer slot (``utils.py:148-154``)."""
    while True:
        for i, t in enumerate(slots):
            if not t.busy():
                return i
        slots[0].done.wait(timeout=0.01)


class GPURec:
    def __init__(self, n, nproj, nz, ncz, theta, center, ndark, nflat, dtype='float32',
                 fbp_filter='parzen', remove_stripe_method='none', fw_sigma=1.0,
                 fw_filter='sym16', fw_level=7, minus_log=True, bright_ratio=1.0,
                 flat_linear=False, max_write_threads=8, dezinger=0, dezinger_threshold=5000,
                 ti_beta=0.022, ti_mask=1.0, file_type='standard', binning=0, vo_all_snr=3.0,
                 vo_all_la_size=61, vo_all_sm_size=21, vo_all_dim=1):
        """``n``: detector width of the data; the reconstruction is ``n x n``,
        or ``2n x 2n`` for ``file_type='double_fov'`` (reader.py:143-148: the
        rotation axis ``center`` is mirrored into the right half when it lies in
        the left one).  With ``binning=b`` (reader.py:137-140) ``n``, ``nz`` and
        ``center`` are the binned sizes; the raw chunks that ``read_chunk``
        returns hold ``lz 2^b`` rows of ``n 2^b`` columns."""
        import torch
        if ncz == 1:
            ncz = 2  # fourierrec pairs two slices (reader.py:113-115)
        self.ni = int(n)
        n = 2 * self.ni if file_type == 'double_fov' else self.ni
        rec_center = center
        if file_type == 'double_fov' and float(center) < self.ni // 2:
            rec_center = self.ni - float(center)
        self.n, self.nproj, self.nz, self.ncz = int(n), int(nproj), int(nz), int(ncz)
        self.ndark, self.nflat = int(ndark), int(nflat)
        self.nzchunk = -(-self.nz // self.ncz)
        self.lzchunk = [min(self.ncz, self.nz - k * self.ncz) for k in range(self.nzchunk)]
        self.pf = ProcFunctions(self.nproj, self.ncz, self.ni, self.ndark, self.nflat, dtype,
                                minus_log=minus_log, bright_ratio=bright_ratio,
                                flat_linear=flat_linear, remove_stripe_method=remove_stripe_method,
                                fw_sigma=fw_sigma, fw_filter=fw_filter, fw_level=fw_level,
                                dezinger=dezinger, dezinger_threshold=dezinger_threshold,
                                ti_beta=ti_beta, ti_mask=ti_mask, file_type=file_type,
                                center=center, binning=binning, vo_all_snr=vo_all_snr,
                                vo_all_la_size=vo_all_la_size,
                                vo_all_sm_size=vo_all_sm_size, vo_all_dim=vo_all_dim)
        self.binning = int(binning)
        self.bf = BackprojFunctions(self.n, self.nproj, self.ncz, theta, rec_center, dtype,
                                    fbp_filter)
        self.stream1 = torch.cuda.Stream()
        self.stream2 = torch.cuda.Stream()
        self.stream3 = torch.cuda.Stream()
        self.max_write_threads = int(max_write_threads)
        from . import _lib
        self.dtype = dtype
        self.tdtype = _lib.torch_dtype(dtype)
        self._sino = torch.empty((self.ncz, self.nproj, self.n), dtype=self.tdtype,
                                 device='cuda')

    def _process(self, data, dark, flat, rec):
        out = self.pf.proc_chunk(data, dark, flat, out=self._sino)
        self.bf.fbp_filter_backprojection(rec, out)

    def recon_resident(self, raw, dark, flat, rec):
        """Reconstruct a GPU-resident volume: ``raw (nzchunk, nproj, ncz, ni)`` uint16,
        ``dark``/``flat`` ``(nzchunk, ndark|nflat, ncz, ni)`` (rows and columns
        times ``2^binning`` when binning), into
        ``rec (nzchunk*ncz, n, n)`` of the plan's dtype, on stream 2."""
        import torch
        with torch.cuda.stream(self.stream2):
            for k in range(self.nzchunk):
                self._process(raw[k], dark[k], flat[k], rec[k * self.ncz:(k + 1) * self.ncz])
        return rec

    def recon_all(self, read_chunk, write_chunk):
        """Conveyor over host data.

        ``read_chunk(k)`` returns ``(data, dark, flat)`` numpy uint16 arrays of
        shapes ``(nproj, lz, ni)``, ``(ndark, lz, ni)``, ``(nflat, lz, ni)`` with
        ``lz = lzchunk[k]`` (``lz 2^b`` rows of ``ni 2^b`` raw columns with
        binning ``b``); ``write_chunk(k, rec)`` receives a ``(lz, n, n)``
        numpy view (float32, or float16 for 16-bit plans) that is valid until it
        returns.
        """
        import torch
        ncz, nproj, n = self.ncz, self.nproj, self.n
        b = self.binning
        rcz, ni = ncz << b, self.ni << b  # raw rows and columns of a chunk
        pin = dict(pin_memory=True)
        item_pinned = {
            'data': torch.zeros((2, nproj, rcz, ni), dtype=torch.uint16, **pin),
            'dark': torch.zeros((2, self.ndark, rcz, ni), dtype=torch.uint16, **pin),
            'flat': torch.ones((2, self.nflat, rcz, ni), dtype=torch.uint16, **pin),
        }
        item_gpu = {k: torch.empty_like(v, device='cuda') for k, v in item_pinned.items()}
        item_gpu['flat'].copy_(item_pinned['flat'])
        # bfloat16 has no numpy type: it is widened on the GPU before the copy
        host_dt = torch.float32 if self.tdtype == torch.bfloat16 else self.tdtype
        rec_pinned = torch.zeros((self.max_write_threads, ncz, n, n), dtype=host_dt, **pin)
        rec_gpu = torch.empty((2, ncz, n, n), dtype=self.tdtype, device='cuda')
        writers = [_WRThread() for _ in range(self.max_write_threads)]
        # the set-up copies above ran on the current stream; order the three
        # conveyor streams after them
        for s in (self.stream1, self.stream2, self.stream3):
            s.wait_stream(torch.cuda.current_stream())
        nzchunk = self.nzchunk
        for k in range(nzchunk + 2):
            if 0 < k < nzchunk + 1:
                with torch.cuda.stream(self.stream2):
                    i = (k - 1) % 2
                    self._process(item_gpu['data'][i], item_gpu['dark'][i], item_gpu['flat'][i],
                                  rec_gpu[i])
            ithread = None
            if k > 1:
                with torch.cuda.stream(self.stream3):
                    ithread = _free_slot(writers)
                    # chunk k-2 was finished on stream 2 before the previous step ended
                    rec_pinned[ithread].copy_(rec_gpu[(k - 2) % 2], non_blocking=True)
            if k < nzchunk:
                data, dark, flat = read_chunk(k)
                lz = self.lzchunk[k] << b
                item_pinned['data'][k % 2, :, :lz].copy_(torch.from_numpy(np.ascontiguousarray(data)))
                item_pinned['dark'][k % 2, :, :lz].copy_(torch.from_numpy(np.ascontiguousarray(dark)))
                item_pinned['flat'][k % 2, :, :lz].copy_(torch.from_numpy(np.ascontiguousarray(flat)))
                if lz < rcz:
                    item_pinned['data'][k % 2, :, lz:].zero_()
                    item_pinned['dark'][k % 2, :, lz:].zero_()
                    item_pinned['flat'][k % 2, :, lz:].copy_(
                        torch.ones((self.nflat, rcz - lz, ni), dtype=torch.int32).to(torch.uint16))
                with torch.cuda.stream(self.stream1):
                    for key in ('data', 'dark', 'flat'):
                        item_gpu[key][k % 2].copy_(item_pinned[key][k % 2], non_blocking=True)
            self.stream3.synchronize()
            if k > 1:
                lz = self.lzchunk[k - 2]
                writers[ithread].run(write_chunk, (k - 2, rec_pinned[ithread, :lz].numpy()))
            self.stream1.synchronize()
            self.stream2.synchronize()
        for t in writers:
            t.join()

    def free(self):
        self.pf.free()
        self.bf.free()
