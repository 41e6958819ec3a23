"""Reconstruction by steps with the reference's structure
(``src/tomocupy/rec_steps.py:63-243`` and
``reconstruction/backproj_parallel.py:315-358``), MI355X-resident.

``GPURecSteps.recon_steps_all(data, dark, flat, write_chunk)`` runs the
reference's three passes over a whole scan:

1. ``proc_sino`` by chunks of rows (outliers, dark/flat, stripe removal),
   ``rec_steps.py:127-192``;
2. ``proc_proj`` by chunks of projections (phase retrieval, projection
   rotation, then minus log), ``rec_steps.py:194-243``;
3. filtered backprojection by chunks of rows
   (``backproj_parallel.py:315-358``).

The reference stages the volume in host memory between the passes and
copies every chunk in and out.  Here the raw scan is uploaded once and the
float32 transmission volume ``(nproj, nz, ni)`` stays in HBM (25 GB at
2048 x 1500 x 2048; 288 GB per MI355X): pass 1 writes its rows, pass 2
filters contiguous projection blocks in place (``ph_apply``), and pass 3
reads each row chunk with the minus log and the swap to sinograms fused
(``pre_minus_log_swap_rows``), then pads for double field of view and runs
the fused filter + backprojection.  Only reconstructed chunks go back to the
host.  Settings the reference reads from ``args``/``params`` are
constructor arguments.
"""
import numpy as np

from . import _lib
from .backproj_functions import BackprojFunctions
from .proc_functions import ProcFunctions
from .retrieve_phase import PhasePlan


class GPURecSteps:
    def __init__(self, n, nproj, nz, ncz, theta, center, ndark, nflat, dtype='float32',
                 fbp_filter='parzen', nproj_per_chunk=8, retrieve_phase_method='none',
                 energy=20.0, propagation_distance=0.0, pixel_size=1.0, retrieve_phase_alpha=0.0,
                 retrieve_phase_delta_beta=1500.0, retrieve_phase_W=2e-4, file_type='standard',
                 phase_batch=1, rotate_proj_angle=0.0, rotate_proj_order=1, **proc_kwargs):
        """``n``, ``nz``, ``center``: binned detector width, rows and axis
        (as ``GPURec``).  Phase retrieval settings in the reference CLI's
        units: energy [keV], propagation distance [mm], pixel size [um], W
        [um] (``proc_functions.py:171-174`` converts them to cm).
        ``phase_batch`` projections share one phase-retrieval launch (one keeps
        the padded spectra, 134 MB at 4096 x 4096, in the 256 MB Infinity
        Cache between the row and column passes)."""
        import torch
        if ncz == 1:
            ncz = 2  # fourierrec pairs two slices (reader.py:113-115)
        self.ni = int(n)
        self.n = 2 * self.ni if file_type == 'double_fov' else self.ni
        rec_center = center
        if file_type == 'double_fov' and float(center) < self.ni // 2:
            rec_center = self.ni - float(center)
        self.nproj, self.nz, self.ncz = int(nproj), int(nz), int(ncz)
        self.ndark, self.nflat = int(ndark), int(nflat)
        self.nzchunk = -(-self.nz // self.ncz)
        self.lzchunk = [min(self.ncz, self.nz - k * self.ncz) for k in range(self.nzchunk)]
        self.ncproj = int(nproj_per_chunk)
        self.binning = int(proc_kwargs.get('binning', 0))
        self.dtype = dtype
        self.pf = ProcFunctions(self.nproj, self.ncz, self.ni, self.ndark, self.nflat, dtype,
                                file_type=file_type, center=center, **proc_kwargs)
        self.bf = BackprojFunctions(self.n, self.nproj, self.ncz, theta, rec_center, dtype,
                                    fbp_filter)
        self.phase = None
        if retrieve_phase_method != 'none':
            self.phase = PhasePlan(self.nz, self.ni, retrieve_phase_method,
                                   pixel_size * 1e-4, propagation_distance / 10, energy,
                                   retrieve_phase_alpha, retrieve_phase_delta_beta,
                                   retrieve_phase_W * 1e-4, batch=int(phase_batch))
        self.rotate_angle, self.rotate_order = float(rotate_proj_angle), int(rotate_proj_order)
        if self.rotate_angle != 0.0 and self.rotate_order not in range(6):
            raise ValueError('rotate_proj_order %d: the interpolation order is 0 to 5'
                             % self.rotate_order)
        self.stream = torch.cuda.Stream()
        self.tdtype = _lib.torch_dtype(dtype)

    def recon_steps_all(self, data, dark, flat, write_chunk):
        """``data (nproj, nz 2^b, ni 2^b)``, ``dark``, ``flat`` numpy uint16
        (the whole scan, as ``read_data_parallel``); ``write_chunk(k, rec)``
        receives each reconstructed row chunk ``(lz, n, n)`` as numpy."""
        import torch
        with torch.cuda.stream(self.stream):
            raw = torch.from_numpy(np.ascontiguousarray(data)).cuda()
            rdark = torch.from_numpy(np.ascontiguousarray(dark)).cuda()
            rflat = torch.from_numpy(np.ascontiguousarray(flat)).cuda()
            self.recon_steps_resident(raw, rdark, rflat, write_chunk=write_chunk)

    def recon_steps_resident(self, raw, dark, flat, rec=None, write_chunk=None):
        """The three passes on a scan already in GPU memory (uint16 tensors);
        chunks go to ``rec (nz, n, n)`` on the GPU and/or to
        ``write_chunk(k, numpy)``.  Runs on torch's current stream."""
        import torch
        vol = self.proc_volume(raw, dark, flat)
        ncz, dev = self.ncz, vol.device
        # 3. minus log + swap per row chunk, [pad360], filter + backprojection
        sino = torch.empty((ncz, self.nproj, self.ni), dtype=self.tdtype, device=dev)
        out = torch.empty((ncz, self.n, self.n), dtype=self.tdtype, device=dev)
        for k in range(self.nzchunk):
            z0, lz = k * ncz, self.lzchunk[k]
            self.pf.log_swap_rows(vol, z0, lz, sino)
            s = self.pf.pad360(sino) if self.pf.file_type == 'double_fov' else sino
            self.bf.fbp_filter_backprojection(out, s)
            if rec is not None:
                rec[z0:z0 + lz].copy_(out[:lz])
            if write_chunk is not None:
                host = out[:lz].float().cpu().numpy() if self.tdtype == torch.bfloat16 \
                    else out[:lz].cpu().numpy()
                write_chunk(k, host)
        return rec

    def sino_row(self, vol, z):
        """The sinogram ``(1, nproj, n)`` of row ``z`` of a processed volume
        (minus log, [pad360]): the try mode's input
        (``backproj_parallel.py:380``)."""
        import torch
        sino = torch.empty((self.ncz, self.nproj, self.ni), dtype=self.tdtype, device=vol.device)
        self.pf.log_swap_rows(vol, z, 1, sino)
        if self.pf.file_type == 'double_fov':
            sino = self.pf.pad360(sino)
        return sino[:1]

    def proc_volume(self, raw, dark, flat):
        """Passes 1 and 2 (``proc_sino`` by row chunks, phase retrieval by
        projection chunks): the float32 transmission volume ``(nproj, nz, ni)``
        on the GPU."""
        import torch
        b = self.binning
        ncz, rcz = self.ncz, self.ncz << b
        dev = raw.device
        vol = torch.empty((self.nproj, self.nz, self.ni), dtype=torch.float32, device=dev)
        cd = torch.zeros((self.nproj, rcz, self.ni << b), dtype=torch.uint16, device=dev)
        ck = torch.zeros((self.ndark, rcz, self.ni << b), dtype=torch.uint16, device=dev)
        cf = torch.ones((self.nflat, rcz, self.ni << b), dtype=torch.uint16, device=dev)
        res = torch.empty((self.nproj, ncz, self.ni), dtype=torch.float32, device=dev)
        # 1. proc_sino by row chunks into the resident volume
        for k in range(self.nzchunk):
            z0, lz = k * ncz, self.lzchunk[k]
            r0, rl = z0 << b, lz << b
            cd[:, :rl].copy_(raw[:, r0:r0 + rl])
            ck[:, :rl].copy_(dark[:, r0:r0 + rl])
            cf[:, :rl].copy_(flat[:, r0:r0 + rl])
            if rl < rcz:  # empty rows for the tail of the last chunk
                cd[:, rl:].zero_()
                ck[:, rl:].zero_()
                cf[:, rl:].fill_(1)
            self.pf.proc_sino(cd, ck, cf, res)
            vol[:, z0:z0 + lz].copy_(res[:, :lz])
        del cd, ck, cf, res
        # 2. proc_proj by projection chunks: phase retrieval, then the
        # rotation of the projections, in place (proc_functions.py:169-184)
        rot = None
        if self.rotate_angle != 0.0:
            rot = torch.empty((self.ncproj, self.nz, self.ni), dtype=torch.float32, device=dev)
        if self.phase is not None or rot is not None:
            for p0 in range(0, self.nproj, self.ncproj):
                blk = vol[p0:p0 + self.ncproj]
                if self.phase is not None:
                    self.phase.apply(blk)
                if rot is not None:
                    ProcFunctions.rotate_proj(blk, self.rotate_angle, self.rotate_order,
                                              out=rot[:blk.shape[0]])
                    blk.copy_(rot[:blk.shape[0]])
        return vol

    def free(self):
        self.pf.free()
        self.bf.free()
        if self.phase is not None:
            self.phase.free()
