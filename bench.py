#!/usr/bin/env python
"""Throughput of the fourierrec reconstruction path on one MI355X.

Workload (default): n=2048 detector columns, nproj=1500 angles in [0, pi),
nz=2048 slices processed in chunks of 8 over 3 HIP streams, raw uint16 counts
with 10 dark and 20 flat frames and column-gain stripes -> dark/flat ->
[Fourier-wavelet stripe removal] -> -log -> ramp filter (centre 1024.5) ->
USFFT backprojection.  All inputs are resident in HBM before the timed region.
One step = the whole nz-slice volume.  Prints one JSON line.

Kernel timing: an untimed pass over the first 16 chunks on ONE stream brackets
every kernel launch with HIP events (isolated kernel durations:
``kernels_untimed_breakdown_step``, ``roofline``) and names the dominant
kernel; the timed steps bracket only that kernel's launches
(``roofline_overlapped_events``: start-to-end times on its stream while the
streams overlap, which include waits behind the other streams' kernels).

``--gpus N`` with N > 1 exits: this build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "slices/s at 2048^2 x1500 proj, FP32 vs FP16 at rel-L2 vs FP64; HBM GB/s"
HBM_PEAK_GBS = 8000.0
ISO_CHUNKS = 16  # chunks of the untimed one-stream breakdown pass
KERNELS = ['k_pre_means', 'k_pre_rows', 'k_fw_rows', 'k_fw_cols', 'k_fw_xsum', 'k_fw_syn_cols',
           'k_fw_syn_rows', 'k_pre_log_swap', 'k_flt_weights', 'k_flt_rows', 'k_flt_prepare',
           'k_fr_prep', 'k_fr_filtprep',
           'k_fr_gscale', 'k_fr_grid_x3', 'k_fr_grid16', 'k_fr_grid_reduce', 'k_fr_rowfft',
           'k_fr_colfft', 'k_pre_median', 'k_pre_dezinger', 'k_ti_mean', 'k_ti_conv', 'k_ti_add',
           'k_pre_pad360', 'k_vo_transpose', 'k_vo_dead', 'k_vo_sort', 'k_vo_xmedian',
           'k_vo_median2d', 'k_vo_large', 'k_vo_unsort']


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--n', type=int, default=2048)
    ap.add_argument('--nproj', type=int, default=1500)
    ap.add_argument('--nz', type=int, default=2048)
    ap.add_argument('--chunk', type=int, default=8)
    ap.add_argument('--streams', type=int, default=3)
    ap.add_argument('--dtype', default='float32', choices=['float32', 'float16', 'bfloat16'])
    ap.add_argument('--stripe', default='fw', choices=['fw', 'none', 'ti', 'vo-all'])
    ap.add_argument('--fw-level', type=int, default=7, help='Fourier-wavelet levels (sym16, sigma 1)')
    ap.add_argument('--vo-all-dim', type=int, default=1, choices=[1, 2])
    ap.add_argument('--phase-batch', type=int, default=1,
                    help='steps form: projections per phase-retrieval launch')
    ap.add_argument('--rotate-proj-angle', type=float, default=0.0,
                    help='steps form: rotation of the projections in degrees')
    ap.add_argument('--rotate-proj-order', type=int, default=1,
                    help='steps form: interpolation order of the rotation (0..5)')
    ap.add_argument('--filter', default='ramp')
    ap.add_argument('--fft', default='static', choices=['static', 'generic'],
                    help='pipeline form: compiled transforms where the size has them (static) '
                         'or the run-time radix plans everywhere (generic)')
    ap.add_argument('--center', type=float, default=None, help='default n/2 + 0.5')
    ap.add_argument('--distinct', type=int, default=8,
                    help='distinct synthetic slices, tiled over nz')
    ap.add_argument('--cpu-baseline', default='on', choices=['on', 'off'])
    ap.add_argument('--cpu-workers', type=int, default=0, help='0 = min(16, cpu count)')
    ap.add_argument('--no-prof', action='store_true', help='do not time kernels with events')
    ap.add_argument('--half', default='float16', choices=['float16', 'bfloat16', 'off'],
                    help='pipeline form at --dtype float32: also time the 16-bit storage form '
                         'on the same workload (the "half" object)')
    ap.add_argument('--conveyor', default='on', choices=['on', 'off'],
                    help='pipeline form: also time GPURec.recon_all over host memory (H2D, '
                         'D2H through pinned buffers)')
    ap.add_argument('--form', default='pipeline', choices=['pipeline', 'batch', 'cg', 'steps'],
                    help='pipeline: the chunked reconstruction (default); batch: 96 independent '
                         'reconstructions of mixed sizes as one HIP graph vs eager launches; '
                         'cg: CG least squares with the autograd projector')
    ap.add_argument('--requests', type=int, default=96, help='batch form: number of requests')
    ap.add_argument('--iters', type=int, default=50, help='cg form: CG iterations')
    ap.add_argument('--tune', action='store_true', help='batch form: autotune each request size')
    ap.add_argument('--copy-calibration-mb', type=float, default=0.0,
                    help='run one device-to-device copy of this size before the timed steps '
                         '(a known-size kernel for HBM counter calibration)')
    return ap.parse_args()


def fw_levels(a, L=32):
    """Band shapes (H, W, Ho, Wo, Hs, Ws) of the sym16 Fourier-wavelet levels."""
    H, W = a.nproj + a.nproj // 8, a.n
    out = []
    for _ in range(a.fw_level):
        Ho, Wo = (H + L - 1) // 2, (W + L - 1) // 2
        out.append((H, W, Ho, Wo, 2 * Ho - L + 2, 2 * Wo - L + 2))
        H, W = Ho, Wo
    return out


def fw_bytes(name, a):
    """Mean external I/O per launch of a Fourier-wavelet kernel (one launch per
    level and chunk, levels launched equally often; correction form: the
    column analysis writes ll only, the column synthesis reads the
    correction of the level below, the last pass reads the raw counts and
    writes the sinograms at storage precision)."""
    ch = a.chunk
    r = 4 if a.dtype == 'float32' else 2
    per = []
    for k, (H, W, Ho, Wo, Hs, Ws) in enumerate(fw_levels(a)):
        if name == 'k_fw_rows':      # read the rows (level 0: uint16 counts), write lo+hi
            rd = 2 * a.nproj * a.n if k == 0 else 4 * H * W
            per.append(ch * (rd + 4 * 2 * H * Wo))
        elif name == 'k_fw_cols':    # read lo+hi, write ll
            per.append(4 * ch * (2 * H * Wo + Ho * Wo))
        elif name == 'k_fw_syn_cols':  # read the correction of the level below, write lo+hi
            rd = 0 if k == len(fw_levels(a)) - 1 else Ho * Wo
            per.append(4 * ch * (rd + 2 * Hs * Wo))
        elif name == 'k_fw_syn_rows':  # read lo+hi (last: + counts), write the correction / sinograms
            if k == 0:
                per.append(ch * (4 * 2 * Hs * Wo + 2 * a.nproj * a.n + r * a.nproj * a.n))
            else:
                per.append(4 * ch * (2 * Hs * Wo + Hs * Ws))
        else:
            return None
    return sum(per) // len(per)


def algorithmic_bytes(name, a, s):
    """External I/O of one launch of a kernel at storage precision (bytes):
    c = bytes of a stored complex value, r = of a stored real value."""
    if name.startswith('k_fw_') and a.stripe == 'fw':
        return fw_bytes(name, a)
    n, nproj, ch = a.n, a.nproj, a.chunk
    r = 4 if a.dtype == 'float32' else 2
    c = 2 * r
    nzc = (ch + 1) // 2
    grp = 4  # complex slices per gridding group
    if name in ('k_fr_grid_x3', 'k_fr_grid16'):  # read the polar spectra, write the 2n x 2n grid
        return grp * (nproj * n * c + (2 * n) ** 2 * c)
    if name == 'k_fr_rowfft':    # read the grid rows, write the pruned half
        return nzc * ((2 * n) ** 2 * c + 2 * n * n * c)
    if name == 'k_fr_colfft':    # read the pruned rows, write two real slices per complex slice
        return nzc * (2 * n * n * c) + ch * n * n * r
    if name in ('k_fr_prep', 'k_fr_filtprep'):  # read the real sinograms, write the spectra
        return ch * nproj * n * r + grp * nproj * n * c
    if name == 'k_flt_rows':     # read + write the sinograms in place
        return 2 * ch * nproj * n * r
    if name == 'k_pre_rows':     # read uint16 counts, write the sinograms
        return ch * nproj * n * (2 + r)
    if name == 'k_pre_log_swap':  # read float32 transmission, write the sinograms
        return ch * nproj * n * (4 + r)
    if name == 'k_vo_sort':      # read a float32 column, write sorted values + uint16 ranks
        return ch * nproj * n * (4 + 4 + 2)
    if name == 'k_vo_xmedian':   # read the sorted image, write its detector median
        return ch * nproj * n * (4 + 4)
    return None


def host_data(a):
    from tomocupy_amd import synth
    theta = synth.angles(a.nproj)
    t0 = time.time()
    d = min(a.distinct, a.chunk)
    data, dark, flat, _ = synth.raw_counts(2, a.n, theta, d, ndark=10, nflat=20, stripes=True)
    reps = -(-a.chunk // d)
    data = np.concatenate([data] * reps, axis=1)[:, :a.chunk]
    dark = np.concatenate([dark] * reps, axis=1)[:, :a.chunk]
    flat = np.concatenate([flat] * reps, axis=1)[:, :a.chunk]
    print('# synthetic data: %d distinct slices in %.1fs' % (d, time.time() - t0),
          file=sys.stderr, flush=True)
    return theta, data, dark, flat


def device_data(a, torch, data, dark, flat):
    nchunks = -(-a.nz // a.chunk)
    t0 = time.time()
    g = torch.from_numpy(data).cuda()
    raw = torch.empty((nchunks,) + tuple(g.shape), dtype=torch.uint16, device='cuda')
    raw[:] = g  # every chunk at its own address (no cache reuse across chunks)
    dk = torch.from_numpy(dark).cuda().expand((nchunks,) + dark.shape).contiguous()
    fl = torch.from_numpy(flat).cuda().expand((nchunks,) + flat.shape).contiguous()
    torch.cuda.synchronize()
    print('# tiled to %d chunks on the GPU in %.1fs' % (nchunks, time.time() - t0),
          file=sys.stderr, flush=True)
    return raw, dk, fl


ACC_PAIRS = 2  # slice pairs of chunk 0 whose oracle outputs are kept for the accuracy check


def acc_pair(a, i):
    """Slices of chunk 0 in the GPU's complex pair number i mod chunk/2."""
    h = max(1, a.chunk // 2)
    c = i % h
    return [c, c + h] if c + h < a.chunk else [c, c]


def _cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return ''


def cpu_baseline(a, data, dark, flat, theta, center):
    """FP64 oracle chain on the host: a process pool over slice pairs (one
    pair per worker per round), one small warm-up round, then the median of
    three timed rounds; plus one pair in a single process with every core
    (the oracle's C gridding is OpenMP-parallel).  Returns the baseline
    object and the oracle reconstructions of the first ACC_PAIRS pairs,
    float32, for the accuracy check.

    Pair i is the GPU's complex pair (c, c + chunk/2), c = i mod chunk/2: the
    reference packs those two slices into one complex slice
    (fourierrec.py:61-68), and since the backprojection of a real sinogram
    is not exactly real (the row shift of divphi breaks the symmetry), each
    slice's output carries a small part of its partner's; the oracle has to
    see the same partner."""
    import multiprocessing as mp
    workers = a.cpu_workers or min(16, os.cpu_count() or 1)
    pairs = workers
    sl = []
    for i in range(pairs):
        idx = acc_pair(a, i)
        sl.append((data[:, idx], dark[:, idx], flat[:, idx]))
    args = [(s, theta, center, a, i < ACC_PAIRS, 1) for i, s in enumerate(sl)]
    # warm-up: module imports and the C gridding library, on a tiny problem
    small = synth_small(a)
    # runs before this process initialises the GPU, so forked workers are plain CPU processes
    ctx = mp.get_context('fork')
    times, outs = [], None
    with ctx.Pool(workers) as pool:
        pool.map(_cpu_pair, [small] * workers)
        for r in range(3):
            t0 = time.time()
            res = pool.map(_cpu_pair, args)
            times.append(time.time() - t0)
            if outs is None:
                outs = [o for o in res if o is not None]
            print('# cpu baseline round %d: %.1fs' % (r, times[-1]), file=sys.stderr, flush=True)
    dt = float(np.median(times))
    with ctx.Pool(1) as pool:
        pool.map(_cpu_pair, [small])
        t0 = time.time()
        pool.map(_cpu_pair, [(sl[0], theta, center, a, False, os.cpu_count() or 1)])
        dt1 = time.time() - t0
    oracle = np.concatenate(outs) if outs else None
    return {'value': round(2 * pairs / dt, 3), 'unit': 'slices/s', 'cores': workers,
            'kind': 'port', 'host': {'cpu_count': os.cpu_count(), 'model': _cpu_model()},
            'rounds_s': [round(t, 2) for t in times],
            'single_process': {'value': round(2 / dt1, 3), 'threads': os.cpu_count(),
                               'note': 'one pair in one process, OpenMP gridding on every '
                                       'thread'},
            'sample': '%d slice pairs (%d slices) of the same workload through the FP64 '
                      'oracle (dark/flat, %s-log, %s filter, USFFT backprojection), one pair '
                      'per process per round; warm-up round on a tiny problem, median of 3 '
                      'rounds (%.1fs)' % (pairs, 2 * pairs,
                                          ('%s stripe removal, ' % a.stripe) if a.stripe != 'none' else '',
                                          a.filter, dt)}, oracle


def synth_small(a):
    from tomocupy_amd import synth
    th = synth.angles(24)
    d, dk, fl, _ = synth.raw_counts(1, 64, th, 2, ndark=2, nflat=2)
    return ((d, dk, fl), th, 32.0, a, False, 1)


def _cpu_pair(arg):
    (data, dark, flat), theta, center, a, keep, threads = arg
    os.environ['OMP_NUM_THREADS'] = str(threads)
    from oracle import filter_fp64 as ff
    from oracle import fourierrec_fp64 as fr
    from oracle import preproc_fp64 as pp
    sino = pp.proc_chain(data, dark, flat, stripe=a.stripe, fw_level=a.fw_level)
    n = data.shape[-1]
    ne = 4 * n
    filt = ff.fbp_filter_center(sino, ff.calc_filter(ne, a.filter), ne, center)
    rec = fr.backprojection(filt, theta)
    return rec.astype(np.float32) if keep else None


def rel_l2(got, ref):
    """Per-slice relative L2 error and max |error| / max |ref| (FP64)."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    num = np.sqrt(((got - ref) ** 2).reshape(len(ref), -1).sum(1))
    den = np.sqrt((ref ** 2).reshape(len(ref), -1).sum(1))
    return num / np.maximum(den, 1e-300), float(np.abs(got - ref).max() / np.abs(ref).max())


def accuracy(a, torch, rec, oracle, pf, bf, raw, dark, flat, theta, center, sino):
    """Accuracy of the GPU output against the FP64 oracle on the same slices.

    * end_to_end: the GPU's slices of the first ACC_PAIRS complex pairs of
      chunk 0 against the oracle chain run on the same uint16 counts with the
      same pairing (the CPU baseline's outputs);
    * stagewise: the GPU's reconstruction of the first pair against the
      oracle filter + backprojection of the GPU's own sinograms of those
      slices.  With photon-starved rays the -log clamp (x <= 0 -> 1) flips
      single sinogram values where the transmission after stripe removal is
      near 0, so the end-to-end figure also carries those flips; the
      stagewise figure isolates the filter and backprojection."""
    from oracle import filter_fp64 as ff
    from oracle import fourierrec_fp64 as fr
    out = {}
    k = len(oracle)
    idx = [z for i in range(k // 2) for z in acc_pair(a, i)]
    got = rec[idx].float().cpu().numpy()
    e, mx = rel_l2(got, oracle)
    out['end_to_end'] = {'slices': idx, 'rel_l2_max': float(e.max()),
                         'rel_l2_median': float(np.median(e)), 'max_abs_over_max': mx}
    s = pf.proc_chunk(raw[0], dark[0], flat[0], out=sino)
    r = torch.empty((a.chunk, a.n, a.n), dtype=rec.dtype, device='cuda')
    bf.fbp_filter_backprojection(r, s)
    torch.cuda.synchronize()
    p0 = acc_pair(a, 0)
    s2 = s[p0].double().cpu().numpy()
    ne = 4 * a.n
    ref = fr.backprojection(ff.fbp_filter_center(s2, ff.calc_filter(ne, a.filter), ne, center),
                            theta)
    e, mx = rel_l2(r[p0].double().cpu().numpy(), ref)
    out['stagewise'] = {'slices': p0, 'rel_l2_max': float(e.max()),
                        'rel_l2_median': float(np.median(e)), 'max_abs_over_max': mx}
    del r
    return out


def bench_batch(a):
    """Batching form: independent reconstructions with n in {256, 512, 768,
    1024, 1536, 2048}, nproj = round(1.5 n) from a random start angle, nz in
    [1, 32], filtered-ready float32 sinograms (one synthetic ellipse slice per
    request, repeated over its nz), Parzen filter.  Value: one HIP graph
    with the per-request stages over 4 streams and one gridding launch for
    the whole batch (float32); beside it the graph of per-request launches
    over 4 streams and eager per-request launches on one stream."""
    import torch
    from tomocupy_amd import synth
    from tomocupy_amd.batch import BatchRec
    sizes = list(synth.BATCH_SIZES)
    t0 = time.time()
    cache = {}
    reqs = synth.batch_requests(a.requests, 2024)
    for r in reqs:
        n = r['n']
        if n not in cache:
            cache[n] = synth.sinograms(4, n, synth.angles(len(r['theta'])), 1)[0].astype(np.float32)
        r['sino'] = cache[n]
    print('# %d requests generated in %.1fs' % (len(reqs), time.time() - t0), file=sys.stderr,
          flush=True)
    opts = {}
    if a.tune:
        from tomocupy_amd.autotune import tune
        for n in sorted({r['n'] for r in reqs}):
            opts[n] = tune(n, int(round(1.5 * n)), 8, verbose=True)
    for r in reqs:
        r['rec_options'] = opts.get(r['n'])
    b = BatchRec(reqs, dtype=a.dtype, fbp_filter='parzen', streams=4)
    td = torch.float32 if a.dtype == 'float32' else torch.float16
    for r, d in zip(reqs, b.inputs()):
        d.copy_(torch.from_numpy(r['sino'])[None].expand(r['nz'], -1, -1).to(td))

    def timed(fn):
        for _ in range(a.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(a.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / a.steps

    t_seq = timed(lambda: b.run_eager(sequential=True))
    b.capture()
    t_graph = timed(b.run)
    slices = b.slices
    t_bgrid = None
    if a.dtype == 'float32':
        # the same batch with one gridding launch for all requests (graph)
        ref = [o.clone() for o in b.outputs()]
        b.free()
        del b
        b = BatchRec(reqs, dtype=a.dtype, fbp_filter='parzen', grid='batched')
        for r, d in zip(reqs, b.inputs()):
            d.copy_(torch.from_numpy(r['sino'])[None].expand(r['nz'], -1, -1).to(td))
        b.capture()
        t_bgrid = timed(b.run)
        same = all(torch.equal(x, y) for x, y in zip(ref, b.outputs()))
        del ref
    b.free()
    # the batching form's value: one gridding launch for the whole batch
    # (float32); 16-bit storage: the per-request graph
    t_val = t_bgrid if t_bgrid is not None else t_graph
    line = {'metric': 'slices/s, %d independent mixed-size reconstructions (filter + '
                      'backprojection), one HIP graph' % len(reqs),
            'value': round(slices / t_val, 2), 'unit': 'slices/s', 'n_gpus': 1,
            'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': round(t_val * 1e3, 3),
            'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'dtype': {'float32': 'f32', 'float16': 'f16', 'bfloat16': 'bf16'}[a.dtype],
            'data': 'synthetic ellipse sinograms, one slice per request size repeated over nz',
            'config': {'workload': 'batching form: %d requests, n in %s, nproj=1.5n, nz in [1,32], '
                                   'parzen; one HIP graph: per-request stages over 4 streams, '
                                   'one gridding launch for all requests (float32)'
                                   % (len(reqs), sizes),
                       'requests': len(reqs), 'slices': slices, 'tuned': bool(a.tune)},
            'per_request_graph': {'value': round(slices / t_graph, 2),
                                  'ms_per_step': round(t_graph * 1e3, 3),
                                  'note': 'every request its own launches, over 4 streams, '
                                          'one HIP graph'},
            'sequential_eager': {'value': round(slices / t_seq, 2),
                                 'ms_per_step': round(t_seq * 1e3, 3),
                                 'note': 'per-request launches from the host, one stream'},
            'graph_speedup': round(t_seq / t_val, 3)}
    if t_bgrid is not None:
        line['batched_grid_bitwise_equal'] = bool(same)
    print(json.dumps(line))


def bench_cg(a):
    """Differentiation form: n=1024, nproj=180 angles in [0, pi), nz=32, FP32;
    CG least squares with a Tikhonov term through the USFFT projector and its
    adjoint; reports iterations/s."""
    import torch
    from tomocupy_amd import FourierRec, synth
    from tomocupy_amd.autograd import cgls
    n, nproj, nz = 1024, 180, 32
    theta = synth.angles(nproj)
    t0 = time.time()
    sino = synth.sinograms(5, n, theta, 4)
    y = torch.from_numpy(np.concatenate([sino] * (nz // 4)).astype(np.float32)).cuda()
    print('# data in %.1fs' % (time.time() - t0), file=sys.stderr, flush=True)
    rec = FourierRec(n, nproj, nz, torch.from_numpy(theta).cuda(), 'float32')
    cgls(rec, y, lam=1e-3, niter=2)
    torch.cuda.synchronize()
    t = time.perf_counter()
    x = cgls(rec, y, lam=1e-3, niter=a.iters)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    line = {'metric': 'CG iterations/s (n=1024, nproj=180, nz=32, FP32)',
            'value': round(a.iters / dt, 2), 'unit': 'iterations/s', 'n_gpus': 1,
            'steps': a.iters, 'warmup': 2, 'ms_per_step': round(dt / a.iters * 1e3, 3),
            'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
            'data': 'synthetic ellipse sinograms (4 distinct slices repeated)',
            'config': {'workload': 'differentiation form: %d CG iterations on (A^T A + 1e-3 I) x '
                                   '= A^T y' % a.iters, 'n': n, 'nproj': nproj, 'nz': nz},
            'x_norm': float(torch.linalg.norm(x))}
    print(json.dumps(line))


def bench_steps(a):
    """recon_steps form: the whole scan resident (n, nproj, nz of the
    pipeline workload, default 2048 x 1500 x 2048), the reference's
    recon_steps passes (proc_sino by row chunks, Paganin phase retrieval by
    projection chunks with its test settings, filtered backprojection by row
    chunks) into a resident reconstruction; reports slices/s."""
    import torch
    from tomocupy_amd import _lib
    from tomocupy_amd.rec_steps import GPURecSteps
    theta, h_data, h_dark, h_flat = host_data(a)
    reps = -(-a.nz // a.chunk)
    t0 = time.time()
    raw = torch.from_numpy(h_data).cuda().repeat(1, reps, 1)[:, :a.nz].contiguous()
    dark = torch.from_numpy(h_dark).cuda().repeat(1, reps, 1)[:, :a.nz].contiguous()
    flat = torch.from_numpy(h_flat).cuda().repeat(1, reps, 1)[:, :a.nz].contiguous()
    rec = torch.empty((a.nz, a.n, a.n), dtype=_lib.torch_dtype(a.dtype), device='cuda')
    center = a.center if a.center is not None else a.n / 2 + 0.5
    kw = dict(energy=20.0, propagation_distance=60.0, pixel_size=1.17, retrieve_phase_alpha=1e-4)
    rs = GPURecSteps(a.n, a.nproj, a.nz, a.chunk, torch.from_numpy(theta).cuda(), center,
                     dark.shape[0], flat.shape[0], a.dtype, a.filter, nproj_per_chunk=8,
                     retrieve_phase_method='paganin', remove_stripe_method=a.stripe,
                     phase_batch=a.phase_batch, rotate_proj_angle=a.rotate_proj_angle,
                     rotate_proj_order=a.rotate_proj_order, **kw)
    torch.cuda.synchronize()
    print('# resident scan and plans in %.1fs' % (time.time() - t0), file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        rs.recon_steps_resident(raw, dark, flat, rec)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(a.steps):
        rs.recon_steps_resident(raw, dark, flat, rec)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / a.steps
    # per-kernel breakdown: one more, untimed pass with HIP events around
    # every launch (the events serialise launches, so not inside the timing)
    _lib.prof_reset()
    _lib.prof_enable(True)
    rs.recon_steps_resident(raw, dark, flat, rec)
    torch.cuda.synchronize()
    _lib.prof_enable(False)
    kern = {}
    for k in ('k_ph_rows_fwd', 'k_ph_cols', 'k_ph_rows_inv', 'k_pre_log_swap', 'k_fw_rows',
              'k_fr_grid_x3', 'k_pre_rows'):
        ms, cnt = _lib.prof_read(k)
        if cnt:
            kern[k] = {'avg_ms': round(ms / cnt, 5), 'launches': cnt, 'total_ms': round(ms, 3)}
    line = {'metric': 'recon_steps slices/s (n=%d, nproj=%d, nz=%d, Paganin)' % (a.n, a.nproj, a.nz),
            'value': round(a.nz / dt, 2), 'unit': 'slices/s', 'n_gpus': 1, 'steps': a.steps,
            'warmup': a.warmup, 'ms_per_step': round(dt * 1e3, 3), 'higher_is_better': True,
            'scaling': 'weak', 'vs_baseline': None, 'dtype': {'float32': 'f32', 'float16': 'f16',
                                                              'bfloat16': 'bf16'}[a.dtype],
            'data': 'synthetic ellipse-phantom counts tiled over nz, resident in HBM',
            'config': {'workload': 'recon_steps form: proc_sino (stripe=%s) by chunks of %d rows, '
                                   'Paganin (alpha 1e-4, 60 mm, 20 keV, 1.17 um) by chunks of 8 '
                                   'projections, %s%s filter + backprojection'
                                   % (a.stripe, a.chunk,
                                      ('rotation %g deg order %d, ' % (a.rotate_proj_angle,
                                                                        a.rotate_proj_order)
                                       if a.rotate_proj_angle else ''), a.filter),
                       'n': a.n, 'nproj': a.nproj, 'nz': a.nz, 'chunk': a.chunk},
            'kernels_untimed_breakdown_step': kern}
    rs.free()
    print(json.dumps(line))


class Pipeline:
    """Per-stream plans and buffers of the chunk pipeline in one storage dtype."""

    def __init__(self, a, torch, dtype, th, center, nchunks, ndark, nflat):
        from tomocupy_amd import BackprojFunctions, ProcFunctions, _lib
        self.a, self.torch, self.dtype = a, torch, dtype
        self.streams = [torch.cuda.Stream() for _ in range(a.streams)]
        td = _lib.torch_dtype(dtype)
        self.pfs, self.bfs, self.sinos = [], [], []
        for st in self.streams:
            with torch.cuda.stream(st):
                self.pfs.append(ProcFunctions(a.nproj, a.chunk, a.n, ndark, nflat, dtype,
                                              remove_stripe_method=a.stripe, fw_sigma=1.0,
                                              fw_filter='sym16', fw_level=a.fw_level,
                                              vo_all_dim=a.vo_all_dim))
                self.bfs.append(BackprojFunctions(a.n, a.nproj, a.chunk, th, center, dtype,
                                                  a.filter, rec_options={'fft': a.fft}))
                self.sinos.append(torch.empty((a.chunk, a.nproj, a.n), dtype=td, device='cuda'))
        self.rec = torch.empty((nchunks * a.chunk, a.n, a.n), dtype=td, device='cuda')
        torch.cuda.synchronize()

    def step(self, raw, dark, flat, nstreams=None, count=None):
        a, torch = self.a, self.torch
        nstreams = nstreams or len(self.streams)
        count = raw.shape[0] if count is None else count
        for k in range(count):
            i = k % nstreams
            with torch.cuda.stream(self.streams[i]):
                out = self.pfs[i].proc_chunk(raw[k], dark[k], flat[k], out=self.sinos[i])
                self.bfs[i].fbp_filter_backprojection(self.rec[k * a.chunk:(k + 1) * a.chunk],
                                                      out)

    def timed(self, raw, dark, flat, steps, warmup, label):
        torch = self.torch
        for _ in range(warmup):
            self.step(raw, dark, flat)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            self.step(raw, dark, flat)
            torch.cuda.synchronize()
            print('# %s step %d done' % (label, i), file=sys.stderr, flush=True)
        return (time.perf_counter() - t0) / steps

    def free(self):
        for pf, bf in zip(self.pfs, self.bfs):
            pf.free()
            bf.free()
        del self.rec


def conveyor(a, torch, theta, h_data, h_dark, h_flat, center, dtype):
    """The reference's three-stream conveyor (GPURec.recon_all) over host
    data: every chunk is copied from a host array into pinned memory,
    uploaded, processed and downloaded into pinned memory (no disk; the
    writer callback does nothing).  Returns slices/s over the whole volume."""
    from tomocupy_amd.rec import GPURec
    g = GPURec(a.n, a.nproj, a.nz, a.chunk, torch.from_numpy(theta).cuda(), center,
               h_dark.shape[0], h_flat.shape[0], dtype, fbp_filter=a.filter,
               remove_stripe_method=a.stripe, fw_sigma=1.0, fw_filter='sym16',
               fw_level=a.fw_level)

    def read_chunk(k):
        return h_data, h_dark, h_flat

    def write_chunk(k, rec):
        pass

    g.recon_all(read_chunk, write_chunk)  # warm-up (plans, pinned buffers)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.recon_all(read_chunk, write_chunk)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    g.free()
    up = a.nproj * a.chunk * a.n * 2
    down = a.chunk * a.n * a.n * (4 if dtype == 'float32' else 2)
    nch = -(-a.nz // a.chunk)
    return {'value': round(a.nz / dt, 2), 'unit': 'slices/s', 'seconds': round(dt, 3),
            'dtype': dtype, 'h2d_bytes': up * nch, 'd2h_bytes': down * nch,
            'pcie_gbs': round((up + down) * nch / dt / 1e9, 2),
            'note': 'GPURec.recon_all: host array -> pinned -> H2D (stream 1), chunk '
                    'processing (stream 2), D2H to pinned (stream 3), writer threads; '
                    'no disk; second of two runs'}


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU only (this build has no multi-GPU path); got --gpus %d'
              % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.form == 'batch':
        return bench_batch(a)
    if a.form == 'cg':
        return bench_cg(a)
    if a.form == 'steps':
        return bench_steps(a)
    center = a.center if a.center is not None else a.n / 2 + 0.5
    theta, h_data, h_dark, h_flat = host_data(a)
    cpu, oracle = None, None
    if a.cpu_baseline == 'on':
        print('# cpu baseline...', file=sys.stderr, flush=True)
        cpu, oracle = cpu_baseline(a, h_data, h_dark, h_flat, theta, center)
        print('# cpu baseline %s' % cpu, file=sys.stderr, flush=True)
    import torch
    from tomocupy_amd import _lib
    raw, dark, flat = device_data(a, torch, h_data, h_dark, h_flat)
    nchunks = raw.shape[0]
    th = torch.from_numpy(theta).cuda()
    t0 = time.time()
    pl = Pipeline(a, torch, a.dtype, th, center, nchunks, dark.shape[1], flat.shape[1])
    print('# plans ready in %.1fs' % (time.time() - t0), file=sys.stderr, flush=True)

    if a.copy_calibration_mb > 0:
        nb = int(a.copy_calibration_mb * 1e6) // 4
        src = torch.ones(nb, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(a.warmup):
        pl.step(raw, dark, flat)
    torch.cuda.synchronize()
    # per-kernel breakdown (untimed): the first ISO_CHUNKS chunks on ONE
    # stream with every kernel bracketed by events, i.e. isolated kernel
    # durations; it names the dominant kernel (largest total) and gives the
    # roofline.  (With several streams an event pair around a launch also
    # spans the wait behind the other streams' kernels, so event durations
    # under overlap are not kernel durations.)
    kern = {}
    dom = None
    roof = None
    sfx = {'float32': 'f32', 'float16': 'f16', 'bfloat16': 'bf16'}[a.dtype]
    if not a.no_prof:
        _lib.prof_only(None)
        _lib.prof_reset()
        _lib.prof_enable(True)
        pl.step(raw, dark, flat, 1, min(nchunks, ISO_CHUNKS))
        torch.cuda.synchronize()
        _lib.prof_enable(False)
        for k in KERNELS:
            ms, cnt = _lib.prof_read(k)
            if cnt:
                kern[k] = {'avg_ms': ms / cnt, 'launches': cnt, 'total_ms': ms}
                ab = algorithmic_bytes(k, a, None)
                if ab is not None:
                    kern[k]['hbm_frac'] = round(ab / (ms / cnt * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        if kern:
            dom = max(kern, key=lambda k: kern[k]['total_ms'])
            ab = algorithmic_bytes(dom, a, None)
            if ab is not None:
                ach = ab / (kern[dom]['avg_ms'] * 1e-3) / 1e9
                traffic = None
                tf = os.path.join(ROOT, 'profiles', 'traffic.json')
                if os.path.exists(tf):
                    traffic = json.load(open(tf)).get(
                        '%s@n%d_nproj%d_chunk%d_%s' % (dom, a.n, a.nproj, a.chunk, sfx))
                # end to end: every kernel's algorithmic bytes over the total
                # kernel time of the pass (kernels without a byte model, the
                # small reductions and table builders, count as time only)
                tot_b = sum((algorithmic_bytes(k, a, None) or 0) * v['launches']
                            for k, v in kern.items())
                tot_ms = sum(v['total_ms'] for v in kern.values())
                e2e = {'achieved': round(tot_b / (tot_ms * 1e-3) / 1e9, 1),
                       'frac': round(tot_b / (tot_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       'bytes_per_chunk': round(tot_b / min(nchunks, ISO_CHUNKS)),
                       'kernel_ms_per_chunk': round(tot_ms / min(nchunks, ISO_CHUNKS), 4)}
                roof = {'bound': 'hbm', 'kernel': dom, 'achieved': round(ach, 1),
                        'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                        'frac': round(ach / HBM_PEAK_GBS, 4), 'traffic': traffic,
                        'traffic_source': 'profiles/traffic.json (rocprofv3 FETCH_SIZE + '
                                          'WRITE_SIZE of this kernel and workload, scaled by '
                                          'a copy kernel of known size)',
                        'algorithmic_bytes_per_launch': ab,
                        'avg_launch_ms': round(kern[dom]['avg_ms'], 5),
                        'launches': kern[dom]['launches'], 'end_to_end': e2e,
                        'timing': 'HIP events around each launch, first %d chunks of the '
                                  'workload on one stream (untimed pass before the timed steps)'
                                  % min(nchunks, ISO_CHUNKS)}
    # timed region: events around the dominant kernel's launches only
    _lib.prof_reset()
    if dom is not None:
        _lib.prof_only(dom)
        _lib.prof_enable(True)
    dt = pl.timed(raw, dark, flat, a.steps, 0, a.dtype)
    _lib.prof_enable(False)
    _lib.prof_only(None)
    value = a.nz / dt

    roof_ov = None
    if dom is not None:
        ms, cnt = _lib.prof_read(dom)
        ab = algorithmic_bytes(dom, a, None)
        if ab is not None and cnt:
            avg_ms = ms / cnt
            roof_ov = {'kernel': dom, 'event_avg_ms': round(avg_ms, 5), 'launches': cnt,
                       'achieved': round(ab / (avg_ms * 1e-3) / 1e9, 1), 'unit': 'GB/s',
                       'note': 'events over the timed %d-stream steps: start-to-end on the '
                               'launch stream, including waits behind other streams'
                               % a.streams}
    acc = None
    if oracle is not None:
        print('# accuracy against the FP64 oracle', file=sys.stderr, flush=True)
        acc = accuracy(a, torch, pl.rec, oracle, pl.pfs[0], pl.bfs[0], raw, dark, flat, theta,
                       center, pl.sinos[0])
    pl.free()
    del pl
    torch.cuda.empty_cache()
    # the 16-bit storage form on the same workload, same run
    half = None
    if a.half != 'off' and a.dtype == 'float32':
        hd = a.half
        p16 = Pipeline(a, torch, hd, th, center, nchunks, dark.shape[1], flat.shape[1])
        dt16 = p16.timed(raw, dark, flat, a.steps, max(1, a.warmup), hd)
        half = {'dtype': hd, 'value': round(a.nz / dt16, 2), 'unit': 'slices/s',
                'ms_per_step': round(dt16 * 1e3, 3), 'speedup_vs_f32': round(dt / dt16, 3)}
        if oracle is not None:
            half['accuracy'] = accuracy(a, torch, p16.rec, oracle, p16.pfs[0], p16.bfs[0], raw,
                                        dark, flat, theta, center, p16.sinos[0])
        p16.free()
        del p16
        torch.cuda.empty_cache()
    conv = None
    if a.conveyor == 'on':
        print('# conveyor', file=sys.stderr, flush=True)
        conv = conveyor(a, torch, theta, h_data, h_dark, h_flat, center, a.dtype)
    line = {
        'metric': METRIC, 'value': round(value, 2), 'unit': 'slices/s', 'n_gpus': 1,
        'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': round(dt * 1e3, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': sfx,
        'data': 'synthetic: %d distinct seeded ellipse-phantom slices (uint16 counts, 10 dark, '
                '20 flat, column-gain stripes) tiled over nz' % min(a.distinct, a.chunk),
        'config': {'workload': 'fourierrec full chunk pipeline: n=%d nproj=%d nz=%d chunk=%d '
                               'streams=%d stripe=%s filter=%s centre=%g'
                               % (a.n, a.nproj, a.nz, a.chunk, a.streams, a.stripe, a.filter,
                                  center),
                   'n': a.n, 'nproj': a.nproj, 'nz': a.nz, 'chunk': a.chunk,
                   'streams': a.streams, 'stripe': a.stripe, 'filter': a.filter,
                   'fw_level': a.fw_level, 'fft': a.fft,
                   'timed_with_kernel_events': 'dominant kernel only' if dom else False},
        'roofline': roof, 'cpu_baseline': cpu, 'accuracy': acc, 'half': half,
        'conveyor': conv, 'roofline_overlapped_events': roof_ov,
        'kernels_untimed_breakdown_step': kern,
    }
    print(json.dumps(line))


if __name__ == '__main__':
    main()
