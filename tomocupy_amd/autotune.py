"""Plan autotuner for the fourierrec path.

For one reconstruction shape (n, nproj, nz, dtype) it times the filtered
backprojection (``fr_filter_backprojection``) under each candidate plan
configuration on synthetic data of that shape and returns the fastest:

* ``grid_chunk``: the longest sample list one gridding wave walks before the
  tile is split into partial tiles that a reduction pass sums (64..2048
  entries); short chunks balance the heavily loaded tiles near the spectrum's
  centre, long ones avoid partial-tile traffic;
* ``fft``: compiled size-specialised transforms (``'static'``, power-of-two
  sizes 1024..8192) or run-time radix plans (``'generic'``).

Both options change only the rounding of the results (the order in which the
partial sums of a heavy tile are added, the radix sequence of a transform);
each configuration is deterministic.  Timings use HIP events on the stream the
kernels run on.  ``tune`` caches its choice per shape in a JSON file
(``TOMOHIP_TUNE_CACHE``, default ``~/.cache/tomocupy_amd/tune.json``).
"""
import json
import os

from . import _lib

GRID_CHUNKS = (128, 256, 512, 1024, 2048)
FFTS = ('static', 'generic')


def _cache_path():
    return os.environ.get('TOMOHIP_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'tomocupy_amd',
                                       'tune.json'))


def _key(n, nproj, nz, dtype):
    return '%d_%d_%d_%s' % (n, nproj, nz, dtype)


def time_config(n, nproj, nz, theta, dtype, options, reps=5, center=None, filter_name='ramp'):
    """Median milliseconds of one filtered backprojection of ``nz`` slices
    with plan ``options`` (a dict of FourierRec keyword arguments)."""
    import torch
    from .backproj_functions import BackprojFunctions
    td = _lib.torch_dtype(dtype)
    center = n / 2 if center is None else center
    bf = BackprojFunctions(n, nproj, nz, theta, center, dtype, filter_name, rec_options=options)
    g = torch.Generator(device='cuda').manual_seed(1)
    data = torch.randn((nz, nproj, n), device='cuda', generator=g).to(td)
    obj = torch.empty((nz, n, n), dtype=td, device='cuda')
    stream = torch.cuda.current_stream()
    bf.fbp_filter_backprojection(obj, data)  # warm-up
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        bf.fbp_filter_backprojection(obj, data)
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    bf.free()
    times.sort()
    return times[len(times) // 2]


def tune(n, nproj, nz, theta=None, dtype='float32', grid_chunks=GRID_CHUNKS, ffts=FFTS,
         reps=5, use_cache=True, verbose=False):
    """Fastest plan options for this shape: ``{'grid_chunk': c, 'fft': f}``.

    ``theta``: angles (numpy or torch); the default is nproj angles uniform in
    [0, pi).  The sweep is ``grid_chunks x ffts``.
    """
    import numpy as np
    import torch
    key = _key(n, nproj, nz, dtype)
    path = _cache_path()
    if use_cache and os.path.exists(path):
        try:
            cached = json.load(open(path)).get(key)
        except (OSError, ValueError):
            cached = None
        if cached:
            return cached['options']
    if theta is None:
        theta = np.arange(nproj, dtype=np.float32) * np.float32(np.pi / nproj)
    th = torch.as_tensor(np.asarray(theta, dtype=np.float32)).cuda()
    results = []
    for f in ffts:
        for c in grid_chunks:
            opts = {'grid_chunk': int(c), 'fft': f}
            ms = time_config(n, nproj, nz, th, dtype, opts, reps=reps)
            results.append((ms, opts))
            if verbose:
                print('tune %s: %s -> %.3f ms' % (key, opts, ms), flush=True)
    results.sort(key=lambda r: r[0])
    best_ms, best = results[0]
    if use_cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            try:
                allc = json.load(open(path))
            except (OSError, ValueError):
                allc = {}
            allc[key] = {'options': best, 'ms': best_ms,
                         'sweep': [{'options': o, 'ms': m} for m, o in results]}
            json.dump(allc, open(path, 'w'), indent=1, sort_keys=True)
        except OSError:
            pass
    return best
