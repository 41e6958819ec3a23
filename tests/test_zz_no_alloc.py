"""The allocation-free preprocessing entry point against the plan-based one.

pre_process_noalloc runs the kernels of pre_process with the dark/flat means
in a static device array instead of plan-owned scratch, so its results must
equal pre_process's bit for bit.  Two calls in a row, the second with a
smaller input (the array keeps the first call's means), plus flat_linear and
16-bit outputs.
"""
import ctypes

import numpy as np
import pytest

from tomocupy_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _inputs(seed, nproj, nz, n, ndark, nflat):
    rng = np.random.default_rng(seed)
    flat = rng.integers(20000, 40000, (nflat, nz, n), dtype=np.uint16)
    dark = rng.integers(90, 110, (ndark, nz, n), dtype=np.uint16)
    data = rng.integers(0, 40000, (nproj, nz, n), dtype=np.uint16)
    data[::7, :, ::5] = 0  # zero counts: the -log clamp rule
    dev = torch.device('cuda')
    return [torch.from_numpy(a).to(dev) for a in (data, dark, flat)]


def _plan_process(data, dark, flat, out_dtype, flat_linear, minus_log):
    L = _lib.lib()
    nproj, nz, n = data.shape
    h = ctypes.c_void_p()
    _lib.check(L.pre_create(ctypes.byref(h), nproj, nz, n, dark.shape[0], flat.shape[0]),
               'pre_create')
    out = torch.empty((nz, nproj, n), dtype=_lib.torch_dtype(out_dtype), device='cuda')
    try:
        _lib.check(L.pre_process(h, data.data_ptr(), dark.data_ptr(), flat.data_ptr(),
                                 out.data_ptr(), _lib.DTYPES[out_dtype], 1.0, flat_linear,
                                 minus_log, _lib.stream_ptr(None)), 'pre_process')
        torch.cuda.synchronize()
    finally:
        L.pre_destroy(h)
    return out


def _noalloc_process(data, dark, flat, out_dtype, flat_linear, minus_log):
    L = _lib.lib()
    nproj, nz, n = data.shape
    out = torch.empty((nz, nproj, n), dtype=_lib.torch_dtype(out_dtype), device='cuda')
    torch.cuda.synchronize()  # the allocation above is not part of the call
    _lib.check(L.pre_process_noalloc(nproj, nz, n, dark.shape[0], flat.shape[0],
                                     data.data_ptr(), dark.data_ptr(), flat.data_ptr(),
                                     out.data_ptr(), _lib.DTYPES[out_dtype], 1.0, flat_linear,
                                     minus_log, _lib.stream_ptr(None)), 'pre_process_noalloc')
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("first,second", [
    # (seed, nproj, nz, n, ndark, nflat, out_dtype, flat_linear, minus_log)
    ((1, 300, 8, 512, 10, 20, 'float32', 0, 1), (2, 120, 4, 256, 3, 5, 'float32', 0, 1)),
    ((3, 200, 16, 640, 4, 6, 'float32', 1, 1), (4, 90, 2, 130, 2, 3, 'float16', 1, 0)),
    ((5, 64, 1024, 1024, 2, 3, 'bfloat16', 0, 1), (6, 33, 3, 70, 1, 1, 'float32', 0, 1)),
])
def test_noalloc_matches_plan_twice(first, second):
    for seed, nproj, nz, n, ndark, nflat, odt, fl, ml in (first, second):
        print('no-alloc case', seed, nproj, nz, n, ndark, nflat, odt, fl, ml, flush=True)
        data, dark, flat = _inputs(seed, nproj, nz, n, ndark, nflat)
        ref = _plan_process(data, dark, flat, odt, fl, ml)
        got = _noalloc_process(data, dark, flat, odt, fl, ml)
        assert torch.equal(got.view(torch.int16) if odt != 'float32' else got.view(torch.int32),
                           ref.view(torch.int16) if odt != 'float32' else ref.view(torch.int32))


def test_noalloc_rejects_oversized_chunk():
    L = _lib.lib()
    p = ctypes.c_void_p(256)
    rc = L.pre_process_noalloc(10, 1024, 1025, 1, 1, p, p, p, p, 0, 1.0, 0, 1, None)
    assert rc == _lib.TOMO_EINVAL and 'static scratch' in _lib.last_error()
