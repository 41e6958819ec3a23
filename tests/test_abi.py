"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, validates arguments before any HIP call, and the host-side
filter design matches the FP64 oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from tomocupy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, 'include', 'tomohip.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b([a-z_][a-z0-9_]*)\s*\(', src)) - {'if', 'sizeof'})


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = header_functions()
    assert 'fr_backprojection' in names and 'flt_filter_center' in names
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_version():
    assert b'gfx950' in _lib.lib().tomo_version()


@pytest.mark.parametrize("n,nproj,nz,dtype,msg", [
    (21, 10, 2, 0, 'even'), (18, 10, 2, 0, '>= 20'), (64, 0, 2, 0, 'nproj'),
    (64, 10, 0, 0, 'nz'), (64, 10, 2, 7, 'dtype'), (8192, 10, 2, 0, 'maximum'),
    (262, 10, 2, 0, 'prime factor'),  # 262 = 2 x 131
    # 2n = 8000 = 2^6 5^3 and 8160 = 2^5 3 5 17: 15 values per thread in the
    # radix-3/5 passes -> 576 and 576 threads in the row pass (limit 512)
    (4000, 10, 2, 0, 'more than 512 threads'), (4080, 10, 2, 0, 'more than 512 threads'),
    (2838, 10, 2, 0, 'more than 512 threads')])  # the smallest rejected width
def test_fr_create_validation(n, nproj, nz, dtype, msg):
    h = ctypes.c_void_p()
    rc = _lib.lib().fr_create(ctypes.byref(h), n, nproj, nz, ctypes.c_void_p(16), dtype, None)
    assert rc == _lib.TOMO_EINVAL
    assert msg in _lib.last_error()
    assert not h.value


def test_null_arguments():
    L = _lib.lib()
    assert L.fr_create(None, 64, 10, 2, ctypes.c_void_p(16), 0, None) == _lib.TOMO_EINVAL
    assert L.fr_create(ctypes.byref(ctypes.c_void_p()), 64, 10, 2, None, 0, None) == _lib.TOMO_EINVAL
    assert L.fr_backprojection(None, None, None, None) == _lib.TOMO_EINVAL
    assert L.fr_forward(None, None, None, None) == _lib.TOMO_EINVAL
    assert L.flt_filter(None, None, None, None) == _lib.TOMO_EINVAL
    assert L.flt_filter_center(None, None, None, 0.0, None, None) == _lib.TOMO_EINVAL
    assert L.pre_process(None, None, None, None, None, 0, 1.0, 0, 1, None) == _lib.TOMO_EINVAL
    assert L.pre_minus_log_swap(None, None, None, 0, 1, None) == _lib.TOMO_EINVAL
    L.fr_destroy(None)
    L.flt_destroy(None)
    L.pre_destroy(None)


def test_fr_batch_validation():
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.fr_batch_create(None, None, 1) == _lib.TOMO_EINVAL
    assert L.fr_batch_create(ctypes.byref(h), None, 1) == _lib.TOMO_EINVAL
    plans = (ctypes.c_void_p * 2)(None, None)
    assert L.fr_batch_create(ctypes.byref(h), plans, 0) == _lib.TOMO_EINVAL
    assert L.fr_batch_create(ctypes.byref(h), plans, 2) == _lib.TOMO_EINVAL
    assert 'plan 0 is NULL' in _lib.last_error()
    assert not h.value
    assert L.fr_batch_filter_backprojection(None, None, None, None, None) == _lib.TOMO_EINVAL
    L.fr_batch_destroy(None)


def test_flt_create_validation():
    h = ctypes.c_void_p()
    L = _lib.lib()
    assert L.flt_create(ctypes.byref(h), 64, 60, 10, 2, 0) == _lib.TOMO_EINVAL  # ne < n
    assert L.flt_create(ctypes.byref(h), 64, 64 * 131, 10, 2, 0) == _lib.TOMO_EINVAL  # 131 | ne
    assert 'prime factors' in _lib.last_error()
    assert L.flt_create(ctypes.byref(h), 64, 256, 10, 2, 9) == _lib.TOMO_EINVAL


def test_pre_create_validation():
    h = ctypes.c_void_p()
    L = _lib.lib()
    assert L.pre_create(ctypes.byref(h), 10, 2, 64, 0, 5) == _lib.TOMO_EINVAL


@pytest.mark.parametrize("name", ['none', 'ramp', 'shepp', 'cosine', 'cosine2', 'hamming',
                                  'hann', 'parzen'])
@pytest.mark.parametrize("ne", [256, 2048, 8192])
def test_calc_filter_matches_oracle(name, ne):
    from oracle import filter_fp64 as ff
    w = _lib.calc_filter(ne, name)
    ref = ff.calc_filter(ne, name)
    # the 12x12 Vandermonde inverse is ill-conditioned (~1e9): two LAPACK-style
    # inverses differ at ~1e-7 of the largest weight in the first few bins (the
    # reference's CuPy/cuSOLVER inverse is no closer to numpy's)
    assert np.max(np.abs(w - ref)) <= 5e-7 * np.max(np.abs(ref))


def test_calc_filter_rejects_unknown():
    with pytest.raises(_lib.TomoError):
        _lib.calc_filter(256, 'bogus')


def test_rotate_proj_validation():
    # argument checks of pre_rotate_proj run before any HIP call
    L = _lib.lib()
    a, b, w = ctypes.c_void_p(16), ctypes.c_void_p(4096), ctypes.c_void_p(1 << 20)
    assert L.pre_rotate_proj(a, b, 1, 8, 8, 3.0, 6, w, None) == _lib.TOMO_EINVAL
    assert 'order' in _lib.last_error()
    assert L.pre_rotate_proj(a, b, 1, 8, 8, 3.0, -1, w, None) == _lib.TOMO_EINVAL
    assert L.pre_rotate_proj(a, b, 1, 8, 8, 3.0, 3, None, None) == _lib.TOMO_EINVAL
    assert 'work' in _lib.last_error()
    assert L.pre_rotate_proj(a, b, 1, 8, 8, 3.0, 2, a, None) == _lib.TOMO_EINVAL
    assert L.pre_rotate_proj(a, b, 1, 8, 8, 3.0, 5, b, None) == _lib.TOMO_EINVAL
    assert L.pre_rotate_proj(a, a, 1, 8, 8, 3.0, 1, None, None) == _lib.TOMO_EINVAL
    assert L.pre_rotate_proj(a, b, 1, 8, 8, float('nan'), 1, None, None) == _lib.TOMO_EINVAL
