"""ctypes binding of libtomohip.so (the C ABI declared in ``include/tomohip.h``).

The library is required: there is no CPU or PyTorch fallback for any compute
entry point.  If ``libtomohip.so`` is missing, :func:`lib` raises and names the
build command.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get('TOMOHIP_LIB', os.path.join(HERE, 'libtomohip.so'))

TOMO_OK, TOMO_EINVAL, TOMO_EHIP, TOMO_ENOMEM, TOMO_EUNSUPPORTED = 0, -1, -2, -3, -4
DTYPES = {'float32': 0, 'float16': 1, 'bfloat16': 2}


def torch_dtype(name):
    """torch element type of the storage precision ``name``."""
    import torch
    return {'float32': torch.float32, 'float16': torch.float16,
            'bfloat16': torch.bfloat16}[name]

_c_i64 = ctypes.c_int64
_c_p = ctypes.c_void_p
_c_int = ctypes.c_int
_c_d = ctypes.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    'tomo_last_error': (ctypes.c_char_p, []),
    'tomo_version': (ctypes.c_char_p, []),
    'tomo_prof_enable': (_c_int, [_c_int]),
    'tomo_prof_only': (_c_int, [ctypes.c_char_p]),
    'tomo_prof_reset': (_c_int, []),
    'tomo_prof_read': (_c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(_c_i64)]),
    'tomo_dot_f64': (_c_int, [_c_p, _c_p, _c_i64, _c_p, _c_p]),
    'tomo_cg_step': (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_p]),
    'tomo_cg_direction': (_c_int, [_c_p, _c_p, _c_i64, _c_p, _c_p, _c_p]),
    'fr_create': (_c_int, [ctypes.POINTER(_c_p), _c_i64, _c_i64, _c_i64, _c_p, _c_int, _c_p]),
    'fr_create_opts': (_c_int, [ctypes.POINTER(_c_p), _c_i64, _c_i64, _c_i64, _c_p, _c_int, _c_int,
                                _c_int, _c_p]),
    'fr_filter_backprojection_prepared': (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_p]),
    'fr_batch_create': (_c_int, [ctypes.POINTER(_c_p), ctypes.POINTER(_c_p), _c_int]),
    'fr_batch_filter_backprojection': (_c_int, [_c_p, ctypes.POINTER(_c_p), ctypes.POINTER(_c_p),
                                                ctypes.POINTER(_c_p), _c_p]),
    'fr_batch_destroy': (None, [_c_p]),
    'flt_prepare': (_c_int, [_c_p, _c_p, _c_d, _c_p, _c_p]),
    'fr_backprojection': (_c_int, [_c_p, _c_p, _c_p, _c_p]),
    'fr_forward': (_c_int, [_c_p, _c_p, _c_p, _c_p]),
    'fr_filter_backprojection': (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_d, _c_p, _c_p]),
    'fr_destroy': (None, [_c_p]),
    'fr_get_sizes': (_c_int, [_c_p, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64),
                              ctypes.POINTER(_c_i64)]),
    'flt_create': (_c_int, [ctypes.POINTER(_c_p), _c_i64, _c_i64, _c_i64, _c_i64, _c_int]),
    'flt_filter': (_c_int, [_c_p, _c_p, _c_p, _c_p]),
    'flt_filter_center': (_c_int, [_c_p, _c_p, _c_p, _c_d, _c_p, _c_p]),
    'flt_destroy': (None, [_c_p]),
    'flt_calc_filter': (_c_int, [_c_i64, ctypes.c_char_p, _c_p]),
    'pre_create': (_c_int, [ctypes.POINTER(_c_p), _c_i64, _c_i64, _c_i64, _c_i64, _c_i64]),
    'pre_process': (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_int, _c_d, _c_int, _c_int, _c_p]),
    'pre_process_noalloc': (_c_int, [_c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p,
                                     _c_p, _c_int, _c_d, _c_int, _c_int, _c_p]),
    'pre_darkflat': (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_d, _c_int, _c_p]),
    'pre_minus_log_swap': (_c_int, [_c_p, _c_p, _c_p, _c_int, _c_int, _c_p]),
    'pre_remove_stripe_fw': (_c_int, [_c_p, _c_p, _c_d, ctypes.c_char_p, _c_int, _c_p]),
    'pre_process_fw': (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_int, _c_d, _c_int, _c_int, _c_d,
                                ctypes.c_char_p, _c_int, _c_p]),
    'pre_remove_outliers': (_c_int, [_c_p, _c_p, _c_i64, _c_int, _c_int, _c_p]),
    'pre_remove_outliers_f32': (_c_int, [_c_p, _c_p, _c_i64, _c_int, _c_int, _c_p]),
    'pre_process_f32': (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_int, _c_d, _c_int, _c_int,
                                 _c_p]),
    'pre_darkflat_f32': (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_d, _c_int, _c_p]),
    'pre_bin': (_c_int, [_c_p, _c_p, _c_p, _c_i64, _c_int, _c_p]),
    'pre_remove_all_stripe': (_c_int, [_c_p, _c_p, _c_d, _c_int, _c_int, _c_int, _c_p]),
    'ph_create': (_c_int, [ctypes.POINTER(_c_p), _c_i64, _c_i64, _c_i64, _c_int, _c_d, _c_d, _c_d,
                           _c_d, _c_d, _c_d]),
    'ph_get_pads': (_c_int, [_c_p, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64),
                             ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64)]),
    'ph_apply': (_c_int, [_c_p, _c_p, _c_i64, _c_p]),
    'ph_destroy': (None, [_c_p]),
    'pre_rotate_proj': (_c_int, [_c_p, _c_p, _c_i64, _c_i64, _c_i64, _c_d, _c_int, _c_p, _c_p]),
    'pre_minus_log_swap_rows': (_c_int, [_c_p, _c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_int, _c_int,
                                         _c_p]),
    'pre_remove_stripe_ti': (_c_int, [_c_p, _c_p, _c_d, _c_d, _c_p]),
    'pre_pad360': (_c_int, [_c_p, _c_p, _c_p, _c_int, _c_d, _c_int, _c_p]),
    'pre_destroy': (None, [_c_p]),
}

_lock = threading.Lock()
_lib = None


class TomoError(RuntimeError):
    """A libtomohip entry point returned a negative status."""

    def __init__(self, what, status, message):
        self.status = status
        super().__init__('%s failed (status %d): %s' % (what, status, message))


class TomoUnsupported(TomoError):
    pass


def lib():
    """Load libtomohip.so once and bind every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIBPATH):
                raise ImportError('libtomohip.so not found at %s; build it with '
                                  '`python -m tomocupy_amd.build`' % LIBPATH)
            L = ctypes.CDLL(LIBPATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def last_error():
    return lib().tomo_last_error().decode(errors='replace')


def check(rc, what):
    if rc != TOMO_OK:
        cls = TomoUnsupported if rc == TOMO_EUNSUPPORTED else TomoError
        raise cls(what, rc, last_error())


def stream_ptr(stream):
    """hipStream_t (as int) for None (torch's current stream), a
    ``torch.cuda.Stream``, any object with ``.ptr`` (CuPy-style, as the
    reference passes), or a raw integer."""
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    if hasattr(stream, 'cuda_stream'):
        return stream.cuda_stream
    if hasattr(stream, 'ptr'):
        return int(stream.ptr)
    raise TypeError('unsupported stream object %r' % (stream,))


def dev_ptr(t, shape, dtype, name):
    """Validate a torch tensor (device, dtype, contiguity, shape) and return its address."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError('%s must be a torch tensor' % name)
    if not t.is_cuda:
        raise ValueError('%s must be a GPU tensor' % name)
    if t.dtype != dtype:
        raise TypeError('%s must be %s, got %s' % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError('%s must be contiguous' % name)
    if tuple(t.shape) != tuple(shape):
        raise ValueError('%s has shape %s, expected %s' % (name, tuple(t.shape), tuple(shape)))
    return t.data_ptr()


def calc_filter(ne, name):
    """Host FP64 filter design (``flt_calc_filter``); returns float64 numpy."""
    out = np.zeros(ne // 2 + 1, dtype=np.float64)
    check(lib().flt_calc_filter(ne, name.encode(), out.ctypes.data), 'flt_calc_filter')
    return out


def prof_enable(on=True):
    lib().tomo_prof_enable(1 if on else 0)


def prof_only(kernel=None):
    """Restrict kernel timing to one kernel name (None: every kernel)."""
    lib().tomo_prof_only(kernel.encode() if kernel else None)


def prof_reset():
    lib().tomo_prof_reset()


def prof_read(kernel):
    """(total milliseconds, launches) of one kernel since the last reset."""
    ms = ctypes.c_double()
    cnt = _c_i64()
    check(lib().tomo_prof_read(kernel.encode(), ctypes.byref(ms), ctypes.byref(cnt)),
          'tomo_prof_read')
    return ms.value, cnt.value
