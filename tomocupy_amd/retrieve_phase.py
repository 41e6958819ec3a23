"""Single-step phase retrieval with the reference's interface.

Mirrors ``src/tomocupy/processing/retrieve_phase.py``: ``paganin_filter``
(:59-104, methods ``paganin`` and ``Gpaganin``) and ``farago_filter``
(:106-141), in place on float32 GPU projections ``(nproj, dy, dz)``; the
work runs in ``ph_apply`` (HIP).  Plans (pads, reciprocal grid, the filter's
maximum, the work buffer) are cached per shape and parameters.  Units as the
reference's functions: pixel size and distance in cm, energy in keV (the
reference's ``proc_proj`` converts its CLI's micrometres and millimetres,
``proc_functions.py:171-174``).  Both dimensions must be even; ``pad`` is
always on (the reference's callers never turn it off).
"""
import ctypes

from . import _lib

METHODS = {'paganin': 0, 'Gpaganin': 1, 'farago': 2}
_plans = {}


class PhasePlan:
    def __init__(self, dy, dz, method='paganin', pixel_size=1e-4, dist=50, energy=20,
                 alpha=1e-3, db=1000, W=2e-4, batch=4):
        if method not in METHODS:
            raise ValueError('method must be one of %s' % sorted(METHODS))
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().ph_create(ctypes.byref(h), int(dy), int(dz), int(batch),
                                        METHODS[method], float(pixel_size), float(dist),
                                        float(energy), float(alpha), float(db), float(W)),
                   'ph_create')
        self._h = h
        self.dy, self.dz = int(dy), int(dz)
        v = [ctypes.c_int64() for _ in range(4)]
        _lib.check(_lib.lib().ph_get_pads(h, *[ctypes.byref(x) for x in v]), 'ph_get_pads')
        self.py, self.pz, self.Ny, self.Nx = (x.value for x in v)

    def apply(self, data):
        import torch
        if data.dim() != 3 or tuple(data.shape[1:]) != (self.dy, self.dz):
            raise ValueError('expected (nproj, %d, %d) projections, got %s'
                             % (self.dy, self.dz, tuple(data.shape)))
        _lib.check(_lib.lib().ph_apply(
            self._h, _lib.dev_ptr(data, tuple(data.shape), torch.float32, 'data'),
            int(data.shape[0]), _lib.stream_ptr(None)), 'ph_apply')
        return data

    def free(self):
        h, self._h = getattr(self, '_h', None), None
        if h is not None and h.value:
            _lib.lib().ph_destroy(h)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _plan(data, method, pixel_size, dist, energy, alpha, db, W):
    key = (tuple(data.shape[1:]), data.device.index, method, float(pixel_size), float(dist),
           float(energy), float(alpha), float(db), float(W))
    p = _plans.get(key)
    if p is None:
        p = _plans[key] = PhasePlan(data.shape[1], data.shape[2], method, pixel_size, dist,
                                    energy, alpha, db, W)
    return p


def paganin_filter(data, pixel_size=1e-4, dist=50, energy=20, alpha=1e-3, method='paganin',
                   db=1000, W=2e-4, pad=True):
    """Paganin (``method='paganin'``) or generalised Paganin (``'Gpaganin'``)
    phase retrieval in place (``retrieve_phase.py:59-104``)."""
    if not pad:
        raise NotImplementedError('pad=False is not supported')
    return _plan(data, method, pixel_size, dist, energy, alpha, db, W).apply(data)


def farago_filter(data, pixel_size=1e-4, dist=50, energy=20, db=1000, method='farago', pad=True):
    """Farago phase retrieval in place (``retrieve_phase.py:106-141``)."""
    if not pad:
        raise NotImplementedError('pad=False is not supported')
    return _plan(data, 'farago', pixel_size, dist, energy, 0.0, db, 2e-4).apply(data)


def clear_plans():
    for p in _plans.values():
        p.free()
    _plans.clear()
