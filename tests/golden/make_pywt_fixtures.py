"""Write PyWavelets fixtures used to pin the Fourier-wavelet stripe-removal oracle.

Run with an interpreter that has PyWavelets (the container's
``/opt/conda/bin/python3.9`` has PyWavelets 1.1.1):

    /opt/conda/bin/python3.9 tests/golden/make_pywt_fixtures.py

Writes ``wavelets.json`` (decomposition/reconstruction taps of the wavelets the
reference's ``--fw-filter`` flag accepts, ``src/tomocupy/config.py:178-182``)
and ``dwt2_sym16.npz`` (seeded inputs with ``pywt.dwt2``/``pywt.idwt2`` outputs
in ``mode='symmetric'``, plus ``pywt.dwt_coeff_len``).  PyWavelets is a
third-party library, not the reference.
"""
import json
import os

import numpy as np
import pywt

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    taps = {}
    for name in ('haar', 'db5', 'sym5', 'sym16'):
        w = pywt.Wavelet(name)
        taps[name] = {'dec_lo': list(w.dec_lo), 'dec_hi': list(w.dec_hi),
                      'rec_lo': list(w.rec_lo), 'rec_hi': list(w.rec_hi)}
    meta = {'pywt_version': pywt.__version__, 'taps': taps}
    with open(os.path.join(HERE, 'wavelets.json'), 'w') as f:
        json.dump(meta, f, indent=1)

    rng = np.random.default_rng(20240101)
    out = {}
    shapes = [(45, 38), (40, 52), (20, 33)]
    for i, shp in enumerate(shapes):
        x = rng.standard_normal(shp)
        cA, (cH, cV, cD) = pywt.dwt2(x, 'sym16', mode='symmetric')
        out['x%d' % i] = x
        out['cA%d' % i] = cA
        out['cH%d' % i] = cH
        out['cV%d' % i] = cV
        out['cD%d' % i] = cD
        c2 = [rng.standard_normal(cA.shape) for _ in range(4)]
        out['in_cA%d' % i], out['in_cH%d' % i], out['in_cV%d' % i], out['in_cD%d' % i] = c2
        out['rec%d' % i] = pywt.idwt2((c2[0], (c2[1], c2[2], c2[3])), 'sym16', mode='symmetric')
    lens = np.array([[N, pywt.dwt_coeff_len(N, 32, mode='symmetric')] for N in range(1, 200)])
    out['coeff_len_sym16'] = lens
    np.savez_compressed(os.path.join(HERE, 'dwt2_sym16.npz'), **out)


if __name__ == '__main__':
    main()
