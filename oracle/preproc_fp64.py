"""FP64 restatement of tomocupy's chunk preprocessing on the fourierrec path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Citations: ``src/tomocupy/processing/proc_functions.py`` (``P``),
``src/tomocupy/processing/remove_stripe.py`` (``R``).

Wavelet taps are PyWavelets 1.1.1 values recorded in
``tests/golden/wavelets.json`` (PyWavelets is a third-party dependency of the
reference, ``setup.py:17``; it is not importable in the test interpreter).

Parity status: ``darkflat_correction`` and ``minus_log`` are closed forms
checked on hand-made cases (clamp, NaN and inf rules); the DWT analysis and
synthesis (``afb1d``/``sfb1d``) are pinned against PyWavelets ``dwt2``/``idwt2``
golden vectors and perfect reconstruction; the damping is a closed form
(``tests/test_oracle_preproc.py``).  ``remove_outliers`` is
``scipy.ndimage.median_filter`` (the CPU twin of the ``cupyx`` routine the
reference calls) plus the replacement rule, pinned by a brute-force window
median on small arrays.  ``downsample`` (binning) is pinned by block sums
through a reshape and by the conservation of the total.  The ``vo-all``
functions (``detect_stripe``, ``rs_sort``, ``rs_large``, ``rs_dead``,
``remove_all_stripe``) use SciPy's ``median_filter``, ``uniform_filter1d`` and
``binary_dilation`` (the CPU twins of the ``cupyx`` routines the reference
calls) and a stable argsort; they are pinned by ``np.polyfit``, closed-form
stripe cases and a pure-Python sort/median/unsort (``tests/test_oracle_vo_all.py``).
"""

import json
import os

import numpy as np

_TAPS = None


def wavelet_taps(name):
    """``(dec_lo, dec_hi, rec_lo, rec_hi)`` of a PyWavelets wavelet, from the fixture."""
    global _TAPS
    if _TAPS is None:
        here = os.path.dirname(os.path.abspath(__file__))
        with open(os.path.join(here, '..', 'tests', 'golden', 'wavelets.json')) as f:
            _TAPS = json.load(f)['taps']
    t = _TAPS[name]
    return tuple(np.array(t[k], dtype=np.float64) for k in ('dec_lo', 'dec_hi', 'rec_lo', 'rec_hi'))


def remove_stripe_ti(data, beta, mask_size):
    """``remove_stripe_ti`` (R:296-309) on ``data (nproj, nz, ni)``, FP64, even
    ``ni`` (the reference's irfft returns ``ni - 1`` samples for odd ``ni``).
    The mask covers the central ``floor(mask_size * ni)`` columns (the
    reference slices with ``mask_size * ni``, a float for its float option)."""
    data = np.asarray(data, dtype=np.float64)
    ni = data.shape[-1]
    if ni % 2:
        raise ValueError('remove_stripe_ti needs an even detector width')
    gamma = beta * ((1 - beta) / (1 + beta)) ** np.abs(np.fft.fftfreq(ni) * ni)
    gamma[0] -= 1
    v = data.mean(axis=0)
    v = v - v[:, 0:1]
    v = np.fft.irfft(np.fft.rfft(v) * np.fft.rfft(gamma), ni)
    mask = np.zeros(v.shape)
    m = int(np.floor(mask_size * ni))
    mask[:, ni // 2 - m // 2:ni // 2 + m // 2] = 1
    return data + v * mask


def pad360(sino, centeri):
    """``ProcFunctions.pad360`` (P:109-124) on sinograms ``(nz, nproj, ni)``
    with the centre rule of ``reader.py:143-148``: returns ``(padded
    (nz, nproj, 2 ni) FP64, centre of the padded sinogram, flipped)``."""
    sino = np.asarray(sino, dtype=np.float64)
    ni = sino.shape[-1]
    flip = centeri < ni // 2
    center = ni - centeri if flip else centeri
    data = (sino[..., ::-1] if flip else sino).copy()
    w = max(1, int(2 * (ni - center)))
    v = np.linspace(1, 0, w, endpoint=False)
    v = v ** 5 * (126 - 420 * v + 540 * v ** 2 - 315 * v ** 3 + 70 * v ** 4)
    data[..., -w:] *= v
    return np.pad(data, ((0, 0), (0, 0), (0, ni)), 'constant'), center, flip


def remove_outliers(data, width, threshold):
    """``ProcFunctions.remove_outliers`` (P:96-107) on a ``(rows, nz, n)``
    array of counts, uint16 or float (binned): median over ``[width, 1,
    width]`` ('reflect' borders), then ``where(data > f and data - f >
    threshold, f, data)``.  Returns a copy."""
    from scipy import ndimage
    data = np.asarray(data)
    if int(width) <= 0:
        return data.copy()
    w = int(width)
    f = ndimage.median_filter(data, [w, 1, w])
    wide = np.float64 if data.dtype.kind == 'f' else np.int64
    d = data.astype(wide)
    fi = f.astype(wide)
    return np.where((d > fi) & (d - fi > threshold), f, data).astype(data.dtype)


def downsample(data, binning):
    """``utils.downsample`` (``src/tomocupy/utils.py:157-167``): ``binning``
    halvings of the detector columns (sums of neighbour pairs), then as many of
    the rows, on ``(frames, rows, cols)``; FP64 sums."""
    data = np.asarray(data, dtype=np.float64)
    for _ in range(int(binning)):
        data = data[:, :, ::2] + data[:, :, 1::2]
    for _ in range(int(binning)):
        data = data[:, ::2] + data[:, 1::2]
    return data


def darkflat_correction(data, dark, flat, bright_ratio=1.0, flat_linear=False):
    """``ProcFunctions.darkflat_correction`` (P:55-74).

    ``data (nproj, nz, n)``, ``dark (ndark, nz, n)``, ``flat (nflat, nz, n)``.
    ``flat0 = mean(flat)/bright_ratio * (1+1e-5) - mean(dark)``;
    ``res = (data - dark0)/flat0``.  With ``flat_linear`` (P:62-66) the flat is
    interpolated per projection between the means of the first half and of the
    second half (skipping the middle frame), as in the reference.
    """
    data = np.asarray(data, dtype=np.float64)
    dark0 = np.asarray(dark, dtype=np.float64)
    flat0 = np.asarray(flat, dtype=np.float64) / bright_ratio
    if flat_linear:
        nproj = data.shape[0]
        f0 = np.mean(flat0[:flat0.shape[0] // 2], axis=0)
        f1 = np.mean(flat0[flat0.shape[0] // 2 + 1:], axis=0)
        v = np.linspace(0, 1, nproj)[:, None, None]
        flat0 = (1 - v) * f0 + v * f1
    else:
        flat0 = np.mean(flat0, axis=0)
    dark0 = np.mean(dark0, axis=0)
    flat0 = flat0 * np.float64(np.float32(1 + 1e-5))
    flat0 = flat0 - dark0
    with np.errstate(divide='ignore', invalid='ignore'):
        return (data - dark0) / flat0


def minus_log(data):
    """``ProcFunctions.minus_log`` (P:76-83): ``x<=0 -> 1``; ``-log``;
    NaN -> 6, +inf -> 0, -inf -> 0.  Returns a new array."""
    data = np.where(data <= 0, 1.0, data)
    with np.errstate(divide='ignore', invalid='ignore'):
        data = -np.log(data)
    return np.nan_to_num(data, nan=6.0, posinf=0.0, neginf=0.0)


def dwt_coeff_len(N, L):
    """PyWavelets ``dwt_coeff_len(N, L, 'symmetric')`` = ``floor((N+L-1)/2)``."""
    return (N + L - 1) // 2


def _reflect_index(idx, l):
    """``_reflect(x, -0.5, l-0.5)`` on integer indices (R:51-68): half-sample
    symmetric extension."""
    minx, maxx = -0.5, l - 0.5
    rng = maxx - minx
    rng2 = 2 * rng
    mod = np.fmod(idx - minx, rng2)
    mod = np.where(mod < 0, mod + rng2, mod)
    out = np.where(mod >= rng, rng2 - mod, mod) + minx
    return out.astype(np.int64)


def afb1d(x, h0, h1, axis):
    """Analysis filter bank (R:91-129): symmetric pad, stride-2 correlation with
    the reversed decomposition taps ``h0``/``h1``.  Returns ``(lo, hi)``."""
    x = np.moveaxis(np.asarray(x, dtype=np.float64), axis, -1)
    N = x.shape[-1]
    L = len(h0)
    outsize = dwt_coeff_len(N, L)
    p = 2 * (outsize - 1) - N + L
    idx = _reflect_index(np.arange(-(p // 2), N + (p + 1) // 2), N)
    xp = x[..., idx]
    lo = np.zeros(x.shape[:-1] + (outsize,))
    hi = np.zeros_like(lo)
    for j in range(L):
        sl = xp[..., j:j + 2 * outsize:2]
        lo += h0[j] * sl
        hi += h1[j] * sl
    return np.moveaxis(lo, -1, axis), np.moveaxis(hi, -1, axis)


def sfb1d(lo, hi, g0, g1, axis):
    """Synthesis filter bank (R:132-156): upsample-by-2 scatter of
    ``g0*lo + g1*hi`` and crop of ``L-2`` samples on each side."""
    lo = np.moveaxis(np.asarray(lo, dtype=np.float64), axis, -1)
    hi = np.moveaxis(np.asarray(hi, dtype=np.float64), axis, -1)
    H = lo.shape[-1]
    L = len(g0)
    size = (H - 1) * 2 + L
    out = np.zeros(lo.shape[:-1] + (size,))
    for i in range(L):
        out[..., i:i + 2 * H:2] += g0[i] * lo + g1[i] * hi
    return np.moveaxis(out[..., L - 2:size - (L - 2)], -1, axis)


def dwt2(x, wname):
    """One level of ``DWTForward.apply`` (R:158-207) on ``(..., H, W)``: rows
    (axis -1) first, then columns (axis -2).  Returns ``ll, (lh, hl, hh)`` where
    the first letter is the row (W) band and the second the column (H) band;
    ``hl`` is the band the damping acts on (``yh[:, :, 1]``)."""
    dec_lo, dec_hi, _, _ = wavelet_taps(wname)
    h0, h1 = dec_lo[::-1], dec_hi[::-1]
    lo, hi = afb1d(x, h0, h1, -1)
    ll, lh = afb1d(lo, h0, h1, -2)
    hl, hh = afb1d(hi, h0, h1, -2)
    return ll, (lh, hl, hh)


def idwt2(ll, bands, wname):
    """One level of ``DWTInverse.apply`` (R:209-251): columns, then rows."""
    _, _, rec_lo, rec_hi = wavelet_taps(wname)
    lh, hl, hh = bands
    lo = sfb1d(ll, lh, rec_lo, rec_hi, -2)
    hi = sfb1d(hl, hh, rec_lo, rec_hi, -2)
    return sfb1d(lo, hi, rec_lo, rec_hi, -1)


def fw_damping(my, sigma):
    """Damping of the ``rfft`` of the stripe band along angles (R:276-280):
    ``-expm1(-yhat^2/(2 sigma^2))`` with ``yhat = ifftshift((arange(-my,my,2)+1)/2)``
    truncated to ``my//2+1`` (``k+0.5`` for even ``my``, ``k`` for odd)."""
    yhat = np.fft.ifftshift((np.arange(-my, my, 2) + 1) / 2)[:my // 2 + 1]
    return -np.expm1(-yhat ** 2 / (2 * sigma ** 2))


def remove_stripe_fw(data, sigma, wname, level):
    """``remove_stripe_fw`` (R:253-293) on ``data (nproj, nz, ni)``, FP64."""
    data = np.asarray(data, dtype=np.float64)
    nproj, nz, ni = data.shape
    nproj_pad = nproj + nproj // 8
    st = (nproj_pad - nproj) // 2
    sli = np.zeros((nz, nproj_pad, ni))
    sli[:, st:st + nproj] = data.swapaxes(0, 1)
    cc = []
    for k in range(level):
        sli, (lh, hl, hh) = dwt2(sli, wname)
        my = hl.shape[-2]
        f = np.fft.rfft(hl, axis=-2)
        f *= fw_damping(my, sigma)[:, None]
        hl = np.fft.irfft(f, my, axis=-2)
        cc.append((lh, hl, hh))
    for k in range(level)[::-1]:
        h, w = cc[k][0].shape[-2:]
        sli = sli[..., :h, :w]
        sli = idwt2(sli, cc[k], wname)
    return sli[:, st:st + nproj, :ni].swapaxes(0, 1)


def rotate_proj(data, angle, order=1):
    """``ProcFunctions.rotate_proj`` (P:126-138): ``ndimage.rotate(data, angle,
    reshape=False, order=order, axes=(2, 1), mode='nearest')`` (SciPy, the CPU
    twin of the ``cupyx`` routine), FP64."""
    from scipy import ndimage
    return ndimage.rotate(np.asarray(data, dtype=np.float64), float(angle), reshape=False,
                          order=order, axes=(2, 1), mode='nearest')


def detect_stripe(listdata, snr):
    """``_detect_stripe_batch`` (R:352-375), Algorithm 4 of Vo et al. 2018, per
    row of ``listdata (nz, ni)``: sort descending, least-squares line through
    the middle half (ranks ``ndrop .. ni-ndrop-2``, ``ndrop = int(ni/4)``),
    noise = the line's drop over the row; a row whose extreme value stands
    ``>= snr`` noise levels off the line gets a mask on the values beyond
    ``snr/2`` noise levels.  FP64 throughout (the reference fits in float32)."""
    listdata = np.asarray(listdata, dtype=np.float64)
    nz, numdata = listdata.shape
    srt = np.sort(listdata, axis=1)[:, ::-1]
    xlist = np.arange(numdata, dtype=np.float64)
    ndrop = int(0.25 * numdata)
    x = xlist[ndrop:-ndrop - 1]
    y = srt[:, ndrop:-ndrop - 1]
    n = x.size
    xm = x.mean()
    ym = y.mean(axis=1)
    sxy = (y * x).sum(axis=1) - n * ym * xm
    sxx = (x ** 2).sum() - n * xm ** 2
    slope = sxy / sxx
    intercept = ym - slope * xm
    numt1 = intercept + slope * xlist[-1]
    noise = np.clip(np.abs(numt1 - intercept), 1e-6, None)
    val1 = np.abs(srt[:, 0] - intercept) / noise
    val2 = np.abs(srt[:, -1] - numt1) / noise
    upper = (intercept + noise * snr * 0.5)[:, None]
    lower = (numt1 - noise * snr * 0.5)[:, None]
    mask = np.zeros((nz, numdata))
    mask = np.where((val1 >= snr)[:, None] & (listdata > upper), 1.0, mask)
    mask = np.where((val2 >= snr)[:, None] & (listdata <= lower), 1.0, mask)
    return mask


def _dilate3(mask):
    """``binary_dilation`` with a (1, 3) structure, border 0 (R:421-423)."""
    from scipy import ndimage
    return ndimage.binary_dilation(mask > 0, structure=np.ones((1, 3), dtype=bool)).astype(
        np.float64)


def _unsort(vals, ids):
    """Put sorted-order values back at the positions ``ids`` they came from,
    along axis 0 (the inverse permutation of R:377-386 applied)."""
    out = np.empty_like(vals)
    np.put_along_axis(out, ids, vals, axis=0)
    return out


def rs_sort(tomo, size, dim=1):
    """``_rs_sort3`` (R:389-401), Algorithm 3 of Vo et al.: sort every detector
    column of ``tomo (nproj, nz, ni)`` over the angles (stable), median-filter
    the sorted image along the detector (``dim=1``: window ``size``; ``dim=2``:
    ``size x size`` over detector and rank), put the values back in the
    original order."""
    from scipy import ndimage
    tomo = np.asarray(tomo, dtype=np.float64)
    ids = np.argsort(tomo, axis=0, kind='stable')
    srt = np.take_along_axis(tomo, ids, axis=0)
    if dim == 1:
        sm = ndimage.median_filter(srt, (1, 1, size))
    else:
        sm = ndimage.median_filter(srt, (size, 1, size))
    return _unsort(sm, ids)


def rs_large(tomo, snr, size, drop_ratio=0.1, norm=True):
    """``_rs_large3`` (R:404-435), Algorithm 5 of Vo et al.: sorted columns,
    their detector-median smoothing (window ``size``), the ratio of the two
    over the central ranks (``ndrop = int(0.5 drop_ratio nproj)`` dropped at
    each end) locates large stripes (``detect_stripe`` + dilation); ``norm``
    divides every column by that ratio, then the located columns take the
    smoothed sorted values in their original order."""
    from scipy import ndimage
    tomo = np.asarray(tomo, dtype=np.float64)
    drop_ratio = max(min(drop_ratio, 0.8), 0)
    nproj = tomo.shape[0]
    ndrop = int(0.5 * drop_ratio * nproj)
    ids = np.argsort(tomo, axis=0, kind='stable')
    srt = np.take_along_axis(tomo, ids, axis=0)
    smooth = ndimage.median_filter(srt, (1, 1, size))
    list1 = np.mean(srt[ndrop:nproj - ndrop], axis=0)
    list2 = np.mean(smooth[ndrop:nproj - ndrop], axis=0)
    fact = list1 / list2
    mask = _dilate3(detect_stripe(fact, snr))
    if norm:
        tomo = tomo / fact[None]
    corrected = _unsort(smooth, ids)
    return np.where(mask[None] > 0, corrected, tomo)


def rs_dead(tomo, snr, size, norm=True):
    """``_rs_dead3`` (R:438-465), Algorithm 6 of Vo et al.: the summed absolute
    difference of every column from its 10-angle running mean
    (``uniform_filter1d``, 'reflect'), over its detector median (window
    ``size``), locates dead/unresponsive columns (``detect_stripe`` +
    dilation, two columns at each edge never); they are replaced by the
    linear interpolation between the nearest unmasked columns; then
    ``rs_large`` (with ``norm``)."""
    from scipy import ndimage
    tomo = np.array(tomo, dtype=np.float64)
    nproj, nz, ni = tomo.shape
    smooth = ndimage.uniform_filter1d(tomo, 10, axis=0)
    listdiff = np.sum(np.abs(tomo - smooth), axis=0)
    bck = ndimage.median_filter(listdiff, (1, size))
    fact = listdiff / bck
    mask = _dilate3(detect_stripe(fact, snr))
    mask[:, 0:2] = 0.0
    mask[:, -2:] = 0.0
    for m in range(nz):
        listx = np.where(mask[m] < 1.0)[0]
        miss = np.where(mask[m] > 0.0)[0]
        if len(miss) > 0:
            matz = tomo[:, m, listx]
            ids = np.searchsorted(listx, miss)
            tomo[:, m, miss] = (matz[:, ids - 1] + (miss - listx[ids - 1]) *
                                (matz[:, ids] - matz[:, ids - 1]) /
                                (listx[ids] - listx[ids - 1]))
    if norm:
        tomo = rs_large(tomo, snr, size)
    return tomo


def remove_all_stripe(tomo, snr=3, la_size=61, sm_size=21, dim=1):
    """``remove_all_stripe`` (R:468-497), the ``vo-all`` method of
    ``proc_sino`` (P:157-159): ``rs_dead`` (with ``rs_large``) then
    ``rs_sort``, on the transmission chunk ``(nproj, nz, ni)``."""
    tomo = rs_dead(tomo, snr, la_size)
    return rs_sort(tomo, sm_size, dim)


def proc_chain(data, dark, flat, *, stripe='none', fw_sigma=1.0, fw_filter='sym16',
               fw_level=7, take_minus_log=True, bright_ratio=1.0, flat_linear=False,
               dezinger=0, dezinger_threshold=5000, ti_beta=0.022, ti_mask=1.0, binning=0,
               vo_snr=3.0, vo_la_size=61, vo_sm_size=21, vo_dim=1):
    """``proc_sino`` then ``proc_proj`` then the swap to sinogram order
    (P:140-198, ``rec.py:172-174``): [binning, as the reader applies it] ->
    [outliers] -> darkflat -> [fw] -> [-log] -> ``(nz, nproj, n)``."""
    if binning > 0:
        data, dark, flat = (downsample(a, binning) for a in (data, dark, flat))
    if dezinger > 0:
        data = remove_outliers(data, dezinger, dezinger_threshold)
        dark = remove_outliers(dark, dezinger, dezinger_threshold)
        flat = remove_outliers(flat, dezinger, dezinger_threshold)
    res = darkflat_correction(data, dark, flat, bright_ratio, flat_linear)
    if stripe == 'fw':
        res = remove_stripe_fw(res, fw_sigma, fw_filter, fw_level)
    elif stripe == 'ti':
        res = remove_stripe_ti(res, ti_beta, ti_mask)
    elif stripe == 'vo-all':
        res = remove_all_stripe(res, vo_snr, vo_la_size, vo_sm_size, vo_dim)
    elif stripe != 'none':
        raise ValueError(stripe)
    if take_minus_log:
        res = minus_log(res)
    return np.ascontiguousarray(res.swapaxes(0, 1))
