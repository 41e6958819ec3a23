"""Pins of the FP64 ``vo-all`` oracle (reference ``processing/remove_stripe.py``
:311-497, Vo et al. 2018 algorithms 3-6) by closed forms, brute force and a
library routine, each chosen so that a plausible slip (window bounds, offsets,
tie order, a dropped normalisation) fails."""
import numpy as np
import pytest

from oracle import preproc_fp64 as pp


def _fit_line(row):
    """np.polyfit over the middle half of the descending sort (library
    least squares, independent of the oracle's closed-form sums)."""
    n = len(row)
    srt = np.sort(row)[::-1]
    nd = n // 4
    x = np.arange(nd, n - nd - 1)
    slope, icpt = np.polyfit(x, srt[nd:n - nd - 1], 1)
    return slope, icpt


def test_detect_stripe_fit_matches_polyfit():
    rng = np.random.default_rng(3)
    n = 64
    for snr in (1.0, 3.0):
        row = np.cumsum(rng.uniform(0.5, 1.5, n))
        rng.shuffle(row)
        row[5] += 2000.0     # one far outlier above
        row[9] -= 3000.0     # one far outlier below
        slope, icpt = _fit_line(row)
        end = icpt + slope * (n - 1)
        noise = abs(end - icpt)
        ref = ((row > icpt + noise * snr / 2) & (abs(row.max() - icpt) / noise >= snr)) | \
              ((row <= end - noise * snr / 2) & (abs(row.min() - end) / noise >= snr))
        got = pp.detect_stripe(row[None], snr)[0]
        assert np.array_equal(got > 0, ref)
        assert got[5] == 1 and got[9] == 1


def test_detect_stripe_window_excludes_the_tails():
    # an exact line except the top ndrop and the bottom ndrop+1 ranks: the fit
    # over ranks ndrop .. n-ndrop-2 is exact, so the thresholds are closed
    # forms; one more rank on either side would bend the line
    n, snr = 40, 3.0
    nd = n // 4
    asc = np.arange(n, dtype=np.float64)      # descending rank i holds n-1-i
    asc[n - nd:] += 1000.0                    # top nd ranks far above
    asc[:nd + 1] -= 1000.0                    # bottom nd+1 ranks far below
    row = asc[::-1].copy()
    got = pp.detect_stripe(row[None], snr)[0]
    # line: value(i) = n-1-i -> intercept n-1, end 0, noise n-1
    upper = (n - 1) + (n - 1) * snr / 2
    lower = 0 - (n - 1) * snr / 2
    ref = (row > upper) | (row <= lower)
    assert np.array_equal(got > 0, ref)
    assert ref.sum() == 2 * nd + 1


def _brute_xmedian(a, w):
    """Pure-Python median along the last axis: window offsets -(w//2) ..
    w-1-w//2, half-sample reflection, rank w//2."""
    n = a.shape[-1]

    def refl(m):
        P = 2 * n
        m %= P
        return P - 1 - m if m >= n else m
    out = np.empty_like(a)
    for idx in np.ndindex(a.shape[:-1]):
        for x in range(n):
            win = sorted(a[idx + (refl(x + j - w // 2),)] for j in range(w))
            out[idx + (x,)] = win[w // 2]
    return out


@pytest.mark.parametrize("w", [3, 4, 9])
def test_rs_sort_brute_force_with_ties(w):
    # columns sorted over angles (stable: equal values keep their angle
    # order), median across the detector, values back in the original order
    rng = np.random.default_rng(w)
    t = rng.integers(0, 4, size=(7, 2, 6)).astype(np.float64)   # many ties
    got = pp.rs_sort(t, w)
    nproj, nz, ni = t.shape
    ref = np.empty_like(t)
    for z in range(nz):
        order = [sorted(range(nproj), key=lambda p: (t[p, z, x], p)) for x in range(ni)]
        srt = np.array([[t[order[x][r], z, x] for x in range(ni)] for r in range(nproj)])
        med = _brute_xmedian(srt, w)
        for x in range(ni):
            for r in range(nproj):
                ref[order[x][r], z, x] = med[r, x]
    assert np.array_equal(got, ref)


def test_rs_sort_is_identity_on_permuted_columns():
    # every column a permutation of one multiset: sorted image is constant
    # across the detector, the median leaves it, the output is the input
    rng = np.random.default_rng(1)
    base = rng.uniform(0.2, 0.9, 50)
    t = np.stack([np.stack([rng.permutation(base) for _ in range(30)], 1)
                  for _ in range(2)], 1)
    assert np.array_equal(pp.rs_sort(t, 11), t)


def test_rs_large_removes_a_gain_stripe_exactly():
    # one column x0 = 1.5 x a permutation of the common multiset: the ratio
    # is 1.5 there and 1 elsewhere, so only x0 (dilated to x0 +- 1) is located;
    # norm restores x0 and the located columns take the common sorted values
    # in their own order -> the stripe-free data
    rng = np.random.default_rng(2)
    base = rng.uniform(0.2, 0.9, 40)
    clean = np.stack([rng.permutation(base) for _ in range(32)], 1)[:, None, :]
    t = clean.copy()
    t[:, 0, 12] *= 1.5
    got = pp.rs_large(t, 3.0, 9)
    assert np.allclose(got, clean, rtol=0, atol=1e-15)
    # without norm the located columns are still replaced
    assert np.allclose(pp.rs_large(t, 3.0, 9, norm=False), clean, atol=1e-15)


def test_rs_dead_interpolates_a_dead_column():
    # a constant (dead) column in noisy data: zero roughness, located with its
    # two neighbours (dilation) and replaced by the line between x0-2 and x0+2
    rng = np.random.default_rng(4)
    t = rng.uniform(0.4, 0.6, size=(60, 1, 48))
    x0 = 20
    t[:, 0, x0] = 0.5
    got = pp.rs_dead(t, 3.0, 11, norm=False)
    left, right = t[:, 0, x0 - 2], t[:, 0, x0 + 2]
    for k, x in enumerate((x0 - 1, x0, x0 + 1), start=1):
        assert np.allclose(got[:, 0, x], left + k * (right - left) / 4, atol=1e-15)
    keep = [x for x in range(48) if abs(x - x0) > 1]
    assert np.array_equal(got[:, 0, keep], t[:, 0, keep])


def test_rs_dead_running_mean_offsets():
    # the 10-angle running mean of rs_dead (uniform_filter1d) covers the
    # offsets -5..4 with half-sample reflection, the window the GPU kernel
    # uses: checked against an explicit window mean
    from scipy import ndimage
    rng = np.random.default_rng(5)
    col = rng.uniform(0, 1, 23)
    sm = ndimage.uniform_filter1d(col, 10)

    def refl(m, n=23):
        m %= 2 * n
        return 2 * n - 1 - m if m >= n else m
    brute = np.array([np.mean([col[refl(p + j)] for j in range(-5, 5)]) for p in range(23)])
    assert np.allclose(sm, brute, atol=1e-15)


def test_remove_all_stripe_chain_order():
    # vo-all = rs_dead (which ends in rs_large) then rs_sort
    rng = np.random.default_rng(6)
    t = rng.uniform(0.3, 0.8, size=(40, 2, 32))
    t[:, 1, 7] *= 1.3
    ref = pp.rs_sort(pp.rs_dead(t, 3, 9), 5, 1)
    assert np.array_equal(pp.remove_all_stripe(t, 3, 9, 5, 1), ref)
    assert not np.array_equal(ref, t)
