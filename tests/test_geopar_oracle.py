"""Pins the CPU oracle of GeoParGen (src/GeoParGen/src/main/main.cpp:71-215,
mrcview.h:43-147, mrcview.cpp:13-127) before the GPU is checked against it.

OpenCV is absent, so its float roundings in cv::normalize, chip * ones and
cv::mean are parity unpinned against an OpenCV binary; what is pinned here:
  - the view orientation and chip layout, against numpy slicing and the
    reference's own loop bounds;
  - the normalisation, against exact rational arithmetic of x * a + b
    rounded once to float, and its end points;
  - the chip profiles, against exactly rounded sums (math.fsum);
  - the NCC, against numpy.corrcoef;
  - the search, against a numpy evaluation of every pair (argmax and
    maximum) and against the invariant of the reference's float running
    maximum: the returned value is the float of the largest NCC, the chosen
    pair rounds to it, and no later pair exceeds it;
  - the regressions, against numpy.polyfit with the same weights and
    against closed forms;
  - the whole pipeline, against volumes holding a tilted slab whose angle,
    thickness and centre are known exactly.
"""
import math
from fractions import Fraction

import numpy as np
import pytest


def test_nchips_and_pairs(orc):
    for w in [0, 1, 50, 51, 99, 100, 101, 150, 151, 1024, 1000]:
        want = len(range(0, w - 50, 50)) if w >= 50 else 0
        assert orc.gp_nchips(w) == want, w
    for n in [0, 1, 30, 31, 32, 80, 300]:
        want = sum(1 for l in range(n) if l + 30 < n for r in range(l + 30, n))
        assert orc.gp_npairs(n) == want, n


def test_view_orientation(orc):
    vol = np.random.default_rng(0).random((7, 5, 6)).astype(np.float32)  # [nz][ny][nx]
    for j in range(6):
        np.testing.assert_array_equal(orc.gp_view(vol, 0, j), vol[:, :, j].T)
    for i in range(7):
        np.testing.assert_array_equal(orc.gp_view(vol, 1, i), vol[i])


def round_f32(fr):
    """Correct rounding (to nearest, ties to even) of a Fraction to float32
    (normal range)."""
    if fr == 0:
        return np.float32(0.0)
    s = -1 if fr < 0 else 1
    a = abs(fr)
    e = math.floor(math.log2(a.numerator) - math.log2(a.denominator))
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    scaled = a / Fraction(2) ** (e - 23)  # in [2^23, 2^24)
    m = scaled.numerator // scaled.denominator
    rem = scaled - m
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and m % 2 == 1):
        m += 1
    return np.float32(s * m * 2.0 ** (e - 23))


def test_normalize_exact(orc):
    rng = np.random.default_rng(1)
    img = (rng.standard_normal(400) * 37 + 5).astype(np.float32)
    got = orc.gp_normalize(img)
    smin, smax = float(img.min()), float(img.max())
    a = np.float32(1.0 / (smax - smin))
    b = np.float32(np.float32(0.0) - np.float32(smin * float(a)))
    for x, g in zip(img[:200], got[:200]):
        want = round_f32(Fraction(float(x)) * Fraction(float(a)) + Fraction(float(b)))
        assert g == want
    # end points: the minimum lands within rounding of 0, the maximum of 1
    assert abs(got.min()) < 1e-6 and abs(got.max() - 1) < 1e-6
    assert np.all(np.diff(got[np.argsort(img, kind="stable")]) >= 0)
    # a constant image has max - min = 0: scale 0, every value 0
    assert np.all(orc.gp_normalize(np.full(9, 3.25, np.float32)) == 0)


def test_profiles_exact_sums(orc):
    rng = np.random.default_rng(2)
    norm = rng.random((33, 170)).astype(np.float32)
    prof = orc.gp_profiles(norm)
    assert prof.shape == (3, 33)
    for q in range(3):
        for r in range(33):
            exact = math.fsum(float(v) for v in norm[r, 50 * q:50 * q + 50])
            assert abs(float(prof[q, r]) - exact) <= abs(exact) * 2 ** -24 * 1.0001


def test_ncc_matches_corrcoef(orc):
    rng = np.random.default_rng(3)
    n = 90
    a = np.cumsum(rng.standard_normal(n)).astype(np.float32)
    for l, r in [(0, 30), (10, 60), (59, 89), (0, 88), (1, 89), (20, 70)]:
        rect = np.ones(n)
        rect[l:r + 1] = 0
        want = abs(np.corrcoef(a.astype(np.float64), rect)[0, 1])
        assert abs(orc.gp_ncc(a, l, r) - want) < 1e-12


def all_pairs(a):
    """NCC of every pair in the reference's scan order, in float64 numpy."""
    n = a.size
    x = a.astype(np.float64) - a.astype(np.float64).mean()
    P = np.concatenate([[0.0], np.cumsum(x)])
    out = []
    for l in range(n - 30):
        r = np.arange(l + 30, n)
        k = r - l + 1
        m = n - k
        s0 = P[r + 1] - P[l]
        num = -s0  # sum of x over ones = -(sum over zeros), since sum x = 0
        with np.errstate(invalid="ignore", divide="ignore"):
            den = np.sqrt((x * x).sum() * (m * k / n))
            out.append(np.abs(num / den))
    return np.concatenate(out)


def pair_of(t, n):
    for l in range(n - 30):
        row = n - 30 - l
        if t < row:
            return l, l + 30 + t
        t -= row
    raise IndexError


@pytest.mark.parametrize("seed", range(6))
def test_search_matches_all_pairs(orc, seed):
    rng = np.random.default_rng(10 + seed)
    n = int(rng.integers(40, 140))
    base = np.convolve(rng.standard_normal(n + 20), np.ones(9) / 9, "same")[10:10 + n]
    a = (base + rng.random() * 3).astype(np.float32)
    z, t, nc, lo, hi = orc.gp_search(a)
    res = all_pairs(a)
    best = int(np.nanargmax(res))
    M = res[best]
    assert abs(float(nc) - M) <= M * 2 ** -23
    second = np.partition(res[np.isfinite(res)], -2)[-2]
    if M - second > 1e-9 * M:
        assert (lo, hi) == pair_of(best, n)
    assert (z, t) == ((lo + hi) // 2, hi - lo)


@pytest.mark.parametrize("seed", range(4))
def test_search_float_maximum_invariant(orc, seed):
    """The reference keeps its running maximum in a float, so a later pair
    wins when its NCC exceeds the float of the best so far.  Invariant: the
    returned value is float(max NCC); the chosen pair's NCC rounds to it;
    no pair after the chosen one exceeds it.  Symmetric profiles make exact
    mirror ties, decided only by rounding."""
    rng = np.random.default_rng(40 + seed)
    h = rng.random(40)
    a = np.concatenate([h, h[::-1]]).astype(np.float32)
    n = a.size
    z, t, nc, lo, hi = orc.gp_search(a)
    res = np.array([orc.gp_ncc(a, l, r) for l in range(n - 30) for r in range(l + 30, n)])
    M = np.nanmax(res)
    assert nc == np.float32(M)
    idx = [(l, r) for l in range(n - 30) for r in range(l + 30, n)].index((lo, hi))
    assert np.float32(res[idx]) == nc
    later = res[idx + 1:]
    assert not np.any(later[np.isfinite(later)] > float(nc))


def test_search_recovers_a_step(orc):
    """An interior step; a step touching either end would tie exactly with
    its complement, the NCC being taken in absolute value."""
    n = 120
    for l0, r0 in [(10, 60), (1, 40), (70, 118), (30, 100)]:
        a = np.full(n, 2.0, np.float32)
        a[l0:r0 + 1] = 7.0
        z, t, nc, lo, hi = orc.gp_search(a)
        assert (lo, hi) == (l0, r0) and nc == np.float32(1.0)
        assert (z, t) == ((l0 + r0) // 2, r0 - l0)


def test_search_degenerate(orc):
    # too short: no pair at all; constant: every NCC is 0/0
    for a in [np.arange(30, dtype=np.float32), np.zeros(80, np.float32)]:
        assert orc.gp_search(a) == (-1, -1, 0.0, -1, -1)


def test_view_result_closed_form(orc):
    nc = np.ones(12, np.float32)
    for d in [0, 1, -3, 7]:
        z = 40 + d * np.arange(12)
        t = np.full(12, 36)
        ang, th, zz = orc.gp_view_result(z, t, nc)
        assert abs(ang - math.degrees(math.atan(d / 50))) < 1e-5
        assert th == 36 and abs(zz - z.mean()) < 1e-5


def test_view_result_matches_weighted_fit(orc):
    rng = np.random.default_rng(5)
    for _ in range(5):
        k = int(rng.integers(3, 21))
        z = rng.integers(20, 200, k)
        t = rng.integers(30, 120, k)
        nc = rng.random(k).astype(np.float32)
        ang, th, zz = orc.gp_view_result(z, t, nc)
        w = (nc / nc.sum(dtype=np.float32)).astype(np.float64)  # weight = w^2
        x = 50 * np.arange(k) + 25
        slope = np.polyfit(x, z, 1, w=w)[0]
        assert abs(ang - math.degrees(math.atan(slope))) < 1e-4
        assert abs(th - (w * w * t).sum() / (w * w).sum()) < 1e-4 * th
        assert abs(zz - (w * w * z).sum() / (w * w).sum()) < 1e-4 * zz


def test_view_result_empty_and_failed(orc):
    a = orc.gp_view_result([], [], np.zeros(0, np.float32))
    assert math.isnan(a[0]) and a[1] == 0 and a[2] == 0
    a = orc.gp_view_result([-1, -1], [-1, -1], np.zeros(2, np.float32))
    assert all(math.isnan(v) for v in a)


def slab_volume(axis, nv, rows, cols, l0, d, t0, seed):
    """A bright slab whose rows step by d per chip: chip q covers rows
    [l0 + q d, l0 + q d + t0]."""
    rng = np.random.default_rng(seed)
    views = np.zeros((nv, rows, cols), np.float32)
    for v in range(nv):
        views[v] = 0.5 + 0.1 * rng.random()
        for q in range(orc_nchips(cols)):
            lo = l0 + q * d
            views[v, lo:lo + t0 + 1, 50 * q:50 * q + 50] = 3.0 + v
    if axis == 0:  # view j, element (r, c) = vol[c][r][j]
        return np.ascontiguousarray(views.transpose(2, 1, 0))
    return views


def orc_nchips(w):
    return len(range(0, w - 50, 50)) if w >= 50 else 0


@pytest.mark.parametrize("axis", [0, 1])
def test_whole_pipeline_tilted_slab(orc, axis):
    rows, cols, nv, l0, d, t0 = 90, 260, 5, 12, 4, 36
    vol = slab_volume(axis, nv, rows, cols, l0, d, t0, 7)
    views, chips, avg = orc.geopar(vol, axis)
    assert views.shape == (nv, 3) and chips.shape == (nv, 5, 4)
    zq = [l0 + q * d + t0 // 2 for q in range(5)]
    for v in range(nv):
        np.testing.assert_array_equal(chips[v, :, 0], zq)
        np.testing.assert_array_equal(chips[v, :, 1], t0)
        np.testing.assert_array_equal(chips[v, :, 2], 1.0)
    want = math.degrees(math.atan(d / 50))
    assert np.allclose(views[:, 0], want, atol=1e-5)
    assert np.allclose(views[:, 1], t0) and np.allclose(views[:, 2], np.mean(zq), atol=1e-5)
    assert abs(avg[0] - want) < 1e-5 and abs(avg[1] - t0) < 1e-5 and abs(avg[2] - np.mean(zq)) < 1e-5


def test_whole_pipeline_matches_parts(orc):
    """orc_geopar = view -> normalize -> profiles -> search -> view result ->
    float averages, composed here from the exported parts."""
    rng = np.random.default_rng(9)
    vol = np.cumsum(rng.standard_normal((130, 64, 4)), axis=1).astype(np.float32)
    views, chips, avg = orc.geopar(vol, 0)
    acc = np.zeros(3, np.float32)
    for j in range(4):
        prof = orc.gp_profiles(orc.gp_normalize(orc.gp_view(vol, 0, j)))
        res = [orc.gp_search(p) for p in prof]
        np.testing.assert_array_equal(chips[j, :, 0], [r[0] for r in res])
        np.testing.assert_array_equal(chips[j, :, 2], [r[2] for r in res])
        vr = orc.gp_view_result([r[0] for r in res], [r[1] for r in res], np.array([r[2] for r in res]))
        np.testing.assert_array_equal(views[j], vr)
        acc = (acc + vr).astype(np.float32)
    np.testing.assert_array_equal(avg, (acc / np.float32(4)).astype(np.float32))
