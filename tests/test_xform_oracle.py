"""Pins the CPU restatement of mrcstack's .xf application
(oracle/xform_oracle.c) before the GPU is checked against it.

OpenCV is not installed here and the reference ships no fixtures for
mrcstack, so agreement with OpenCV's own binary is parity unpinned.  What is
pinned, independently of the oracle's formulas:
  - both kernel tables against numpy evaluations of the Keys cubic
    (A = -0.75) and of sinc(d) sinc(d / 4);
  - the inverse matrix against numpy's inverse of mrcstack's matrix;
  - the fixed-point source coordinate against the exact affine map to within
    the 1/32-pixel quantisation;
  - whole images against closed forms: identity and integer shifts (an exact
    copy inside, the border rule at the edges, the border value outside),
    quarter turns, constant images, and ramps, whose result is the first
    moment of the kernel at the quantised fraction.
"""
import numpy as np
import pytest

A = -0.75


def keys(d):
    d = np.abs(d)
    return np.where(d <= 1, (A + 2) * d**3 - (A + 3) * d**2 + 1,
                    np.where(d < 2, A * d**3 - 5 * A * d**2 + 8 * A * d - 4 * A, 0.0))


def lanczos(t):
    """Normalised 8-tap Lanczos-4 weights at fraction t (taps -3 .. 4)."""
    d = t - np.arange(-3, 5)
    w = np.sinc(d) * np.sinc(d / 4.0)
    return w / w.sum()


def first_moment(tab):
    """sum_k w_k * (k - o): the value a kernel gives on the ramp f(i) = i."""
    K = tab.shape[1]
    o = 1 if K == 4 else 3
    return (tab.astype(np.float64) * (np.arange(K) - o)).sum(axis=1)


def mrcstack_matrix(xf, W, H):
    """mrcstack's forward matrix (main.cpp:56-60) in float32 arithmetic."""
    a11, a12, a21, a22, dx, dy = (np.float32(v) for v in xf)
    cx, cy = np.float32(W * 0.5), np.float32(H * 0.5)
    return np.array([[a11, a12, cx + dx - a11 * cx - a12 * cy],
                     [a21, a22, cy + dy - a21 * cx - a22 * cy]], np.float64)


def rand_stack(P, H, W, seed=0):
    return np.random.default_rng(seed).standard_normal((P, H, W)).astype(np.float32)


# ---------------------------------------------------------------- tables

def test_cubic_table(orc):
    t = orc.xf_table(False)
    f = np.arange(32) / 32.0
    want = keys(f[:, None] - np.arange(-1, 3)[None, :])
    np.testing.assert_allclose(t, want, atol=3e-7)
    np.testing.assert_array_equal(t[0], [0, 1, 0, 0])
    assert np.all(np.abs(t.astype(np.float64).sum(axis=1) - 1) < 2e-7)
    # Keys' kernel with A = -0.75 is not linear-exact: its first moment is
    # t - 2 (A + 1/2) t (1 - t) (1 - 2 t)
    np.testing.assert_allclose(first_moment(t), f - 2 * (A + 0.5) * f * (1 - f) * (1 - 2 * f), atol=5e-7)


def test_lanczos_table(orc):
    t = orc.xf_table(True)
    for i in range(32):
        np.testing.assert_allclose(t[i], lanczos(i / 32.0), atol=2e-6)
    assert abs(t[0, 3] - 1) < 1.2e-7 and np.all(np.abs(np.delete(t[0], 3)) < 1e-29)
    np.testing.assert_allclose(t[16], t[16][::-1], atol=1e-7)


# ---------------------------------------------------------------- geometry

XFS = [
    (1, 0, 0, 1, 0, 0),
    (0.9986, -0.0523, 0.0523, 0.9986, 4.37, -2.91),
    (0.0872, -0.9962, 0.9962, 0.0872, -11.2, 6.6),
    (1.03, 0.01, -0.02, 0.97, 0.5, 0.25),
    (-0.7071, -0.7071, 0.7071, -0.7071, 1.0, 2.0),
]


@pytest.mark.parametrize("xf", XFS)
def test_inverse_matrix(orc, xf):
    W, H = 96, 80
    F = mrcstack_matrix(xf, W, H)
    M = orc.xf_inverse(np.array(xf, np.float32), W, H).reshape(2, 3)
    want = np.linalg.inv(np.vstack([F, [0, 0, 1]]))[:2]
    np.testing.assert_allclose(M, want, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("xf", XFS)
def test_source_coordinate_quantisation(orc, xf):
    """ix + fx/32 is the exact source coordinate rounded to 1/32, up to the
    two 1/1024 roundings of the row and column terms."""
    W, H = 64, 48
    M = orc.xf_inverse(np.array(xf, np.float32), W, H)
    rng = np.random.default_rng(1)
    for x, y in rng.integers(0, [W, H], size=(200, 2)):
        ix, iy, t = orc.xf_coord(M, x, y)
        u = M[0] * x + M[1] * y + M[2]
        v = M[3] * x + M[4] * y + M[5]
        assert abs(ix + (t % 32) / 32.0 - u) <= 1 / 64 + 1 / 1024 + 1e-12
        assert abs(iy + (t // 32) / 32.0 - v) <= 1 / 64 + 1 / 1024 + 1e-12


# ---------------------------------------------------------------- images

def shifted_expect(src, dx, dy, cv, K):
    """Closed form of an integer shift with fraction 0: the kernel is a delta
    at the centre tap.  Inside: an exact copy; at the edges the border rule
    gives cv + (s - cv) when the centre tap is inside and cv otherwise."""
    H, W = src.shape
    o = 1 if K == 4 else 3
    out = np.full_like(src, cv)
    for y in range(H):
        for x in range(W):
            cx_, cy_ = x - dx, y - dy
            sx, sy = cx_ - o, cy_ - o
            inside = 0 <= sx <= W - K and 0 <= sy <= H - K
            if inside:
                out[y, x] = src[cy_, cx_]
            elif 0 <= cx_ < W and 0 <= cy_ < H:
                out[y, x] = np.float32(cv + np.float32(src[cy_, cx_] - cv))
    return out


@pytest.mark.parametrize("dx,dy", [(0, 0), (3, -2), (-7, 5), (40, 0)])
def test_integer_shift_cubic(orc, dx, dy):
    src = rand_stack(1, 30, 36, seed=dx + 17)[0] + np.float32(2.5)
    cv = np.float32(orc.image_mean(src))
    assert abs(float(cv) - src.astype(np.float64).mean()) < 1e-6
    out, cvs = orc.apply_xf(src[None], np.array([[1, 0, 0, 1, dx, dy]], np.float32), mode=1)
    assert cvs[0] == cv
    np.testing.assert_array_equal(out[0], shifted_expect(src, dx, dy, cv, 4))


def test_integer_shift_lanczos(orc):
    src = rand_stack(1, 30, 36, seed=5)[0]
    out, cvs = orc.apply_xf(src[None], np.array([[1, 0, 0, 1, 4, -3]], np.float32), mode=2)
    want = shifted_expect(src, 4, -3, cvs[0], 8)
    # the centre weight is 1e30 * (1 / sum): within one ulp of 1
    np.testing.assert_allclose(out[0], want, rtol=3e-7, atol=1e-30)


def test_fully_outside_is_border_value(orc):
    src = rand_stack(1, 20, 24, seed=9)[0]
    out, cvs = orc.apply_xf(src[None], np.array([[1, 0, 0, 1, 30, 0]], np.float32), mode=1)
    assert np.all(out[0] == cvs[0])


def test_quarter_turn(orc):
    """(0, -1, 1, 0) about (W/2, H/2), W = H: out[y, x] = src[W - x, y]."""
    W = 32
    src = rand_stack(1, W, W, seed=3)[0]
    out, _ = orc.apply_xf(src[None], np.array([[0, -1, 1, 0, 0, 0]], np.float32), mode=1)
    for y in range(2, W - 3):
        for x in range(4, W - 1):
            assert out[0, y, x] == src[W - x, y]


@pytest.mark.parametrize("mode", [1, 2])
def test_constant_image(orc, mode):
    src = np.full((1, 40, 40), 3.25, np.float32)
    out, cvs = orc.apply_xf(src, np.array([[0.97, 0.2, -0.2, 0.97, 1.3, -0.7]], np.float32), mode=mode)
    assert cvs[0] == np.float32(3.25)
    np.testing.assert_allclose(out, 3.25, rtol=2e-6)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("axis", [0, 1])
def test_ramp_gives_kernel_first_moment(orc, mode, axis):
    """On f = column index (or row index) an interior pixel is the integer
    tap origin plus the kernel's first moment at its fraction in that axis.
    Swapped row/column tables or a shifted tap origin break this."""
    W, H = 72, 64
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    src = (xx if axis == 0 else yy)[None].copy()
    xf = np.array([[0.995, 0.0998, -0.0998, 0.995, 1.37, -0.81]], np.float32)
    out, _ = orc.apply_xf(src, xf, mode=mode)
    tab = orc.xf_table(mode == 2)
    m1 = first_moment(tab)
    M = orc.xf_inverse(xf[0], W, H)
    K = tab.shape[1]
    o = K // 2 - 1
    n = 0
    for y in range(H):
        for x in range(W):
            ix, iy, t = orc.xf_coord(M, x, y)
            if not (0 <= ix - o <= W - K and 0 <= iy - o <= H - K):
                continue
            want = ix + m1[t % 32] if axis == 0 else iy + m1[t // 32]
            assert abs(out[0, y, x] - want) < 2e-4, (x, y)
            n += 1
    assert n > W * H // 2
