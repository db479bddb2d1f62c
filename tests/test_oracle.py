"""Pins the CPU oracle (oracle/tiltrec_oracle.c) before anything is checked
against it.  Every check here is independent of the oracle's own formulas:
closed forms, an explicitly built FP64 system matrix, adjointness between the
splat and the gather, rotation matrices, and numpy's FFT.

The reference holds no tests, fixtures or golden vectors for TiltRec, and its
CPU binaries need MPI and OpenCV/FFTW, so they are not built here; these
checks are what pin the restatement.
"""
import numpy as np
import pytest

PI_180 = np.float32(0.01745329252)


def hat(t):
    return np.maximum(0.0, 1.0 - np.abs(t))


def sysmat(coef1, X, rows, T, org):
    """FP64 system matrix of one tilt, built from the hat-function form of a
    bilinear splat: voxel (x, y, z) lands on pixel (i, j) with weight
    hat(u - i) * hat(v - j), (u, v) the linear map of the coefficients.
    Rows index pixels j*X + i, columns voxels (y*T + z)*X + x."""
    ox, oy, oz = (float(v) for v in org)
    a0, a1, a2, a3, b0, b1, b2, b3 = (float(c) for c in coef1)
    y, z, x = np.meshgrid(np.arange(rows), np.arange(T), np.arange(X), indexing="ij")
    Xc = (x.astype(np.float32) - np.float32(ox)).astype(np.float64)
    Yc = (y.astype(np.float32) - np.float32(oy)).astype(np.float64)
    Zc = (z.astype(np.float32) - np.float32(oz)).astype(np.float64)
    u = (a0 + a1 * Xc + a2 * Yc + a3 * Zc + ox).ravel()
    v = (b0 + b1 * Xc + b2 * Yc + b3 * Zc + oy).ravel()
    j, i = np.meshgrid(np.arange(rows), np.arange(X), indexing="ij")
    i = i.ravel()[:, None]
    j = j.ravel()[:, None]
    return hat(u[None, :] - i) * hat(v[None, :] - j)


def tilt_set(P, span=60.0):
    return np.linspace(-span, span, P).astype(np.float32)


# ---------------------------------------------------------------- coefficients

def test_coeffs_plain_geometry(orc):
    ang = tilt_set(7)
    c = orc.coeffs_mpi(ang)
    r = (ang * PI_180).astype(np.float32)
    np.testing.assert_array_equal(c[:, 1], np.cos(r).astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(c[:, 3], (-np.sin(r).astype(np.float32)).astype(np.float64))
    assert np.all(c[:, [0, 2, 4, 5, 7]] == 0.0)
    assert np.all(c[:, 6] == 1.0)


@pytest.mark.parametrize("offset,pitch,zshift", [(3.0, 0.0, 0.0), (0.0, 5.0, 0.0), (2.5, -4.0, 7.0)])
def test_coeffs_match_rotation_matrices(orc, offset, pitch, zshift):
    """DecorateCoefficients (main-mpi.cpp:110-183) equals row-vector times
    Rx(alpha) @ Ry(beta), alpha = -pitch, beta = offset, then the z shift."""
    ang = tilt_set(5)
    c = orc.coeffs_mpi(ang, offset, pitch, zshift)
    al = float(-(np.float32(pitch) * PI_180))
    be = float(np.float32(offset) * PI_180)
    Rx = np.array([[1, 0, 0], [0, np.cos(al), np.sin(al)], [0, -np.sin(al), np.cos(al)]])
    Ry = np.array([[np.cos(be), 0, -np.sin(be)], [0, 1, 0], [np.sin(be), 0, np.cos(be)]])
    M = Rx @ Ry
    r = (ang * PI_180).astype(np.float32)
    a = np.stack([np.cos(r).astype(np.float32), np.zeros_like(r), -np.sin(r).astype(np.float32)], 1).astype(np.float64)
    b = np.tile([0.0, 1.0, 0.0], (len(ang), 1))
    na, nb = a @ M, b @ M
    np.testing.assert_allclose(c[:, 1:4], na, rtol=0, atol=1e-15)
    np.testing.assert_allclose(c[:, 5:8], nb, rtol=0, atol=1e-15)
    np.testing.assert_allclose(c[:, 0], na[:, 2] * -zshift, atol=1e-12)
    np.testing.assert_allclose(c[:, 4], nb[:, 2] * -zshift, atol=1e-12)


def test_coeffs_cuda_is_float_version(orc):
    ang = tilt_set(9)
    cd = orc.coeffs_mpi(ang, 1.5, 3.0, 2.0)
    cf = orc.coeffs_cuda(ang, 1.5, 3.0, 2.0)
    np.testing.assert_allclose(cf.astype(np.float64), cd, rtol=0, atol=2e-6)


# ---------------------------------------------------------------- projection

def test_reproject_zero_tilt_is_column_sum(orc):
    """At 0 degrees u = x exactly: value = sum over z, weight = T."""
    rng = np.random.default_rng(1)
    rows, T, X = 3, 6, 10
    vol = rng.standard_normal((rows, T, X)).astype(np.float32)
    c = orc.coeffs_mpi(np.array([0.0], np.float32))[0]
    rv, rw = orc.reproject_f64(c, vol, orc.origin(X, T, T / 2))
    np.testing.assert_allclose(rv, vol.astype(np.float64).sum(1), rtol=1e-13, atol=1e-12)
    np.testing.assert_array_equal(rw, np.full((rows, X), float(T)))


@pytest.mark.parametrize("offset,pitch,zshift", [(0.0, 0.0, 0.0), (2.0, 0.0, 1.5), (0.0, 6.0, 0.0), (1.0, -8.0, 2.0)])
def test_reproject_matches_system_matrix(orc, offset, pitch, zshift):
    rng = np.random.default_rng(2)
    rows, T, X = 5, 6, 9
    org = orc.origin(X, T, T / 2)
    vol = rng.standard_normal((rows, T, X))
    for ang in (-57.0, -31.0, -2.0, 0.0, 13.0, 44.0, 60.0):
        c = orc.coeffs_mpi(np.array([ang], np.float32), offset, pitch, zshift)[0]
        W = sysmat(c, X, rows, T, org)
        rv, rw = orc.reproject_f64(c, vol, org)
        np.testing.assert_allclose(rv.ravel(), W @ vol.ravel(), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(rw.ravel(), W.sum(1), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("pitch", [0.0, 7.0])
def test_backprojection_is_adjoint(orc, pitch):
    rng = np.random.default_rng(3)
    rows, T, X = 6, 7, 11
    org = orc.origin(X, T, T / 2)
    vol = rng.standard_normal((rows, T, X))
    img = rng.standard_normal((rows, X))
    for ang in (-45.0, 3.0, 52.0):
        c = orc.coeffs_mpi(np.array([ang], np.float32), 0.5, pitch, 0.0)[0]
        rv, _ = orc.reproject_f64(c, vol, org)
        s, cw = orc.backproject_sums_f64(c, img, T, org)
        assert abs(np.vdot(rv, img) - np.vdot(vol, s)) < 1e-10 * np.abs(vol).sum()
        W = sysmat(c, X, rows, T, org)
        np.testing.assert_allclose(cw.ravel(), W.sum(0), rtol=1e-12, atol=1e-12)


def test_reproject_f32_close_to_f64(orc):
    rng = np.random.default_rng(4)
    rows, T, X = 4, 16, 40
    org = orc.origin(X, T, T / 2)
    vol = rng.random((rows, T, X)).astype(np.float32)
    c = orc.coeffs_mpi(np.array([37.0], np.float32))[0]
    a, aw = orc.reproject_f32(c, vol, org)
    b, bw = orc.reproject_f64(c, vol, org)
    assert orc.rel_l2(a, b) < 2e-7
    assert orc.rel_l2(aw, bw) < 2e-7


# ---------------------------------------------------------------- methods

def test_bpt_of_ones_is_tilt_count(orc):
    """A stack of ones back-projects to P wherever a voxel's taps are in bounds."""
    P, rows, T, X = 9, 3, 8, 24
    ang = tilt_set(P, 40.0)
    c = orc.coeffs_mpi(ang)
    proj = np.ones((P, rows, X), np.float32)
    org = orc.origin(X, T, T / 2)
    v64 = orc.bpt_f64(c, proj, T, org)
    v32 = orc.bpt_f32(c, proj, T, org)
    interior = v64[:, :, 8:16]
    np.testing.assert_allclose(interior, float(P), rtol=1e-14)
    np.testing.assert_allclose(v32[:, :, 8:16], float(P), rtol=1e-6)
    assert np.all(v64 <= P + 1e-12)


def _matrix_sirt(Ws, ps, x, iters, gamma):
    for _ in range(iters):
        val = np.zeros_like(x)
        wt = np.zeros_like(x)
        for W, p in zip(Ws, ps):
            rw = W.sum(1)
            rv = W @ x
            q = np.where(rw != 0, rv / np.where(rw != 0, rw, 1), rv)
            val += W.T @ (p - q)
            wt += W.sum(0)
        x = x + gamma * np.where(wt != 0, val / np.where(wt != 0, wt, 1), 0)
    return x


def _matrix_sart(Ws, ps, x, iters, gamma):
    for _ in range(iters):
        for W, p in zip(Ws, ps):
            rw = W.sum(1)
            rv = W @ x
            r = p - np.where(rw != 0, rv / np.where(rw != 0, rw, 1), rv)
            s = W.T @ r
            c = W.sum(0)
            x = x + gamma * np.where(c != 0, s / np.where(c != 0, c, 1), 0)
    return x


@pytest.mark.parametrize("pitch", [0.0, 5.0])
def test_sirt_sart_match_matrix_form(orc, pitch):
    rng = np.random.default_rng(5)
    P, rows, T, X = 5, 4, 6, 8
    ang = tilt_set(P, 50.0)
    c = orc.coeffs_mpi(ang, 0.0, pitch, 0.0)
    org = orc.origin(X, T, T / 2)
    proj = rng.random((P, rows, X))
    Ws = [sysmat(c[k], X, rows, T, org) for k in range(P)]
    ps = [proj[k].ravel() for k in range(P)]
    x0 = np.zeros(rows * T * X)
    ref = _matrix_sirt(Ws, ps, x0, 3, 0.8)
    got = orc.sirt_f64(c, proj, T, org, 3, 0.8)
    np.testing.assert_allclose(got.ravel(), ref, rtol=1e-11, atol=1e-11)
    ref = _matrix_sart(Ws, ps, x0, 2, 0.7)
    got = orc.sart_f64(c, proj, T, org, 2, 0.7)
    np.testing.assert_allclose(got.ravel(), ref, rtol=1e-11, atol=1e-11)
    g32 = orc.sirt_f32(c, proj.astype(np.float32), T, org, 3, 0.8)
    assert orc.rel_l2(g32, _matrix_sirt(Ws, ps, x0, 3, 0.8).reshape(g32.shape)) < 1e-6


def _cut(values, frac):
    """A threshold midway between two neighbouring distinct positive values,
    about `frac` of the way up their sorted list, so that rounding in how a
    sum was formed cannot move a value across it."""
    v = np.unique(values[values > 0])
    k = int(frac * (len(v) - 1))
    while k + 1 < len(v) and v[k + 1] - v[k] < 1e-9 * v[k + 1]:
        k += 1
    return 0.5 * (v[k] + v[k + 1])


def _matrix_sirt_thr(Ws, ps, x, iters, gamma, tu, tr):
    for _ in range(iters):
        val = np.zeros_like(x)
        wt = np.zeros_like(x)
        for W, p in zip(Ws, ps):
            rw = W.sum(1)
            q = np.where(rw > tr, (W @ x) / np.where(rw > tr, rw, 1), 0.0)
            val += W.T @ (p - q)
            wt += W.sum(0)
        x = x + gamma * np.where(wt > tu, val / np.where(wt > tu, wt, 1), 0)
    return x


def _matrix_sart_thr(Ws, ps, x, iters, gamma, tu, tr):
    for _ in range(iters):
        for W, p in zip(Ws, ps):
            rw = W.sum(1)
            r = p - np.where(rw > tr, (W @ x) / np.where(rw > tr, rw, 1), 0.0)
            c = W.sum(0)
            x = x + gamma * np.where(c > tu, (W.T @ r) / np.where(c > tu, c, 1), 0)
    return x


@pytest.mark.parametrize("pitch,rules", [(0.0, 0), (5.0, 0), (5.0, 1)])
def test_thresholds_match_masked_matrix_forms(orc, pitch, rules):
    """The update and residual thresholds (TiltRec-cuda's 1e-5 / 1e-6 / 1e-7
    tests, basic.cu:170,184,409,434) equal the matrix forms with every
    weight sum at or below its threshold masked out.  The thresholds here
    sit inside the spread of a thick volume's weight sums so that the masks
    cut many voxels and pixels; rules 1 puts the y origin at Y/2 as the CUDA
    binaries do (main-cuda.cpp:259-261)."""
    rng = np.random.default_rng(15)
    P, rows, T, X = 5, 4, 12, 8
    ang = tilt_set(P, 50.0)
    c = orc.coeffs_mpi(ang, 0.0, pitch, 0.0)
    org = orc.rules_origin(X, rows, T, rules)
    proj = rng.random((P, rows, X))
    Ws = [sysmat(c[k], X, rows, T, org) for k in range(P)]
    ps = [proj[k].ravel() for k in range(P)]
    C = sum(W.sum(0) for W in Ws)
    c1 = np.concatenate([W.sum(0) for W in Ws])
    rw = np.concatenate([W.sum(1) for W in Ws])
    tu_sirt, tu_sart, tr = _cut(C, 0.3), _cut(c1, 0.3), _cut(rw, 0.3)
    assert ((C > 0) & (C <= tu_sirt)).sum() > 10 and ((rw > 0) & (rw <= tr)).sum() > 5
    x0 = np.zeros(rows * T * X)
    ref = _matrix_sirt_thr(Ws, ps, x0, 3, 0.8, tu_sirt, tr)
    got = orc.sirt_f64(c, proj, T, org, 3, 0.8, thr=(tu_sirt, tr))
    np.testing.assert_allclose(got.ravel(), ref, rtol=1e-11, atol=1e-11)
    assert not np.allclose(ref, _matrix_sirt_thr(Ws, ps, x0, 3, 0.8, 0.0, 0.0))  # the masks matter
    ref = _matrix_sart_thr(Ws, ps, x0, 2, 0.7, tu_sart, tr)
    got = orc.sart_f64(c, proj, T, org, 2, 0.7, thr=(tu_sart, tr))
    np.testing.assert_allclose(got.ravel(), ref, rtol=1e-11, atol=1e-11)
    bref = sum(np.where(W.sum(0) > tu_sart, (W.T @ p) / np.where(W.sum(0) > tu_sart, W.sum(0), 1), 0)
               for W, p in zip(Ws, ps))
    np.testing.assert_allclose(orc.bpt_f64(c, proj, T, org, thr=(tu_sart, 0.0)).ravel(), bref, rtol=1e-12, atol=1e-12)
    g32 = orc.sirt_f32(c, proj.astype(np.float32), T, org, 3, 0.8, thr=(tu_sirt, tr))
    assert orc.rel_l2(g32, _matrix_sirt_thr(Ws, ps, x0, 3, 0.8, tu_sirt, tr).reshape(g32.shape)) < 1e-6


def test_cuda_thresholds_known_answer(orc):
    """A stack of ones and a tilt at which corner voxel (0, T-1) reaches
    column 0 only, with weight w in (1e-6, 1e-5]: BPT adds s / c = 1 for that
    tilt under TiltRec-mpi's exact test and nothing under TiltRec-cuda's
    c > 1e-5 (basic.cu:409).  With only tilts that cannot see the voxel
    besides, SIRT's weight sum is w as well: the voxel moves by the residual
    under the exact test and stays 0 under wtvol > 10e-6 (basic.cu:184)."""
    from tests.helpers import edge_tilt, voxel_u
    X, T, rows = 32, 40, 2
    te = edge_tilt(orc, X, T, 1e-6, 1e-5)
    ang = np.array([te, 20.0, 30.0, 45.0, 60.0], np.float32)
    c = orc.coeffs_mpi(ang)
    u = voxel_u(c, 0, T - 1, X, T)
    assert -1 < u[0] <= -1 + 1e-5 and np.all(u[1:] < -1)   # only tilt 0 sees the voxel
    org = orc.origin(X, T, T / 2)
    ones = np.ones((len(ang), rows, X), np.float32)
    mpi = orc.bpt_f32(c, ones, T, org)
    cuda = orc.bpt_f32(c, ones, T, org, thr=orc.CUDA_THRESHOLDS["BPT"])
    assert np.all(mpi[:, T - 1, 0] == 1.0) and np.all(cuda[:, T - 1, 0] == 0.0)
    same = np.ones_like(mpi, bool)
    same[:, T - 1, 0] = False
    np.testing.assert_array_equal(mpi[same], cuda[same])
    proj = np.random.default_rng(3).random((len(ang), rows, X)).astype(np.float32) + 0.5
    s_mpi = orc.sirt_f32(c, proj, T, org, 1)
    s_cuda = orc.sirt_f32(c, proj, T, org, 1, thr=orc.CUDA_THRESHOLDS["SIRT"])
    assert np.all(np.abs(s_mpi[:, T - 1, 0]) > 0.1) and np.all(s_cuda[:, T - 1, 0] == 0.0)


def test_sirt_initial_volume_is_continued(orc):
    """--initial seeds the iteration: k then m iterations equal k+m."""
    rng = np.random.default_rng(6)
    P, rows, T, X = 6, 2, 5, 12
    c = orc.coeffs_mpi(tilt_set(P))
    org = orc.origin(X, T, T / 2)
    proj = rng.random((P, rows, X))
    a = orc.sirt_f64(c, proj, T, org, 4)
    b = orc.sirt_f64(c, proj, T, org, 2, vol=orc.sirt_f64(c, proj, T, org, 2))
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13)


def test_sirt_residual_decreases(orc):
    """On consistent data (projections of a volume) SIRT with gamma <= 1
    lowers the weighted residual every iteration."""
    rng = np.random.default_rng(7)
    P, rows, T, X = 11, 2, 8, 24
    c = orc.coeffs_mpi(tilt_set(P))
    org = orc.origin(X, T, T / 2)
    truth = rng.random((rows, T, X))
    proj = np.stack([orc.reproject_f64(c[k], truth, org)[0] / np.maximum(orc.reproject_f64(c[k], truth, org)[1], 1e-300) for k in range(P)])
    prev = None
    v = None
    for it in range(6):
        v = orc.sirt_f64(c, proj, T, org, 1, 1.0, vol=v)
        res = 0.0
        for k in range(P):
            rv, rw = orc.reproject_f64(c[k], v, org)
            q = np.where(rw != 0, rv / np.where(rw != 0, rw, 1), 0)
            res += float(np.sum(rw * (proj[k] - q) ** 2))
        if prev is not None:
            assert res < prev
        prev = res


# ---------------------------------------------------------------- R-weighting

def rweight_numpy(rows):
    """filter_prj.cpp:652-736 restated with numpy's FFT in float64."""
    rows = np.asarray(rows, np.float64)
    Xn = rows.shape[-1]
    N = Xn + Xn // 10
    nxp = N - Xn + 1
    j = np.arange(Xn, N)
    ext = np.concatenate([rows, ((j - Xn + 1) / nxp) * rows[..., :1] + ((N - j) / nxp) * rows[..., -1:]], -1)
    E = np.fft.rfft(ext, axis=-1)
    K = int(np.floor(np.float32(N) * np.float32(0.05)))
    sig = float(np.float32(N) * np.float32(0.5))
    k = np.arange(E.shape[-1])
    w = np.where(k == 0, 0.2, np.where(k <= K, k, K * np.exp(-((k - K) ** 2) / sig ** 2)))
    y = np.fft.irfft(E * w, n=N, axis=-1) * N
    return y[..., :Xn] / Xn


@pytest.mark.parametrize("X", [16, 23, 64, 100])
def test_rweight_matches_numpy_fft(orc, X):
    rng = np.random.default_rng(8)
    rows = rng.standard_normal((5, X))
    np.testing.assert_allclose(orc.rweight_f64(rows), rweight_numpy(rows), rtol=1e-10, atol=1e-10)


def test_rweight_constant_row_known_answer(orc):
    """A constant row c filters to 0.2 * c * N / X (only the DC bin survives)."""
    for X, c in ((256, 3.0), (50, -1.25)):
        N = X + X // 10
        out = orc.rweight_f64(np.full((2, X), c))
        np.testing.assert_allclose(out, 0.2 * c * N / X, rtol=1e-12)
    np.testing.assert_allclose(orc.rweight_f64(np.full((1, 256), 1.0)), 0.21953125, rtol=1e-12)


# ---------------------------------------------------------------- ADMM

def _matrix_admm(Ws, ps, outer, cgiter, mu, soft):
    """TiltRecZ-mpi's ADMM (main-mpiz.cpp:603-1013) on explicit FP64 system
    matrices: op(x) = sum W^T D W x + mu x with D = 1/(W 1) where |W 1| >
    1e-5, htb = sum W^T b, CG with the reference's alpha/beta and true
    residual, soft thresholding, the persistent htb update."""
    n = Ws[0].shape[1]
    D = []
    for W in Ws:
        w = W.sum(axis=1)
        D.append(np.where(np.abs(w) > 10e-6, 1.0 / np.where(w == 0, 1, w), 0.0))

    def op(x):
        out = mu * x
        for W, d in zip(Ws, D):
            out = out + W.T @ (d * (W @ x))
        return out

    htb = sum(W.T @ p for W, p in zip(Ws, ps))
    vol = np.zeros(n)
    u = np.zeros(n)
    dd = np.zeros(n)
    for _ in range(outer):
        htb = htb + (u - dd) * mu
        r = htb - op(vol)
        p = r.copy()
        for _ in range(cgiter):
            h = op(p)
            d0, d2 = r @ r, h @ p
            alpha = d0 / d2 if abs(d2) > 10e-6 else 0.0
            vol = vol + alpha * p
            r = htb - op(vol)
            beta = (r @ h) / d2 if abs(d2) > 10e-6 else 0.0
            p = r + beta * p
        t = vol + dd
        u = np.where(t < -soft, t + soft, np.where(t > soft, t - soft, 0.0))
        dd = dd + vol - u
    return vol


@pytest.mark.parametrize("pitch", [0.0, 6.0])
def test_admm_matches_matrix_form(orc, pitch):
    rng = np.random.default_rng(11)
    P, rows, T, X = 5, 4, 6, 9
    c = orc.coeffs_mpi(tilt_set(P, 55.0), 0.0, pitch, 0.0)
    org = orc.origin(X, T, T / 2)
    proj = rng.random((P, rows, X))
    Ws = [sysmat(c[k], X, rows, T, org) for k in range(P)]
    ps = [proj[k].ravel() for k in range(P)]
    ref = _matrix_admm(Ws, ps, 3, 4, 0.5, 0.02)
    got = orc.admm_f64(c, proj, T, org, 3, 4, 0.5, 0.02)
    np.testing.assert_allclose(got.ravel(), ref, rtol=1e-9, atol=1e-11)
    g32 = orc.admm_f32(c, proj.astype(np.float32), T, org, 3, 4, 0.5, 0.02)
    assert orc.rel_l2(g32, ref.reshape(g32.shape)) < 1e-5


def test_admm_soft_zero_keeps_d_zero_and_one_outer_is_cg(orc):
    """With soft = 0, u = vol + d and d stays 0.  One outer iteration is
    then cgiter steps of the reference's CG on (sum W^T D W + mu I) v = htb
    from v = 0; with cgiter = 1 this is one steepest-descent step,
    v = (r.r / r.Ar) r with r = htb."""
    rng = np.random.default_rng(12)
    P, rows, T, X = 4, 3, 5, 8
    c = orc.coeffs_mpi(tilt_set(P, 40.0))
    org = orc.origin(X, T, T / 2)
    proj = rng.random((P, rows, X))
    Ws = [sysmat(c[k], X, rows, T, org) for k in range(P)]
    mu = 0.3
    htb = sum(W.T @ proj[k].ravel() for k, W in enumerate(Ws))
    A = mu * np.eye(htb.size)
    for W in Ws:
        w = W.sum(axis=1)
        A = A + W.T @ np.diag(np.where(np.abs(w) > 1e-5, 1 / np.where(w == 0, 1, w), 0)) @ W
    step = (htb @ htb) / (htb @ A @ htb) * htb
    got = orc.admm_f64(c, proj, T, org, 1, 1, mu, 0.0)
    np.testing.assert_allclose(got.ravel(), step, rtol=1e-10, atol=1e-12)
