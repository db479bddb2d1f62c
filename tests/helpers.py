"""Shared inputs for the parity tests: seeded tilt series small enough for
the CPU oracle to finish in seconds."""
import numpy as np


def tilts(P, span=60.0):
    return np.linspace(-span, span, P).astype(np.float32)


def smooth_volume(rows, T, X, seed):
    """A blobby volume (sum of Gaussians) in TiltRec-mpi layout [rows][T][X]."""
    rng = np.random.default_rng(seed)
    y, z, x = np.meshgrid(np.arange(rows), np.arange(T), np.arange(X), indexing="ij")
    v = np.zeros((rows, T, X))
    for _ in range(6):
        c = rng.random(3) * [rows, T, X]
        s = 1.5 + rng.random(3) * [rows / 3 + 1, T / 4 + 1, X / 6 + 1]
        v += rng.uniform(0.2, 1.0) * np.exp(-(((y - c[0]) / s[0]) ** 2 + ((z - c[1]) / s[1]) ** 2 + ((x - c[2]) / s[2]) ** 2))
    return v.astype(np.float32)


def noisy_stack(orc, coef, vol, org, seed, noise=0.1):
    """Projections [P][rows][X] of `vol` (ray-weight normalised), plus noise."""
    rng = np.random.default_rng(seed)
    P = coef.shape[0]
    rows, T, X = vol.shape
    out = np.zeros((P, rows, X), np.float32)
    for k in range(P):
        rv, rw = orc.reproject_f64(coef[k], vol, org)
        out[k] = np.where(rw > 0, rv / np.where(rw > 0, rw, 1), 0)
    out += noise * float(np.abs(out).mean()) * rng.standard_normal(out.shape).astype(np.float32)
    return out


def to_layout_z(vol_y):
    """[Y][T][X] -> [T][Y][X]"""
    return np.ascontiguousarray(np.transpose(vol_y, (1, 0, 2)))


def voxel_u(coef, x, z, X, T):
    """Detector column u of voxel (x, z) at pitch 0 for each row of `coef`,
    in ValCoef's double operations (main-mpi.cpp:204-226)."""
    ox, oz = np.float32(X * 0.5), np.float32(T * 0.5)
    Xc = float(np.float32(np.float32(x) - ox))
    Zc = float(np.float32(np.float32(z) - oz))
    c = np.asarray(coef, np.float64)
    return ((c[:, 0] + c[:, 1] * Xc) + c[:, 2] * 0.0 + c[:, 3] * Zc) + float(ox)


def edge_tilt(orc, X, T, wmin, wmax, zshift=0.0):
    """A float tilt angle near 3 degrees at which corner voxel (0, T-1) of an
    X-wide, T-thick volume maps just right of detector column -1: its only
    in-bounds tap (column 0) weighs w = frac(u) with wmin < w <= wmax.  That
    weight is then the voxel's whole weight sum for the tilt, which is what
    TiltRec-cuda's thresholds test (basic.cu:184,409,434).  The root of
    u = -1 is bracketed by bisection, then neighbouring float angles are
    scanned for a qualifying weight."""
    f = lambda a: voxel_u(orc.coeffs_mpi(np.array([a], np.float32), 0.0, 0.0, zshift), 0, T - 1, X, T)[0] + 1.0
    lo, hi = 0.5, 30.0
    assert f(lo) > 0 > f(hi)
    for _ in range(60):
        m = 0.5 * (lo + hi)
        lo, hi = (m, hi) if f(m) > 0 else (lo, m)
    base = np.float32(lo)
    cand = [base]
    up = dn = base
    for _ in range(4000):
        up = np.nextafter(up, np.float32(90))
        dn = np.nextafter(dn, np.float32(0))
        cand += [up, dn]
    ang = np.array(cand, np.float32)
    u = voxel_u(orc.coeffs_mpi(ang, 0.0, 0.0, zshift), 0, T - 1, X, T)
    w = (u - np.floor(u)).astype(np.float32)
    ok = (np.floor(u) == -1) & (w > wmin) & (w <= wmax)
    assert ok.any(), "no float angle gives the wanted tap weight"
    return np.float32(ang[ok][np.argmax(w[ok])])
