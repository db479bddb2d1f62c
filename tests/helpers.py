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
