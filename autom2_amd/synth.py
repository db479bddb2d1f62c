"""Synthetic tilt series of the benchmark recipes (libtrec_synth.so).

Recipe: 40 random ellipsoids (semi-axes 4-30 px, density 0.2-1.0) and 10
gold beads (radius 3 px, density 5.0), lengths multiplied by `scale`;
exact line integrals in TiltRec's detector convention; Gaussian noise with
sigma = 0.5 x the mean projection value; per-tilt shifts of +-4 px and
in-plane rotations of +-0.5 degrees.  Seed = 1000 + configuration index.
"""
import ctypes
import os

import numpy as np

from . import LIB_DIR

_lib = None

# name: (X = Y, P, angle step, T, scale, seed)
CONFIGS = {
    "wbp256": dict(X=256, Y=256, P=41, step=3.0, T=64, scale=1.0, seed=1000),
    "sirt1024": dict(X=1024, Y=1024, P=61, step=2.0, T=256, scale=4.0, seed=1001),
    "sirt2048": dict(X=2048, Y=2048, P=61, step=2.0, T=512, scale=8.0, seed=1002),
}


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(LIB_DIR, "libtrec_synth.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build with __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.trec_synth_series.restype = ctypes.c_int
        L.trec_synth_series.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def angles(P, step):
    return (np.arange(P, dtype=np.float64) * step - (P - 1) * step / 2).astype(np.float32)


def tilt_series(X, Y, T, P, step, scale=1.0, seed=1000, n_ellipsoids=40, n_beads=10, noise=0.5,
                shift=4.0, rot=0.5, with_volume=False):
    """Returns (angles [P], stack [P][Y][X] float32, xf [P][6], phantom or None)."""
    ang = angles(P, step)
    stack = np.empty((P, Y, X), np.float32)
    xf = np.empty((P, 6), np.float32)
    vol = np.empty((Y, T, X), np.float32) if with_volume else None
    rc = lib().trec_synth_series(X, Y, T, P, ang.ctypes.data, n_ellipsoids, n_beads, scale, noise, shift, rot,
                                 seed, stack.ctypes.data, xf.ctypes.data,
                                 vol.ctypes.data if vol is not None else None)
    if rc != 0:
        raise ValueError("bad synthetic series arguments")
    return ang, stack, xf, vol


def config_series(name, **kw):
    c = CONFIGS[name]
    return tilt_series(c["X"], c["Y"], c["T"], c["P"], c["step"], c["scale"], c["seed"], **kw)
