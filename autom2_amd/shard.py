"""Y-split of one TiltRec reconstruction over several devices (host-side plan).

TiltRec-mpi splits the volume rows over MPI ranks (main-mpi.cpp:710-740):
rank k reconstructs rows [start, start + length) from the same rows of
every projection, widened by add_left / add_right = ceil(|tan pitch| * T)
rows when the specimen is pitched, and never exchanges volume rows.  This
module plans the same split so that every shard's owned rows equal the
whole reconstruction's:

* At pitch 0 the detector row of voxel (x, y, z) is y itself, so a shard of
  rows [y0, y1) reads only projection rows [y0, y1): no halo and no
  exchange.  The FP32 forms are then bit-identical to the whole volume;
  the FP16 / BF16 forms need the stack's global RMS (one all-reduced double
  of sums of squares) so that every shard scales its copy alike.
* With a pitch, v = b0 + b1 X + b2 Y + b3 Z + oy moves over x and z, so the
  owned voxels read detector rows J = [j0, j1) beyond [y0, y1), and the
  forward projection of those rows sums volume rows of the neighbours.  A
  shard therefore works on rows R = [ya, yb) of both volume and stack,
  R covering J and every voxel row that reaches J.  One-pass methods (BPT,
  WBP, FBP) read R's projection rows once.  SIRT needs the neighbours'
  current rows R \\ [y0, y1) before every iteration (exchange(), a halo of
  a few rows of T x X floats per side); SART before every tilt.

A shard's sub-problem is an ordinary reconstruction of yb - ya rows with the
y origin moved to oy - ya (ValCoef, main-mpi.cpp:204-226, in row indices
relative to ya).  Owned ranges are whole 16-row blocks, the unit of the
row-blocked kernels.
"""
import math
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Shard:
    rows: tuple      # (y0, y1): volume rows this shard owns and writes
    work: tuple      # (ya, yb): volume and projection rows of its sub-problem
    detector: tuple  # (j0, j1): projection rows the owned voxels read

    @property
    def halo(self):
        """Rows above and below the owned ones the sub-problem carries."""
        return self.rows[0] - self.work[0], self.work[1] - self.rows[1]


def _v_bounds(coef, X, T, org, ys):
    """Least and greatest detector row v of each volume row in ys over all
    x in [0, X), z in [0, T) and tilts (v is linear in x and z, so its
    extremes are at the ends)."""
    ox, oy, oz = (float(o) for o in org)
    b = np.asarray(coef, np.float64)[:, 4:8]
    ys = np.asarray(ys, np.float64)[:, None]
    lo = np.full(ys.shape[0], np.inf)
    hi = np.full(ys.shape[0], -np.inf)
    for x in (0.0, X - 1.0):
        for z in (0.0, T - 1.0):
            v = b[:, 0] + b[:, 1] * (x - ox) + b[:, 2] * (ys - oy) + b[:, 3] * (z - oz) + oy
            lo = np.minimum(lo, v.min(axis=1))
            hi = np.maximum(hi, v.max(axis=1))
    return lo, hi


def y_shards(ny, nshards, coef, X, T, org, block=16, margin=1):
    """Plans the split of ny volume rows over nshards devices for the tilts'
    [P][8] coefficients (a0..a3, b0..b3) and origin (ox, oy, oz).  margin:
    rows added on each side of every computed range against the rounding of
    the float ValCoef."""
    if nshards < 1:
        raise ValueError("nshards must be at least 1")
    if ny < 1:
        raise ValueError("ny must be at least 1")
    nb = (ny + block - 1) // block
    bounds = [min(ny, block * (nb * k // nshards)) for k in range(nshards + 1)]
    b = np.asarray(coef, np.float64)[:, 4:8]
    flat = bool(np.all(b[:, 0] == 0) and np.all(b[:, 1] == 0) and np.all(b[:, 2] == 1) and np.all(b[:, 3] == 0))
    lo, hi = _v_bounds(coef, X, T, org, np.arange(ny))
    out = []
    for k in range(nshards):
        y0, y1 = bounds[k], bounds[k + 1]
        if y1 <= y0 or flat:
            out.append(Shard((y0, y1), (y0, y1), (y0, y1)))
            continue
        # in-range taps of the owned voxels: floor(v) and floor(v) + 1
        j0 = max(0, math.floor(lo[y0:y1].min()) - margin)
        j1 = min(ny, math.floor(hi[y0:y1].max()) + 2 + margin)
        # voxel rows with a tap in [j0, j1): v in [j0 - 1, j1)
        reach = np.nonzero((hi >= j0 - 1 - margin) & (lo < j1 + margin))[0]
        ya = max(0, min(j0, y0, int(reach.min())))
        yb = min(ny, max(j1, y1, int(reach.max()) + 1))
        out.append(Shard((y0, y1), (ya, yb), (j0, j1)))
    return out


def sub_origin(shard, org):
    """Origin of a shard's sub-problem: rows counted from shard.work[0]."""
    ox, oy, oz = org
    return (ox, np.float32(float(oy) - shard.work[0]), oz)


def exchange(shards):
    """Transfers that refresh every shard's halo from the owners before an
    iterative pass: (src, dst, y_lo, y_hi), rows [y_lo, y_hi) of the whole
    volume owned by shard src and carried by shard dst."""
    out = []
    for d, sd in enumerate(shards):
        for s, ss in enumerate(shards):
            if s == d:
                continue
            lo = max(ss.rows[0], sd.work[0])
            hi = min(ss.rows[1], sd.work[1])
            if hi > lo:
                out.append((s, d, lo, hi))
    return out
