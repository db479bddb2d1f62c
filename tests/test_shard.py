"""The y-split of one reconstruction over several devices (autom2_amd/shard.py),
checked on the CPU: the plan against brute-force tap sets, and a two-rank
gloo run of the oracle per shard, with the halo exchange before every SIRT
iteration, against the whole volume.

TiltRec-mpi's own split (main-mpi.cpp:710-740) widens each rank's rows by
ceil(|tan pitch| * T) but never exchanges them, so its iterative results at a
pitch depend on the rank count; the plan here reproduces the single-volume
result instead.
"""
import os
import socket

import numpy as np
import pytest

from autom2_amd import shard as sh

PI_180 = np.float32(0.01745329252)


def _geometry(orc, pitch, offset=0.0, zshift=0.0, P=9, X=40, Y=64, T=12):
    c = orc.coeffs_mpi(np.linspace(-60, 60, P).astype(np.float32), offset, pitch, zshift)
    return c, orc.origin(X, T, Y / 2), X, Y, T


def _tap_rows(coef, X, T, org, y, Y):
    """Detector rows (in [0, Y)) tapped by any voxel of volume row y, with
    ValCoef's float coordinate differences (main-mpi.cpp:204-226)."""
    ox, oy, oz = (np.float32(o) for o in org)
    x = np.arange(X, dtype=np.float32)
    z = np.arange(T, dtype=np.float32)
    Xc = (x - ox).astype(np.float64)[None, :]
    Zc = (z - oz).astype(np.float64)[:, None]
    Yc = float(np.float32(y) - oy)
    out = set()
    for b0, b1, b2, b3 in np.asarray(coef, np.float64)[:, 4:8]:
        v = b0 + b1 * Xc + b2 * Yc + b3 * Zc + float(oy)
        f = np.floor(v).astype(np.int64).ravel()
        for j in np.unique(np.concatenate([f, f + 1])):
            if 0 <= j < Y:
                out.add(int(j))
    return out


def test_flat_geometry_needs_no_halo(orc):
    c, org, X, Y, T = _geometry(orc, 0.0)
    shards = sh.y_shards(Y, 3, c, X, T, org)
    assert [s.rows for s in shards] == [(0, 16), (16, 32), (32, 64)]
    assert all(s.work == s.rows == s.detector for s in shards)
    assert sh.exchange(shards) == []


@pytest.mark.parametrize("pitch,offset,zshift,n", [(8.0, 0.0, 0.0, 2), (-15.0, 2.5, 1.0, 3), (25.0, -1.0, -2.0, 4)])
def test_plan_covers_every_dependency(orc, pitch, offset, zshift, n):
    """Owned voxels' taps lie in the detector range, and every voxel row that
    taps the detector range lies in the work range (brute force over all
    voxels and tilts)."""
    c, org, X, Y, T = _geometry(orc, pitch, offset, zshift)
    shards = sh.y_shards(Y, n, c, X, T, org)
    assert shards[0].rows[0] == 0 and shards[-1].rows[1] == Y
    assert all(a.rows[1] == b.rows[0] for a, b in zip(shards, shards[1:]))
    taps = [_tap_rows(c, X, T, org, y, Y) for y in range(Y)]
    for s in shards:
        need = set().union(*(taps[y] for y in range(*s.rows)))
        assert need and min(need) >= s.detector[0] and max(need) < s.detector[1]
        reach = [y for y in range(Y) if taps[y] & need]
        assert min(reach) >= s.work[0] and max(reach) < s.work[1]
        assert s.work[0] <= s.detector[0] and s.detector[1] <= s.work[1]
    # the halo is no wider than a few rows beyond TiltRec-mpi's margin
    margin = int(np.ceil(abs(np.tan(np.radians(pitch))) * T))
    assert max(max(s.halo) for s in shards) <= 2 * margin + 6


def test_exchange_is_what_each_halo_lacks(orc):
    c, org, X, Y, T = _geometry(orc, 20.0, 1.0)
    shards = sh.y_shards(Y, 3, c, X, T, org)
    xfer = sh.exchange(shards)
    for d, s in enumerate(shards):
        got = sorted((lo, hi) for src, dst, lo, hi in xfer if dst == d)
        rows = set(range(*s.work)) - set(range(*s.rows))
        assert set().union(*(set(range(lo, hi)) for lo, hi in got)) == rows
    for src, dst, lo, hi in xfer:
        assert shards[src].rows[0] <= lo < hi <= shards[src].rows[1]


def _free_port():
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def _run_rank(rank, world, port, case, outdir):
    """One rank: its shard of the case through the oracle, halos refreshed
    from the owners over gloo before every iteration."""
    import torch
    import torch.distributed as dist
    from oracle import orc

    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist.init_process_group("gloo", rank=rank, world_size=world)
    try:
        c, proj, T, org, method, iters = _case(orc, case)
        Y = proj.shape[1]
        shards = sh.y_shards(Y, world, c, proj.shape[2], T, org)
        me = shards[rank]
        ya, yb = me.work
        sorg = sh.sub_origin(me, org)
        sproj = np.ascontiguousarray(proj[:, ya:yb])
        if method == "BPT":
            vol = orc.bpt_f32(c, sproj, T, sorg)
        else:
            vol = np.zeros((yb - ya, T, proj.shape[2]), np.float32)
            xfer = sh.exchange(shards)
            for _ in range(iters):
                reqs = []
                for src, dst, lo, hi in xfer:
                    if src == rank:
                        buf = torch.from_numpy(np.ascontiguousarray(vol[lo - ya:hi - ya]))
                        reqs.append(dist.isend(buf, dst))
                    elif dst == rank:
                        buf = torch.from_numpy(vol[lo - ya:hi - ya])  # a view: received in place
                        reqs.append(dist.irecv(buf, src))
                for r in reqs:
                    r.wait()
                vol = orc.sirt_f32(c, sproj, T, sorg, 1, vol=vol)
        y0, y1 = me.rows
        np.save(os.path.join(outdir, f"rank{rank}.npy"), vol[y0 - ya:y1 - ya])
    finally:
        dist.destroy_process_group()


def _case(orc, case):
    pitch, offset, method, iters = case
    c, org, X, Y, T = _geometry(orc, pitch, offset, P=11, X=48, Y=64, T=10)
    rng = np.random.default_rng(int(abs(pitch) * 10) + iters)
    proj = rng.random((c.shape[0], Y, X), dtype=np.float32)
    return c, proj, T, org, method, iters


@pytest.mark.parametrize("case", [(0.0, 0.0, "SIRT", 4), (0.0, 1.5, "BPT", 1), (12.0, 1.5, "BPT", 1),
                                  (12.0, 1.5, "SIRT", 4), (-20.0, 0.0, "SIRT", 3)])
def test_two_ranks_equal_the_whole_volume(orc, tmp_path, case):
    import torch.multiprocessing as mp

    world = 2
    mp.spawn(_run_rank, args=(world, _free_port(), case, str(tmp_path)), nprocs=world, join=True)
    got = np.concatenate([np.load(tmp_path / f"rank{r}.npy") for r in range(world)])
    c, proj, T, org, method, iters = _case(orc, case)
    whole = orc.bpt_f32(c, proj, T, org) if method == "BPT" else orc.sirt_f32(c, proj, T, org, iters)
    assert got.shape == whole.shape
    if case[0] == 0.0:
        # detector row == volume row: the shard's taps are the whole's exactly
        assert np.array_equal(got, whole)
    else:
        # the shard's origin oy - ya moves the rounding of ValCoef's double
        # v + oy by ulps, which reaches the float fractions now and then
        err = np.abs(got - whole).max(axis=(1, 2)) / np.abs(whole).max()
        assert err.max() < 1e-6, err


def test_missing_halo_is_caught(orc):
    """Planted mistake: shards without their halo rows (TiltRec-mpi's split
    with no margin) differ from the whole volume near the cut."""
    c, proj, T, org, method, iters = _case(orc, (12.0, 1.5, "SIRT", 4))
    Y = proj.shape[1]
    shards = sh.y_shards(Y, 2, c, proj.shape[2], T, org)
    whole = orc.sirt_f32(c, proj, T, org, iters)
    parts = []
    for s in shards:
        y0, y1 = s.rows
        bare = sh.Shard(s.rows, s.rows, s.rows)
        parts.append(orc.sirt_f32(c, np.ascontiguousarray(proj[:, y0:y1]), T, sh.sub_origin(bare, org), iters))
    err = np.abs(np.concatenate(parts) - whole).max() / np.abs(whole).max()
    assert err > 1e-3
