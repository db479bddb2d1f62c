"""Batching, HIP-graph replay and the serving loop on the GPU.

A job's volume must be bit for bit the same alone, in a batch, in a replayed
graph, and in the server's changing mix: every kernel computes each job from
its own data in a fixed order.
"""
import time

import numpy as np
import pytest

from tests.helpers import noisy_stack, smooth_volume, tilts

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def trec(torch_cuda):
    from autom2_amd import tiltrec
    tiltrec.lib()
    return tiltrec


def make_series(orc, P, Y, T, X, seed, span=60.0):
    ang = tilts(P, span)
    coef = orc.coeffs_mpi(ang)
    stack = noisy_stack(orc, coef, smooth_volume(Y, T, X, seed), orc.origin(X, T, T / 2), seed + 1)
    return ang, stack


MIX = [  # (P, Y, T, X, method, storage)
    (9, 20, 12, 40, "SIRT,4,1.0", "f32"),
    (7, 33, 8, 64, "WBP", "f32"),
    (11, 16, 24, 72, "SIRT,6,0.8", "f16"),
    (5, 18, 10, 30, "SART,2,0.6", "f32"),
    (13, 17, 16, 48, "BPT", "f32"),
    (9, 24, 20, 56, "SIRT,3,1.0", "bf16"),
]


def alone(torch, trec, ctx, ang, stack, T, method, storage):
    P, Y, X = stack.shape
    g = trec.make_geom(X, Y, P, T)
    out = torch.empty((Y, T, X), dtype=torch.float32, device="cuda")
    ctx.reconstruct(g, ang, torch.from_numpy(stack).cuda(), trec.make_method(method, storage=storage), out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_batch_equals_alone(torch_cuda, trec, orc):
    ctx = trec.Context(0)
    series, refs = [], []
    for k, (P, Y, T, X, method, sto) in enumerate(MIX):
        ang, stack = make_series(orc, P, Y, T, X, 40 + k)
        refs.append(alone(torch_cuda, trec, ctx, ang, stack, T, method, sto))
        series.append(trec.Series(trec.make_geom(X, Y, P, T), trec.make_method(method, storage=sto), ang,
                                  torch_cuda.from_numpy(stack).cuda(),
                                  torch_cuda.empty((Y, T, X), dtype=torch_cuda.float32, device="cuda")))
    ctx.reconstruct_batch(series)
    torch_cuda.cuda.synchronize()
    for s, ref, m in zip(series, refs, MIX):
        np.testing.assert_array_equal(s.volume.cpu().numpy(), ref, err_msg=str(m))
    ctx.close()


def test_graph_replays_match_eager(torch_cuda, trec, orc, monkeypatch):
    """The batch path runs as a cached HIP graph; replay it several times with
    eager work and synchronisations in between and check every replay."""
    ctx = trec.Context(0)
    P, Y, T, X = 9, 20, 12, 40
    ang, stack = make_series(orc, P, Y, T, X, 60)
    s = trec.Series(trec.make_geom(X, Y, P, T), trec.make_method("SIRT,5,1.0"), ang,
                    torch_cuda.from_numpy(stack).cuda(), torch_cuda.empty((Y, T, X), device="cuda"))
    monkeypatch.setenv("TREC_NO_GRAPH", "1")
    ctx.reconstruct_batch([s])
    torch_cuda.cuda.synchronize()
    eager = s.volume.cpu().numpy()
    monkeypatch.delenv("TREC_NO_GRAPH")
    for rep in range(4):
        s.volume.fill_(float(rep + 7))                  # eager work between replays
        junk = torch_cuda.randn(1 << 20, device="cuda").sum()
        torch_cuda.cuda.synchronize()
        ctx.reconstruct_batch([s])
        torch_cuda.cuda.synchronize()
        np.testing.assert_array_equal(s.volume.cpu().numpy(), eager, err_msg=f"replay {rep}")
        del junk
    ctx.close()


def test_sart_graph_path_matches_oracle(torch_cuda, trec, orc):
    """SART with many stages takes the graph path in trec_reconstruct."""
    ctx = trec.Context(0)
    P, Y, T, X = 17, 16, 8, 36
    ang, stack = make_series(orc, P, Y, T, X, 70)
    got = alone(torch_cuda, trec, ctx, ang, stack, T, "SART,5,0.5", "f32")   # 85 stages
    ref = orc.sart_f32(orc.coeffs_mpi(ang), stack, T, orc.origin(X, T, T / 2), 5, 0.5)
    assert orc.rel_l2(got, ref) < 1e-5
    ctx.close()


def test_server_mix_equals_alone(torch_cuda, trec, orc):
    ctx = trec.Context(0)
    reqs = []
    for k, (P, Y, T, X, method, sto) in enumerate(MIX * 2):
        ang, stack = make_series(orc, P, Y, T, X, 80 + k)
        reqs.append((ang, stack, T, method, sto, alone(torch_cuda, trec, ctx, ang, stack, T, method, sto)))
    ctx.close()
    srv = trec.Server(0, budget=1 << 30, max_active=5)
    ids = []
    outs = []
    for ang, stack, T, method, sto, _ in reqs:
        P, Y, X = stack.shape
        out = np.zeros((Y, T, X), np.float32)
        outs.append(out)
        ids.append(srv.submit(trec.make_geom(X, Y, P, T), trec.make_method(method, storage=sto), ang, stack, out))
        time.sleep(0.002)  # staggered arrivals join a running mix
    recs = [srv.wait(i) for i in ids]
    srv.close()
    for (ang, stack, T, method, sto, ref), out, rec in zip(reqs, outs, recs):
        assert rec["status"] == 0
        assert rec["submit_s"] <= rec["admit_s"] <= rec["done_s"]
        np.testing.assert_array_equal(out, ref, err_msg=method)


def test_server_admission_by_memory(torch_cuda, trec, orc):
    """With a budget that holds one request at a time the server still
    completes all of them, admitting each after the previous one freed its
    memory; a request larger than the whole budget is refused at submit."""
    P, Y, T, X = 9, 32, 16, 64
    g = trec.make_geom(X, Y, P, T)
    m = trec.make_method("SIRT,3,1.0")
    fp = trec.request_footprint(g, m)
    srv = trec.Server(0, budget=int(fp * 1.5), max_active=4)
    outs, ids = [], []
    stacks = []
    for k in range(3):
        ang, stack = make_series(orc, P, Y, T, X, 90 + k)
        stacks.append(stack)
        out = np.zeros((Y, T, X), np.float32)
        outs.append(out)
        ids.append(srv.submit(g, m, ang, stack, out))
    recs = [srv.wait(i) for i in ids]
    for a, b in zip(recs, recs[1:]):
        assert b["admit_s"] >= a["done_s"] - 1e-3   # one at a time
    big = trec.make_geom(4 * X, 4 * Y, P, T)
    with pytest.raises(trec.TrecError) as ei:
        srv.submit(big, m, tilts(P), np.zeros((P, 4 * Y, 4 * X), np.float32),
                   np.zeros((4 * Y, T, 4 * X), np.float32))
    assert ei.value.status == trec.ERR_BUSY
    srv.close()
    for stack, out in zip(stacks, outs):
        ref = orc.sirt_f32(orc.coeffs_mpi(tilts(P)), stack, T, orc.origin(X, T, T / 2), 3, 1.0)
        assert orc.rel_l2(out, ref) < 1e-5


def test_batch_graph_follows_geometry(torch_cuda, trec, orc):
    """The cached graph is keyed on the per-tilt geometry: the same buffers
    reconstructed with another pitch, offset or set of angles must not
    replay the first call's kernel choices (row windows, candidate counts,
    the general or row path)."""
    ctx = trec.Context(0)
    P, Y, T, X = 9, 24, 12, 48
    stack = np.random.default_rng(5).random((P, Y, X)).astype(np.float32)
    dstack = torch_cuda.from_numpy(stack).cuda()
    vol = torch_cuda.empty((Y, T, X), device="cuda")
    cases = [(0.0, 4.0, tilts(P)), (0.0, 12.0, tilts(P)), (2.0, 4.0, tilts(P)), (0.0, 0.0, tilts(P)),
             (0.0, 4.0, tilts(P, 75.0)), (0.0, 0.0, tilts(P, 80.0))]
    for offset, pitch, ang in cases:
        g = trec.make_geom(X, Y, P, T, offset, pitch, 0.0)
        ctx.reconstruct_batch([trec.Series(g, trec.make_method("SIRT,3,1.0"), ang, dstack, vol)])
        torch_cuda.cuda.synchronize()
        ref = orc.sirt_f32(orc.coeffs_mpi(ang, offset, pitch, 0.0), stack, T, orc.origin(X, T, T / 2), 3, 1.0)
        err = orc.rel_l2(vol.cpu().numpy(), ref)
        print(f"offset {offset} pitch {pitch} span {ang[-1]}: rel L2 {err:.2e}")
        assert err < 1e-5, (offset, pitch, ang[-1])
    ctx.close()


def test_calls_on_two_streams_share_the_workspace(torch_cuda, trec, orc):
    """Back-to-back calls on one context but different streams: the second
    call's descriptors and workspace must wait for the first call's kernels."""
    torch = torch_cuda
    ctx = trec.Context(0)
    a = make_series(orc, 31, 64, 64, 256, 110)
    b = make_series(orc, 9, 20, 12, 40, 111)
    refs = [alone(torch, trec, ctx, a[0], a[1], 64, "SIRT,8,1.0", "f32"),
            alone(torch, trec, ctx, b[0], b[1], 12, "SIRT,2,1.0", "f16")]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(3):
        outs = []
        for (ang, stack), T, meth, sto, st in ((a, 64, "SIRT,8,1.0", "f32", s1), (b, 12, "SIRT,2,1.0", "f16", s2)):
            P, Y, X = stack.shape
            d = torch.from_numpy(stack).cuda()
            o = torch.empty((Y, T, X), device="cuda")
            torch.cuda.synchronize()
            ctx.reconstruct(trec.make_geom(X, Y, P, T), ang, d, trec.make_method(meth, storage=sto), o, stream=st)
            outs.append((o, d))
        torch.cuda.synchronize()
        for (o, _), ref in zip(outs, refs):
            np.testing.assert_array_equal(o.cpu().numpy(), ref, err_msg=f"rep {rep}")
    ctx.close()


def test_server_steep_tilts_and_thin_volume(torch_cuda, trec, orc):
    """Requests whose real angles take the general path (tilts past about 78
    degrees) are sized with those angles at submit: a thin SART volume and a
    SIRT volume, each checked against the oracle."""
    srv = trec.Server(0, budget=256 << 20, max_active=4)
    cases = [(9, 16, 3, 40, "SART,2,0.7", 80.0), (11, 20, 10, 36, "SIRT,3,1.0", 82.0)]
    ids, outs, data = [], [], []
    for k, (P, Y, T, X, meth, span) in enumerate(cases):
        ang, stack = make_series(orc, P, Y, T, X, 120 + k, span)
        out = np.zeros((Y, T, X), np.float32)
        outs.append(out)
        data.append((ang, stack))
        ids.append(srv.submit(trec.make_geom(X, Y, P, T), trec.make_method(meth), ang, stack, out))
    recs = [srv.wait(i) for i in ids]
    srv.close()
    for (P, Y, T, X, meth, span), (ang, stack), out, rec in zip(cases, data, outs, recs):
        assert rec["status"] == 0
        it, gam = int(meth.split(",")[1]), float(meth.split(",")[2])
        fn = orc.sart_f32 if meth.startswith("SART") else orc.sirt_f32
        ref = fn(orc.coeffs_mpi(ang), stack, T, orc.origin(X, T, T / 2), it, gam)
        assert orc.rel_l2(out, ref) < 1e-5, meth


def test_server_admm_equals_alone(torch_cuda, trec, orc):
    """ADMM requests run through the server as one stage each (the whole CG
    recursion), beside SIRT requests; each volume equals trec_reconstruct's
    alone, bit for bit, on the row path and with a pitch (general path)."""
    ctx = trec.Context(0)
    cases = [(9, 20, 12, 40, "ADMM,2,3,0.5,0.001", 0.0), (7, 18, 10, 36, "SIRT,3,1.0", 0.0),
             (9, 16, 10, 32, "ADMM,2,2,0.8,0.0", 4.0)]
    refs, data = [], []
    for k, (P, Y, T, X, meth, pitch) in enumerate(cases):
        ang, stack = make_series(orc, P, Y, T, X, 130 + k)
        g = trec.make_geom(X, Y, P, T, 0.0, pitch, 0.0)
        out = torch_cuda.empty((Y, T, X), device="cuda")
        ctx.reconstruct(g, ang, torch_cuda.from_numpy(stack).cuda(), trec.make_method(meth), out)
        torch_cuda.cuda.synchronize()
        refs.append(out.cpu().numpy())
        data.append((g, ang, stack))
    ctx.close()
    srv = trec.Server(0, budget=512 << 20, max_active=4)
    ids, outs = [], []
    for (P, Y, T, X, meth, pitch), (g, ang, stack) in zip(cases, data):
        out = np.zeros((Y, T, X), np.float32)
        outs.append(out)
        ids.append(srv.submit(g, trec.make_method(meth), ang, stack, out))
    recs = [srv.wait(i) for i in ids]
    srv.close()
    for c, out, ref, rec in zip(cases, outs, refs, recs):
        assert rec["status"] == 0
        np.testing.assert_array_equal(out, ref, err_msg=c[4])


def test_device_allocation_failures_are_reported(torch_cuda, trec):
    """Allocation failures come back as TREC_ERR_NOMEM and leave the context
    and the device usable: a workspace larger than the device (ADMM keeps
    seven volume-sized vectors and does not run in slabs), and a server arena
    of 1 PiB."""
    torch = torch_cuda
    ctx = trec.Context(0)
    X, Y, T, P = 4096, 4096, 1024, 61
    stack = torch.zeros((P, Y, X), device="cuda")
    vol = torch.empty((Y, T, X), device="cuda")
    g = trec.make_geom(X, Y, P, T)
    m = trec.make_method("ADMM,1,1,0.5,0.0")
    print(f"workspace {trec.workspace_size(g, m) / 2**30:.0f} GiB")
    with pytest.raises(trec.TrecError) as ei:   # refused by the budget (free device memory)
        ctx.reconstruct(g, tilts(P), stack, m, vol)
    assert ei.value.status == trec.ERR_NOMEM
    ctx.set_workspace_limit(1 << 62)            # no budget: hipMalloc itself fails
    with pytest.raises(trec.TrecError) as ei:
        ctx.reconstruct(g, tilts(P), stack, m, vol)
    assert ei.value.status == trec.ERR_NOMEM
    print(ei.value)
    ctx.set_workspace_limit(0)
    del stack, vol
    torch.cuda.empty_cache()
    s = torch.ones((5, 16, 32), device="cuda")
    v = torch.empty((16, 8, 32), device="cuda")
    ctx.reconstruct(trec.make_geom(32, 16, 5, 8), tilts(5), s, trec.make_method("BPT"), v)
    torch.cuda.synchronize()
    assert float(v[:, :, 12:20].min()) == 5.0
    ctx.close()
    with pytest.raises(trec.TrecError) as ei:
        trec.Server(0, budget=1 << 50, max_active=2)
    assert ei.value.status == trec.ERR_NOMEM
