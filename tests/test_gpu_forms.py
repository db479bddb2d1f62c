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
