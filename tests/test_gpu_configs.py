"""Parity at the benchmark shapes.

At pitch 0 every volume row depends only on the same row of every
projection, so the CPU oracle evaluated on a few rows is exact for those
rows of the full-size GPU reconstruction.  The oracle runs on the sampled
rows in parallel threads (ctypes releases the GIL).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WORKERS = min(8, os.cpu_count() or 4)


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


def gpu_run(torch, trec, ctx, ang, stack, T, method, storage="f32"):
    P, Y, X = stack.shape
    g = trec.make_geom(X, Y, P, T)
    out = torch.empty((Y, T, X), dtype=torch.float32, device="cuda")
    ctx.reconstruct(g, ang, torch.from_numpy(stack).cuda(), trec.make_method(method, storage=storage), out)
    torch.cuda.synchronize()
    return out


def rows_oracle(fn, rows, stack):
    with ThreadPoolExecutor(WORKERS) as ex:
        return list(ex.map(lambda y: fn(stack[:, y:y + 1, :]), rows))


def test_config_wbp256_full(torch_cuda, trec, orc):
    """41 x 256^2 -> 256^2 x 64, WBP (the CPU-runnable configuration)."""
    from autom2_amd import synth
    ang, stack, _, _ = synth.config_series("wbp256")
    T = 64
    ctx = trec.Context(0)
    got = gpu_run(torch_cuda, trec, ctx, ang, stack, T, "WBP").cpu().numpy()
    ctx.close()
    coef = orc.coeffs_mpi(ang)
    org = orc.origin(256, T, T / 2)
    rows = list(range(0, 256, 8))
    ref = rows_oracle(lambda s: orc.bpt_f32(coef, orc.rweight_f64(s.astype(np.float64)).astype(np.float32), T, org),
                      rows, stack)
    e = orc.rel_l2(got[rows], np.concatenate(ref))
    print(f"wbp256: rel L2 vs oracle on {len(rows)} rows {e:.3e}")
    assert e < 1e-5


def test_config_sirt1024_rows(torch_cuda, trec, orc):
    """61 x 1024^2 -> 1024^2 x 256, SIRT x20, FP32 (the benchmark workload)."""
    from autom2_amd import synth
    ang, stack, _, _ = synth.config_series("sirt1024")
    T = 256
    ctx = trec.Context(0)
    vol = gpu_run(torch_cuda, trec, ctx, ang, stack, T, "SIRT,20,1.0")
    rows = [0, 333, 512, 1023]
    got = vol[rows].cpu().numpy()
    del vol
    ctx.close()
    coef = orc.coeffs_mpi(ang)
    org = orc.origin(1024, T, T / 2)
    ref = np.concatenate(rows_oracle(lambda s: orc.sirt_f32(coef, s, T, org, 20, 1.0), rows, stack))
    e = orc.rel_l2(got, ref)
    print(f"sirt1024 x20: rel L2 vs FP32 oracle on rows {rows}: {e:.3e}")
    assert e < 1e-5


@pytest.mark.timeout(400)
def test_precision_forms_sirt2048(torch_cuda, trec, orc):
    """61 x 2048^2 -> 2048^2 x 512: FP32, FP16 and BF16 storage against the
    FP64 oracle at iterations 1, 10 and 30 (two rows)."""
    from autom2_amd import synth
    ang, stack, _, _ = synth.config_series("sirt2048")
    T = 512
    rows = [700, 1400]
    coef = orc.coeffs_mpi(ang)
    org = orc.origin(2048, T, T / 2)
    ref64 = {}
    vols = [None] * len(rows)
    done = 0
    with ThreadPoolExecutor(WORKERS) as ex:
        for it in (1, 10, 30):
            n = it - done  # continue each row's FP64 iteration from the last checkpoint
            vols = list(ex.map(lambda yv: orc.sirt_f64(coef, stack[:, yv[0]:yv[0] + 1, :], T, org, n, 1.0,
                                                       vol=yv[1]), zip(rows, vols)))
            done = it
            ref64[it] = np.concatenate(vols)
    ctx = trec.Context(0)
    for sto, bound in (("f32", 1e-5), ("f16", 1e-3), ("bf16", 5e-3)):
        for it in (1, 10, 30):
            vol = gpu_run(torch_cuda, trec, ctx, ang, stack, T, f"SIRT,{it},1.0", storage=sto)
            got = vol[rows].cpu().numpy()
            del vol
            e = orc.rel_l2(got, ref64[it])
            print(f"sirt2048 {sto} x{it}: rel L2 vs FP64 oracle {e:.3e}")
            assert e < bound, (sto, it, e)
    ctx.close()


def batch_mix(n, seed, methods):
    """The batching configuration's series: size {512, 1024, 2048} px square,
    tilts {41, 61, 81} at 3, 2 and 1.5 degree steps, thickness 128..512."""
    rng = np.random.default_rng(seed)
    steps = {41: 3.0, 61: 2.0, 81: 1.5}
    out = []
    for _ in range(n):
        X = int(rng.choice([512, 1024, 2048]))
        P = int(rng.choice([41, 61, 81]))
        T = int(rng.integers(128, 513))
        m = str(rng.choice(methods))
        ang = (np.arange(P) * steps[P] - (P - 1) * steps[P] / 2).astype(np.float32)
        out.append((X, P, T, m, ang))
    return out


def smooth_stack(torch, P, X, seed):
    """A device stack of smooth random projections plus noise: low-resolution
    random fields upsampled bilinearly, scaled to ~50, with 10 % noise."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    base = torch.rand((P, 1, max(2, X // 64), max(2, X // 64)), device="cuda", generator=g)
    s = torch.nn.functional.interpolate(base, size=(X, X), mode="bilinear", align_corners=True)[:, 0]
    s = 50.0 * s + 5.0 * torch.rand((P, X, X), device="cuda", generator=g)
    return s.contiguous()


def oracle_row(orc, m, coef, row, T, X, f64=False):
    """The oracle's reconstruction of one volume row from that row of every tilt."""
    org = orc.origin(X, T, T / 2)
    name = m.split(",")[0]
    if name == "WBP":
        filt = orc.rweight_f64(row.astype(np.float64))
        return orc.bpt_f64(coef, filt, T, org) if f64 else orc.bpt_f32(coef, filt.astype(np.float32), T, org)
    it, gam = int(m.split(",")[1]), float(m.split(",")[2])
    return (orc.sirt_f64 if f64 else orc.sirt_f32)(coef, row, T, org, it, gam)


@pytest.mark.timeout(600)
def test_config_batching_mix(torch_cuda, trec, orc):
    """The batching configuration: 16 series of the mix (WBP and SIRT x10, FP32)
    in one trec_reconstruct_batch call, one HIP graph; one sampled row of
    every volume against the FP32 oracle."""
    torch = torch_cuda
    mix = batch_mix(16, 1003, ["WBP", "SIRT,10,1.0"])
    series, rows = [], []
    for k, (X, P, T, m, ang) in enumerate(mix):
        stack = smooth_stack(torch, P, X, 2000 + k)
        vol = torch.empty((X, T, X), device="cuda")
        series.append(trec.Series(trec.make_geom(X, X, P, T), trec.make_method(m), ang, stack, vol))
        rows.append((7 * k + 13) % X)
    ctx = trec.Context(0)
    ctx.reconstruct_batch(series)
    torch.cuda.synchronize()
    got = [s.volume[y].cpu().numpy() for s, y in zip(series, rows)]
    srows = [s.stack[:, y:y + 1, :].cpu().numpy() for s, y in zip(series, rows)]
    del series
    ctx.close()
    torch.cuda.empty_cache()

    def ref(k):
        X, P, T, m, ang = mix[k]
        return oracle_row(orc, m, orc.coeffs_mpi(ang), srows[k], T, X)[0]

    with ThreadPoolExecutor(16) as ex:
        refs = list(ex.map(ref, range(len(mix))))
    for k, ((X, P, T, m, _), g, r) in enumerate(zip(mix, got, refs)):
        e = orc.rel_l2(g, r)
        print(f"series {k}: {P} x {X}^2 -> thickness {T}, {m}: rel L2 vs FP32 oracle {e:.3e}")
        assert e < 1e-5, k


@pytest.mark.timeout(600)
def test_config_slab_volume_4096(torch_cuda, trec, orc):
    """The batching configuration's large volume: 61 x 4096^2 -> 4096^2 x 1000
    FP32, SIRT x10, reconstructed slab-wise on one GPU (a 16 GB workspace
    budget, so about 50 rows per slab) into a 67 GB device volume; two rows
    against the FP32 oracle."""
    torch = torch_cuda
    X, P, T = 4096, 61, 1000
    ang = (np.arange(P) * 2.0 - 60.0).astype(np.float32)
    stack = smooth_stack(torch, P, X, 4096)
    g = trec.make_geom(X, X, P, T)
    m = trec.make_method("SIRT,10,1.0")
    ctx = trec.Context(0)
    ctx.set_workspace_limit(16 << 30)
    info = ctx.plan(g, ang, m)
    print(f"4096^2 x 1000 plan: {info}")
    assert info["slabs"] > 1 and info["workspace"] <= 16 << 30
    vol = torch.empty((X, T, X), device="cuda")
    ctx.reconstruct(g, ang, stack, m, vol)
    torch.cuda.synchronize()
    rows = [5, 3333]
    got = [vol[y].cpu().numpy() for y in rows]
    srows = [stack[:, y:y + 1, :].cpu().numpy() for y in rows]
    del vol, stack
    ctx.close()
    torch.cuda.empty_cache()
    coef = orc.coeffs_mpi(ang)
    with ThreadPoolExecutor(2) as ex:
        refs = list(ex.map(lambda s: oracle_row(orc, "SIRT,10,1.0", coef, s, T, X)[0], srows))
    for y, gr, r in zip(rows, got, refs):
        e = orc.rel_l2(gr, r)
        print(f"4096^2 x 1000 SIRT x10, row {y}: rel L2 vs FP32 oracle {e:.3e}")
        assert e < 1e-5


@pytest.mark.timeout(600)
def test_config_serving_fp16(torch_cuda, trec, orc):
    """The serving configuration: requests of the batching mix's sizes with
    WBP, SIRT x10 or SIRT x20, FP16 storage, arriving as a Poisson stream
    (2 per second) at the continuous-batching server; every other request
    returns a half volume (MRC mode 12) and, for SIRT, also takes a half
    stack (WBP keeps a float stack: its ramp filter raises the stack's half
    rounding to about 3e-3 of the volume).  One row of every served volume
    against the FP64 oracle on the float stack, at 1e-3: half input and
    output are part of the form's error."""
    import time
    torch = torch_cuda
    mix = batch_mix(16, 1004, ["WBP", "SIRT,10,1.0", "SIRT,20,1.0"])
    stacks = [smooth_stack(torch, P, X, 3000 + k).cpu().numpy() for k, (X, P, T, m, ang) in enumerate(mix)]
    torch.cuda.empty_cache()
    gaps = np.random.default_rng(1005).exponential(0.5, len(mix))
    srv = trec.Server(0, budget=128 << 30, max_active=16)
    ids, outs = [], []
    t0 = time.perf_counter()
    for k, ((X, P, T, m, ang), st) in enumerate(zip(mix, stacks)):
        time.sleep(float(gaps[k]))
        half = k % 2 == 1
        out = np.empty((X, T, X), np.float16 if half else np.float32)
        outs.append(out)
        half_in = half and not m.startswith("WBP")
        ids.append(srv.submit(trec.make_geom(X, X, P, T), trec.make_method(m, storage="f16"), ang,
                              st.astype(np.float16) if half_in else st, out))
    recs = [srv.wait(i) for i in ids]
    srv.close()
    print(f"served {len(mix)} requests in {time.perf_counter() - t0:.1f} s")
    rows = [(11 * k + 5) % mix[k][0] for k in range(len(mix))]

    def ref(k):
        X, P, T, m, ang = mix[k]
        return oracle_row(orc, m, orc.coeffs_mpi(ang), stacks[k][:, rows[k]:rows[k] + 1, :], T, X, f64=True)[0]

    with ThreadPoolExecutor(16) as ex:
        refs = list(ex.map(ref, range(len(mix))))
    for k, ((X, P, T, m, _), out, rec, r) in enumerate(zip(mix, outs, recs, refs)):
        assert rec["status"] == 0
        e = orc.rel_l2(out[rows[k]].astype(np.float32), r)
        io = ("half out" if m.startswith("WBP") else "half in and out") if k % 2 else "float in and out"
        print(f"request {k}: {P} x {X}^2 -> thickness {T}, {m}, FP16, {io}: rel L2 vs FP64 oracle {e:.3e}")
        assert e < 1e-3, k
