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
