"""trec_filter_rweight_static: the R-weighting filter for callers that cannot
allocate, checked through the C ABI against trec_filter_rweight on the same
rows, on two calls in a row (the second with a smaller input), plus its size
limit."""
import ctypes

import numpy as np
import pytest

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


def _static_vs_arena(torch, ctx, rows_np):
    a = torch.from_numpy(rows_np).cuda()
    b = a.clone()
    ctx.filter_rweight(a)
    ctx.filter_rweight_static(b)
    torch.cuda.synchronize()
    want, got = a.cpu().numpy(), b.cpu().numpy()
    err = float(np.linalg.norm((got - want).astype(np.float64)) / np.linalg.norm(want.astype(np.float64)))
    ndiff = int((got.view(np.int32) != want.view(np.int32)).sum())
    print(f"rows {rows_np.shape}: rel L2 {err:.3e}, {ndiff} of {got.size} values differ in any bit")
    return err


def test_static_rweight_matches_the_allocating_entry_point(torch_cuda, trec, orc):
    torch = torch_cuda
    rng = np.random.default_rng(11)
    ctx = trec.Context(0)
    # two calls in a row, the second smaller: 61 rows of 1024 (N = 1126), then
    # 7 rows of 300 (N = 330), then the widest rows the filter holds
    big = rng.standard_normal((61, 1024)).astype(np.float32)
    small = rng.standard_normal((7, 300)).astype(np.float32)
    assert _static_vs_arena(torch, ctx, big) < 1e-6
    assert _static_vs_arena(torch, ctx, small) < 1e-6
    widest = rng.standard_normal((2, 7448)).astype(np.float32)
    assert _static_vs_arena(torch, ctx, widest) < 1e-6
    # against the FP64 filter as well, at the oracle's tolerance for WBP
    b = torch.from_numpy(small).cuda()
    ctx.filter_rweight_static(b)
    torch.cuda.synchronize()
    assert orc.rel_l2(b.cpu().numpy(), orc.rweight_f64(small)) < 1e-6
    ctx.close()


def test_static_rweight_rejects_rows_beyond_its_array(torch_cuda, trec):
    torch = torch_cuda
    ctx = trec.Context(0)
    rows = torch.zeros((1, 7449), device="cuda")
    st = trec.lib().trec_filter_rweight_static(ctx.handle, ctypes.c_void_p(rows.data_ptr()), 7449, 1, None)
    assert st == trec.ERR_INVALID
    assert b"7448" in trec.lib().trec_last_error(ctx.handle)
    ctx.close()
