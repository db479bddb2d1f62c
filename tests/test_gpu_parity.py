"""HIP library vs the CPU oracle on identical seeded inputs.

Tolerances: the plain FP32 form must be within relative L2 1e-5 of the FP32
restatement of TiltRec-mpi; the achieved figures are
printed.  The 16-bit storage forms are checked against the FP64 oracle at
1e-3.  Every call goes through the C ABI (autom2_amd.tiltrec).
"""
import numpy as np
import pytest

from tests.helpers import noisy_stack, smooth_volume, tilts, to_layout_z

pytestmark = pytest.mark.gpu

TOL_F32 = 1e-5


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


@pytest.fixture(scope="module")
def ctx(trec):
    c = trec.Context(0)
    yield c
    c.close()


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def case(orc, P=9, Y=20, T=12, X=40, span=60.0, seed=0, offset=0.0, pitch=0.0, zshift=0.0):
    ang = tilts(P, span)
    coef = orc.coeffs_mpi(ang, offset, pitch, zshift)
    org = orc.origin(X, T, T / 2)
    vol = smooth_volume(Y, T, X, seed)
    stack = noisy_stack(orc, coef, vol, org, seed + 1)
    return ang, coef, org, stack


def run(torch, trec, ctx, stack, ang, T, method, initial=None, offset=0.0, pitch=0.0, zshift=0.0,
        rules=0, layout=0, storage=0):
    P, Y, X = stack.shape
    g = trec.make_geom(X, Y, P, T, offset, pitch, zshift, rules)
    m = trec.make_method(method, storage=storage, layout=layout)
    s = dev(torch, stack)
    out = torch.empty((Y, T, X) if layout == 0 else (T, Y, X), dtype=torch.float32, device="cuda")
    init = dev(torch, initial) if initial is not None else None
    ctx.reconstruct(g, ang, s, m, out, initial=init)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def report(name, got, ref, ref64=None):
    from oracle.orc import rel_l2
    e = rel_l2(got, ref)
    msg = f"{name}: rel L2 vs FP32 oracle {e:.3e}"
    if ref64 is not None:
        msg += f", vs FP64 oracle {rel_l2(got, ref64):.3e}"
    print(msg)
    return e


@pytest.mark.parametrize("shape", [(9, 20, 12, 40), (7, 16, 5, 33), (13, 3, 16, 64)])
def test_bpt(torch_cuda, trec, ctx, orc, shape):
    P, Y, T, X = shape
    ang, coef, org, stack = case(orc, P, Y, T, X)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "BPT")
    ref = orc.bpt_f32(coef, stack, T, org)
    ref64 = orc.bpt_f64(coef, stack.astype(np.float64), T, org)
    assert report("BPT", got, ref, ref64) < 1e-6


@pytest.mark.parametrize("iters", [1, 5])
def test_sirt(torch_cuda, trec, ctx, orc, iters):
    P, Y, T, X = 11, 18, 10, 48
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=3)
    got = run(torch_cuda, trec, ctx, stack, ang, T, f"SIRT,{iters},1.0")
    ref = orc.sirt_f32(coef, stack, T, org, iters, 1.0)
    ref64 = orc.sirt_f64(coef, stack.astype(np.float64), T, org, iters, 1.0)
    assert report(f"SIRT x{iters}", got, ref, ref64) < TOL_F32


def test_sirt_z_split_matches(torch_cuda, trec, ctx, orc, monkeypatch):
    """The forward projection split into z chunks gives the same volume."""
    P, Y, T, X = 9, 16, 24, 40
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=4)
    ref = orc.sirt_f32(coef, stack, T, org, 3, 0.9)
    for nzc in ("1", "3", "24"):
        monkeypatch.setenv("TREC_NZC", nzc)
        got = run(torch_cuda, trec, ctx, stack, ang, T, "SIRT,3,0.9")
        assert report(f"SIRT nzc={nzc}", got, ref) < TOL_F32


def test_sirt_initial(torch_cuda, trec, ctx, orc):
    P, Y, T, X = 8, 17, 9, 30
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=5)
    init = smooth_volume(Y, T, X, 9)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "SIRT,2,1.0", initial=init)
    ref = orc.sirt_f32(coef, stack, T, org, 2, 1.0, vol=init)
    assert report("SIRT initial", got, ref) < TOL_F32


def test_sart(torch_cuda, trec, ctx, orc):
    P, Y, T, X = 7, 16, 8, 36
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=6)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "SART,2,0.5")
    ref = orc.sart_f32(coef, stack, T, org, 2, 0.5)
    ref64 = orc.sart_f64(coef, stack.astype(np.float64), T, org, 2, 0.5)
    assert report("SART", got, ref, ref64) < TOL_F32


def test_wbp(torch_cuda, trec, ctx, orc):
    P, Y, T, X = 9, 16, 10, 50
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=7)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "WBP")
    filt = orc.rweight_f64(stack.astype(np.float64))
    ref64 = orc.bpt_f64(coef, filt, T, org)
    ref32 = orc.bpt_f32(coef, filt.astype(np.float32), T, org)
    assert report("WBP", got, ref32, ref64) < TOL_F32


def test_filter_api(torch_cuda, trec, ctx, orc):
    """Widths through both filter kernels: the source-ordered one (one and
    two passes of 1024 outputs) and, past about 6200 pixels, the tap-ordered
    one."""
    rng = np.random.default_rng(8)
    for X in (2, 7, 64, 100, 256, 1024, 1100, 2300, 6500):
        rows = (rng.standard_normal((10, X)) + np.linspace(0, 2, X)).astype(np.float32)
        print(f"R-weighting X={X}", flush=True)
        t = dev(torch_cuda, rows)
        ctx.filter_rweight(t)
        torch_cuda.cuda.synchronize()
        ref = orc.rweight_f64(rows.astype(np.float64))
        e = orc.rel_l2(t.cpu().numpy(), ref)
        print(f"  rel L2 vs FP64 oracle {e:.3e}")
        assert e < 1e-5


def test_layout_z(torch_cuda, trec, ctx, orc):
    P, Y, T, X = 9, 20, 12, 40
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=9)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "SIRT,2,1.0", layout=1)
    ref = to_layout_z(orc.sirt_f32(coef, stack, T, org, 2, 1.0))
    assert report("SIRT layout Z", got, ref) < TOL_F32


@pytest.mark.parametrize("offset,zshift", [(3.0, 0.0), (0.0, 2.5), (-2.0, -1.5)])
def test_offset_zshift_rows_path(torch_cuda, trec, ctx, orc, offset, zshift):
    P, Y, T, X = 9, 16, 12, 40
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=10, offset=offset, zshift=zshift)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "SIRT,3,1.0", offset=offset, zshift=zshift)
    ref = orc.sirt_f32(coef, stack, T, org, 3, 1.0)
    assert report(f"SIRT offset={offset} zshift={zshift}", got, ref) < TOL_F32


@pytest.mark.parametrize("method", ["BPT", "SIRT,3,1.0", "SART,1,0.8", "WBP"])
def test_pitch_general_path(torch_cuda, trec, ctx, orc, method):
    P, Y, T, X = 7, 12, 8, 28
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=11, pitch=6.0, offset=1.0)
    got = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=6.0, offset=1.0)
    name = method.split(",")[0]
    if name == "BPT":
        ref = orc.bpt_f32(coef, stack, T, org)
    elif name == "SIRT":
        ref = orc.sirt_f32(coef, stack, T, org, 3, 1.0)
    elif name == "SART":
        ref = orc.sart_f32(coef, stack, T, org, 1, 0.8)
    else:
        ref = orc.bpt_f32(coef, orc.rweight_f64(stack.astype(np.float64)).astype(np.float32), T, org)
    assert report(f"pitch {method}", got, ref) < TOL_F32


# Offset 0 with a pitch: v does not depend on x, so the forward takes the
# row-window kernels (three rows per slot: 16 detector rows per thread up to
# ~19.2 degrees of pitch, 8 up to ~25.8, 4 up to ~33.6; the splat beyond) and
# the back-projection the per-voxel row kernel.  X > 256 and Y not a multiple
# of the row block exercise the partial work-groups; a 70 degree span needs
# 6 candidate columns per row.  19.1/19.3, 25.8/25.9 and 33.5 sit on either
# side of the rows-per-thread limits, (1 - cos pitch)(RJ + 2) <= 1.
# A volume thicker than it is wide (T = 40, X = 48) projects its corners
# past the detector's columns and, with the pitch, its end rows past its rows.
@pytest.mark.parametrize("method,pitch,span,layout,shape", [
    ("SIRT,3,1.0", 0.5, 60.0, 0, (9, 37, 14, 300)), ("SIRT,3,1.0", 20.0, 70.0, 0, (9, 37, 14, 300)),
    ("SIRT,3,1.0", 25.0, 60.0, 1, (9, 37, 14, 300)), ("SIRT,3,1.0", 35.0, 60.0, 0, (9, 37, 14, 300)),
    ("SART,1,0.8", 4.0, 60.0, 0, (9, 37, 14, 300)), ("BPT", 12.0, 66.0, 1, (9, 37, 14, 300)),
    ("SIRT,3,1.0", 20.0, 60.0, 0, (9, 21, 40, 48)), ("BPT", 8.0, 60.0, 1, (9, 21, 40, 48)),
    ("SIRT,3,1.0", 19.1, 60.0, 0, (9, 37, 14, 300)), ("SIRT,3,1.0", 19.3, 60.0, 0, (9, 37, 14, 300)),
    ("SIRT,3,1.0", 25.8, 60.0, 0, (9, 37, 14, 300)), ("SIRT,3,1.0", 25.9, 60.0, 0, (9, 37, 14, 300)),
    ("SIRT,3,1.0", 33.5, 60.0, 0, (9, 37, 14, 300))])
def test_pitch_rows_path(torch_cuda, trec, ctx, orc, method, pitch, span, layout, shape):
    P, Y, T, X = shape
    ang, coef, org, stack = case(orc, P, Y, T, X, span=span, seed=21, pitch=pitch)
    got = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=pitch, layout=layout)
    if layout == 1:
        got = got.transpose(1, 0, 2)
    name = method.split(",")[0]
    if name == "BPT":
        ref = orc.bpt_f32(coef, stack, T, org)
    elif name == "SIRT":
        ref = orc.sirt_f32(coef, stack, T, org, 3, 1.0)
    else:
        ref = orc.sart_f32(coef, stack, T, org, 1, 0.8)
    assert report(f"pitch {pitch} {method} span {span} layout {layout}", got, ref) < TOL_F32
    again = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=pitch, layout=layout)
    if layout == 1:
        again = again.transpose(1, 0, 2)
    np.testing.assert_array_equal(again, got)  # deterministic


@pytest.mark.parametrize("pitch,span,shape", [(0.0, 60.0, (7, 18, 10, 36)), (5.0, 60.0, (7, 18, 10, 36)),
                                              (20.0, 70.0, (7, 37, 10, 300)), (25.0, 60.0, (7, 37, 10, 300)),
                                              (35.0, 60.0, (5, 21, 9, 70))])
def test_forward_project_api(torch_cuda, trec, ctx, orc, pitch, span, shape):
    P, Y, T, X = shape
    ang = tilts(P, span)
    coef = orc.coeffs_mpi(ang, 0.0, pitch, 0.0)
    org = orc.origin(X, T, T / 2)
    vol = smooth_volume(Y, T, X, 12)
    g = trec.make_geom(X, Y, P, T, 0.0, pitch, 0.0)
    rv = torch_cuda.empty((P, Y, X), dtype=torch_cuda.float32, device="cuda")
    rw = torch_cuda.empty_like(rv)
    ctx.forward_project(g, ang, dev(torch_cuda, vol), rv, rw)
    torch_cuda.cuda.synchronize()
    ref = np.stack([orc.reproject_f32(coef[k], vol, org)[0] for k in range(P)])
    refw = np.stack([orc.reproject_f32(coef[k], vol, org)[1] for k in range(P)])
    assert report("forward value", rv.cpu().numpy(), ref) < 1e-6
    assert report("forward weight", rw.cpu().numpy(), refw) < 1e-6


# FP16 storage must stay within 1e-3 of the FP64 oracle.  BF16 keeps 8
# mantissa bits, so the stored projections alone carry ~2^-9 relative error
# and the volume inherits it (~2e-3): its bound is looser and its error is
# reported, not held to the FP16 bar.
@pytest.mark.parametrize("storage,bound", [("f16", 1e-3), ("bf16", 5e-3)])
def test_16bit_storage_vs_fp64(torch_cuda, trec, ctx, orc, storage, bound):
    P, Y, T, X = 11, 16, 10, 48
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=13)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "SIRT,10,1.0", storage=storage)
    ref64 = orc.sirt_f64(coef, stack.astype(np.float64), T, org, 10, 1.0)
    e = orc.rel_l2(got, ref64)
    print(f"SIRT x10 {storage}: rel L2 vs FP64 {e:.3e}")
    assert e < bound


def _cuda_ref(orc, method, coef, stack, T, org, rules):
    """The FP32 oracle under either binary's thresholds."""
    name = method.split(",")[0]
    thr = orc.CUDA_THRESHOLDS[name] if rules == 1 else (0.0, 0.0)
    if name == "BPT":
        return orc.bpt_f32(coef, stack, T, org, thr=thr)
    if name == "WBP":
        return orc.bpt_f32(coef, orc.rweight_f64(stack.astype(np.float64)).astype(np.float32), T, org, thr=thr)
    if name == "FBP":
        return orc.fbp_f32(coef, stack, T, org, thr=thr)
    it, gam = int(method.split(",")[1]), float(method.split(",")[2])
    fn = orc.sirt_f32 if name == "SIRT" else orc.sart_f32
    return fn(coef, stack, T, org, it, gam, thr=thr)


@pytest.mark.parametrize("method", ["BPT", "WBP", "FBP", "SIRT,3,1.0", "SART,2,0.7"])
def test_cuda_rules_tiny_weights(torch_cuda, trec, ctx, orc, method):
    """TiltRec-cuda's thresholds where they change the volume.  One tilt maps
    corner voxel (0, T-1) just right of detector column -1, so its only tap
    (column 0) weighs w in (1e-6, 1e-5] (SART: (1e-7, 1e-6]); the other
    tilts cannot see that voxel at all.  Under TREC_RULES_CUDA the voxel
    skips that tilt (c > 1e-5, basic.cu:409; SIRT wtvol > 10e-6, :184; SART
    c > 10e-7, :434); under TREC_RULES_MPI it takes the full s / c.  Both
    rules must match the oracle at those voxels and over the volume."""
    from tests.helpers import edge_tilt, voxel_u
    X, T, Y = 32, 40, 4
    name = method.split(",")[0]
    lo, hi = (1e-7, 1e-6) if name == "SART" else (1e-6, 1e-5)
    te = edge_tilt(orc, X, T, lo, hi)
    ang = np.array([te, 20.0, 27.0, 35.0, 45.0, 52.0, 60.0], np.float32)
    coef = orc.coeffs_mpi(ang)
    u = voxel_u(coef, 0, T - 1, X, T)
    assert -1 < u[0] <= -1 + hi and np.all(u[1:] < -1)
    org = orc.origin(X, T, T / 2)
    stack = np.random.default_rng(31).random((len(ang), Y, X)).astype(np.float32) + 0.5
    refs = {r: _cuda_ref(orc, method, coef, stack, T, org, r) for r in (0, 1)}
    corner = (slice(None), T - 1, 0)
    gap = np.abs(refs[0][corner] - refs[1][corner])
    assert np.all(gap > 1e-3 * np.abs(refs[1]).max()), f"the rules do not differ at the corner: {gap}"
    for rules in (0, 1):
        got = run(torch_cuda, trec, ctx, stack, ang, T, method, rules=rules)
        e = report(f"{method} rules {rules}", got, refs[rules])
        assert e < TOL_F32
        np.testing.assert_allclose(got[corner], refs[rules][corner], rtol=1e-5, atol=1e-5,
                                   err_msg=f"corner voxels, rules {rules}")


@pytest.mark.parametrize("method", ["BPT", "WBP", "SIRT,3,1.0", "SART,1,0.7"])
def test_cuda_rules_pitch_origin(torch_cuda, trec, ctx, orc, method):
    """With a pitch, TiltRec-cuda's y origin (Y/2, main-cuda.cpp:259-261)
    differs from TiltRec-mpi's (T/2, main-mpi.cpp:743-745): the general path
    under TREC_RULES_CUDA must follow Y/2 and the CUDA thresholds."""
    P, Y, T, X = 9, 24, 10, 40
    ang = tilts(P)
    coef = orc.coeffs_mpi(ang, 0.0, 4.0, 0.0)
    stack = noisy_stack(orc, coef, smooth_volume(Y, T, X, 41), orc.origin(X, T, Y / 2), 42)
    ref_cuda = _cuda_ref(orc, method, coef, stack, T, orc.rules_origin(X, Y, T, 1), 1)
    ref_mpi = _cuda_ref(orc, method, coef, stack, T, orc.rules_origin(X, Y, T, 0), 1)
    assert orc.rel_l2(ref_mpi, ref_cuda) > 1e-3  # the origin matters here
    got = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=4.0, rules=1)
    assert report(f"{method} pitch 4 rules CUDA", got, ref_cuda) < TOL_F32


def test_bad_arguments_fail_loudly(torch_cuda, trec, ctx):
    g = trec.make_geom(0, 4, 3, 4)
    m = trec.make_method("BPT")
    s = torch_cuda.zeros((3, 4, 4), device="cuda")
    out = torch_cuda.zeros((4, 4, 4), device="cuda")
    with pytest.raises(trec.TrecError) as ei:
        ctx.reconstruct(g, [0, 1, 2], s, m, out)
    assert ei.value.status == trec.ERR_INVALID
    g = trec.make_geom(4, 4, 3, 4)
    with pytest.raises(trec.TrecError) as ei:
        ctx.reconstruct(g, [0, 1, 2], s, trec.make_method("ADMM,5,3,1.0,-0.1"), out)
    assert ei.value.status == trec.ERR_INVALID


@pytest.mark.parametrize("method", ["SIRT,3,1.0", "SART,1,0.5"])
def test_half_volume_copy_edges(torch_cuda, trec, ctx, orc, method):
    """Half storage projects a half copy of the volume kept as (c, c+1) pair
    entries: built at preparation, rewritten by every back-projection with
    the entries at 64-voxel tile seams written in halves by two tiles, plus
    an entry before cell 0.  Spikes in the starting volume at the detector's
    edges and at the tile seam (x = 63, 64) make any entry written wrong or
    left stale show up far above the half-precision error."""
    P, Y, T, X = 9, 16, 10, 70
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=21)
    init = smooth_volume(Y, T, X, 22)
    for x in (0, 1, 62, 63, 64, 65, 68, 69):
        init[:, :, x] += 3.0
    got = run(torch_cuda, trec, ctx, stack, ang, T, method, initial=init, storage="f16")
    f32 = run(torch_cuda, trec, ctx, stack, ang, T, method, initial=init)
    name, it, gamma = method.split(",")
    fn = orc.sirt_f64 if name == "SIRT" else orc.sart_f64
    ref64 = fn(coef, stack.astype(np.float64), T, org, int(it), float(gamma), vol=init.astype(np.float64))
    e64, e32 = orc.rel_l2(got, ref64), orc.rel_l2(got, f32)
    print(f"{method} half copy: rel L2 vs FP64 {e64:.3e}, vs float storage {e32:.3e}")
    assert e64 < 1e-3 and e32 < 1e-3


# The fused forward stages each tilt group's band union in LDS: 352 cells
# when every union of the job fits (the small shapes above), else 400, and
# a work-group whose union exceeds even that gathers from L2 directly.
# Thick volumes widen the unions: +-70 degrees at T = 200 reaches 372 cells
# (and 6 candidates per layer), +-60 degrees at T = 400 reaches 437.
@pytest.mark.parametrize("shape,span", [((9, 16, 200, 64), 70.0), ((9, 16, 400, 96), 60.0)])
def test_sirt_wide_unions(torch_cuda, trec, ctx, orc, shape, span):
    P, Y, T, X = shape
    ang, coef, org, stack = case(orc, P, Y, T, X, span=span, seed=33)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "SIRT,2,1.0")
    ref = orc.sirt_f32(coef, stack, T, org, 2, 1.0)
    assert report(f"SIRT x2 T={T} span {span}", got, ref) < TOL_F32


def test_f16_wide_unions(torch_cuda, trec, ctx, orc):
    """The half forward with the 400-cell band (unions of 372 cells)."""
    P, Y, T, X = 9, 16, 200, 64
    ang, coef, org, stack = case(orc, P, Y, T, X, span=70.0, seed=34)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "SIRT,3,1.0", storage="f16")
    ref64 = orc.sirt_f64(coef, stack.astype(np.float64), T, org, 3, 1.0)
    e = orc.rel_l2(got, ref64)
    print(f"SIRT x3 f16 T={T} span 70: rel L2 vs FP64 {e:.3e}")
    assert e < 1e-3


@pytest.mark.parametrize("shape", [(5, 3, 4, 5), (7, 2, 6, 9), (3, 5, 3, 4)])
def test_pitch_tiny_shapes(torch_cuda, trec, ctx, orc, shape):
    """Pitched SIRT and BPT where every work-group is partial: fewer rows
    than a row block, as few columns as candidates, a few layers."""
    P, Y, T, X = shape
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=35, pitch=5.0)
    got = run(torch_cuda, trec, ctx, stack, ang, T, "SIRT,2,1.0", pitch=5.0)
    assert report(f"pitch 5 SIRT x2 {shape}", got, orc.sirt_f32(coef, stack, T, org, 2, 1.0)) < TOL_F32
    got = run(torch_cuda, trec, ctx, stack, ang, T, "BPT", pitch=5.0)
    assert report(f"pitch 5 BPT {shape}", got, orc.bpt_f32(coef, stack, T, org)) < TOL_F32


@pytest.mark.parametrize("method,storage,layout", [("SIRT,3,1.0", "f32", 0), ("SIRT,3,1.0", "f16", 1),
                                                   ("SART,1,0.7", "f32", 0), ("WBP", "f32", 1),
                                                   ("BPT", "f32", 0)])
def test_slabs_equal_whole_volume(torch_cuda, trec, orc, method, storage, layout):
    """A workspace limit below the whole job's workspace makes trec_reconstruct
    run the volume in y slabs, one after another (the last one partial here);
    the volume must equal the one-slab result bit for bit and the oracle."""
    torch = torch_cuda
    P, Y, T, X = 9, 83, 12, 40
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=51)
    init = smooth_volume(Y, T, X, 52)
    ctx = trec.Context(0)
    whole = run(torch, trec, ctx, stack, ang, T, method, initial=to_layout_z(init) if layout else init,
                layout=layout, storage=storage)
    g = trec.make_geom(X, Y, P, T)
    m = trec.make_method(method, storage=storage, layout=layout)
    ws = trec.workspace_size(g, m)
    ctx.set_workspace_limit(ws // 3)
    info = ctx.plan(g, ang, m)
    print(f"{method} {storage} layout {layout}: {info}")
    assert info["slabs"] >= 3 and info["workspace"] <= ws // 3
    sl = run(torch, trec, ctx, stack, ang, T, method, initial=to_layout_z(init) if layout else init,
             layout=layout, storage=storage)
    ctx.close()
    np.testing.assert_array_equal(sl, whole)
    if storage == "f32":
        name = method.split(",")[0]
        if name in ("SIRT", "SART"):
            it, gam = int(method.split(",")[1]), float(method.split(",")[2])
            ref = (orc.sirt_f32 if name == "SIRT" else orc.sart_f32)(coef, stack, T, org, it, gam, vol=init)
        elif name == "WBP":
            ref = orc.bpt_f32(coef, orc.rweight_f64(stack.astype(np.float64)).astype(np.float32), T, org, vol=init)
        else:
            ref = orc.bpt_f32(coef, stack, T, org, vol=init)
        got = np.transpose(sl, (1, 0, 2)) if layout else sl
        assert report(f"{method} in slabs", got, ref) < TOL_F32


@pytest.mark.parametrize("storage,shapes", [("f32", [(6, 4), (4, 4), (8, 4), (6, 2), (8, 2)]),
                                            ("f16", [(8, 4), (6, 4), (12, 4), (8, 2)])])
def test_back_shapes_equal(torch_cuda, trec, orc, monkeypatch, storage, shapes):
    """The autotuner's SIRT back-projection shapes (tilts staged per pass,
    voxels per thread along z) change the staging and the tile, not the sums:
    every shape gives the default's volume bit for bit, at a size with
    partial tiles in x and z."""
    P, Y, T, X = 11, 20, 37, 100
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=61)
    ctx = trec.Context(0)
    vols = []
    for g, zt in shapes:
        monkeypatch.setenv("TREC_BACK_SHAPE", f"{g},{zt}")
        info = ctx.plan(trec.make_geom(X, Y, P, T), ang, trec.make_method("SIRT,3,1.0", storage=storage))
        assert (info["back_tilts"], info["back_voxels_z"]) == (g, zt)
        vols.append(run(torch_cuda, trec, ctx, stack, ang, T, "SIRT,3,1.0", storage=storage))
    ctx.close()
    for (g, zt), v in zip(shapes[1:], vols[1:]):
        np.testing.assert_array_equal(v, vols[0], err_msg=f"shape {g},{zt}")
    if storage == "f32":
        assert report("SIRT x3, tuned shapes", vols[0], orc.sirt_f32(coef, stack, T, org, 3, 1.0)) < TOL_F32


# The general forward kernels one by one (TREC_GENERAL_FORWARD forces one
# where it applies): the splat (any pitch and offset with a1 >= 1/3), the
# row windows (offset 0, and a pitch with an offset while |b1| over a wave's
# columns stays under about half a row) and the per-pixel gather (everything
# else), values and ray weights against Reproject at 1e-6, each twice with
# equal bits.  With a 0.5 degree pitch and a 1 degree offset the +60 degree
# tilt has a1 = 0.485 (5 candidate columns); 5 and 4 degrees sit near the row
# windows' limit on |b1|.
@pytest.mark.parametrize("kind", ["splat", "rows", "gather"])
@pytest.mark.parametrize("pitch,offset,span,shape", [
    (0.5, 0.0, 60.0, (7, 37, 10, 300)), (12.0, 0.0, 66.0, (7, 37, 10, 300)), (3.0, 2.0, 60.0, (7, 37, 12, 300)),
    (6.0, -4.0, 70.0, (5, 40, 9, 130)), (35.0, 5.0, 60.0, (5, 21, 9, 70)), (4.0, 1.0, 60.0, (5, 9, 40, 20)),
    (0.5, 1.0, 60.0, (9, 37, 12, 300)), (5.0, 4.0, 60.0, (7, 37, 12, 300))])
def test_general_forward_kernels(torch_cuda, trec, ctx, orc, monkeypatch, kind, pitch, offset, span, shape):
    monkeypatch.setenv("TREC_GENERAL_FORWARD", kind)
    P, Y, T, X = shape
    ang = tilts(P, span)
    coef = orc.coeffs_mpi(ang, offset, pitch, 0.0)
    org = orc.origin(X, T, T / 2)
    vol = smooth_volume(Y, T, X, 17)
    g = trec.make_geom(X, Y, P, T, offset, pitch, 0.0)
    ref = [orc.reproject_f32(coef[k], vol, org) for k in range(P)]
    outs = []
    for _ in range(2):
        rv = torch_cuda.empty((P, Y, X), dtype=torch_cuda.float32, device="cuda")
        rw = torch_cuda.empty_like(rv)
        ctx.forward_project(g, ang, dev(torch_cuda, vol), rv, rw)
        torch_cuda.cuda.synchronize()
        outs.append((rv.cpu().numpy(), rw.cpu().numpy()))
    tag = f"{kind} pitch {pitch} offset {offset} span {span} {shape}"
    assert report(f"forward value, {tag}", outs[0][0], np.stack([r[0] for r in ref])) < 1e-6
    assert report(f"forward weight, {tag}", outs[0][1], np.stack([r[1] for r in ref])) < 1e-6
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


# A context replays the graph it captured while the call, its buffers and the
# forcing variables are unchanged.  The same SIRT call on the same buffers
# under the row windows, then under the splat, must give the splat's result
# (that of a fresh context), not a replay of the row windows' graph.
def test_graph_follows_forced_kernel(torch_cuda, trec, ctx, orc, monkeypatch):
    torch = torch_cuda
    P, Y, T, X = 9, 37, 12, 300
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=62, pitch=0.5, offset=1.0)
    g = trec.make_geom(X, Y, P, T, 1.0, 0.5, 0.0)
    m = trec.make_method("SIRT,3,1.0", layout=0)
    s = dev(torch, stack)
    out = torch.empty((Y, T, X), dtype=torch.float32, device="cuda")
    res = {}
    for kind in ("rows", "splat"):
        monkeypatch.setenv("TREC_GENERAL_FORWARD", kind)
        ctx.reconstruct(g, ang, s, m, out)
        torch.cuda.synchronize()
        res[kind] = out.cpu().numpy()
    c2 = trec.Context(0)
    try:
        c2.reconstruct(g, ang, s, m, out)  # still the splat
        torch.cuda.synchronize()
        fresh = out.cpu().numpy()
    finally:
        c2.close()
    print(f"row windows and splat differ in {int((res['rows'] != res['splat']).sum())} voxels")
    ref = orc.sirt_f32(coef, stack, T, org, 3, 1.0)
    assert report("splat after the row windows", res["splat"], ref) < TOL_F32
    np.testing.assert_array_equal(res["splat"], fresh)


# A pitch with an offset (b1 != 0): v moves along x, so the back-projection
# stages 3 or 4 detector rows per wave (|b1| up to 2/511), or reads its taps
# from memory beyond.  X = 600 spans two waves of 512 voxels per row.
@pytest.mark.parametrize("method", ["SIRT,3,1.0", "SART,1,0.8", "BPT"])
@pytest.mark.parametrize("pitch,offset", [(3.0, 2.0), (4.0, 3.0), (6.0, 3.0)])
def test_pitch_offset_paths(torch_cuda, trec, ctx, orc, method, pitch, offset):
    P, Y, T, X = 7, 21, 10, 600
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=41, pitch=pitch, offset=offset)
    got = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=pitch, offset=offset)
    name = method.split(",")[0]
    if name == "BPT":
        ref = orc.bpt_f32(coef, stack, T, org)
    elif name == "SIRT":
        ref = orc.sirt_f32(coef, stack, T, org, 3, 1.0)
    else:
        ref = orc.sart_f32(coef, stack, T, org, 1, 0.8)
    b1 = abs(coef[0, 5])
    assert report(f"pitch {pitch} offset {offset} (b1 span {511 * b1:.2f} rows) {method}", got, ref) < TOL_F32
    again = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=pitch, offset=offset)
    np.testing.assert_array_equal(again, got)


# The back-projection's group-shared window: the four waves of a work-group
# take four layers of one volume row (layout Y, T a multiple of 4) and stage
# the union of their windows once per tilt.  It must give the bits of every
# wave staging its own window (TREC_GENERAL_BACK=wave), and the oracle's
# values.  Two, three and four staged rows per wave (offset 0, |b1| spans of
# 0.93 and 1.87 rows); X = 600 leaves a partial tile of 512 voxels per row.
@pytest.mark.parametrize("method", ["SIRT,3,1.0", "SART,1,0.8", "BPT"])
@pytest.mark.parametrize("pitch,offset", [(0.5, 0.0), (15.0, 0.0), (3.0, 2.0), (4.0, 3.0)])
def test_general_back_shared_window(torch_cuda, trec, ctx, orc, monkeypatch, method, pitch, offset):
    P, Y, T, X = 7, 21, 12, 600
    ang, coef, org, stack = case(orc, P, Y, T, X, seed=43, pitch=pitch, offset=offset)
    got = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=pitch, offset=offset)
    name = method.split(",")[0]
    if name == "BPT":
        ref = orc.bpt_f32(coef, stack, T, org)
    elif name == "SIRT":
        ref = orc.sirt_f32(coef, stack, T, org, 3, 1.0)
    else:
        ref = orc.sart_f32(coef, stack, T, org, 1, 0.8)
    assert report(f"shared window pitch {pitch} offset {offset} {method}", got, ref) < TOL_F32
    monkeypatch.setenv("TREC_GENERAL_BACK", "wave")
    own = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=pitch, offset=offset)
    np.testing.assert_array_equal(got, own)


# The row windows with per-voxel row weights (a pitch with an offset): both
# layouts, a negative offset with a z shift, SART's few-tilt launches (4 rows
# per thread and layer chunks), partial pixel and row blocks (X = 300 or 600,
# Y = 29 or 37), spans to +-66 degrees (up to 6 candidate columns), against
# the FP32 oracle and bit for bit on a second run.  Small offsets take the
# wave's rows; 8 with -6, 10 with 5 and 12 with 8 degrees each lane's (a
# shift of up to 4 rows); lane=1 forces each lane's rows everywhere
# (TREC_GENERAL_ROWS_LANE=1).
@pytest.mark.parametrize("lane", ["auto", "1"])
@pytest.mark.parametrize("method,pitch,offset,span,layout,zshift,shape", [
    ("SIRT,3,1.0", 0.5, 1.0, 60.0, 1, 0.0, (9, 37, 14, 300)),
    ("SIRT,3,1.0", 3.0, -2.0, 66.0, 0, 2.0, (9, 29, 14, 300)),
    ("SART,1,0.8", 2.0, 2.0, 60.0, 1, 0.0, (9, 37, 14, 300)),
    ("SART,1,0.8", 5.0, 4.0, 60.0, 0, -1.5, (7, 21, 12, 600)),
    ("SIRT,3,1.0", 8.0, -6.0, 60.0, 0, 0.0, (9, 37, 14, 300)),
    ("SIRT,3,1.0", 10.0, 5.0, 60.0, 1, 1.0, (9, 37, 14, 300)),
    ("SART,1,0.8", 12.0, 8.0, 50.0, 0, 0.0, (7, 29, 12, 300))])
def test_offset_row_windows(torch_cuda, trec, ctx, orc, monkeypatch, lane, method, pitch, offset, span, layout, zshift,
                            shape):
    if lane == "1":
        monkeypatch.setenv("TREC_GENERAL_ROWS_LANE", "1")
    P, Y, T, X = shape
    ang, coef, org, stack = case(orc, P, Y, T, X, span=span, seed=61, pitch=pitch, offset=offset, zshift=zshift)
    got = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=pitch, offset=offset, zshift=zshift, layout=layout)
    if layout == 1:
        got = got.transpose(1, 0, 2)
    if method.startswith("SIRT"):
        ref = orc.sirt_f32(coef, stack, T, org, 3, 1.0)
    else:
        ref = orc.sart_f32(coef, stack, T, org, 1, 0.8)
    tag = f"offset rows lane={lane} {method} pitch {pitch} offset {offset} span {span} layout {layout} zshift {zshift}"
    assert report(tag, got, ref) < TOL_F32
    again = run(torch_cuda, trec, ctx, stack, ang, T, method, pitch=pitch, offset=offset, zshift=zshift, layout=layout)
    if layout == 1:
        again = again.transpose(1, 0, 2)
    np.testing.assert_array_equal(again, got)
