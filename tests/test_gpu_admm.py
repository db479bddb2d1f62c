"""ADMM (TiltRecZ's ADMM, main-mpiz.cpp:603-1013) on the GPU against the CPU
oracle.

The GPU runs the reference's recursion: htb = A^T b, then per outer
iteration the htb update, CG on op(v) = sum A^T (A v / w) + mu v, the soft
threshold and the d update, with the CG scalars summed in double on the
device.  The FP32 oracle and the GPU order their float sums differently;
the CG recursion carries those roundings forward, and the FP32 oracle
itself is within 2.4e-6 of the FP64 oracle at 5 x 4 iterations.  The bar is
the plain FP32 one, relative L2 1e-5.
"""
import os
import subprocess

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


def case(orc, P, Y, T, X, seed):
    ang = tilts(P)
    coef = orc.coeffs_mpi(ang)
    org = orc.origin(X, T, T / 2)
    stack = noisy_stack(orc, coef, smooth_volume(Y, T, X, seed), org, seed + 1)
    return ang, coef, org, stack


@pytest.mark.parametrize("params,layout,rules", [("ADMM,1,1,0.5,0.0", 0, 0), ("ADMM,3,3,0.5,0.01", 1, 1),
                                                 ("ADMM,5,4,0.2,0.02", 0, 0)])
def test_admm_matches_oracle(torch_cuda, trec, ctx, orc, params, layout, rules):
    torch = torch_cuda
    P, Y, T, X = 9, 18, 10, 40
    ang, coef, org, stack = case(orc, P, Y, T, X, 50)
    outer, cg, mu, soft = trec.parse_admm(params)
    g = trec.make_geom(X, Y, P, T, rules=rules)
    m = trec.make_method(params, layout=layout)
    out = torch.empty((Y, T, X) if layout == 0 else (T, Y, X), dtype=torch.float32, device="cuda")
    ctx.reconstruct(g, ang, dev(torch, stack), m, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = orc.admm_f32(coef, stack, T, org, outer, cg, mu, soft)
    ref64 = orc.admm_f64(coef, stack.astype(np.float64), T, org, outer, cg, mu, soft)
    if layout == 1:
        ref, ref64 = to_layout_z(ref), to_layout_z(ref64)
    e = orc.rel_l2(got, ref)
    print(f"{params} layout {layout}: rel L2 vs FP32 oracle {e:.3e}, vs FP64 {orc.rel_l2(got, ref64):.3e}")
    assert e < TOL_F32


def test_admm_initial_volume(torch_cuda, trec, ctx, orc):
    torch = torch_cuda
    P, Y, T, X = 7, 17, 9, 33
    ang, coef, org, stack = case(orc, P, Y, T, X, 60)
    init = smooth_volume(Y, T, X, 61) * np.float32(0.5)
    g = trec.make_geom(X, Y, P, T)
    m = trec.make_method("ADMM,2,3,0.4,0.01")
    out = torch.empty((Y, T, X), dtype=torch.float32, device="cuda")
    ctx.reconstruct(g, ang, dev(torch, stack), m, out, initial=dev(torch, init))
    torch.cuda.synchronize()
    ref = orc.admm_f32(coef, stack, T, org, 2, 3, 0.4, 0.01, vol=init)
    e = orc.rel_l2(out.cpu().numpy(), ref)
    print(f"ADMM from an initial volume: rel L2 vs FP32 oracle {e:.3e}")
    assert e < TOL_F32


# The general path -- a pitch, a pitch with an offset, tilts past about 78
# degrees, or pitch 0 forced onto it -- runs the same recursion with its own
# operator: the general forward projection, the quotient -R / w (0 where w
# is at most 10e-6) and the back-projection -sum s + mu x.
@pytest.mark.parametrize("pitch,offset,span,layout,force", [
    (5.0, 0.0, 60.0, 0, False),  # row-window forward, staged back-projection
    (4.0, 2.0, 60.0, 1, False),  # pitch and offset: per-pixel gather, per-voxel back-projection
    (0.0, 0.0, 80.0, 0, False),  # tilts past about 78 degrees
    (0.0, 0.0, 60.0, 1, True)])  # pitch 0 forced onto the general path
def test_admm_general_path(torch_cuda, trec, ctx, orc, monkeypatch, pitch, offset, span, layout, force):
    torch = torch_cuda
    P, Y, T, X = 9, 18, 10, 40
    ang = tilts(P, span)
    coef = orc.coeffs_mpi(ang, offset, pitch, 0.0)
    org = orc.origin(X, T, T / 2)
    stack = noisy_stack(orc, coef, smooth_volume(Y, T, X, 80), org, 81)
    if force:
        monkeypatch.setenv("TREC_FORCE_GENERAL", "1")
    g = trec.make_geom(X, Y, P, T, offset, pitch, 0.0)
    m = trec.make_method("ADMM,3,3,0.5,0.01", layout=layout)
    out = torch.empty((Y, T, X) if layout == 0 else (T, Y, X), dtype=torch.float32, device="cuda")
    print(f"ADMM general path: pitch {pitch} offset {offset} span {span} layout {layout}", flush=True)
    ctx.reconstruct(g, ang, dev(torch, stack), m, out)
    torch.cuda.synchronize()
    ref = orc.admm_f32(coef, stack, T, org, 3, 3, 0.5, 0.01)
    ref64 = orc.admm_f64(coef, stack.astype(np.float64), T, org, 3, 3, 0.5, 0.01)
    if layout == 1:
        ref, ref64 = to_layout_z(ref), to_layout_z(ref64)
    got = out.cpu().numpy()
    e = orc.rel_l2(got, ref)
    print(f"  rel L2 vs FP32 oracle {e:.3e}, vs FP64 {orc.rel_l2(got, ref64):.3e}")
    assert e < TOL_F32


def test_cli_admm(tmp_path, torch_cuda, orc):
    from autom2_amd import BIN_DIR, mrc
    P, Y, T, X = 9, 20, 12, 40
    ang, coef, org, stack = case(orc, P, Y, T, X, 70)
    mrc.write(str(tmp_path / "stack.mrc"), stack)
    (tmp_path / "tilt.tlt").write_text("\n".join(f"{a:.2f}" for a in ang) + "\n")
    r = subprocess.run([os.path.join(BIN_DIR, "TiltRecZ-mpi"), "--input", str(tmp_path / "stack.mrc"), "--output",
                        str(tmp_path / "v.mrc"), "--tiltfile", str(tmp_path / "tilt.tlt"), "--geometry",
                        f"0,0,0,{T}", "--method", "ADMM,2,2,0.5,0.01"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    h, vol = mrc.read(str(tmp_path / "v.mrc"))
    assert (h["nx"], h["ny"], h["nz"]) == (X, Y, T)
    stack_file = mrc.read(str(tmp_path / "stack.mrc"))[1]
    ang_file = np.loadtxt(tmp_path / "tilt.tlt", dtype=np.float32)
    ref = to_layout_z(orc.admm_f32(orc.coeffs_mpi(ang_file), stack_file, T, org, 2, 2, 0.5, 0.01))
    assert orc.rel_l2(vol, ref) < TOL_F32
