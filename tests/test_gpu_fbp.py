"""FBP (Shepp-Logan spatial filter + back-projection) and the --f2b byte
output on the GPU against the CPU oracle.

FBP is the reference's ApplyFilterInplace mode 0 followed by BPT
(main-mpi.cpp:835-840, main-cuda.cpp:185-189).  The filter is checked
against the FP64 oracle (relative L2 1e-6: float multiply-adds over 251
taps); the reconstruction against the FP32 back-projection of the filtered
stack (1e-5, the plain FP32 bar).  The byte conversion is checked bit for
bit given the moments, and the moments to one float rounding.
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


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("X", [1, 7, 40, 257, 1024])
def test_filter_fbp_api(torch_cuda, ctx, orc, mode, X):
    rng = np.random.default_rng(X + 10 * mode)
    rows = (rng.standard_normal((13, X)) + np.linspace(0, 3, X)).astype(np.float32)
    t = dev(torch_cuda, rows)
    ctx.filter_fbp(t, mode)
    torch_cuda.cuda.synchronize()
    got = t.cpu().numpy()
    ref = orc.fbp_filter_f64(rows.astype(np.float64), mode)
    e = orc.rel_l2(got, ref)
    print(f"FBP filter mode {mode} X={X}: rel L2 vs FP64 oracle {e:.3e}")
    assert e < 1e-6


def recon(torch, trec, ctx, stack, ang, T, method, pitch=0.0, rules=0, layout=0):
    P, Y, X = stack.shape
    g = trec.make_geom(X, Y, P, T, 0.0, pitch, 0.0, rules)
    m = trec.make_method(method, layout=layout)
    out = torch.empty((Y, T, X) if layout == 0 else (T, Y, X), dtype=torch.float32, device="cuda")
    ctx.reconstruct(g, ang, dev(torch, stack), m, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("pitch,rules,layout", [(0.0, 0, 0), (0.0, 1, 1), (4.0, 0, 0)])
def test_fbp_reconstruction(torch_cuda, trec, ctx, orc, pitch, rules, layout):
    """Row-blocked path (pitch 0, both rules and layouts) and the general
    path (pitch 4)."""
    P, Y, T, X = 9, 18, 12, 46
    ang = tilts(P)
    coef = orc.coeffs_mpi(ang, 0.0, pitch, 0.0)
    oy = T / 2 if rules == 0 else Y / 2
    org = orc.origin(X, T, oy)
    stack = noisy_stack(orc, orc.coeffs_mpi(ang), smooth_volume(Y, T, X, 31), orc.origin(X, T, T / 2), 32)
    got = recon(torch_cuda, trec, ctx, stack, ang, T, "FBP", pitch, rules, layout)
    filt = orc.fbp_filter_f64(stack.astype(np.float64), 0)
    # TiltRec-cuda's back-projection skips c <= 1e-5 (basic.cu:409), the
    # oracle c == 0; no voxel here has 0 < c <= 1e-5, so one oracle serves both
    ref = orc.bpt_f32(coef, filt.astype(np.float32), T, org)
    ref64 = orc.bpt_f64(coef, filt, T, org)
    if layout == 1:
        ref, ref64 = to_layout_z(ref), to_layout_z(ref64)
    e = orc.rel_l2(got, ref)
    print(f"FBP pitch {pitch} rules {rules} layout {layout}: rel L2 vs FP32 oracle {e:.3e}, "
          f"vs FP64 {orc.rel_l2(got, ref64):.3e}")
    assert e < TOL_F32


@pytest.mark.parametrize("n", [1, 3, 1000, 123457])
def test_float_to_byte(torch_cuda, trec, ctx, orc, n):
    torch = torch_cuda
    v = (np.random.default_rng(n).standard_normal(n) * 3 + 1).astype(np.float32)
    out = torch.empty(n, dtype=torch.uint8, device="cuda")
    mean, sd = ctx.float_to_byte(dev(torch, v), out)
    got = out.cpu().numpy()
    _, (m64, s64) = orc.float_to_byte(v)
    assert abs(mean - m64) <= abs(m64) * 1.2e-7 + 1e-30
    assert abs(sd - s64) <= abs(s64) * 1.2e-7 + 1e-30
    want = orc.bytes_from_stats(v, mean, sd)
    assert np.array_equal(got, want)


def test_cli_fbp_and_f2b(tmp_path, torch_cuda, orc):
    from autom2_amd import BIN_DIR, mrc
    P, Y, T, X = 9, 20, 12, 40
    ang = tilts(P)
    coef = orc.coeffs_mpi(ang)
    org = orc.origin(X, T, T / 2)
    stack = noisy_stack(orc, coef, smooth_volume(Y, T, X, 40), org, 41)
    mrc.write(str(tmp_path / "stack.mrc"), stack)
    (tmp_path / "tilt.tlt").write_text("\n".join(f"{a:.2f}" for a in ang) + "\n")
    base = [os.path.join(BIN_DIR, "TiltRecZ-cuda"), "--input", str(tmp_path / "stack.mrc"), "--tiltfile",
            str(tmp_path / "tilt.tlt"), "--geometry", f"0,0,0,{T}", "--method", "FBP"]
    r = subprocess.run(base + ["--output", str(tmp_path / "f.mrc")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    h, vol = mrc.read(str(tmp_path / "f.mrc"))
    assert (h["nx"], h["ny"], h["nz"], h["mode"]) == (X, Y, T, 2)
    ang_file = np.loadtxt(tmp_path / "tilt.tlt", dtype=np.float32)
    stack_file = mrc.read(str(tmp_path / "stack.mrc"))[1]
    filt = orc.fbp_filter_f64(stack_file.astype(np.float64), 0).astype(np.float32)
    ref = to_layout_z(orc.bpt_f32(orc.coeffs_mpi(ang_file), filt, T, orc.origin(X, T, Y / 2)))
    assert orc.rel_l2(vol, ref) < TOL_F32
    r = subprocess.run(base + ["--output", str(tmp_path / "b.mrc"), "--f2b"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = open(tmp_path / "b.mrc", "rb").read()
    hb = np.frombuffer(raw[:1024], mrc.HEADER_DTYPE)[0]
    assert (hb["nx"], hb["ny"], hb["nz"], hb["mode"]) == (X, Y, T, 0)
    b = np.frombuffer(raw[1024:], np.uint8).reshape(T, Y, X)
    want, _ = orc.float_to_byte(vol)
    # the float volumes agree to ~1e-7, so a byte may truncate the other way
    assert np.abs(b.astype(int) - want.astype(int)).max() <= 1
    assert (b == want).mean() > 0.99
