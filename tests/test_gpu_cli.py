"""The TiltRec-compatible CLI end to end on the GPU: MRC stack and tilt
file in, MRC volume out, against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from autom2_amd import BIN_DIR, mrc
from tests.helpers import noisy_stack, smooth_volume, tilts

pytestmark = pytest.mark.gpu


def write_inputs(d, orc, P=9, Y=20, T=12, X=40, seed=20):
    ang = tilts(P)
    coef = orc.coeffs_mpi(ang)
    org = orc.origin(X, T, T / 2)
    stack = noisy_stack(orc, coef, smooth_volume(Y, T, X, seed), org, seed + 1)
    mrc.write(str(d / "stack.mrc"), stack)
    with open(d / "tilt.tlt", "w") as f:
        f.write("\n".join(f"{a:.2f}" for a in ang) + "\n")
    return ang, coef, org, stack


@pytest.mark.parametrize("name,method", [("TiltRec-mpi", "SIRT,4,1.0"), ("TiltRecZ-mpi", "SIRT,4,1.0"),
                                         ("TiltRec-mpi", "WBP"), ("TiltRec-mpi", "SART,1,0.5")])
def test_cli_end_to_end(tmp_path, orc, name, method):
    ang, coef, org, stack = write_inputs(tmp_path, orc)
    P, Y, X = stack.shape
    T = 12
    out = tmp_path / "vol.mrc"
    r = subprocess.run([os.path.join(BIN_DIR, name), "--input", str(tmp_path / "stack.mrc"), "--output", str(out),
                        "--tiltfile", str(tmp_path / "tilt.tlt"), "--geometry", f"0,0,0,{T}", "--method", method],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    h, vol = mrc.read(str(out))
    kind = method.split(",")[0]
    if kind == "SIRT":
        ref = orc.sirt_f32(coef, stack, T, org, 4, 1.0)
    elif kind == "SART":
        ref = orc.sart_f32(coef, stack, T, org, 1, 0.5)
    else:
        ref = orc.bpt_f32(coef, orc.rweight_f64(stack.astype(np.float64)).astype(np.float32), T, org)
    if name.startswith("TiltRecZ"):
        assert (h["nx"], h["ny"], h["nz"]) == (X, Y, T)
        ref = np.transpose(ref, (1, 0, 2))
    else:
        assert (h["nx"], h["ny"], h["nz"]) == (X, T, Y)
    e = orc.rel_l2(vol, ref)
    print(f"{name} {method}: rel L2 {e:.3e}")
    assert e < 1e-5
    amin, amax, amean = mrc.header_stats(vol)
    assert h["amin"] == amin and h["amax"] == amax


def test_cli_initial_volume(tmp_path, orc):
    ang, coef, org, stack = write_inputs(tmp_path, orc, seed=30)
    P, Y, X = stack.shape
    T = 12
    init = smooth_volume(Y, T, X, 31)
    mrc.write(str(tmp_path / "init.mrc"), init)
    out = tmp_path / "vol.mrc"
    r = subprocess.run([os.path.join(BIN_DIR, "TiltRec-mpi"), "-i", str(tmp_path / "stack.mrc"), "-o", str(out),
                        "-t", str(tmp_path / "tilt.tlt"), "-g", f"0,0,0,{T}", "-m", "SIRT,2,1.0",
                        "--inital", str(tmp_path / "init.mrc")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _, vol = mrc.read(str(out))
    assert orc.rel_l2(vol, orc.sirt_f32(coef, stack, T, org, 2, 1.0, vol=init)) < 1e-5


@pytest.mark.parametrize("name,geometry", [("TiltRec-mpi", (1.5, 4.0, 2.0)), ("TiltRecZ-mpi", (0.0, 3.0, -1.0))])
def test_cli_geometry(tmp_path, orc, name, geometry):
    """--geometry with an offset, a pitch and a z shift, as a GeoParGen
    estimate is fed back: the general path through the CLI (with offset 0,
    its row-window kernels; with both, the per-pixel gather)."""
    off, pitch, zs = geometry
    P, Y, T, X = 9, 20, 12, 40
    ang = tilts(P)
    coef = orc.coeffs_mpi(ang, off, pitch, zs)
    org = orc.origin(X, T, T / 2)
    stack = noisy_stack(orc, coef, smooth_volume(Y, T, X, 40), org, 41)
    mrc.write(str(tmp_path / "stack.mrc"), stack)
    (tmp_path / "tilt.tlt").write_text("\n".join(f"{a:.2f}" for a in ang) + "\n")
    out = tmp_path / "vol.mrc"
    r = subprocess.run([os.path.join(BIN_DIR, name), "--input", str(tmp_path / "stack.mrc"), "--output", str(out),
                        "--tiltfile", str(tmp_path / "tilt.tlt"), "--geometry", f"{off},{pitch},{zs},{T}",
                        "--method", "SIRT,3,1.0"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    h, vol = mrc.read(str(out))
    ref = orc.sirt_f32(coef, stack, T, org, 3, 1.0)
    if name.startswith("TiltRecZ"):
        ref = np.transpose(ref, (1, 0, 2))
    e = orc.rel_l2(vol, ref)
    print(f"{name} --geometry {off},{pitch},{zs},{T}: rel L2 {e:.3e}")
    assert e < 1e-5


def test_cli_terminated_mid_reconstruction(tmp_path):
    """SIGTERM while the CLI's SIRT is on the GPU: the process ends without
    writing a volume (the reference writes per slab, so a killed run can
    leave a partial file), and the GPU serves the next reconstruction."""
    import signal
    import time

    import torch
    from autom2_amd import tiltrec as trec
    P, Y, X, T = 61, 512, 512, 256
    rng = np.random.default_rng(9)
    mrc.write(str(tmp_path / "stack.mrc"), rng.random((P, Y, X), dtype=np.float32))
    with open(tmp_path / "tilt.tlt", "w") as f:
        f.write("\n".join(f"{a:.2f}" for a in tilts(P)) + "\n")
    out = tmp_path / "vol.mrc"
    # several seconds of GPU work: 512^2 x 256 x 61 x 2500 voxel-projection updates
    p = subprocess.Popen([os.path.join(BIN_DIR, "TiltRec-mpi"), "--input", str(tmp_path / "stack.mrc"),
                          "--output", str(out), "--tiltfile", str(tmp_path / "tilt.tlt"), "--geometry",
                          f"0,0,0,{T}", "--method", "SIRT,2500,1.0"], stdout=subprocess.PIPE, text=True)
    line = p.stdout.readline()  # printed once the stack is read, before the upload
    assert line.startswith("TiltRec-hip:"), line
    time.sleep(1.0)
    assert p.poll() is None, "the reconstruction ended before the signal"
    p.send_signal(signal.SIGTERM)
    rc = p.wait(timeout=30)
    print(f"CLI ended with {rc}")
    assert rc == -signal.SIGTERM
    assert not out.exists()
    ctx = trec.Context(0)
    s = torch.ones((5, 16, 32), device="cuda")
    v = torch.empty((16, 8, 32), device="cuda")
    ctx.reconstruct(trec.make_geom(32, 16, 5, 8), tilts(5), s, trec.make_method("BPT"), v)
    torch.cuda.synchronize()
    assert float(v[:, :, 12:20].min()) == 5.0  # interior voxels see every tilt of a stack of ones
    ctx.close()
