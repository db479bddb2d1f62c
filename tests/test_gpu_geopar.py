"""GeoParGen on the GPU (trec_geopar) against the CPU oracle.

Per chip, the chosen pair, z, thickness and NCC must equal the oracle's bit
for bit: the GPU evaluates the reference's NCC loop exactly for every pair
that can win under its float running maximum.  Per view, thickness and z
are bit-exact; the angle goes through atan, whose last bit may differ
between the device's and the host's libm, so it is held to one float ulp.
The averages are float sums over views and follow from those.
"""
import math
import os
import subprocess

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


@pytest.fixture(scope="module")
def ctx(trec):
    c = trec.Context(0)
    yield c
    c.close()


def slab_noise_volume(shape, axis, seed, tilt=3, thick=40, noise=0.3):
    """[nz][ny][nx] volume: a bright slab whose centre row drifts along the
    chip axis, plus smooth noise."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    rows = np.arange(ny)[:, None]
    cols = nz if axis == 0 else nx
    centre = ny / 2 + tilt * (np.arange(cols) - cols / 2) / 50.0
    slab = (np.abs(rows - centre[None, :]) < thick / 2).astype(np.float32)  # [ny][cols]
    vol = np.empty(shape, np.float32)
    if axis == 0:  # view j, element (r, c) = vol[c][r][j]
        vol[:] = slab.T[:, :, None]
    else:
        vol[:] = slab[None, :, :]
    vol += (noise * np.cumsum(rng.standard_normal(shape), axis=1) / np.sqrt(ny)).astype(np.float32)
    return vol


def run(torch, trec, ctx, vol, axis, exhaustive=False):
    nz, ny, nx = vol.shape
    nv, cols = (nx, nz) if axis == 0 else (nz, nx)
    nc = len(range(0, cols - 50, 50)) if cols >= 50 else 0
    d = torch.from_numpy(np.ascontiguousarray(vol)).cuda()
    views = torch.empty((nv, 3), dtype=torch.float32, device="cuda")
    chips = torch.empty((nv, max(nc, 1), 4), dtype=torch.float32, device="cuda")
    avg, nex = ctx.geopar(d, axis, views, chips, exhaustive=exhaustive)
    torch.cuda.synchronize()
    return views.cpu().numpy(), chips.cpu().numpy()[:, :nc], np.array(avg, np.float32), nex


def check_against_oracle(orc, vol, axis, views, chips, avg):
    v_ref, c_ref, a_ref = orc.geopar(vol, axis)
    np.testing.assert_array_equal(chips, c_ref)
    np.testing.assert_array_equal(views[:, 1:], v_ref[:, 1:])
    a, b = views[:, 0], v_ref[:, 0]
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    close = np.abs(a - b) <= np.spacing(np.abs(b).astype(np.float32))
    assert np.all(same | close)
    for k in range(3):
        if math.isnan(a_ref[k]):
            assert math.isnan(avg[k])
        else:
            assert abs(avg[k] - a_ref[k]) <= 1e-6 * max(1.0, abs(a_ref[k]))
    return (a == b).mean()


@pytest.mark.parametrize("axis,shape", [(0, (170, 80, 24)), (0, (260, 97, 7)), (1, (9, 64, 211)),
                                         (1, (5, 131, 120))])
def test_matches_oracle(torch_cuda, trec, ctx, orc, axis, shape):
    print(f"geopar axis {axis} shape {shape}", flush=True)
    vol = slab_noise_volume(shape, axis, sum(shape))
    views, chips, avg, nex = run(torch_cuda, trec, ctx, vol, axis)
    exact = check_against_oracle(orc, vol, axis, views, chips, avg)
    print(f"  {chips.shape[0] * chips.shape[1]} chips, {nex} exhaustive, angles bit-equal {exact:.3f}")
    assert nex == 0  # the window decided every chip


def test_mirror_ties(torch_cuda, trec, ctx, orc):
    """Profiles symmetric in the row: every NCC has a mirror twin equal in
    exact arithmetic, so only rounding and the float maximum decide."""
    rng = np.random.default_rng(5)
    nz, ny, nx = 160, 90, 12
    half = rng.random((nz, ny // 2, nx)).astype(np.float32)
    vol = np.concatenate([half, half[:, ::-1]], axis=1)
    print("geopar mirror ties", flush=True)
    views, chips, avg, nex = run(torch_cuda, trec, ctx, vol, 0)
    check_against_oracle(orc, vol, 0, views, chips, avg)
    views2, chips2, avg2, nex2 = run(torch_cuda, trec, ctx, vol, 0, exhaustive=True)
    np.testing.assert_array_equal(chips2, chips)
    assert nex2 == chips.shape[0] * chips.shape[1]


def test_slab_closed_form(torch_cuda, trec, ctx):
    """A slab stepping d rows per chip, of thickness t0: angle atan(d / 50),
    thickness t0, z the mean chip centre."""
    rows, cols, nv, l0, d, t0 = 120, 410, 16, 10, 3, 40
    views = np.full((nv, rows, cols), 0.25, np.float32)
    nc = len(range(0, cols - 50, 50))
    for q in range(nc):
        lo = l0 + q * d
        views[:, lo:lo + t0 + 1, 50 * q:50 * q + 50] = 2.0
    vol = np.ascontiguousarray(views.transpose(2, 1, 0))
    print("geopar slab closed form", flush=True)
    v, c, avg, _ = run(torch_cuda, trec, ctx, vol, 0)
    zq = [l0 + q * d + t0 // 2 for q in range(nc)]
    assert abs(avg[0] - math.degrees(math.atan(d / 50))) < 1e-5
    assert abs(avg[1] - t0) < 1e-5 and abs(avg[2] - np.mean(zq)) < 1e-4
    np.testing.assert_array_equal(c[:, :, 0], np.broadcast_to(zq, (nv, nc)))


def test_degenerate(torch_cuda, trec, ctx, orc):
    cases = [(np.full((120, 60, 5), 2.5, np.float32), "constant volume"),
             (np.random.default_rng(1).random((120, 30, 5)).astype(np.float32), "30 rows: no pair"),
             (np.random.default_rng(2).random((50, 60, 5)).astype(np.float32), "50 columns: no chip")]
    for vol, name in cases:
        print(f"geopar degenerate: {name}", flush=True)
        views, chips, avg, _ = run(torch_cuda, trec, ctx, vol, 0)
        check_against_oracle(orc, vol, 0, views, chips, avg)


def test_errors(torch_cuda, trec, ctx):
    torch = torch_cuda
    vol = torch.zeros((60, 40, 8), device="cuda")
    vol[3, 4, 5] = float("nan")
    with pytest.raises(trec.TrecError):
        ctx.geopar(vol)
    with pytest.raises(trec.TrecError):
        ctx.geopar(torch.zeros((60, 40, 8), device="cuda"), axis=2)
    with pytest.raises(trec.TrecError):
        ctx.geopar(torch.zeros((2, 4097, 60), device="cuda"))
    # the context still works after the errors
    (a, t, z), _ = ctx.geopar(torch.from_numpy(slab_noise_volume((120, 70, 6), 0, 3)).cuda())
    assert np.isfinite([a, t, z]).all()


def test_full_size_fast_equals_exhaustive(torch_cuda, trec, ctx, orc):
    """The bench shape: TiltRec-cuda's output of a 1024 x 1024 series at
    thickness 300, MRC (1024, 300, 1024): 1024 x views of 300 rows, 20 chips
    each.  The windowed search equals the exhaustive one bit for bit on
    every chip; two views are checked against the oracle."""
    torch = torch_cuda
    g = torch.Generator(device="cuda").manual_seed(7)
    nz, ny, nx = 1024, 300, 1024
    r = torch.arange(ny, device="cuda", dtype=torch.float32)[None, :, None]
    c = torch.arange(nz, device="cuda", dtype=torch.float32)[:, None, None]
    x = torch.arange(nx, device="cuda", dtype=torch.float32)[None, None, :]
    centre = 150 + 0.05 * (c - 512) + 0.01 * (x - 512)
    vol = ((r - centre).abs() < 55).float() + 0.4 * torch.randn((nz, ny, nx), device="cuda", generator=g)
    views = torch.empty((nx, 3), device="cuda")
    chips = torch.empty((nx, 20, 4), device="cuda")
    chips2 = torch.empty_like(chips)
    print("geopar full size: windowed", flush=True)
    avg, nex = ctx.geopar(vol, 0, views, chips)
    print(f"  averages {avg}, {nex} of {nx * 20} chips exhaustive", flush=True)
    print("geopar full size: exhaustive", flush=True)
    avg2, nex2 = ctx.geopar(vol, 0, None, chips2, exhaustive=True)
    assert torch.equal(chips, chips2) and avg == avg2 and nex2 == nx * 20
    assert nex < nx * 20 // 100
    host = vol.cpu().numpy()
    for j in (3, 700):
        sub = np.ascontiguousarray(host[:, :, j:j + 1])
        _, c_ref, _ = orc.geopar(sub, 0)
        np.testing.assert_array_equal(chips[j].cpu().numpy(), c_ref[0])
    assert abs(avg[0] - math.degrees(math.atan(0.05))) < 0.5 and abs(avg[1] - 110) < 4


def test_cli(tmp_path, torch_cuda, orc):
    from autom2_amd import BIN_DIR, mrc
    vol = slab_noise_volume((210, 90, 14), 0, 77)
    mrc.write(str(tmp_path / "temp.mrc"), vol)
    exe = os.path.join(BIN_DIR, "GeoParGen")
    print("GeoParGen CLI", flush=True)
    r = subprocess.run([exe, str(tmp_path / "temp.mrc")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = dict(l.split(" : ") for l in r.stdout.splitlines() if l.startswith("average "))
    _, _, a_ref = orc.geopar(vol, 0)
    for key, want in zip(("average angle", "average thick", "average z"), a_ref):
        assert abs(float(lines[key]) - want) <= 1e-5 * max(1, abs(want)), key
    r = subprocess.run([exe, str(tmp_path / "temp.mrc"), "y"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "average z : " in r.stdout
    r = subprocess.run([exe, str(tmp_path / "temp.mrc"), "q"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "invalid parameter!" in r.stdout
    r = subprocess.run([exe, str(tmp_path / "missing.mrc")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 4 and "file read failed!" in r.stdout
