"""MRC2014 data on the GPU: trec_convert against numpy, and the TiltRec CLI
reading every MRC mode and byte order into the same reconstruction, then
writing half volumes and MRC2014 headers."""
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import tilts

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


def bits_equal(a, b):
    """Equal bit patterns, any NaN matching any NaN."""
    a, b = np.asarray(a), np.asarray(b)
    ia = a.view(np.uint32 if a.itemsize == 4 else np.uint16)
    ib = b.view(np.uint32 if b.itemsize == 4 else np.uint16)
    return bool(np.all((ia == ib) | (np.isnan(a) & np.isnan(b))))


def test_convert_to_float(torch_cuda, trec, ctx):
    torch = torch_cuda
    cases = [
        (trec.DT_I8, np.arange(-128, 128, dtype=np.int8), lambda a: a.astype(np.float32)),
        (trec.DT_U8, np.arange(256, dtype=np.uint8), lambda a: a.astype(np.float32)),
        (trec.DT_U8_DIV255, np.arange(256, dtype=np.uint8), lambda a: a.astype(np.float32) / np.float32(255)),
        (trec.DT_I16, np.arange(-32768, 32768, dtype=np.int32).astype(np.int16), lambda a: a.astype(np.float32)),
        (trec.DT_U16, np.arange(65536, dtype=np.int32).astype(np.uint16), lambda a: a.astype(np.float32)),
        (trec.DT_F16, np.arange(65536, dtype=np.int32).astype(np.uint16).view(np.float16),
         lambda a: a.astype(np.float32)),
    ]
    for dt, src, ref in cases:
        print(f"convert type {dt} -> float", flush=True)
        raw = torch.from_numpy(src.view(np.uint8).copy()).cuda()
        out = torch.empty(src.size, dtype=torch.float32, device="cuda")
        ctx.convert(raw, dt, out, trec.DT_F32)
        torch.cuda.synchronize()
        assert bits_equal(out.cpu().numpy(), ref(src)), dt


def test_convert_float_to_half(torch_cuda, trec, ctx):
    torch = torch_cuda
    rng = np.random.default_rng(3)
    h = np.arange(65536, dtype=np.int32).astype(np.uint16).view(np.float16)
    fin = h[np.isfinite(h)].astype(np.float32)
    mids = ((fin[:-1].astype(np.float64) + fin[1:]) / 2).astype(np.float32)  # exact ties between halves
    vals = np.concatenate([fin, mids, rng.standard_normal(100000).astype(np.float32) * 1000,
                           np.float32([65504, 65519.99, 65520, 1e6, -1e6, 1e-8, -1e-8, 0.0, -0.0,
                                       np.inf, -np.inf, np.nan])]).astype(np.float32)
    print("convert float -> half", flush=True)
    src = torch.from_numpy(vals).cuda()
    out = torch.empty(vals.size, dtype=torch.float16, device="cuda")
    ctx.convert(src, trec.DT_F32, out, trec.DT_F16)
    torch.cuda.synchronize()
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16)  # numpy rounds to nearest even, overflow to inf
    assert bits_equal(out.cpu().numpy(), want)


def test_convert_errors(torch_cuda, trec, ctx):
    torch = torch_cuda
    a = torch.zeros(16, device="cuda")
    with pytest.raises(trec.TrecError):
        ctx.convert(a, trec.DT_F32, a, trec.DT_F32)  # overlap
    with pytest.raises(trec.TrecError):
        ctx.convert(a, trec.DT_F32, torch.zeros(16, dtype=torch.int16, device="cuda"), trec.DT_I16)


def write_mode(path, data, mode, big_endian=False):
    from autom2_amd import mrc
    nz, ny, nx = data.shape
    h = mrc.new_header(nx, ny, nz)
    h["mode"] = mode
    if mode == 101:
        rb = (nx + 1) // 2
        flat = data.reshape(nz * ny, nx).astype(np.uint8)
        packed = np.zeros((nz * ny, rb), np.uint8)
        for x in range(nx):
            packed[:, x // 2] |= flat[:, x] << (4 * (x & 1))
        payload = packed.tobytes()
    else:
        dt = np.dtype(mrc.MODES[mode]).newbyteorder(">" if big_endian else "<")
        payload = data.astype(dt).tobytes()
    hdr = h
    if big_endian:
        hdr = h.astype(mrc.HEADER_DTYPE.newbyteorder(">"))
        hdr["stamp"] = b"\x11\x11\x00\x00"
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(payload)


def test_cli_modes_and_half_output(tmp_path, torch_cuda):
    from autom2_amd import BIN_DIR, mrc
    P, Y, X, T = 7, 16, 40, 12
    rng = np.random.default_rng(9)
    stack = rng.integers(0, 16, (P, Y, X)).astype(np.float32)  # exact in every mode
    ang = tilts(P)
    (tmp_path / "t.tlt").write_text("\n".join(f"{a:.2f}" for a in ang) + "\n")
    exe = os.path.join(BIN_DIR, "TiltRecZ-cuda")

    def run(inp, out, *extra):
        r = subprocess.run([exe, "--input", str(inp), "--output", str(out), "--tiltfile", str(tmp_path / "t.tlt"),
                            "--geometry", f"0,0,0,{T}", "--method", "SIRT,3,1.0", *extra],
                           capture_output=True, text=True, timeout=120)
        print(r.stdout[-300:], r.stderr[-300:])
        assert r.returncode == 0
        return open(out, "rb").read()

    mrc.write(str(tmp_path / "f.mrc"), stack)
    print("CLI: float stack", flush=True)
    ref = run(tmp_path / "f.mrc", tmp_path / "ref.mrc")
    for mode, be in [(0, False), (1, False), (6, False), (12, False), (101, False), (2, True), (1, True),
                     (12, True)]:
        name = f"m{mode}{'be' if be else ''}"
        write_mode(str(tmp_path / f"{name}.mrc"), stack, mode, be)
        print(f"CLI: stack mode {mode}, big-endian {be}", flush=True)
        got = run(tmp_path / f"{name}.mrc", tmp_path / f"{name}_out.mrc")
        assert got[1024:] == ref[1024:], name
    print("CLI: half output with an MRC2014 header", flush=True)
    run(tmp_path / "f.mrc", tmp_path / "h.mrc", "--output-mode", "12", "--mrc2014")
    h, vol = mrc.read(str(tmp_path / "h.mrc"))
    _, vref = mrc.read(str(tmp_path / "ref.mrc"))
    assert int(h["mode"]) == 12 and os.path.getsize(tmp_path / "h.mrc") == 1024 + 2 * vref.size
    np.testing.assert_array_equal(vol, vref.astype(np.float16).astype(np.float32))
    raw = open(tmp_path / "h.mrc", "rb").read(1024)
    assert raw[208:212] == b"MAP " and raw[212:214] == b"\x44\x44"
    assert np.frombuffer(raw[108:112], "<i4")[0] == 20140
    assert tuple(np.frombuffer(raw[28:40], "<i4")) == (X, Y, T)
    amin, amax, amean = np.frombuffer(raw[76:88], "<f4")
    v = vol.astype(np.float64)
    assert (amin, amax) == (np.float32(v.min()), np.float32(v.max()))
    assert abs(amean - v.mean()) <= 1e-6 * max(1, abs(v.mean()))
    assert abs(np.frombuffer(raw[216:220], "<f4")[0] - v.std()) <= 1e-5 * v.std()
    r = subprocess.run([exe, "--input", str(tmp_path / "f.mrc"), "--output", str(tmp_path / "x.mrc"), "--tiltfile",
                        str(tmp_path / "t.tlt"), "--geometry", f"0,0,0,{T}", "--method", "BPT", "--f2b",
                        "--output-mode", "12"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
