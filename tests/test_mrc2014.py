"""MRC2014 completeness of the Python MRC module (autom2_amd/mrc.py): byte
order, mode 101, complex modes, and the MRC2014 header fields at the byte
offsets the MRC2014 specification gives (MAP at 208, machine stamp at 212,
NVERSION at word 28 = byte 108, RMS at 216, NLABL at 220)."""
import os

import numpy as np
import pytest

from autom2_amd import mrc


def write_be(path, data, mode, stamp=True):
    """A big-endian MRC file written field by field with numpy's '>' types."""
    nz, ny, nx = data.shape
    h = mrc.new_header(nx, ny, nz)
    h["mode"] = mode
    hb = h.astype(mrc.HEADER_DTYPE.newbyteorder(">"))
    if stamp:
        hb["stamp"] = b"\x11\x11\x00\x00"
    dt = np.dtype(mrc.MODES[mode]).newbyteorder(">")
    with open(path, "wb") as f:
        f.write(hb.tobytes())
        f.write(np.ascontiguousarray(data).astype(dt).tobytes())


@pytest.mark.parametrize("mode,stamp", [(2, True), (2, False), (1, True), (1, False), (6, True), (12, False)])
def test_big_endian(tmp_path, mode, stamp):
    rng = np.random.default_rng(mode)
    data = (rng.standard_normal((3, 5, 7)) * 100).astype(mrc.MODES[mode])
    p = str(tmp_path / "be.mrc")
    write_be(p, data, mode, stamp)
    h, got = mrc.read(p)
    assert (int(h["nx"]), int(h["ny"]), int(h["nz"]), int(h["mode"])) == (7, 5, 3, mode)
    np.testing.assert_array_equal(got, data.astype(np.float32))


def test_mode_101(tmp_path):
    rng = np.random.default_rng(1)
    nz, ny, nx = 2, 3, 7  # odd width: each row padded to a whole byte
    vals = rng.integers(0, 16, (nz, ny, nx)).astype(np.uint8)
    rb = (nx + 1) // 2
    packed = np.zeros((nz * ny, rb), np.uint8)
    flat = vals.reshape(nz * ny, nx)
    for x in range(nx):
        packed[:, x // 2] |= flat[:, x] << (4 * (x & 1))
    h = mrc.new_header(nx, ny, nz)
    h["mode"] = 101
    p = str(tmp_path / "n.mrc")
    with open(p, "wb") as f:
        f.write(h.tobytes())
        f.write(packed.tobytes())
    _, got = mrc.read(p)
    np.testing.assert_array_equal(got, vals.astype(np.float32))


def test_complex_mode_refused(tmp_path):
    h = mrc.new_header(4, 4, 1)
    h["mode"] = 4
    p = str(tmp_path / "c.mrc")
    with open(p, "wb") as f:
        f.write(h.tobytes())
        f.write(np.zeros(32, np.float32).tobytes())
    with pytest.raises(ValueError):
        mrc.read(p)


@pytest.mark.parametrize("mode", [2, 12])
def test_mrc2014_header(tmp_path, mode):
    rng = np.random.default_rng(2)
    data = (rng.standard_normal((4, 6, 9)) * 3 + 1).astype(np.float32)
    p = str(tmp_path / "v.mrc")
    mrc.write(p, data, mode=mode, mrc2014=True)
    n = data.size
    assert os.path.getsize(p) == 1024 + n * (4 if mode == 2 else 2)
    raw = open(p, "rb").read(1024)
    assert raw[208:212] == b"MAP " and raw[212:216] == b"\x44\x44\x00\x00"
    assert np.frombuffer(raw[108:112], "<i4")[0] == 20140
    assert tuple(np.frombuffer(raw[0:16], "<i4")) == (9, 6, 4, mode)
    assert tuple(np.frombuffer(raw[28:40], "<i4")) == (9, 6, 4)       # mx, my, mz
    assert tuple(np.frombuffer(raw[40:52], "<f4")) == (9.0, 6.0, 4.0)  # cell lengths, 1 A per voxel
    stored = data.astype(np.float16 if mode == 12 else np.float32).astype(np.float64)
    amin, amax, amean = np.frombuffer(raw[76:88], "<f4")
    rms = np.frombuffer(raw[216:220], "<f4")[0]
    assert (amin, amax) == (np.float32(stored.min()), np.float32(stored.max()))
    assert abs(amean - stored.mean()) < 1e-6 and abs(rms - stored.std()) < 1e-6
    h, back = mrc.read(p)
    np.testing.assert_array_equal(back, stored.astype(np.float32))


def test_reference_header_unchanged(tmp_path):
    """Without mrc2014 the header stays TiltRec's: no MAP stamp, mx = 1,
    statistics of section nz/2 with amax starting at FLT_MIN."""
    data = -np.ones((3, 2, 2), np.float32)
    p = str(tmp_path / "r.mrc")
    mrc.write(p, data)
    raw = open(p, "rb").read(1024)
    assert raw[208:216] == b"\0" * 8
    assert tuple(np.frombuffer(raw[28:40], "<i4")) == (1, 1, 1)
    amin, amax, amean = np.frombuffer(raw[76:88], "<f4")
    assert amin == -1 and amax == mrc.FLT_MIN and amean == -1
