"""MRC format and the TiltRec command line (no GPU needed)."""
import os
import subprocess

import numpy as np
import pytest

from autom2_amd import BIN_DIR, mrc

CLI = os.path.join(BIN_DIR, "tiltrec")


def test_header_layout_offsets():
    """Offsets of mrcheader.h:24-101 (packed, 1024 bytes)."""
    d = mrc.HEADER_DTYPE
    expect = {"nx": 0, "mode": 12, "mx": 28, "xlen": 40, "alpha": 52, "mapc": 64, "amin": 76, "ispg": 88,
              "next": 92, "creatid": 96, "nint": 128, "min2": 136, "idtype": 160, "tiltangles": 172,
              "xorg": 196, "cmap": 208, "stamp": 212, "rms": 216, "nlabl": 220, "labels": 224}
    for k, v in expect.items():
        assert d.fields[k][1] == v, k
    assert d.itemsize == 1024


def test_round_trip_and_header_stats(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 7, 9)).astype(np.float32)
    p = str(tmp_path / "a.mrc")
    mrc.write(p, a)
    h, b = mrc.read(p)
    np.testing.assert_array_equal(a, b)
    assert os.path.getsize(p) == 1024 + 4 * a.size
    mid = a[2]
    assert h["amin"] == mid.min() and h["amax"] == mid.max()
    assert abs(h["amean"] - mid.astype(np.float64).mean()) < 1e-6
    assert (h["mx"], h["my"], h["mz"], h["mapc"], h["mapr"], h["maps"], h["creatid"]) == (1, 1, 1, 1, 2, 3, 1000)
    assert h["alpha"] == 90.0 and h["nlabl"] == 0 and h["next"] == 0


def test_all_negative_section_keeps_flt_min_max(tmp_path):
    """UpdateHeader starts amax at FLT_MIN, so an all-negative middle section
    reports amax = FLT_MIN (mrcstack.cpp:111-112)."""
    p = str(tmp_path / "n.mrc")
    mrc.write(p, -np.ones((3, 2, 2), np.float32))
    assert mrc.read_header(p)["amax"] == mrc.FLT_MIN


@pytest.mark.parametrize("mode,dtype", [(0, np.int8), (1, np.int16), (6, np.uint16), (12, np.float16)])
def test_read_integer_and_half_modes(tmp_path, mode, dtype):
    h = mrc.new_header(4, 3, 2)
    h["mode"] = mode
    h["next"] = 16
    vals = np.arange(-8, 16).astype(dtype) if mode != 6 else np.arange(24).astype(dtype)
    p = str(tmp_path / "m.mrc")
    with open(p, "wb") as f:
        f.write(h.tobytes())
        f.write(b"\0" * 16)
        f.write(vals.tobytes())
    _, d = mrc.read(p)
    np.testing.assert_array_equal(d.ravel(), vals.astype(np.float32))


def run_cli(args, env=None, name=None):
    exe = CLI
    if name:
        exe = os.path.join(os.path.dirname(CLI), name)
        if not os.path.exists(exe):
            os.symlink("tiltrec", exe)
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([exe] + args, capture_output=True, text=True, env=e, timeout=60)


BASE = ["--input", "in.mrc", "--output", "out.mrc", "--tiltfile", "t.tlt", "--geometry", "0,0,0,300"]


def test_cli_parses_reference_argv():
    import json
    r = run_cli(BASE + ["--method", "SIRT,20,1.0", "--dry-run"])
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)
    assert (o["method"], o["iterations"], o["gamma"], o["thickness"]) == ("SIRT", 20, 1.0, 300)
    assert (o["layout"], o["resolved_rules"]) == ("y", "cuda")
    r = run_cli(["-i", "a", "-o", "b", "-t", "c", "-g", "1.5,-3,2,64", "-m", "WBP", "-a", "z", "--dry-run"])
    o = json.loads(r.stdout)
    assert (o["offset"], o["pitch"], o["zshift"], o["layout"]) == (1.5, -3.0, 2.0, "z")
    r = run_cli(BASE + ["--method", "BPT", "--inital", "init.mrc", "--dry-run"])
    assert json.loads(r.stdout)["initial"] == "init.mrc"


def test_cli_program_name_selects_orientation_and_rules():
    import json
    for name, layout, rules in (("TiltRecZ-cuda", "z", "cuda"), ("TiltRec-mpi", "y", "mpi"),
                                ("TiltRecZ-mpi", "z", "mpi"), ("TiltRec-cuda", "y", "cuda")):
        r = run_cli(BASE + ["--method", "BPT", "--dry-run"], name=name)
        o = json.loads(r.stdout)
        assert (o["layout"], o["resolved_rules"]) == (layout, rules), name


@pytest.mark.parametrize("bad", [
    ["--method", "SIRT"],                  # reference would read uninitialised values
    ["--method", "SIRT,10"],
    ["--method", "FOO"],
    ["--method", "WBP,1"],
    ["--method", "BPT", "--axis", "x"],
    ["--method", "BPT", "--storage", "f8"],
    ["--method", "BPT", "--bogus"],
])
def test_cli_rejects_bad_options(bad):
    r = run_cli(BASE + bad)
    assert r.returncode == 2, (bad, r.stdout, r.stderr)


def test_cli_rejects_bad_geometry():
    r = run_cli(["-i", "a", "-o", "b", "-t", "c", "-g", "0,0,0", "-m", "BPT"])
    assert r.returncode == 2
    r = run_cli(["-i", "a", "-o", "b", "-t", "c", "-g", "0,0,x,10", "-m", "BPT"])
    assert r.returncode == 2


def test_cli_non_zero_rank_exits_cleanly():
    r = run_cli(BASE + ["--method", "BPT"], env={"OMPI_COMM_WORLD_RANK": "2"})
    assert r.returncode == 0 and "rank 0" in r.stdout


def test_cli_missing_input_fails(tmp_path):
    r = run_cli(["-i", str(tmp_path / "none.mrc"), "-o", str(tmp_path / "o.mrc"), "-t", "t", "-g", "0,0,0,8",
                 "-m", "BPT"])
    assert r.returncode == 4 and "cannot open" in r.stderr


def test_cli_parses_admm_and_fbp():
    import json
    r = run_cli(BASE + ["--method", "ADMM,5,3,1.0,0.1", "--dry-run"])
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert (d["method"], d["iterations"], d["cgiter"], d["gamma"], d["soft"]) == ("ADMM", 5, 3, 1.0, 0.1)
    r = run_cli(BASE + ["--method", "FBP", "--f2b", "--dry-run"])
    assert r.returncode == 0 and json.loads(r.stdout)["f2b"] is True
