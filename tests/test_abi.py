"""The C-ABI library loads without a GPU and exports every function that
include/tiltrec.h declares; the Python binding covers all of them."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tiltrec.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(trec_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared()
    for n in ("trec_reconstruct", "trec_filter_rweight", "trec_forward_project", "trec_backproject",
              "trec_sirt_iterate", "trec_reconstruct_batch", "trec_server_create"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from autom2_amd import tiltrec
    lib = ctypes.CDLL(tiltrec.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    from autom2_amd import tiltrec
    bound = {s[0] for s in tiltrec.SIGNATURES}
    assert set(declared()) == bound


def test_host_only_calls_work_without_gpu():
    """Coefficients, workspace sizes and status strings need no device."""
    import numpy as np
    from autom2_amd import tiltrec
    from oracle import orc
    ang = np.linspace(-60, 60, 7).astype(np.float32)
    g = tiltrec.make_geom(64, 32, 7, 16, 1.5, 3.0, 2.0)
    c = tiltrec.coefficients(g, ang)
    np.testing.assert_array_equal(c, orc.coeffs_mpi(ang, 1.5, 3.0, 2.0))
    assert tiltrec.workspace_size(g, tiltrec.make_method("SIRT,5,1.0")) > 0
    assert tiltrec.lib().trec_status_string(tiltrec.ERR_BUSY) == b"busy"
    assert tiltrec.lib().trec_version().startswith(b"tiltrec-hip")
    with pytest.raises(tiltrec.TrecError):
        tiltrec.workspace_size(tiltrec.make_geom(1, 1, 1, 1), tiltrec.make_method("BPT"))


def test_context_without_device_reports_device_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from autom2_amd import tiltrec
    with pytest.raises(tiltrec.TrecError) as ei:
        tiltrec.Context(0)
    assert ei.value.status == tiltrec.ERR_DEVICE


def test_method_and_geometry_strings():
    from autom2_amd import tiltrec as t
    assert t.parse_method("SIRT,20,1.0") == (t.SIRT, 20, 1.0)
    assert t.parse_method("SART,5,0.5") == (t.SART, 5, 0.5)
    assert t.parse_method("WBP") == (t.WBP, 0, 1.0)
    assert t.parse_geometry("0,0,0,300") == (0.0, 0.0, 0.0, 300)
    for bad in ("SIRT", "FOO,1,1", "WBP,3", "SIRT,1"):
        with pytest.raises(ValueError):
            t.parse_method(bad)
