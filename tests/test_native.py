"""Host-only C++ pieces of the library, built with g++ under the address and
undefined-behaviour sanitizers and run on the CPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_range_allocator_under_sanitizers(tmp_path):
    exe = tmp_path / "range_alloc_test"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "autom2_amd", "csrc"),
                           os.path.join(ROOT, "tests", "native", "range_alloc_test.cpp"), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stdout + out.stderr
