"""The C++ adapter (wavefront-renderer_amd/host/wavefront_host.hpp) driven by a C++ program that
restates the reference's GoogleTest cases against it (host_selftest.cpp)."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HOST = ROOT / "wavefront-renderer_amd" / "host"
LIB = ROOT / "wavefront-renderer_amd" / "lib"


def build_selftest(out: Path) -> Path:
    exe = out / "host_selftest"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", f"-I{HOST}",
                    str(HOST / "host_selftest.cpp"), f"-L{LIB}", "-lwavefront_hip",
                    f"-Wl,-rpath,{LIB}", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(exe)],
                   check=True)
    return exe


def test_cpp_adapter_builds(tmp_path):
    assert build_selftest(tmp_path).exists()


@pytest.mark.gpu
def test_cpp_adapter_reference_tests(tmp_path):
    exe = build_selftest(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
