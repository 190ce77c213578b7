"""The C++ host mirror (include/candela_rt.hpp) compiles against the library and behaves like RayIntersector<T>."""
import os
import subprocess

import pytest

from conftest import ROOT


def _build(tmp_path):
    exe = str(tmp_path / "host_api_demo")
    lib_dir = os.path.join(ROOT, "candela_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", os.path.join(ROOT, "tests", "cpp",
                    "host_api_demo.cpp"), f"-L{lib_dir}", "-lcandela_rt", f"-Wl,-rpath,{lib_dir}", "-o", exe],
                   check=True)
    return exe


def test_host_api_cpu(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "OK" in out.stdout, out.stdout + out.stderr


@pytest.mark.gpu
def test_host_api_gpu(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe, "--gpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "OK" in out.stdout, out.stdout + out.stderr
