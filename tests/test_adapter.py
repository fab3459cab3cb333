"""The C++ adapter (include/vgi_foreground.hpp) compiled and run as a program: without a
device it must fail cleanly through LastStatus(); on the GPU, RenderList over two
primitives with model matrices must equal the oracle's voxelization of each."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "adapter_test.cpp")


def _build(tmp_path):
    exe = str(tmp_path / "adapter_test")
    libdirs = [os.path.join(ROOT, "final184_amd"), os.path.join(ROOT, "oracle", "_build")]
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe]
    for d in libdirs:
        cmd += ["-L", d, f"-Wl,-rpath,{d}"]
    cmd += ["-lvgi", "-lvgi_oracle"]
    subprocess.run(cmd, check=True)
    return exe


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device path")
def test_adapter_without_device(tmp_path):
    out = subprocess.run([_build(tmp_path), "nodevice"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ADAPTER NODEVICE OK" in out.stdout


@pytest.mark.gpu
def test_adapter_on_gpu(tmp_path):
    print("case adapter_test gpu", flush=True)
    out = subprocess.run([_build(tmp_path), "gpu"], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ADAPTER GPU OK" in out.stdout
