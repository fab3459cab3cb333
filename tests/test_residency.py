"""The serving loop's residency policy (final184_amd/csrc/residency.hpp), compiled with g++
and run on the host: LRU, budget, pins, protection, same-shape recycling, waiting without
side effects, and randomized invariants."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_residency_policy(tmp_path):
    exe = str(tmp_path / "residency_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-fsanitize=address,undefined",
                    os.path.join(ROOT, "tests", "cpp", "residency_test.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "RESIDENCY OK" in out.stdout
