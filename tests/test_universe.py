"""The Universe-shaped C++ wrapper (include/pm_universe.hpp) over the C ABI.

CPU: the header compiles and links against libpm.so with g++ alone (no HIP
headers), as a reference caller would build it.  GPU: tests/cpp/universe_check
(built by the build step) constructs a universe, clears it, adds two galaxies
(one with a black hole), checks the addGalaxy initial-condition law on the
read-back state, generates and pops frames, and stops.
"""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "cosmossimulationwithcuda_amd")


def test_universe_header_compiles_with_gxx():
    src = os.path.join(ROOT, "tests", "cpp", "universe_check.cpp")
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "universe_check")
        r = subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), src,
                            "-L" + LIB_DIR, "-lpm", "-pthread", "-o", out], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_universe_drives_the_engine():
    exe = os.path.join(ROOT, "tests", "bin", "universe_check")
    assert os.path.exists(exe), "build first: python -m cosmossimulationwithcuda_amd.build"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "universe ok" in r.stdout, r.stdout + r.stderr
