"""The integer restatement of the reference's int64-through-double division
(piranha_amd/csrc/ring_div.h, functors.cuh:33-45) against the host's IEEE-754
binary64 division: the same header the GPU kernels compile, built for the host
with g++, checked on edge operands, every power of two and random operands of
every bit-length pairing."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "piranha_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("ring_div") / "ring_div_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "native", "ring_div_check.cpp"),
                    "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_signed_div_matches_binary64(checker, seed):
    r = subprocess.run([checker, "1500000", str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 mismatches" in r.stdout
