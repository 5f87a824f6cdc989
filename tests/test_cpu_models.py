"""CPU models of engine kernels, built with g++ and the address and
undefined-behaviour sanitizers and run on the host (no GPU).

radix_pass_model.cpp restates one LSD digit pass of the engine's radix sort
(radix_hist_kernel, the offset scan, radix_scatter_kernel in
sparqsim_amd/csrc/engine/sort.hip) on lockstep 64-lane waves with bounds checks
on every LDS and global access and a race detector on the LDS arrays, runs the
waves of each block in several orders, and checks the stable order against
std::stable_sort -- for the committed kernel and for two re-orderings of it
(digit offsets loaded at kernel start; all write-out LDS reads before the
stores).  Two planted mistakes (a missing barrier, a rank off by one) must be
caught."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_models", "radix_pass_model.cpp")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")


def build(tmp_path, *defines):
    exe = str(tmp_path / ("model" + "".join("_" + d for d in defines)))
    cmd = ["g++", "-std=c++17", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           *[f"-D{d}" for d in defines], SRC, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_radix_pass_model_in_bounds_race_free_stable(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "every access in bounds, no LDS race" in r.stdout


@pytest.mark.parametrize("plant,message", [("PLANT_MISSING_BARRIER", "another wave"),
                                           ("PLANT_RANK_OFF_BY_ONE", "MODEL FAILURE")])
def test_radix_pass_model_catches_planted_mistakes(tmp_path, plant, message):
    exe = build(tmp_path, plant)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert message in r.stderr, r.stderr[-2000:]
