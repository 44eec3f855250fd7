"""The seeded initial network (netinit.hpp, behind tecno_init_net) against
PyTorch's own draws: tests/golden/fixtures_init.npz holds the weights and
biases that torch.manual_seed(seed), one nn.Linear per layer and
nn.init.xavier_uniform_ of every weight produce -- the trainer's
initialisation sequence (DSP-WENO_network_trainer.py:54-60, 296) -- for the
shipped 5-5-5-5-5 widths at seed 6008 and two other widths / seeds (one
above 2^32).  Host build with g++, no GPU; bitwise."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dsp-weno_amd", "csrc")
FIX = np.load(os.path.join(ROOT, "tests", "golden", "fixtures_init.npz"))
NCASE = sum(1 for k in FIX.files if k.startswith("widths"))

HARNESS = r"""
#include <cstdio>
#include <cstdlib>
#include "netinit.hpp"
int main(int argc, char** argv) {
  const uint64_t seed = std::strtoull(argv[1], nullptr, 10);
  std::vector<int32_t> w;
  for (int i = 2; i < argc; ++i) w.push_back(std::atoi(argv[i]));
  std::vector<double> W, b;
  tecno::init_net_torch((int)w.size() - 1, w.data(), seed, W, b);
  for (double x : W) std::printf("%a\n", x);
  for (double x : b) std::printf("%a\n", x);
  return 0;
}
"""


@pytest.fixture(scope="module")
def exe():
    with tempfile.TemporaryDirectory() as d:
        src, out = os.path.join(d, "h.cpp"), os.path.join(d, "h")
        with open(src, "w") as f:
            f.write(HARNESS)
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, src, "-o", out], check=True)
        yield out


@pytest.mark.parametrize("k", range(NCASE))
def test_init_net_matches_pytorch_draws(exe, k):
    widths = [int(x) for x in FIX[f"widths{k}"]]
    seed = int(FIX[f"seed{k}"])
    out = subprocess.run([exe, str(seed)] + [str(x) for x in widths], check=True,
                         capture_output=True, text=True).stdout.split()
    vals = np.array([float.fromhex(x) for x in out])
    W, b = FIX[f"W{k}"], FIX[f"b{k}"]
    assert len(vals) == len(W) + len(b)
    assert np.array_equal(vals[:len(W)], W)
    assert np.array_equal(vals[len(W):], b)


def test_init_net_seed_changes_draws(exe):
    a = subprocess.run([exe, "6008", "5", "5", "5"], check=True, capture_output=True,
                       text=True).stdout.split()
    b = subprocess.run([exe, "6009", "5", "5", "5"], check=True, capture_output=True,
                       text=True).stdout.split()
    assert a != b
