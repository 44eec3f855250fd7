"""The seeded initial network (netinit.hpp, behind tecno_init_net) against
PyTorch's own draws: tests/golden/fixtures_init.npz holds the weights and
biases that torch.manual_seed(seed), one nn.Linear per layer and
nn.init.xavier_uniform_ of every weight produce -- the trainer's
initialisation sequence (DSP-WENO_network_trainer.py:54-60, 296) -- for the
shipped 5-5-5-5-5 widths at seed 6008 and two other widths / seeds (one
above 2^32).  The same fixtures hold torch.randperm permutations under
seeded generators and the row orders of the first three epochs of a
DataLoader(shuffle=True) built around that network, as the trainer builds it
(:348, 351, 384-392): randperm_torch and tecno_trainer_epoch_seeds (through
the C ABI, host-only) reproduce them.  Host build with g++, no GPU; bitwise."""
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
  if (argv[1][0] == 'p') {  // p seed n: head 64, tail 64, sum i * p[i] mod 2^64
    const uint64_t seed = std::strtoull(argv[2], nullptr, 10);
    const int64_t n = std::strtoll(argv[3], nullptr, 10);
    std::vector<int> p;
    tecno::randperm_torch(seed, n, p);
    uint64_t sum = 0;
    for (int64_t i = 0; i < n; ++i) sum += (uint64_t)i * (uint64_t)p[(size_t)i];
    for (int64_t i = 0; i < n; ++i)
      if (i < 64 || i >= n - 64 || n <= 5000) std::printf("%d\n", p[(size_t)i]);
    std::printf("%llu\n", (unsigned long long)sum);
    return 0;
  }
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


def _perm(exe, seed, n):
    out = subprocess.run([exe, "p", str(seed), str(n)], check=True, capture_output=True,
                         text=True).stdout.split()
    return np.array([int(x) for x in out[:-1]]), int(out[-1])


@pytest.mark.parametrize("k", range(sum(1 for k in FIX.files if k.startswith("perm_n"))))
def test_randperm_matches_pytorch(exe, k):
    n, seed = int(FIX[f"perm_n{k}"]), int(FIX[f"perm_seed{k}"])
    p, total = _perm(exe, seed, n)
    if f"perm{k}" in FIX.files:
        assert np.array_equal(p, FIX[f"perm{k}"])
    else:
        assert np.array_equal(p[:64], FIX[f"perm_head{k}"])
        assert np.array_equal(p[64:], FIX[f"perm_tail{k}"])
        assert total == int(FIX[f"perm_sum{k}"])


def test_trainer_epoch_orders_match_dataloader(exe):
    """tecno_trainer_epoch_seeds (C ABI, host-only) + randperm_torch give the
    row order of each of the trainer's DataLoader epochs."""
    import tecno
    orders = FIX["epoch_orders"]
    seeds = tecno.trainer_epoch_seeds([5, 5, 5, 5, 5], 6008, len(orders))
    for e, sd in enumerate(seeds):
        p, _ = _perm(exe, sd, orders.shape[1])
        assert np.array_equal(p, orders[e]), e
