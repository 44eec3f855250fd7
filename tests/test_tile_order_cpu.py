"""Host build of the XCD-aware tile order of the stage kernels (xcd_tile in
fv_kernels.hpp, compiled with g++ under TECNO_HOST_BUILD): for every grid
the map from workgroup id to tile is a bijection, each XCD (workgroup id
mod 8) takes its tiles in runs of whole tile rows / planes, and those runs
come from all over the grid (the load of a tile depends on the flow in it)."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dsp-weno_amd", "csrc")

HARNESS = r"""
#define TECNO_HOST_BUILD 1
#include <cstdio>
#include <vector>
#include "fv_kernels.hpp"
using namespace tecno;
int main() {
  const unsigned grids[][3] = {{128, 137, 1}, {1, 1, 1}, {3, 5, 1}, {7, 7, 1}, {16, 16, 64},
                               {5, 3, 7}, {64, 2, 1}, {128, 128, 1}, {32, 69, 1}, {17, 1, 1},
                               {32, 32, 128}, {9, 11, 13}, {256, 1, 1}, {13, 2, 1}};
  for (auto& g : grids) {
    const unsigned T = g[0] * g[1] * g[2];
    std::vector<int> seen(T, 0);
    int bad = 0;
    for (unsigned l = 0; l < T; ++l) {
      const unsigned t = xcd_tile(l, g[0], g[1], g[2]);
      if (t >= T) { bad = 1; break; }
      seen[t]++;
    }
    for (unsigned t = 0; t < T && !bad; ++t) bad |= seen[t] != 1;
    // XCD 0 of the 2048^2 grid of 16 x 15 tiles: consecutive workgroups of an
    // XCD walk along tile rows; the rows it gets span the grid
    int contiguous = 0, steps = 0;
    unsigned rmin = ~0u, rmax = 0;
    for (unsigned k = 0; 8 * (k + 1) < T; ++k) {
      const unsigned a = xcd_tile(8 * k, g[0], g[1], g[2]), b = xcd_tile(8 * (k + 1), g[0], g[1], g[2]);
      contiguous += (b == a + 1);
      ++steps;
      const unsigned row = a / g[0];
      rmin = row < rmin ? row : rmin;
      rmax = row > rmax ? row : rmax;
    }
    std::printf("grid %u %u %u bijection %d contiguous %d/%d rows %u..%u\n", g[0], g[1], g[2],
                !bad, contiguous, steps, rmin, rmax);
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness_output():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "h.cpp")
        exe = os.path.join(d, "h")
        with open(src, "w") as f:
            f.write(HARNESS)
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = {}
    for line in out.splitlines():
        w = line.split()
        rows[tuple(int(x) for x in w[1:4])] = dict(
            bij=w[5] == "1", cont=tuple(int(x) for x in w[7].split("/")),
            rows=tuple(int(x) for x in w[9].split("..")))
    return rows


def test_tile_order_is_a_bijection(harness_output):
    assert len(harness_output) == 14
    for g, r in harness_output.items():
        assert r["bij"], g


def test_tile_order_runs_and_spread(harness_output):
    # 2048^2 with 16 x 15 tiles: 128 x 137 tiles, chunks of two tile rows
    r = harness_output[(128, 137, 1)]
    c, n = r["cont"]
    assert c >= 0.99 * n            # XCD 0 walks whole rows
    assert r["rows"][0] <= 2 and r["rows"][1] >= 120   # from the bottom to the top of the grid
    # 3D: chunks of two tile planes
    r = harness_output[(16, 16, 64)]
    c, n = r["cont"]
    assert c >= 0.99 * n
