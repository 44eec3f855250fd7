"""Host build of the mixed-precision face numerics (fv_mixed.hpp compiled
with g++ under TECNO_HOST_BUILD, no GPU): the fused vertex combination
against the vertex lists of compute_vertices, the rounding-aware sign rule,
the residual-corrected quotient, the per-cell-logarithm log-mean against
the oracle's logavg (misc_func.h:14-32), and the FP32 SP-WENO jumps against
the oracle's sp_weno (limiter.h:193-252)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_ctypes as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dsp-weno_amd", "csrc")

HARNESS = r"""
#define TECNO_HOST_BUILD 1
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <random>
#include "fv_mixed.hpp"
using namespace tecno;
int main() {
  std::mt19937 g(11);
  std::uniform_real_distribution<float> U(-3.0f, 3.0f);
  const float sp[] = {1.0f, 1.0f + 1e-8f, 1.0f - 1e-8f, 1.0f + 2e-8f, 0.0f, -1.0f, 2.0f,
                      1.0000001f, 0.9999999f, 0.5f, -0.5f};
  double dmax = 0.0;
  for (int it = 0; it < 300000; ++it) {
    float tm = U(g), tp = U(g), md = std::fabs(U(g)) * 0.1f;
    if (it % 7 == 0) tm = sp[it % 11];
    if (it % 5 == 0) tp = sp[(it / 5) % 11];
    if (it % 13 == 0) md = 0.0f;
    float lg[5], xv[5], yv[5], s[5], C1a, C2a, C1b, C2b;
    for (int i = 0; i < 5; ++i) lg[i] = U(g);
    fast::vertices(tm, tp, md, xv, yv);
    fast::head_c(lg, xv, yv, C1a, C2a);
    float mx = 0.0f;
    for (int i = 0; i < 5; ++i) mx = fmaxf(mx, lg[i]);
    for (int i = 0; i < 5; ++i) s[i] = fast::exp(lg[i] - mx);
    fast::vertex_combine(tm, tp, md, s, C1b, C2b);
    dmax = std::fmax(dmax, std::fmax(std::fabs(C1a - C1b), std::fabs(C2a - C2b)));
  }
  std::printf("combine %.6e\n", dmax);
  // sign rule: tiny wrong-signed jump -> 0; large wrong-signed kept; right sign kept
  std::printf("sign %g %g %g\n", fast::sign_round(-1e-9f, 1.0f, 1.5f, 1.2f),
              fast::sign_round(-0.1f, 1.0f, 1.5f, 1.2f), fast::sign_round(0.2f, 1.0f, 1.5f, 1.2f));
  // quotient: exact for representable quotients (equal / opposite jumps)
  int exact = 1;
  for (int it = 0; it < 100000; ++it) {
    float d = U(g) * std::pow(10.0f, (float)(it % 20 - 10));
    if (d == 0.0f) continue;
    exact &= fast::quot(d, d) == 1.0f && fast::quot(-d, d) == -1.0f && fast::quot(2.0f * d, d) == 2.0f;
  }
  std::printf("quot %d\n", exact);
  // log-mean from per-cell logarithms, for the caller to compare with the oracle
  std::mt19937_64 h(5);
  std::uniform_real_distribution<double> V(0.2, 5.0);
  FILE* f = std::fopen(std::getenv("LOGAVG_OUT"), "w");
  for (int it = 0; it < 20000; ++it) {
    double a = V(h), b = (it % 3 == 0) ? a * (1.0 + (V(h) - 2.6) * 1e-3) : V(h);
    double ib;
    const double m2 = logavg_ln(a, b, std::log(a), std::log(b), &ib);
    std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g\n", a, b, logavg_ln(a, b, std::log(a), std::log(b)),
                 logavg_fast(a, b), m2, ib);
  }
  std::fclose(f);
  // record logarithm: largest error in ulp of std::log over 24 decades, near 1,
  // at powers of two and at the boundary of the mantissa range; NaN for x <= 0
  double ulp_max = 0.0, abs_near1 = 0.0;
  std::uniform_real_distribution<double> E(-40.0, 40.0), W(-1e-3, 1e-3);
  for (int it = 0; it < 400000; ++it) {
    double x = std::exp(E(h));
    if (it % 4 == 1) x = 1.0 + W(h);
    if (it % 64 == 2) x = std::ldexp(1.0, it % 200 - 100);
    if (it % 64 == 3) x = std::ldexp(0.70710678118654752 * (1.0 + W(h) * 1e-9), it % 40 - 20);
    const double a = log64(x), b = std::log(x);
    if (b != 0.0) ulp_max = std::fmax(ulp_max, std::fabs(a - b) / (std::nextafter(std::fabs(b), INFINITY) - std::fabs(b)));
    else ulp_max = std::fmax(ulp_max, std::fabs(a) * 1e300);
    if (std::fabs(x - 1.0) < 1e-3) abs_near1 = std::fmax(abs_near1, std::fabs(a - b));
  }
  std::printf("log64 %.3f %.3e %d %d %d\n", ulp_max, abs_near1, (int)std::isnan(log64(0.0)),
              (int)std::isnan(log64(-2.0)), (int)std::isnan(log64(NAN)));
  // FP32 SP-WENO jump (plain and sp_wenoc correction) on random, smooth,
  // stepped and degenerate stencils, for the caller to compare with the
  // oracle's FP64 sp_weno
  FILE* q = std::fopen(std::getenv("SPW_OUT"), "w");
  std::uniform_real_distribution<double> X(0.0, 6.283185307179586), Hs(-3.0, -0.5);
  for (int it = 0; it < 60000; ++it) {
    double u[4];
    const int kind = it % 6;
    const double x0 = X(h), hh = std::pow(10.0, Hs(h));
    for (int k = 0; k < 4; ++k) {
      if (kind == 0) u[k] = V(h) - 2.6;
      else if (kind == 1) u[k] = std::sin(x0 + k * hh);
      else if (kind == 2) u[k] = (k >= 1 + it % 3) ? 1.0 + 0.1 * std::sin(x0 + k * hh) : 0.125;
      else if (kind == 3) u[k] = 0.5 + k * hh;                      // equal jumps
      else if (kind == 4) u[k] = (k % 2) ? 0.5 + hh : 0.5;           // alternating jumps
      else u[k] = (k == it % 4) ? 0.5 + hh : 0.5;                   // one jump of three nonzero
    }
    for (int k = 0; k < 4; ++k) u[k] = (double)(float)u[k];
    const float d0 = (float)(u[1] - u[0]), d1 = (float)(u[2] - u[1]), d2 = (float)(u[3] - u[2]);
    const float z1 = (float)u[1], z2 = z1 + d1;
    // the FP32-state form (sp_weno<float> on the cell values)
    float l0, r0, l1, r1;
    sp_weno<float>((float)u[0], (float)u[1], (float)u[2], (float)u[3], 0, l0, r0);
    sp_weno<float>((float)u[0], (float)u[1], (float)u[2], (float)u[3], 1, l1, r1);
    std::fprintf(q, "%.17g %.17g %.17g %.17g %.9g %.9g %.9g %.9g\n", u[0], u[1], u[2], u[3],
                 fast::spw_jump(d0, d1, d2, z1, z2, 0), fast::spw_jump(d0, d1, d2, z1, z2, 1),
                 (double)r0 - (double)l0, (double)r1 - (double)l1);
  }
  std::fclose(q);
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
        la = os.path.join(d, "logavg.txt")
        sw = os.path.join(d, "spw.txt")
        out = subprocess.run([exe], check=True, capture_output=True, text=True,
                             env=dict(os.environ, LOGAVG_OUT=la, SPW_OUT=sw)).stdout
        vals = np.loadtxt(la)
        spw = np.loadtxt(sw)
    return out, vals, spw


def test_vertex_combine_matches_vertex_lists(harness_output):
    out, _, _ = harness_output
    d = float(out.split("combine")[1].split()[0])
    assert d <= 1e-6


def test_sign_round(harness_output):
    out, _, _ = harness_output
    vals = [float(x) for x in out.split("sign")[1].split()[:3]]
    assert vals[0] == 0.0 and vals[1] == pytest.approx(-0.1) and vals[2] == pytest.approx(0.2)


def test_quotient_exact_for_representable(harness_output):
    out, _, _ = harness_output
    assert out.split("quot")[1].split()[0] == "1"


def test_logavg_from_cell_logs_matches_oracle(harness_output):
    _, vals, _ = harness_output
    ref = orc.logavg(vals[:, 0], vals[:, 1])
    rel = np.abs(vals[:, 2] - ref) / np.abs(ref)
    assert rel.max() <= 4e-15


def test_logavg_fast_and_inverse_match_oracle(harness_output):
    """logavg_fast (the FP32 paths' FP64 flux: one reciprocal for f, log64 in
    the logarithm branch) and logavg_ln's mean-and-inverse form against the
    oracle's logavg (misc_func.h:14-32), on pairs that take both branches."""
    _, vals, _ = harness_output
    ref = orc.logavg(vals[:, 0], vals[:, 1])
    assert (np.abs(vals[:, 3] - ref) / ref).max() <= 8e-15
    assert (np.abs(vals[:, 4] - ref) / ref).max() <= 8e-15
    assert (np.abs(vals[:, 5] * ref - 1.0)).max() <= 8e-15


def test_record_log_within_few_ulp(harness_output):
    """log64 (the per-cell logarithm of the mixed path's records) against the
    C library's log: at most 2 ulp anywhere, 1e-18 absolute near 1, NaN for
    x <= 0 and NaN input.  (The 2-ulp cases are x just above 1, where
    m + 1 rounds.)"""
    out, _, _ = harness_output
    f = out.split("log64")[1].split()
    assert float(f[0]) <= 2.25
    assert float(f[1]) <= 1e-18
    assert f[2:5] == ["1", "1", "1"]


@pytest.mark.parametrize("form", ["jump", "state"])
@pytest.mark.parametrize("corr", [0, 1])
def test_fp32_spweno_jump_matches_oracle(harness_output, corr, form):
    """The FP32 SP-WENO jump against the oracle's FP64 sp_weno
    (limiter.h:193-252; corr 1 = sp_wenoc) on random, smooth, stepped and
    degenerate stencils (equal, alternating and single jumps, where a ratio is
    exactly +-1): fast::spw_jump of the mixed paths (cell jumps taken in FP64,
    rounded to FP32) and sp_weno<float> of the FP32-state paths.  The jump
    r - l is within 1e-6 (jump form) / 2e-6 (state form) of the stencil's
    jump sum; on the degenerate stencils a wrong region branch is off by a
    quarter of it or more.  The state form's l and r are FP32 values, so its
    bound adds 4 FP32 ulp of the stencil's largest value."""
    _, _, spw = harness_output
    u = np.ascontiguousarray(spw[:, :4])
    l, r = orc.sp_weno(u, corr)
    ref = r - l
    scale = np.abs(np.diff(u, axis=1)).sum(axis=1)
    col = (4 if form == "jump" else 6) + corr
    err = np.abs(spw[:, col] - ref)
    bound = 1e-6 * scale + 1e-30
    if form == "state":
        bound = 2e-6 * scale + 4 * np.finfo(np.float32).eps * np.abs(u).max(axis=1)
    assert (err <= bound).all(), (err / bound).max()
