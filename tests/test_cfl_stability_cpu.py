"""Stability of the fully discrete scheme (TeCNO flux + SP-WENO + SSPRK3)
on the grid-scale checkerboard, on the oracle (the reference's arithmetic).

SP-WENO (limiter.h:103-252) reconstructs, on an alternating stencil, an
interface jump twice the cell jump, so the TeCNO dissipation term
(kepes_tecno4_roe.cc:141-176) damps the checkerboard (-1)^(i+j+k) of a wave
of speed |lambda| at rate 4 |lambda| / h per direction.  With the time step
of compute_dt (fv.cc:365-382, dt = cfl h / max(|u| + c)) a resting
acoustic checkerboard is multiplied per SSPRK3 step by
g(z) = 1 + z + z^2/2 + z^3/6 with z = -4 cfl d in d dimensions, and
|g| <= 1 needs z >= -2.5127: cfl <= 0.628 / d (0.31 in 2D, 0.21 in 3D).
The tests measure g on the oracle and compare it with this closed form."""
import numpy as np
import pytest

import oracle_ctypes as orc
from tecno import make_cfg

Z_STAB = -2.5127453266183286  # real root of 1 + z + z^2/2 + z^3/6 = -1


def ssprk3_gain(z):
    return 1 + z + z * z / 2 + z ** 3 / 6


def test_ssprk3_real_axis_bound():
    assert abs(ssprk3_gain(Z_STAB) + 1) < 1e-12
    assert abs(ssprk3_gain(Z_STAB + 1e-3)) < 1 < abs(ssprk3_gain(Z_STAB - 1e-3))


def test_spweno_jump_on_alternating_stencil_is_twice_the_cell_jump():
    st = np.array([[1.0, -1.0, 1.0, -1.0], [-1.0, 1.0, -1.0, 1.0]]) * 1e-3
    for corr in (0, 1, 2):
        l, r = orc.sp_weno(st, corr)
        assert np.allclose((r - l) / (st[:, 2] - st[:, 1]), 2.0, rtol=1e-12)


def checkerboard_gain(cfl, dims, n=8, steps=6):
    """Per-step growth of a resting acoustic checkerboard (rho = 1,
    p = 100/1.4 so c = 10; dp = 1e-8 p chi, drho = dp / c^2) on an n^d
    periodic mesh, from the last two steps"""
    cfg = make_cfg(n=(n, n if dims >= 2 else 2, n if dims == 3 else 2), dims=dims,
                   hi=(2 * np.pi,) * 3, recon="sp_weno", flux="kepes_tecno4_roe", bc="periodic",
                   cfl=cfl, fix_zmom=1)
    s = orc.OracleSolver(cfg)
    if dims == 1:
        chi = (-1.0) ** np.arange(s.ncells)
    else:
        nz = s.ncells // (n * n)
        i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(nz), indexing="ij")
        chi = ((-1.0) ** (i + j + k)).transpose(2, 1, 0).reshape(-1)
    p0 = 100 / 1.4
    base = np.zeros((s.nvars, s.ncells))
    base[0], base[-1] = 1.0, p0
    P = base.copy()
    P[-1] += 1e-8 * p0 * chi
    P[0] += 1e-8 * p0 * chi / 100
    s.set_state(P)
    amp = []
    for _ in range(steps):
        s.step()
        d = s.get_state() - base
        amp.append(np.sqrt((d * d).sum()))
    return amp[-1] / amp[-2]


@pytest.mark.parametrize("dims,cfl", [(1, 0.6), (1, 0.65), (2, 0.3), (2, 0.35), (3, 0.2),
                                      (3, 0.25), (3, 0.3)])
def test_checkerboard_gain_matches_closed_form(dims, cfl):
    """Measured growth equals |g(-4 cfl d)| to 1e-3: stable below
    cfl = 0.628 / d, unstable above (3D: gain 0.82 at 0.2, 2.0 at 0.25;
    CFL 0.5 is outside the bound in 2D and 3D)."""
    g = abs(ssprk3_gain(-4 * cfl * dims))
    got = checkerboard_gain(cfl, dims)
    assert abs(got - g) <= 1e-3 * max(1.0, g), (got, g)
    assert (got > 1) == (-4 * cfl * dims < Z_STAB)
