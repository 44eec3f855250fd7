"""Batched ensembles of scalar problems: every member ends bitwise (FP64 and
FP32 arithmetic on an FP64 state) or to FP32 rounding (FP32 state) where the
same problem run alone ends, with and without the captured hipGraph, for
every tile shape the autotuner may pick."""
import numpy as np
import pytest

from tecno import FP32, FP32_STATE64, FP64, Ensemble, Solver, make_cfg

pytestmark = pytest.mark.gpu


def problems():
    """mixed 1D/2D, Burgers/linear advection, FP64/FP32, uneven sizes"""
    P = []
    for n in (100, 173):
        P.append((make_cfg(system="scalar", flux="bur_tecno4", recon="sp_weno", n=(n, 1, 1), dims=1,
                           lo=(-np.pi, 0, 0), hi=(np.pi, 1, 1), bc="periodic"), 0.6))
    P.append((make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno", n=(64, 64, 1), dims=2,
                       lo=(-1, -1, 0), hi=(1, 1, 1), vel=(1, 1, 0), bc="periodic"), 0.3))
    P.append((make_cfg(system="scalar", flux="bur_tecno4", recon="sp_weno", n=(50, 37, 1), dims=2,
                       lo=(-1, -1, 0), hi=(1, 1, 1), bc="periodic"), 0.2))
    P.append((make_cfg(system="scalar", flux="lin_tecno4", recon="sp_wenoc", n=(48, 40, 1), dims=2,
                       lo=(-1, -1, 0), hi=(1, 1, 1), vel=(1, 1, 0), bc="periodic",
                       precision=FP32), 0.25))
    P.append((make_cfg(system="scalar", flux="bur_tecno4", recon="sp_weno", n=(211, 1, 1), dims=1,
                       lo=(-np.pi, 0, 0), hi=(np.pi, 1, 1), bc="periodic", precision=FP32), 0.5))
    P.append((make_cfg(system="scalar", flux="lin_tecno4", recon="sp_wenoc", n=(56, 44, 1), dims=2,
                       lo=(-1, -1, 0), hi=(1, 1, 1), vel=(1, 1, 0), bc="periodic",
                       precision=FP32_STATE64), 0.25))
    P.append((make_cfg(system="scalar", flux="bur_tecno4", recon="sp_weno", n=(157, 1, 1), dims=1,
                       lo=(-np.pi, 0, 0), hi=(np.pi, 1, 1), bc="periodic",
                       precision=FP32_STATE64), 0.5))
    return P


def initial(cfg):
    n0, n1 = cfg.n[0], cfg.n[1]
    x = cfg.lo[0] + (np.arange(n0) + 0.5) * (cfg.hi[0] - cfg.lo[0]) / n0
    if cfg.dims == 1:
        return np.sin(x)
    y = cfg.lo[1] + (np.arange(n1) + 0.5) * (cfg.hi[1] - cfg.lo[1]) / n1
    X, Y = np.meshgrid(x, y)
    return (np.sin(np.pi * X) * np.sin(np.pi * Y) + 0.05 * np.cos(3 * np.pi * X + 2 * np.pi * Y)).ravel()


def alone(cfg, t):
    s = Solver(cfg)
    s.set_state(initial(cfg))
    s.advance(t, 2)
    return s.get_state(), s.time()


@pytest.mark.parametrize("use_graph", [True, False])
def test_ensemble_members_equal_single_runs(use_graph):
    P = problems()
    ref = [alone(c, t) for c, t in P]
    ens = Ensemble([c for c, _ in P])
    for m, (c, _) in zip(ens.members, P):
        m.set_state(initial(c))
    cu = ens.advance([t for _, t in P], t_stamps=2, use_graph=use_graph)
    assert cu > 0
    for i, (m, (c, t)) in enumerate(zip(ens.members, P)):
        st, (t_ref, it_ref) = ref[i]
        got = m.get_state()
        tm, itm = m.time()
        assert itm == it_ref and tm == t_ref
        if c.precision in (FP64, FP32_STATE64):
            assert np.array_equal(got, st), i
        else:
            assert np.abs(got - st).max() <= 1e-6 * np.abs(st).max(), i


def test_autotune_picks_a_candidate_and_keeps_results():
    P = problems()
    ref = [alone(c, t) for c, t in P]
    ens = Ensemble([c for c, _ in P])
    for m, (c, _) in zip(ens.members, P):
        m.set_state(initial(c))
    ens.autotune(steps=2)
    groups = ens.groups()
    assert len(groups) == 6
    for g in groups:
        assert g["tile"] in g["tuned_ms"]
        assert all(ms > 0 for ms in g["tuned_ms"].values())
        assert g["tuned_ms"][g["tile"]] == min(g["tuned_ms"].values())
    ens.advance([t for _, t in P], t_stamps=2)
    for i, (m, (c, t)) in enumerate(zip(ens.members, P)):
        st, _ = ref[i]
        if c.precision in (FP64, FP32_STATE64):
            assert np.array_equal(m.get_state(), st), i
