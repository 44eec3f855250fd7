"""Navier-Stokes with wall and neumann boundaries (which the reference
rejects, parameter.cc:564-568) in the oracle, pinned by invariants:

* a wall is a mirror plane: the sub-box [0, pi]^3 with walls reproduces the
  periodic [0, 2 pi]^3 run of the mirrored state;
* a neumann boundary of a state uniform along its axis changes nothing:
  the run equals the periodic one;
both with every node counted once (fix_edges; the reference's periodic loops
add the nodes of periodic edges twice, fv.cc:821-1156, which the default
keeps)."""
import numpy as np

import oracle_ctypes as orc
from cases import mirror_full, ns_box_cfg, tgv_sub_box
from tecno import TS_RK1


def _run(cfg, state, steps):
    o = orc.OracleSolver(cfg, None, faithful=True)
    assert o.ncells == state[0].size
    o.set_state(state.reshape(5, -1))
    dts = [o.step() for _ in range(steps)]
    return o.get_state().reshape(state.shape), dts


def _rel(a, b):
    """max difference per variable over that variable's magnitude (at least
    1e-12: w of a z-uniform state stays 0)"""
    scale = np.maximum(np.abs(b).max(axis=(1, 2, 3), keepdims=True), 1e-12)
    return float((np.abs(a - b) / scale).max())


def test_walls_are_mirror_planes():
    m = 4
    sub = tgv_sub_box(m, m, np.random.default_rng(2))
    full = mirror_full(sub, (0, 1, 2))
    a, da = _run(ns_box_cfg((m, m, m), (np.pi,) * 3, "wall", fix_edges=1), sub, 3)
    b, db = _run(ns_box_cfg((2 * m,) * 3, (2 * np.pi,) * 3, "periodic", fix_edges=1), full, 3)
    assert np.allclose(da, db, rtol=1e-14, atol=0)
    # the periodic run stays mirror symmetric, and its octant is the wall run
    assert _rel(b, mirror_full(b[:, :m, :m, :m], (0, 1, 2))) <= 1e-13
    assert _rel(a, b[:, :m, :m, :m]) <= 1e-13


def test_wall_and_neumann_edges():
    """walls in x and y, neumann in z, on a z-uniform state: the edges where a
    wall meets a neumann boundary read ghost cells composed of both maps"""
    m, kz = 4, 3
    sub = tgv_sub_box(m, kz, np.random.default_rng(4), z_uniform=True)
    full = mirror_full(sub, (0, 1))
    bc = ["wall", "wall", "wall", "wall", "neumann", "neumann"]
    a, _ = _run(ns_box_cfg((m, m, kz), (np.pi, np.pi, kz * np.pi / m), bc, fix_edges=1), sub, 3)
    b, _ = _run(ns_box_cfg((2 * m, 2 * m, kz), (2 * np.pi, 2 * np.pi, kz * np.pi / m), "periodic",
                           fix_edges=1), full, 3)
    assert _rel(a, b[:, :, :m, :m]) <= 1e-13
    assert np.abs(a - a[:, :1]).max() == 0.0  # stays z-uniform


def test_periodic_edges_counted_twice_by_default():
    """fix_edges = 0 is the reference's arithmetic (the golden 3D NS runs hold
    it); on a state with viscous stress at the domain edges one forward-Euler
    stage differs from fix_edges = 1 only in the cells along the edges."""
    m = 4
    st = mirror_full(tgv_sub_box(m, m, np.random.default_rng(5)), (0, 1, 2))
    a, _ = _run(ns_box_cfg((2 * m,) * 3, (2 * np.pi,) * 3, "periodic", fix_edges=0,
                           time_scheme=TS_RK1), st, 1)
    b, _ = _run(ns_box_cfg((2 * m,) * 3, (2 * np.pi,) * 3, "periodic", fix_edges=1,
                           time_scheme=TS_RK1), st, 1)
    diff = np.abs(a - b).max(axis=0)
    n = 2 * m
    edge = np.zeros((n, n, n), bool)
    e = [0, n - 1]
    for k in range(n):
        for j in range(n):
            for i in range(n):
                edge[k, j, i] = ((i in e) + (j in e) + (k in e)) >= 2
    assert diff[~edge].max() == 0.0
    assert diff[edge].max() > 0.0
