"""GPU diagnostics against the reference's own compute_global and
find_vorticity (tests/golden/fixtures_globals.npz, written by the reference
build through tests/golden/make_globals_fixtures.py), and the FP64 audit sums
against the oracle (tor_audit, itself pinned to the same fixtures by
test_globals_cpu.py).

Bars: total entropy, kinetic energy and enstrophy 1e-12 relative per step;
KE_diss_pr and KE_diss_stress (sums with cancellation between cells) 1e-11
of the largest term's scale, i.e. of the sum of the absolute values, bounded
here by 1e-10 relative to the value's magnitude over the run; vorticity 1e-12
relative to its largest magnitude.  Sources: compute_global.cc:10-89,
fv.cc:756-764 (KE_diss_pr), fv.cc:792-1156 (KE_diss_stress), find_vort.cc:9-29
with nodal_avg_and_der.cc:16-82 and the ghost fill of fv.cc:121-361."""
import os

import numpy as np
import pytest

import oracle_ctypes as orc
from cases import GOLDEN, case, collapse, euler_cfg, euler_prim_for, scalar_cfg
from tecno import Solver

pytestmark = pytest.mark.gpu

GLOBALS = np.load(os.path.join(GOLDEN, "fixtures_globals.npz"))
KEYS = sorted({k.split("/")[0] for k in GLOBALS.files})


def gcase(key):
    return {k.split("/", 1)[1]: GLOBALS[k] for k in GLOBALS.files if k.split("/", 1)[0] == key}


def _solver_for(key):
    if key.startswith("sc"):
        d = case("scalar", key)
        cfg, final, stamps = scalar_cfg(key)
        s = Solver(cfg)
        s.set_state(collapse(d["init"], cfg.dims).reshape(-1))
        return s, cfg, final, stamps
    d = case("euler", key)
    cfg, final, stamps = euler_cfg(key)
    s = Solver(cfg)
    if key.endswith("dsp"):
        s.load_net(orc.NET_FILE)
    init = np.asarray(d["init"])
    s.set_state(init.reshape(5, -1) if cfg.dims == 3 else euler_prim_for(init, 2))
    return s, cfg, final, stamps


@pytest.mark.parametrize("key", KEYS)
def test_globals_match_reference(key):
    g = gcase(key)
    rows = g["rows"]
    s, cfg, final, stamps = _solver_for(key)
    scalar = key.startswith("sc")
    # scale of the dissipation rates over the run (for their absolute bars)
    kdp_scale = np.abs(rows[:, 3]).max() if not scalar else 0.0
    kds_scale = np.abs(rows[:, 4]).max() if not scalar else 0.0
    for k in range(rows.shape[0]):
        gl = s.globals()
        assert abs(gl.time - rows[k, 0]) <= 1e-14 * max(1.0, abs(rows[k, 0]))
        assert abs(gl.total_entropy - rows[k, 1]) <= 1e-12 * abs(rows[k, 1]), (k, gl.total_entropy)
        if not scalar:
            assert abs(gl.kinetic_energy - rows[k, 2]) <= 1e-12 * abs(rows[k, 2])
            assert abs(gl.ke_diss_pr - rows[k, 3]) <= 1e-10 * kdp_scale, (k, gl.ke_diss_pr, rows[k, 3])
            assert abs(gl.ke_diss_stress - rows[k, 4]) <= 1e-10 * max(kds_scale, 1e-300) or \
                (kds_scale == 0.0 and gl.ke_diss_stress == 0.0), (k, gl.ke_diss_stress, rows[k, 4])
            assert abs(gl.enstrophy - rows[k, 5]) <= 1e-12 * abs(rows[k, 5]), (k, gl.enstrophy, rows[k, 5])
        s.advance(final, stamps, max_iter=k + 1)


@pytest.mark.parametrize("key", [k for k in KEYS if not k.startswith("sc")])
def test_vorticity_matches_reference(key):
    """nodal vorticity before the first and after the last step, every node
    plane of the reference's storage (the boundary planes of non-periodic 2D
    grids take the reference's periodic edge ghosts)"""
    g = gcase(key)
    s, cfg, final, stamps = _solver_for(key)
    ref0, ref1 = g["vort0"], g["vort1"]
    scale = np.abs(ref0).max()
    w0 = s.vorticity()
    assert w0.shape == ref0.shape
    assert np.abs(w0 - ref0).max() <= 1e-12 * scale, np.abs(w0 - ref0).max() / scale
    s.advance(final, stamps, max_iter=g["rows"].shape[0])
    w1 = s.vorticity()
    assert np.abs(w1 - ref1).max() <= 1e-12 * np.abs(ref1).max()


@pytest.mark.parametrize("key", KEYS)
def test_audit_matches_oracle(key):
    """tecno_audit_state against tor_audit on the same state, 3 steps: total
    entropy, kinetic energy and the conserved sums to 1e-12 relative"""
    s, cfg, final, stamps = _solver_for(key)
    scalar = key.startswith("sc")
    if scalar:
        d = case("scalar", key)
        st = collapse(d["init"], cfg.dims).reshape(-1)
        o = orc.OracleSolver(cfg, None, faithful=False)
        o.set_state(st)
    else:
        d = case("euler", key)
        init = np.asarray(d["init"])
        o = orc.OracleSolver(cfg, orc.load_net() if key.endswith("dsp") else None, faithful=False)
        o.set_state(init.reshape(5, -1) if cfg.dims == 3 else collapse(init, 2).reshape(5, -1))
    for k in range(3):
        a, r = s.audit(), o.audit()
        got = [a.total_entropy, a.kinetic_energy, a.mass, a.momentum[0], a.momentum[1],
               a.momentum[2], a.energy]
        for i, (x, y) in enumerate(zip(got, r[:7])):
            if scalar and i not in (0, 2):
                continue
            scale = max(abs(y), abs(r[2]) if 3 <= i <= 5 else 0.0)
            assert abs(x - y) <= 1e-12 * scale, (k, i, x, y)
        s.advance(final, stamps, max_iter=k + 1)
        o.advance(final, stamps, max_iter=k + 1)
