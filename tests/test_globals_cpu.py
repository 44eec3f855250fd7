"""The oracle's audit sums (tor_audit) against the reference's own
compute_global (enflo/src/compute_global.cc:10-89; enflo_scalar
compute_global.cc:10-38) evaluated by the reference build at the first stage
of every step (tests/golden/fixtures_globals.npz, written by
tests/golden/make_globals_fixtures.py): total entropy and kinetic energy to
1e-13 relative, on the reference's Nz = 2 storage (faithful) and on the
collapsed 2D grid, whose one z layer stands for the reference's two."""
import numpy as np
import pytest

import oracle_ctypes as orc
from cases import GOLDEN, case, collapse, euler_cfg, scalar_cfg

import os

GLOBALS = np.load(os.path.join(GOLDEN, "fixtures_globals.npz"))
KEYS = sorted({k.split("/")[0] for k in GLOBALS.files})


def gcase(key):
    return {k.split("/", 1)[1]: GLOBALS[k] for k in GLOBALS.files if k.split("/", 1)[0] == key}


@pytest.mark.parametrize("key", KEYS)
@pytest.mark.parametrize("faithful", [True, False])
def test_oracle_audit_matches_reference_globals(key, faithful):
    g = gcase(key)
    rows = g["rows"]
    if key.startswith("sc"):
        d = case("scalar", key)
        cfg, final, stamps = scalar_cfg(key)
        init = np.asarray(d["init"])
        st = init.reshape(-1) if faithful else collapse(init, cfg.dims).reshape(-1)
        net = None
    else:
        d = case("euler", key)
        cfg, final, stamps = euler_cfg(key)
        init = np.asarray(d["init"])
        st = init.reshape(5, -1) if (faithful or cfg.dims == 3) else collapse(init, 2).reshape(5, -1)
        net = orc.load_net() if key.endswith("dsp") else None
    o = orc.OracleSolver(cfg, net, faithful=faithful)
    o.set_state(st)
    for k in range(rows.shape[0]):
        a = o.audit()
        assert abs(a[0] - rows[k, 1]) <= 1e-13 * abs(rows[k, 1]), (k, a[0], rows[k, 1])
        if not key.startswith("sc"):
            assert abs(a[1] - rows[k, 2]) <= 1e-13 * abs(rows[k, 2]), (k, a[1], rows[k, 2])
        o.advance(final, stamps, max_iter=k + 1)


def test_globals_fixture_shapes():
    """every systems case carries 6 columns per step and the vorticity of all
    (Nx+1)(Ny+1)(Nz+1) nodes; KE_diss_stress is non-zero exactly for the
    Navier-Stokes cases"""
    for key in KEYS:
        g = gcase(key)
        if key.startswith("sc"):
            assert g["rows"].shape[1] == 2
            continue
        nx, ny, nz = (int(v) for v in g["shape"])
        assert g["rows"].shape[1] == 6
        assert g["vort0"].shape == (3, nz + 1, ny + 1, nx + 1)
        assert (np.abs(g["rows"][:, 4]) > 0).all() == key.startswith("ns3d")
