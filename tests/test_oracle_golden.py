"""Pin the oracle (oracle/liboracle.so) against the reference's own outputs.

The golden vectors in tests/golden/ were produced by the reference solvers
built from their sources (tests/golden/make_fixtures.py).  The oracle follows
the reference's operation order, so on the same host it must agree bit for
bit; where a test runs the oracle in its reduced-dimension mode (no sweeps
along collapsed axes) the tolerance is stated beside the assertion.
"""
import math

import numpy as np
import pytest

import oracle_ctypes as orc
from cases import case, euler_cfg, scalar_cfg, collapse, fixtures
from tecno import make_cfg

NET = None


def net():
    global NET
    if NET is None:
        NET = orc.load_net()
    return NET


# ---------------------------------------------------------------- point level
@pytest.mark.parametrize("tag,rec", [("spweno", "sp_weno"), ("spwenoc", "sp_wenoc"),
                                     ("spwenocorr2", "sp_weno_corr2"), ("dsp", "sp_weno_dl"),
                                     ("first", "first"), ("second", "second"),
                                     ("minmod", "minmod"), ("minabs", "minabs")])
def test_reconstruct_matches_reference(tag, rec):
    st = case("point", "recon_stencils")["st"]
    ref = case("point", f"recon_{tag}")
    cfg = make_cfg(system="scalar", flux="lin_tecno4", recon=rec, n=(8, 2, 2))
    l, r = orc.reconstruct(cfg, st, net())
    np.testing.assert_array_equal(l, ref["l"])
    np.testing.assert_array_equal(r, ref["r"])


@pytest.mark.parametrize("k", [2, 3])
def test_eno_matches_reference(k):
    ref = case("point", f"recon_eno{k}")
    cfg = make_cfg(system="scalar", flux="lin_tecno4", recon=f"eno {k}", n=(8, 2, 2))
    l, r = orc.reconstruct(cfg, ref["st"])
    np.testing.assert_array_equal(l, ref["l"])
    np.testing.assert_array_equal(r, ref["r"])


@pytest.mark.parametrize("flux", ["lin_ec", "lin_ec4", "lin_tecno", "lin_tecno4", "bur_ec",
                                  "bur_ec4", "bur_tecno", "bur_tecno4"])
def test_scalar_flux_matches_reference(flux):
    st = case("point", "sflux_stencils")["st"]
    F = case("point", f"sflux_{flux}")["F"]
    cfg = make_cfg(system="scalar", flux=flux, recon="sp_weno", n=(8, 6, 4), lo=(-1, -2, 0),
                   hi=(1, 1, 0.5), vel=(1.25, -0.5, 0.75))
    h = [2 / 8, 3 / 6, 0.5 / 4]
    for ax in range(3):
        np.testing.assert_array_equal(orc.flux(cfg, st, None, ax, h)[:, 0], F[ax])


@pytest.mark.parametrize("flux", ["lin_tecno4", "bur_tecno"])
@pytest.mark.parametrize("rec", ["second", "minmod", "minabs"])
def test_scalar_flux_with_muscl_matches_reference(flux, rec):
    st = case("point", "sflux_stencils")["st"]
    F = case("point", f"sfluxrec_{flux}_{rec}")["F"]
    cfg = make_cfg(system="scalar", flux=flux, recon=rec, n=(8, 6, 4), lo=(-1, -2, 0),
                   hi=(1, 1, 0.5), vel=(1.25, -0.5, 0.75))
    h = [2 / 8, 3 / 6, 0.5 / 4]
    for ax in range(3):
        np.testing.assert_array_equal(orc.flux(cfg, st, None, ax, h)[:, 0], F[ax])


def test_limited_slopes_known_values():
    """minmod / minabs on hand-worked pairs (limiter.h:20-52): equal signs
    take the smaller magnitude, opposite signs give 0 for minmod and the
    smaller-magnitude argument for minabs, ties keep the first argument."""
    cases = [(1.0, 3.0, 0.5, 0.5), (-4.0, -2.0, -1.0, -1.0), (2.0, -1.0, 0.0, -0.5),
             (-1.0, 0.5, 0.0, 0.25), (0.0, 5.0, 0.0, 0.0), (2.0, -2.0, 0.0, 1.0)]
    for a, b, mm, ma in cases:
        assert orc.limited_slope(a, b, False) == mm
        assert orc.limited_slope(a, b, True) == ma


def test_prim_to_ent_matches_reference():
    cfg = make_cfg()
    for tag in ("2d", "3d"):
        d = case("point", f"ent_{tag}")
        V = orc.prim_to_ent(cfg, d["P"].reshape(-1, 5))
        np.testing.assert_array_equal(V, d["V"].reshape(-1, 5))


def _eflux_keys():
    return sorted({k.split("/")[0] for k in fixtures("point").files if k.startswith("eflux_")})


@pytest.mark.parametrize("key", _eflux_keys())
def test_euler_flux_matches_reference(key):
    parts = key.split("_")
    dim = parts[1]
    ec1 = key.endswith("_ec1")
    rtag = parts[-2] if ec1 else parts[-1]
    flux = "_".join(parts[2:-2] if ec1 else parts[2:-1])
    rec = {"spweno": "sp_weno", "spwenoc": "sp_wenoc", "dsp": "sp_weno_dl", "first": "first",
           "second": "second", "minmod": "minmod", "minabs": "minabs"}[rtag]
    d = case("point", f"ent_{dim}")
    F = case("point", key)["F"]
    cfg = make_cfg(flux=flux, recon=rec, ec1_fix=int(ec1))
    h = [1 / 8, 2 / 6, 0.5 / 4]
    for ax in range(3):
        out = orc.flux(cfg, d["P"], d["V"], ax, h, net())
        np.testing.assert_array_equal(out, F[ax], err_msg=f"axis {ax}")


# ---------------------------------------------------------------- logavg
def test_logavg_closed_form_and_series():
    """logavg(a,b) = (b-a)/(ln b - ln a) (closed form, checked in extended
    precision away from a=b) and the a=b special case."""
    rng = np.random.default_rng(3)
    a = rng.uniform(0.1, 10, 2000)
    b = a * np.exp(rng.uniform(-3, 3, 2000))
    got = orc.logavg(a, b)
    la, lb = np.log(np.longdouble(a)), np.log(np.longdouble(b))
    exact = (np.longdouble(b) - np.longdouble(a)) / (lb - la)
    far = np.abs(np.log(b / a)) > 0.3
    np.testing.assert_allclose(got[far], exact[far].astype(np.float64), rtol=1e-14)
    # series branch (u < 1e-2) — truncation error O(u^4) relative
    near = ~far
    np.testing.assert_allclose(got[near], exact[near].astype(np.float64), rtol=3e-9)
    np.testing.assert_array_equal(orc.logavg(a, a), a)
    # symmetry
    np.testing.assert_allclose(orc.logavg(b, a), got, rtol=1e-15)


# ---------------------------------------------------------------- whole runs
SCALAR_KEYS = sorted({k.split("/")[0] for k in fixtures("scalar").files})
EULER_KEYS = sorted({k.split("/")[0] for k in fixtures("euler").files})


def _run_oracle(cfg, final, stamps, init, steps, faithful, netp=None):
    s = orc.OracleSolver(cfg, netp, faithful=faithful)
    s.set_state(init)
    out, dts = {}, []
    done = 0
    for k in steps:
        s.advance(final, stamps, max_iter=int(k))
        out[int(k)] = s.get_state()
    return s, out


@pytest.mark.parametrize("key", SCALAR_KEYS)
def test_scalar_runs_match_reference_faithful(key):
    d = case("scalar", key)
    cfg, final, stamps = scalar_cfg(key)
    _, out = _run_oracle(cfg, final, stamps, d["init"], d["steps"], True, net())
    for i, k in enumerate(d["steps"]):
        got = out[int(k)].reshape(d["states"][i].shape)
        if key.startswith("ns3d"):
            # the viscous node fluxes are summed into each cell in another
            # order than the reference's boundary-split node loops
            np.testing.assert_allclose(got, d["states"][i], rtol=0,
                                       atol=1e-13 * np.abs(d["states"][i]).max(), err_msg=f"step {k}")
        else:
            np.testing.assert_array_equal(got, d["states"][i], err_msg=f"step {k}")


@pytest.mark.parametrize("key", SCALAR_KEYS)
def test_scalar_runs_reduced_dims(key):
    """Collapsed axes are not swept in reduced mode.  Linear advection along a
    collapsed axis has zero flux, so those runs agree bit for bit; Burgers
    fluxes along a collapsed axis cancel only to rounding, so the bound there
    is 1e-13 relative to the solution scale."""
    d = case("scalar", key)
    cfg, final, stamps = scalar_cfg(key)
    dims = cfg.dims
    init = collapse(d["init"], dims)
    _, out = _run_oracle(cfg, final, stamps, init, d["steps"], False, net())
    for i, k in enumerate(d["steps"]):
        ref = collapse(d["states"][i], dims).reshape(1, -1)
        if "bur" in key:
            np.testing.assert_allclose(out[int(k)], ref, rtol=0, atol=1e-13 * np.abs(ref).max())
        else:
            np.testing.assert_array_equal(out[int(k)], ref, err_msg=f"step {k}")


@pytest.mark.parametrize("key", EULER_KEYS)
def test_euler_runs_match_reference_faithful(key):
    d = case("euler", key)
    cfg, final, stamps = euler_cfg(key)
    _, out = _run_oracle(cfg, final, stamps, d["init"], d["steps"], True, net())
    for i, k in enumerate(d["steps"]):
        got = out[int(k)].reshape(d["states"][i].shape)
        if key.startswith("ns3d"):
            # the viscous node fluxes are summed into each cell in another
            # order than the reference's boundary-split node loops
            np.testing.assert_allclose(got, d["states"][i], rtol=0,
                                       atol=1e-13 * np.abs(d["states"][i]).max(), err_msg=f"step {k}")
        else:
            np.testing.assert_array_equal(got, d["states"][i], err_msg=f"step {k}")


@pytest.mark.parametrize("key", [k for k in EULER_KEYS if k.startswith("eu2d")])
def test_euler_2d_reduced_dims_bitwise(key):
    """A "2D" reference run is Nz = 2 with z faces.  With w = 0 and z-uniform
    data every z-face flux is exactly zero except the z-momentum pressure term,
    which cancels exactly, so a grid without the z sweep is bit-identical."""
    d = case("euler", key)
    cfg, final, stamps = euler_cfg(key)
    init = collapse(d["init"], 2)
    _, out = _run_oracle(cfg, final, stamps, init, d["steps"], False, net())
    for i, k in enumerate(d["steps"]):
        ref = collapse(d["states"][i], 2).reshape(5, -1)
        np.testing.assert_array_equal(out[int(k)], ref, err_msg=f"step {k}")


def test_linadv_sin4_l1_error():
    """Config 1 (linadv sin^4 x, lin_tecno4 + sp_weno, Nx=400, t=0.5): the
    reference run's own L1 error 1.75268e-05 (65 steps)."""
    d = case("scalar", "linadv400_spweno")
    cfg, final, stamps = scalar_cfg("linadv400_spweno")
    s = orc.OracleSolver(cfg, None, faithful=False)
    s.set_state(collapse(d["init"], 1))
    it, t = s.advance(final, stamps)
    assert it == 65 and t == 0.5
    x = -math.pi + (np.arange(400) + 0.5) * 2 * math.pi / 400
    l1 = np.sum(np.abs(s.get_state()[0] - np.sin(x - t) ** 4)) * 2 * math.pi / 400
    assert abs(l1 - 1.75268e-05) < 5e-11


def test_sign_property_of_sp_weno():
    """SP-WENO's defining invariant: the reconstructed jump u+ - u- has the
    sign of the cell jump u2 - u1 (or vanishes)."""
    st = case("point", "recon_stencils")["st"]
    for corr in (0, 1):
        l, r = orc.sp_weno(st, corr)
        jump = st[:, 2] - st[:, 1]
        rj = r - l
        ok = (jump * rj >= 0) | (np.abs(rj) < 1e-12 * (1 + np.abs(st).max(axis=1)))
        assert ok.mean() > 0.999


def test_dsp_softmax_vertices_sign_property():
    """DSP-WENO weights are a convex combination of feasible vertices, so the
    sign property holds for any network output (checked on the shipped net and
    on a random one)."""
    st = case("point", "recon_stencils")["st"]
    st = st[np.abs(st[:, 2] - st[:, 1]) > 1e-9]
    rng = np.random.default_rng(5)
    w = [5, 5, 5, 5, 5]
    rnd = orc.net_from_arrays(w, rng.normal(size=100), rng.normal(size=20))
    for n in (net(), rnd):
        l, r = orc.nn_forward(n, st)
        jump = st[:, 2] - st[:, 1]
        rj = r - l
        ok = (jump * rj >= 0) | (np.abs(rj) < 1e-10 * (1 + np.abs(st).max(axis=1)))
        assert ok.mean() > 0.999
