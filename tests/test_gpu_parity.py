"""GPU parity: libtecno.so (HIP, gfx950) against the reference's golden
vectors and the oracle, through the C ABI.

Bars (written beside each assertion):
  * FP64 path: the reference's arithmetic in the reference's order, compiled
    without FMA contraction; results agree to 1e-12 relative per step (the
    remaining differences are the last-ulp differences of log/exp/tanh between
    the device math library and glibc).
  * FP32 path: stated per test against the FP64 oracle.
"""
import math

import numpy as np
import pytest

import oracle_ctypes as orc
from cases import (case, collapse, euler_cfg, euler_prim_for, fixtures, mirror_full, ns_box_cfg,
                   scalar_cfg, tgv_sub_box)
from tecno import (FP32, FP64, Solver, make_cfg, MLP_BF16, MLP_FP16, MLP_FP32, MLP_FP64,
                   TIME_SCHEME)

pytestmark = pytest.mark.gpu

NET_FILE = orc.NET_FILE


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)
    return float((np.abs(a - b) / scale).max())


def solver_for(cfg, dsp):
    s = Solver(cfg)
    if dsp:
        s.load_net(NET_FILE)
    return s


# ---------------------------------------------------------------- golden runs
SCALAR_KEYS = sorted({k.split("/")[0] for k in fixtures("scalar").files
                      if not k.startswith("sc3d")})
EULER_KEYS = sorted({k.split("/")[0] for k in fixtures("euler").files if k.startswith("eu2d")})


@pytest.mark.parametrize("key", SCALAR_KEYS)
def test_scalar_golden_fp64(key):
    d = case("scalar", key)
    cfg, final, stamps = scalar_cfg(key)
    s = solver_for(cfg, key.endswith("dsp"))
    s.set_state(collapse(d["init"], cfg.dims).reshape(-1))
    done = 0
    for i, k in enumerate(d["steps"]):
        s.advance(final, stamps, max_iter=int(k))
        t, it = s.time()
        assert it == int(k)
        assert abs(t - d["times"][int(k) - 1]) <= 1e-15 * max(1.0, abs(t))
        ref = collapse(d["states"][i], cfg.dims).reshape(1, -1)
        got = s.get_state()
        # 1e-12 relative per step
        tol = 1e-12 * int(k)
        if "bur" in key:
            tol = max(tol, 1e-13)  # collapsed-axis Burgers fluxes cancel only to rounding
        assert rel_err(got, ref) <= tol, (key, k, rel_err(got, ref))


@pytest.mark.parametrize("key", EULER_KEYS)
def test_euler2d_golden_fp64(key):
    d = case("euler", key)
    cfg, final, stamps = euler_cfg(key)
    s = solver_for(cfg, key.endswith("dsp"))
    s.set_state(euler_prim_for(d["init"], 2))
    for i, k in enumerate(d["steps"]):
        s.advance(final, stamps, max_iter=int(k))
        ref = euler_prim_for(d["states"][i], 2)
        got = s.get_state()
        # 1e-12 relative per step (per variable, scaled by its max magnitude)
        assert rel_err(got, ref) <= 1e-12 * int(k), (key, k, rel_err(got, ref))


EULER3_KEYS = sorted({k.split("/")[0] for k in fixtures("euler").files
                      if k.startswith("eu3d") or k.startswith("ns3d")})


@pytest.mark.parametrize("key", EULER3_KEYS)
def test_euler3d_ns_golden_fp64(key):
    """3D Euler and Navier-Stokes (ES / KE viscous fluxes, constant,
    Sutherland and power-law viscosity, sp_weno / sp_weno_dl) Taylor-Green
    runs of the reference: 1e-12 relative per step."""
    d = case("euler", key)
    cfg, final, stamps = euler_cfg(key)
    s = solver_for(cfg, key.endswith("dsp"))
    s.set_state(np.asarray(d["init"]).reshape(5, -1))
    for i, k in enumerate(d["steps"]):
        s.advance(final, stamps, max_iter=int(k))
        ref = np.asarray(d["states"][i]).reshape(5, -1)
        got = s.get_state()
        assert rel_err(got, ref) <= 1e-12 * int(k), (key, k, rel_err(got, ref))


def test_ns3d_per_step_vs_oracle_fix_zmom():
    """3D Navier-Stokes with the z-momentum dissipation fix (no reference run
    exists for it): GPU against the oracle, 3 steps, 1e-12 relative."""
    cfg, final, stamps = euler_cfg("ns3d_es_tg", fix_zmom=1)
    d = case("euler", "ns3d_es_tg")
    init = np.asarray(d["init"]).reshape(5, -1)
    o = orc.OracleSolver(cfg, None, faithful=True)
    o.set_state(init)
    s = Solver(cfg)
    s.set_state(init)
    for k in range(3):
        o.step(1e300)
        s.step(1e300)
        ref = o.get_state()
        assert rel_err(s.get_state(), ref) <= 1e-12 * (k + 1)


@pytest.mark.parametrize("key", ["ns3d_es_tg", "ns3d_es_suth"])
def test_ns3d_fix_visc_vs_oracle(key):
    """ns_es with the K21, K31, K32 viscous blocks kept (fix_visc; no
    reference run exists for it, the oracle's flux is pinned by the
    closed-form entropy production in test_visc_matrix_cpu.py): GPU against
    the oracle, 3 steps, 1e-12 relative, and the fix changes the state"""
    cfg, final, stamps = euler_cfg(key, fix_visc=1)
    d = case("euler", key)
    init = np.asarray(d["init"]).reshape(5, -1)
    o = orc.OracleSolver(cfg, None, faithful=True)
    o.set_state(init)
    s = Solver(cfg)
    s.set_state(init)
    for k in range(3):
        o.step(1e300)
        s.step(1e300)
        assert rel_err(s.get_state(), o.get_state()) <= 1e-12 * (k + 1)
    ref = np.asarray(d["states"][1]).reshape(5, -1) if int(d["steps"][1]) == 3 else None
    if ref is not None:
        assert rel_err(s.get_state(), ref) > 1e-10


@pytest.mark.parametrize("ts", ["rk1", "jameson_rk4", "rk4"])
@pytest.mark.parametrize("system", ["scalar", "euler"])
def test_time_schemes_per_step_vs_oracle(ts, system):
    """rk1, Jameson's RK4 and classical RK4 (update_solution, fv.cc:1160-1198
    and enflo_scalar fv.cc:459-499; the two solvers round RK4's last stage
    differently): 2D grids of 45 x 38, random states, every step within
    1e-12 relative of the oracle and the same dt."""
    rng = np.random.default_rng(19)
    nx, ny = 45, 38
    if system == "scalar":
        cfg = make_cfg(system="scalar", n=(nx, ny, 2), dims=2, lo=(-1, -1, 0), hi=(1, 1, 1),
                       flux="lin_tecno4", recon="sp_weno", vel=(0.8, -0.6, 0.0), bc="periodic",
                       time_scheme=TIME_SCHEME[ts])
        st = rng.normal(size=nx * ny)
        full = st
    else:
        cfg = make_cfg(n=(nx, ny, 2), dims=2, recon="sp_weno", bc="periodic",
                       time_scheme=TIME_SCHEME[ts])
        st = _random_euler_state(rng, nx, ny, False)
        full = np.zeros((5, nx * ny))
        full[[0, 1, 2, 4]] = st
    o = orc.OracleSolver(cfg, None, faithful=False)
    o.set_state(full)
    s = Solver(cfg)
    s.set_state(st)
    for step in range(4):
        dto = o.step()
        dtg = s.step()
        assert abs(dtg - dto) <= 1e-14 * dto
        ref = o.get_state() if system == "scalar" else o.get_state()[[0, 1, 2, 4]]
        assert rel_err(s.get_state(), ref) <= 1e-12 * (step + 1), (step, rel_err(s.get_state(), ref))


@pytest.mark.parametrize("ts", ["jameson_rk4", "rk4"])
def test_time_schemes_mixed_path_close_to_fp64_oracle(ts):
    """The FP32-dissipation headline kernel under the four-stage schemes,
    smooth flow, against the FP64 oracle step by step: the first step within
    1e-8 relative everywhere (the stage coefficients and the accumulator);
    after 8 steps 99% of the cells within 1e-8 and all within 2e-5.  The
    FP32 limiter ratios cannot resolve the reference's |theta - 1| < 1e-13
    tests (FP32 spacing near 1 is 6e-8): where a smooth profile has theta
    within ~1e-7 of 1 the two paths take different SP-WENO branches at one
    face, a local O(1e-6) deviation (seen on a ~10-cell patch)."""
    nx, ny = 64, 48
    cfg = make_cfg(n=(nx, ny, 2), dims=2, recon="sp_weno", bc="periodic",
                   time_scheme=TIME_SCHEME[ts])
    prim = _random_euler_state(None, nx, ny, True)
    o = orc.OracleSolver(cfg, None, faithful=False)
    full = np.zeros((5, nx * ny))
    full[[0, 1, 2, 4]] = prim
    o.set_state(full)
    s = Solver(make_cfg(n=(nx, ny, 2), dims=2, recon="sp_weno", bc="periodic", precision=FP32,
                        sign_fallback=1, time_scheme=TIME_SCHEME[ts]))
    s.set_state(prim)
    for k in range(8):
        o.step()
        s.step()
        ref = o.get_state()[[0, 1, 2, 4]]
        e = (np.abs(s.get_state() - ref) / np.abs(ref).max(axis=1, keepdims=True)).max(axis=0)
        if k == 0:
            assert e.max() <= 1e-8, e.max()
    assert np.quantile(e, 0.99) <= 1e-8, np.quantile(e, 0.99)
    assert e.max() <= 2e-5, e.max()


@pytest.mark.parametrize("bc,fix_edges", [
    (["wall", "wall", "neumann", "neumann", "periodic", "periodic"], 0),
    (["neumann", "neumann", "periodic", "periodic", "wall", "wall"], 0),
    (["wall", "wall", "wall", "wall", "neumann", "neumann"], 0),
    ("periodic", 1)])
def test_ns3d_boundaries_per_step_vs_oracle(bc, fix_edges):
    """Navier-Stokes with wall / neumann boundaries (an extension: the
    reference allows periodic only) and with fix_edges: GPU against the
    oracle, whose extension the CPU tests pin by mirror and uniformity
    invariants; 3 steps, 1e-12 relative per step."""
    rng = np.random.default_rng(31)
    n = (10, 8, 6)
    cfg = ns_box_cfg(n, (2.0, 1.6, 1.2), bc, fix_edges=fix_edges)
    h = 0.2
    z, y, x = np.meshgrid(*[(np.arange(k) + 0.5) * h for k in n[::-1]], indexing="ij")
    st = np.stack([1.0 + 0.1 * np.sin(2 * x + y), np.sin(3 * y) * np.cos(2 * z), np.cos(2 * x) * np.sin(z),
                   0.5 * np.sin(x + z), 5.0 + 0.2 * np.cos(x - y)]) + rng.normal(size=(5,) + x.shape) * 0.01
    init = st.reshape(5, -1)
    o = orc.OracleSolver(cfg, None, faithful=True)
    o.set_state(init)
    s = Solver(cfg)
    s.set_state(init)
    for k in range(3):
        dto = o.step()
        dtg = s.step()
        assert abs(dtg - dto) <= 1e-14 * dto
        assert rel_err(s.get_state(), o.get_state()) <= 1e-12 * (k + 1), (k, rel_err(s.get_state(), o.get_state()))


def test_ns3d_walls_are_mirror_planes_gpu():
    """The GPU run on [0, pi]^3 with walls equals the octant of the periodic
    [0, 2 pi]^3 run of the mirrored state (every node counted once)."""
    m = 8
    sub = tgv_sub_box(m, m, np.random.default_rng(2))
    full = mirror_full(sub, (0, 1, 2))
    a = Solver(ns_box_cfg((m, m, m), (math.pi,) * 3, "wall", fix_edges=1))
    b = Solver(ns_box_cfg((2 * m,) * 3, (2 * math.pi,) * 3, "periodic", fix_edges=1))
    a.set_state(sub.reshape(5, -1))
    b.set_state(full.reshape(5, -1))
    for k in range(4):
        a.step()
        b.step()
        ga = a.get_state().reshape(sub.shape)
        gb = b.get_state().reshape(full.shape)[:, :m, :m, :m]
        assert rel_err(ga.reshape(5, -1), gb.reshape(5, -1)) <= 1e-12 * (k + 1)


@pytest.mark.parametrize("rec", ["sp_weno", "sp_weno_dl"])
def test_euler_1d_shock_tube_vs_oracle(rec):
    """The reference's 1D Euler examples run as Nx x 2 x 2 grids (problem_dim
    1): Lax's shock tube (rho, u, p = 0.445, 0.698, 3.528 | 0.5, 0, 0.571 on
    [-5, 5], Nx = 200, neumann in x, periodic in y), CFL 0.4 - 20 steps of
    the GPU FP64 path within 1e-12 relative per step of the oracle."""
    nx = 200
    cfg = make_cfg(n=(nx, 2, 2), dims=2, lo=(-5, 0, 0), hi=(5, 1, 1), recon=rec, cfl=0.4,
                   bc=["neumann", "neumann", "periodic", "periodic", "periodic", "periodic"])
    x = -5 + (np.arange(nx) + 0.5) * 10 / nx
    left = x < 0
    rho = np.where(left, 0.445, 0.5)
    u = np.where(left, 0.698, 0.0)
    p = np.where(left, 3.528, 0.571)
    prim = np.stack([np.tile(rho, 2), np.tile(u, 2), np.zeros(2 * nx), np.tile(p, 2)])
    net = orc.load_net() if rec == "sp_weno_dl" else None
    o = orc.OracleSolver(cfg, net, faithful=False)
    full = np.zeros((5, 2 * nx))
    full[[0, 1, 2, 4]] = prim
    o.set_state(full)
    s = solver_for(cfg, rec == "sp_weno_dl")
    s.set_state(prim)
    for k in range(20):
        dto = o.step()
        dtg = s.step()
        assert abs(dtg - dto) <= 1e-14 * dto
        ref = o.get_state()[[0, 1, 2, 4]]
        assert rel_err(s.get_state(), ref) <= 1e-12 * (k + 1), (k, rel_err(s.get_state(), ref))
    st = s.get_state()
    assert st[0].min() > 0 and st[3].min() > 0 and np.array_equal(st[:, :nx], st[:, nx:])


def test_linadv_sin4_l1_gpu():
    """Config 1: lin_tecno4 + sp_weno, Nx = 400, t = 0.5 -> L1 = 1.75268e-05
    (the reference's own figure), 65 steps."""
    d = case("scalar", "linadv400_spweno")
    cfg, final, stamps = scalar_cfg("linadv400_spweno")
    s = Solver(cfg)
    s.set_state(collapse(d["init"], 1).reshape(-1))
    it, t = s.advance(final, stamps)
    assert it == 65 and t == 0.5
    x = -math.pi + (np.arange(400) + 0.5) * 2 * math.pi / 400
    # the L1 of a collapsed run counts the reference's Ny*Nz = 4 duplicate
    # cells per held cell, as the reference's sum does (dx*dy*dz*4 = dx)
    l1 = s.l1_error(np.sin(x - t) ** 4)[0]
    assert abs(l1 - 1.75268e-05) < 5e-11, l1


# ---------------------------------------------------------------- per-step vs oracle
def _random_euler_state(rng, nx, ny, smooth):
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    if smooth:
        rho = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
        u = 0.5 * np.sin(2 * np.pi * Y)
        v = 0.3 * np.cos(2 * np.pi * X)
        p = 1.0 + 0.2 * np.cos(2 * np.pi * (X + Y))
    else:
        # cell-to-cell noise; stronger noise makes the reference itself (and the
        # oracle) produce negative pressure with sp_weno_dl
        rho = rng.uniform(0.7, 1.3, (ny, nx))
        u = rng.uniform(-0.3, 0.3, (ny, nx))
        v = rng.uniform(-0.3, 0.3, (ny, nx))
        p = rng.uniform(0.7, 1.3, (ny, nx))
    return np.stack([rho, u, v, p]).reshape(4, -1)


@pytest.mark.parametrize("rec", ["sp_weno", "sp_wenoc", "eno 2", "eno 3", "sp_weno_dl", "first",
                                 "second", "minmod", "minabs"])
@pytest.mark.parametrize("smooth", [True, False])
def test_euler2d_per_step_vs_oracle(rec, smooth):
    """Non-multiple-of-tile grid (70 x 37), periodic x / neumann y: every step
    of the GPU FP64 path is within 1e-12 relative of the oracle."""
    rng = np.random.default_rng(42)
    nx, ny = 70, 37
    cfg = make_cfg(n=(nx, ny, 2), dims=2, recon=rec, bc=["periodic", "periodic", "neumann", "neumann",
                                                           "periodic", "periodic"])
    prim = _random_euler_state(rng, nx, ny, smooth)
    net = orc.load_net() if rec == "sp_weno_dl" else None
    o = orc.OracleSolver(cfg, net, faithful=False)
    full = np.zeros((5, nx * ny))
    full[[0, 1, 2, 4]] = prim
    o.set_state(full)
    s = solver_for(cfg, rec == "sp_weno_dl")
    s.set_state(prim)
    for step in range(4):
        dto = o.step()
        dtg = s.step()
        assert abs(dtg - dto) <= 1e-14 * dto
        ref = o.get_state()[[0, 1, 2, 4]]
        got = s.get_state()
        assert rel_err(got, ref) <= 1e-12 * (step + 1), (step, rel_err(got, ref))


@pytest.mark.parametrize("flux", ["kepes_tecno_roe", "kepes_tecno4_rusanov", "kepes_tecno_rusanov",
                                  "roe_tecno4_roe", "roe_tecno_roe", "kep_tecno_roe", "kepec4",
                                  "roe", "kep4", "keps", "keps4"])
def test_euler2d_flux_variants_vs_oracle(flux):
    rng = np.random.default_rng(7)
    nx, ny = 40, 33
    cfg = make_cfg(n=(nx, ny, 2), dims=2, recon="sp_weno", flux=flux, bc="periodic")
    prim = _random_euler_state(rng, nx, ny, False)
    o = orc.OracleSolver(cfg, None, faithful=False)
    full = np.zeros((5, nx * ny))
    full[[0, 1, 2, 4]] = prim
    o.set_state(full)
    s = Solver(cfg)
    s.set_state(prim)
    for step in range(3):
        o.step()
        s.step()
    assert rel_err(s.get_state(), o.get_state()[[0, 1, 2, 4]]) <= 3e-12


@pytest.mark.parametrize("flux,rec", [("kepes_tecno_roe", "eno 2"), ("kepes_tecno_roe", "eno 3"),
                                      ("roe_tecno4_roe", "eno 3"), ("kepes_tecno4_rusanov", "eno 2"),
                                      ("kepes_tecno_roe", "sp_weno_dl"),
                                      ("roe_tecno4_roe", "sp_weno_dl"),
                                      ("kepes_tecno_roe", "sp_wenoc"), ("kepes_tecno_roe", "second"),
                                      ("kepes_tecno_roe", "first")])
def test_euler2d_generic_face_reconstructions_vs_oracle(flux, rec):
    """FP64 2D Euler, the TeCNO fluxes other than kepes_tecno4_roe with ENO k = 2, 3,
    DSP-WENO, sp_wenoc and the general face's remaining schemes, three steps
    against the oracle (the ENO, DSP-WENO and SP-WENO family runs take the
    generic face built with that reconstruction alone)."""
    rng = np.random.default_rng(17)
    nx, ny = 36, 29
    cfg = make_cfg(n=(nx, ny, 2), dims=2, recon=rec, flux=flux, bc="periodic")
    dsp = rec == "sp_weno_dl"
    # smooth data for the network (its FP64 tanh / exp against the C library's)
    prim = _random_euler_state(rng, nx, ny, dsp)
    o = orc.OracleSolver(cfg, orc.load_net() if dsp else None, faithful=False)
    full = np.zeros((5, nx * ny))
    full[[0, 1, 2, 4]] = prim
    o.set_state(full)
    s = solver_for(cfg, dsp)
    s.set_state(prim)
    for step in range(3):
        o.step()
        s.step()
    assert rel_err(s.get_state(), o.get_state()[[0, 1, 2, 4]]) <= 3e-12
    s.close()


@pytest.mark.parametrize("flux", ["lin_tecno4", "bur_tecno4", "lin_tecno", "bur_ec4"])
@pytest.mark.parametrize("rec", ["sp_weno", "eno 3", "sp_weno_dl", "second", "minmod", "minabs"])
def test_scalar2d_per_step_vs_oracle(flux, rec):
    rng = np.random.default_rng(3)
    nx, ny = 50, 41
    cfg = make_cfg(system="scalar", n=(nx, ny, 2), dims=2, lo=(-1, -1, 0), hi=(1, 1, 1), flux=flux,
                   recon=rec, vel=(1.0, -0.6, 0.0),
                   bc=["periodic", "periodic", "wall", "neumann", "periodic", "periodic"])
    u = rng.normal(size=nx * ny)
    net = orc.load_net() if rec == "sp_weno_dl" else None
    o = orc.OracleSolver(cfg, net, faithful=False)
    o.set_state(u)
    s = solver_for(cfg, rec == "sp_weno_dl")
    s.set_state(u)
    for step in range(4):
        o.step()
        s.step()
        assert rel_err(s.get_state(), o.get_state()) <= 1e-12 * (step + 1)


# ---------------------------------------------------------------- seams
@pytest.mark.parametrize("tag,rec", [("spweno", "sp_weno"), ("spwenoc", "sp_wenoc"),
                                     ("spwenocorr2", "sp_weno_corr2"), ("dsp", "sp_weno_dl"),
                                     ("first", "first"), ("second", "second"),
                                     ("minmod", "minmod"), ("minabs", "minabs")])
def test_reconstruct_seam_golden(tag, rec):
    st = case("point", "recon_stencils")["st"]
    ref = case("point", f"recon_{tag}")
    s = solver_for(make_cfg(system="scalar", flux="lin_tecno4", recon=rec, n=(8, 2, 2), dims=1),
                   rec == "sp_weno_dl")
    l, r = s.reconstruct_batch(st)
    scale = 1.0 + np.abs(st).max(axis=1)
    assert (np.abs(l - ref["l"]) / scale).max() <= 1e-13
    assert (np.abs(r - ref["r"]) / scale).max() <= 1e-13


@pytest.mark.parametrize("k", [2, 3])
def test_eno_seam_golden(k):
    ref = case("point", f"recon_eno{k}")
    s = Solver(make_cfg(system="scalar", flux="lin_tecno4", recon=f"eno {k}", n=(8, 2, 2), dims=1))
    l, r = s.reconstruct_batch(ref["st"])
    np.testing.assert_array_equal(l, ref["l"])
    np.testing.assert_array_equal(r, ref["r"])


def _eflux2d_keys():
    return sorted({k.split("/")[0] for k in fixtures("point").files if k.startswith("eflux_2d_")})


@pytest.mark.parametrize("key", _eflux2d_keys())
def test_euler_flux_seam_golden(key):
    parts = key.split("_")
    ec1 = key.endswith("_ec1")
    rtag = parts[-2] if ec1 else parts[-1]
    flux = "_".join(parts[2:-2] if ec1 else parts[2:-1])
    rec = {"spweno": "sp_weno", "spwenoc": "sp_wenoc", "dsp": "sp_weno_dl", "first": "first",
           "second": "second", "minmod": "minmod", "minabs": "minabs"}[rtag]
    d = case("point", "ent_2d")
    F = case("point", key)["F"]
    cfg = make_cfg(flux=flux, recon=rec, ec1_fix=int(ec1), n=(8, 6, 4), lo=(0, 0, 0), hi=(1, 2, 0.5))
    s = solver_for(cfg, rec == "sp_weno_dl")
    for ax in range(2):
        out = s.flux_batch(d["P"], d["V"], ax)
        scale = np.abs(F[ax]).max(axis=0) + 1e-300
        assert (np.abs(out - F[ax]) / scale).max() <= 1e-12, ax


def test_mlp_forward_seam_vs_oracle():
    """FP64 SIMT network: 1e-13 of the oracle.  FP32 SIMT network: the vertex
    region tests (thresholds 1e-8 around theta = 1) are below FP32 resolution,
    so stencils at a region boundary may pick the neighbouring region; at least
    99.5% of the stencils agree to 2e-5 and every FP32 result keeps the sign
    property."""
    st = case("point", "recon_stencils")["st"]
    st = st[np.abs(st[:, 2] - st[:, 1]) > 1e-12]
    net = orc.load_net()
    lo, ro = orc.nn_forward(net, st)
    scale = 1.0 + np.abs(st).max(axis=1)
    s = Solver(make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno_dl", n=(8, 2, 2), dims=1,
                        mlp_dtype=MLP_FP64, precision=FP64))
    s.load_net(NET_FILE)
    l, r = s.mlp_forward_batch(st)
    assert (np.abs(l - lo) / scale).max() <= 1e-13
    assert (np.abs(r - ro) / scale).max() <= 1e-13
    s = Solver(make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno_dl", n=(8, 2, 2), dims=1,
                        mlp_dtype=MLP_FP32, precision=FP32))
    s.load_net(NET_FILE)
    l, r = s.mlp_forward_batch(st)
    err = np.maximum(np.abs(l - lo), np.abs(r - ro)) / scale
    assert np.mean(err <= 2e-5) >= 0.995
    jump = st[:, 2] - st[:, 1]
    ok = (jump * (r - l) >= 0) | (np.abs(r - l) <= 1e-6 * scale)
    assert ok.all()


@pytest.mark.parametrize("mode,tol", [(MLP_BF16, 0.008), (MLP_FP16, 0.001)])
def test_mlp_mfma_seam_vs_oracle(mode, tol):
    """Matrix-core network (bf16 / fp16 operands, fp32 accumulation): the
    reconstructed values differ from the FP64 network by at most tol times the
    stencil's total variation for 99% of stencils and 3 tol for all of them
    (8 and 11 significant bits in the operands), and every result keeps the
    sign property (convex vertex weights are feasible whatever the network
    outputs)."""
    st = case("point", "recon_stencils")["st"]
    st = st[np.abs(st[:, 2] - st[:, 1]) > 1e-12]
    lo, ro = orc.nn_forward(orc.load_net(), st)
    s = Solver(make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno_dl", n=(8, 2, 2), dims=1,
                        mlp_dtype=mode, precision=FP32))
    s.load_net(NET_FILE)
    l, r = s.mlp_forward_batch(st)
    tv = np.abs(np.diff(st, axis=1)).sum(axis=1) + 1e-300
    err = np.maximum(np.abs(l - lo), np.abs(r - ro)) / tv
    assert np.isfinite(l).all() and np.isfinite(r).all()
    assert np.mean(err <= tol) >= 0.99, np.quantile(err, [0.5, 0.9, 0.99])
    assert err.max() <= 3 * tol, err.max()
    jump = st[:, 2] - st[:, 1]
    scale = 1.0 + np.abs(st).max(axis=1)
    assert ((jump * (r - l) >= 0) | (np.abs(r - l) <= 1e-6 * scale)).all()


@pytest.mark.parametrize("mode,tol99", [(MLP_BF16, 1e-5), (MLP_FP16, 5e-7)])
def test_mfma_stage_close_to_simt(mode, tol99):
    """The stage kernel with the network on matrix cores tracks the FP32 SIMT
    network over 10 steps of a smooth flow: per cell, the largest relative
    difference over the variables is below tol99 for 99% of the cells and
    below 1e-4 everywhere (bf16 / fp16 operands carry 8 / 11 significant
    bits; the few cells above tol99 sit where the network output is steep)."""
    rng = np.random.default_rng(12)
    nx, ny = 64, 48
    prim = _random_euler_state(rng, nx, ny, True)
    out = {}
    for m in (MLP_FP32, mode):
        s = solver_for(make_cfg(n=(nx, ny, 2), dims=2, recon="sp_weno_dl", bc="periodic", precision=FP32,
                                mlp_dtype=m, sign_fallback=1), True)
        s.set_state(prim)
        s.run_steps(10)
        out[m] = s.get_state()
    a, b = out[mode], out[MLP_FP32]
    d = (np.abs(a - b) / np.abs(b).max(axis=-1, keepdims=True)).max(axis=0)
    assert np.isfinite(a).all()
    assert np.quantile(d, 0.99) <= tol99, np.quantile(d, [0.5, 0.99, 1.0])
    assert d.max() <= 1e-4, d.max()


# ---------------------------------------------------------------- invariants
def test_conservation_periodic_fp64():
    """Periodic domain: the FP64 conserved sums move only by rounding."""
    rng = np.random.default_rng(9)
    nx, ny = 96, 64
    cfg = make_cfg(n=(nx, ny, 2), dims=2, recon="sp_weno", bc="periodic")
    s = Solver(cfg)
    s.set_state(_random_euler_state(rng, nx, ny, False))
    a0 = s.audit()
    s.run_steps(20)
    a1 = s.audit()
    for f in ("mass", "energy"):
        assert abs(getattr(a1, f) - getattr(a0, f)) <= 1e-12 * abs(getattr(a0, f))
    for d in range(2):
        assert abs(a1.momentum[d] - a0.momentum[d]) <= 1e-12 * abs(a0.mass)
    assert a1.n_nonfinite == 0


def test_entropy_decreases_fp64():
    """TeCNO with sign-preserving reconstruction is entropy stable: total
    entropy sum(-rho s/(gamma-1)) does not increase on a periodic domain."""
    rng = np.random.default_rng(10)
    nx, ny = 64, 64
    cfg = make_cfg(n=(nx, ny, 2), dims=2, recon="sp_weno", bc="periodic")
    s = Solver(cfg)
    s.set_state(_random_euler_state(rng, nx, ny, False))
    e_prev = s.audit().total_entropy
    for _ in range(5):
        s.run_steps(4)
        e = s.audit().total_entropy
        assert e <= e_prev + 1e-12 * abs(e_prev)
        e_prev = e


@pytest.mark.parametrize("rec", ["sp_weno", "sp_wenoc", "sp_weno_dl"])
def test_fp32_path_close_to_fp64_oracle(rec):
    """Mixed path (FP64 state and entropy-conservative flux, FP32 dissipation):
    smooth flow, 10 steps, within 1e-8 relative of the FP64 oracle."""
    rng = np.random.default_rng(11)
    nx, ny = 64, 48
    cfg = make_cfg(n=(nx, ny, 2), dims=2, recon=rec, bc="periodic")
    prim = _random_euler_state(rng, nx, ny, True)
    net = orc.load_net() if rec == "sp_weno_dl" else None
    o = orc.OracleSolver(cfg, net, faithful=False)
    full = np.zeros((5, nx * ny))
    full[[0, 1, 2, 4]] = prim
    o.set_state(full)
    s = solver_for(make_cfg(n=(nx, ny, 2), dims=2, recon=rec, bc="periodic", precision=FP32,
                            sign_fallback=1), rec == "sp_weno_dl")
    s.set_state(prim)
    for _ in range(10):
        o.step()
    s.run_steps(10)
    assert rel_err(s.get_state(), o.get_state()[[0, 1, 2, 4]]) <= 1e-8


def _vortex(n, t=0.0):
    h = 10.0 / n
    x = -5 + (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(x, x)
    Xs = (X - t + 5) % 10 - 5
    g, beta = 1.4, 5.0
    r2 = Xs ** 2 + Y ** 2
    T = 1.0 - (g - 1) * beta ** 2 / (8 * g * math.pi ** 2) * np.exp(1 - r2)
    return np.stack([T ** (1 / (g - 1)), 1.0 - beta / (2 * math.pi) * Y * np.exp(0.5 * (1 - r2)),
                     beta / (2 * math.pi) * Xs * np.exp(0.5 * (1 - r2)), T ** (g / (g - 1))]).reshape(4, -1)


@pytest.mark.parametrize("rec,mlp", [("sp_weno", MLP_FP32), ("sp_weno_dl", MLP_FP32),
                                     ("sp_weno_dl", MLP_BF16), ("sp_weno_dl", MLP_FP16)])
def test_vortex_l1_and_order_mixed_matches_fp64(rec, mlp):
    """Isentropic vortex (strength 5, periodic [-5,5]^2, t = 1): the mixed
    path's L1(rho) error (FP32 SIMT or bf16/fp16 matrix-core network) is within
    2% of the FP64 path's on every mesh and the observed convergence orders
    agree to 0.05."""
    errs = {FP64: [], FP32: []}
    ns = [64, 128, 256]
    for prec in (FP64, FP32):
        for n in ns:
            s = solver_for(make_cfg(n=(n, n, 2), dims=2, lo=(-5, -5, 0), hi=(5, 5, 1), recon=rec,
                                    bc="periodic", precision=prec, sign_fallback=int(prec == FP32),
                                    mlp_dtype=mlp if prec == FP32 else MLP_FP64),
                           rec == "sp_weno_dl")
            s.set_state(_vortex(n))
            it, t = s.advance(1.0, 1)
            errs[prec].append(s.l1_error(_vortex(n, t))[0])
    for a, b in zip(errs[FP32], errs[FP64]):
        assert abs(a / b - 1) <= 0.02, (errs, rec)
    o64 = np.log2(np.array(errs[FP64][:-1]) / np.array(errs[FP64][1:]))
    o32 = np.log2(np.array(errs[FP32][:-1]) / np.array(errs[FP32][1:]))
    assert np.all(np.abs(o64 - o32) <= 0.05), (o64, o32)
    assert np.all(o64 > 2.5)
