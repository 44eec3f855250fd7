"""3D Euler / Navier-Stokes with FP32 reconstruction on the FP64 state
(precision FP32 in 3D): the five characteristic jumps of each face in FP32
jump form, SP-WENO or DSP-WENO (FP32 SIMT, bf16 / fp16 matrix-core
network); records, entropy-conservative flux, eigensystem, dissipation
product and viscous terms in FP64.  Against the FP64 oracle (the
reference's arithmetic) on a smooth 3D Taylor-Green-type state and against
the FP64 GPU path on a Taylor-Green vortex run."""
import math

import numpy as np
import pytest

import oracle_ctypes as orc
from tecno import FP32, FP64, MLP_BF16, MLP_FP16, MLP_FP32, MLP_FP64, Solver, make_cfg

pytestmark = pytest.mark.gpu
NET = orc.NET_FILE


def smooth_state(n, seed):
    """Smooth periodic state without symmetries: each variable a sum of three
    Fourier modes of random wave vectors (components 0..2) and phases."""
    rng = np.random.default_rng(seed)
    h = [2 * math.pi / k for k in n]
    x, y, z = ((np.arange(k) + 0.5) * hh for k, hh in zip(n, h))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")

    def field(amp):
        f = np.zeros_like(X)
        for _ in range(3):
            k = rng.integers(0, 3, size=3)
            if not k.any():
                k[0] = 1
            f += amp * rng.uniform(0.3, 1.0) * np.sin(k[0] * X + k[1] * Y + k[2] * Z +
                                                      rng.uniform(0, 2 * math.pi))
        return f

    return np.stack([1.0 + field(0.05), field(0.3), field(0.3), field(0.3),
                     10.0 + field(0.3)]).reshape(5, -1)


def cfg3(n, rec, prec, mlp, model="ns_es", cfl=0.5):
    return make_cfg(n=n, dims=3, lo=(0, 0, 0), hi=(2 * math.pi,) * 3, recon=rec, cfl=cfl,
                    flux="kepes_tecno4_roe", bc="periodic", model=model, fix_zmom=1,
                    viscosity=("constant", 1.0 / 1600), precision=prec, mlp_dtype=mlp,
                    sign_fallback=1 if prec == FP32 else 0)


def dev(a, b):
    """per-entry deviation relative to each variable's largest magnitude"""
    return np.abs(a - b) / np.abs(b).max(axis=1, keepdims=True)


# (reconstruction, network, max, 99th percentile, median) of the deviation
# from the FP64 oracle after 5 steps; measured on MI355X: SP-WENO and the
# FP32 network ~1e-8 / 5e-9 / 6e-10 (a face whose SP-WENO region test falls
# the other way in FP32 moves a few cells by up to 2e-3 on finer meshes);
# bf16 network 7.2e-3 / 5.8e-4 / 1.7e-5, fp16 6.1e-3 / 1.5e-4 / 1.9e-6
CASES = [("sp_weno", MLP_FP64, 1e-2, 1e-5, 1e-8), ("sp_weno_dl", MLP_FP32, 1e-2, 1e-5, 1e-8),
         ("sp_weno_dl", MLP_BF16, 2e-2, 2e-3, 1e-4), ("sp_weno_dl", MLP_FP16, 2e-2, 5e-4, 1e-5)]


@pytest.mark.parametrize("model", ["euler", "ns_es"])
@pytest.mark.parametrize("rec,mlp,bmax,b99,bmed", CASES)
def test_fp32_3d_close_to_fp64_oracle(rec, mlp, bmax, b99, bmed, model):
    """Smooth 16 x 12 x 10 state without symmetries, 5 SSPRK3 steps, the
    sign fallback off on both sides (the oracle has none): the FP32
    reconstruction path against the FP64 oracle.  (On a state with exact
    symmetries -- Taylor-Green products on a grid -- ratios of equal jumps
    are exactly 1 in FP64 and not in FP32, and the DSP-WENO region tree takes
    other branches at those faces.)"""
    n = (16, 12, 10)
    init = smooth_state(n, 5)
    net = orc.load_net() if rec == "sp_weno_dl" else None
    o = orc.OracleSolver(cfg3(n, rec, FP64, MLP_FP64, model), net, faithful=False)
    o.set_state(init)
    c = cfg3(n, rec, FP32, mlp, model)
    c.sign_fallback = 0
    s = Solver(c)
    if net is not None:
        s.load_net(NET)
    s.set_state(init)
    for _ in range(5):
        o.step()
    s.run_steps(5)
    d = dev(s.get_state(), o.get_state())
    print(rec, mlp, model, d.max(), np.quantile(d, 0.99), np.median(d))
    assert 0 < d.max() <= bmax
    assert np.quantile(d, 0.99) <= b99
    assert np.median(d) <= bmed


def test_fp32_3d_matrix_core_network_close_to_simt():
    """bf16 and fp16 matrix-core networks against the FP32 SIMT network on
    the same state, sign fallback on: the same method, differing by the
    16-bit weights and activations."""
    n = (16, 12, 10)
    init = smooth_state(n, 5)
    out = {}
    for mlp in (MLP_FP32, MLP_BF16, MLP_FP16):
        s = Solver(cfg3(n, "sp_weno_dl", FP32, mlp))
        s.load_net(NET)
        s.set_state(init)
        s.run_steps(5)
        out[mlp] = s.get_state()
        s.close()
    for mlp, b99 in ((MLP_BF16, 2e-3), (MLP_FP16, 5e-4)):
        d = dev(out[mlp], out[MLP_FP32])
        print(mlp, d.max(), np.quantile(d, 0.99))
        assert d.max() <= 2e-2 and np.quantile(d, 0.99) <= b99


def tgv32():
    h = 2 * math.pi / 32
    x = (np.arange(32) + 0.5) * h
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return np.stack([np.ones_like(X), np.sin(X) * np.cos(Y) * np.cos(Z),
                     -np.cos(X) * np.sin(Y) * np.cos(Z), np.zeros_like(X),
                     100 / 1.4 + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2) / 16]
                    ).reshape(5, -1)


@pytest.mark.parametrize("rec,mlp", [("sp_weno", MLP_FP64), ("sp_weno_dl", MLP_BF16),
                                     ("sp_weno_dl", MLP_FP32)])
def test_tgv_fp32_3d_within_fp64_sensitivity(rec, mlp):
    """Taylor-Green vortex (Re 1600, Mach 0.1) on 32^3 to t = 1 at CFL 0.2:
    the 32^3 flow is under-resolved and amplifies small differences (an FP64
    run from the initial state times (1 + 1e-9 noise) drifts from the
    unperturbed one by up to 4e-4 of the kinetic energy at t = 1).  The FP32
    path's kinetic energy stays within 5x that FP64 drift plus 1e-6 (DSP-WENO:
    5e-6) of the FP64 values at t = 0.25, 0.5, 0.75, 1, its entropy production (total
    entropy minus its initial value) within 5x the drift plus 0.1% (measured
    with the FP32 network: 0.026% more at t = 1), and no state is
    non-finite."""
    prim = tgv32()
    rng = np.random.default_rng(1)
    runs = {}
    for tag, prec, m, pert in (("fp64", FP64, MLP_FP64, 0.0), ("pert", FP64, MLP_FP64, 1e-9),
                               ("fp32", FP32, mlp, 0.0)):
        c = cfg3((32, 32, 32), rec, prec, m, cfl=0.2)
        c.sign_fallback = 0  # as the FP64 path: the fallback adds dissipation of its own
        s = Solver(c)
        if rec == "sp_weno_dl":
            s.load_net(NET)
        s.set_state(prim * (1 + pert * rng.standard_normal(prim.shape)))
        ke, ent = [], []
        s0 = s.audit().total_entropy
        for k in range(4):
            s.advance(0.25 * (k + 1), 1)
            a = s.audit()
            ke.append(a.kinetic_energy)
            ent.append(a.total_entropy)
        runs[tag] = (np.array(ke), np.array(ent) - s0, s.audit().n_nonfinite)
        s.close()
    ke64, e64, _ = runs["fp64"]
    drift = np.abs(runs["pert"][0] - ke64)
    edrift = np.abs(runs["pert"][1] - e64)
    ke32, e32, bad = runs["fp32"]
    print(rec, mlp, np.abs(ke32 - ke64), drift, np.abs(e32 - e64), edrift, e64)
    assert bad == 0
    # DSP-WENO in FP32 changes the reconstruction itself, not only its
    # rounding: the region tree's |theta - 1| > 1e-8 tests resolve below the
    # FP32 spacing near 1 (1.2e-7), so faces within a few ulp of a region
    # boundary take another vertex polygon, and the bf16 network rounds its
    # weights to an 8-bit significand.  Its kinetic energy is allowed 5e-6 of
    # the FP64 values on top of the drift (FP32 network: 3.5e-6 at t = 0.75)
    rel = 5e-6 if rec == "sp_weno_dl" else 1e-6
    assert np.all(np.abs(ke32 - ke64) <= 5 * drift + rel * ke64)
    assert np.all(np.abs(e32 - e64) <= 5 * edrift + 1e-3 * np.abs(e64))
