#!/usr/bin/env python3
"""Headline benchmark: 2D Euler Kelvin-Helmholtz shear layer on 2048^2
(periodic [0,1]^2, rho = 2 / u = -0.5 for |y-0.5| < 0.25, rho = 1 / u = 0.5
elsewhere, p = 2.5, v = 0.01 sum_{k=1..8} a_k sin(2 pi k x), a_k ~ N(0,1)
seeded), kepes_tecno4_roe TeCNO flux, SSPRK3 at CFL 0.5.

A step is one full SSPRK3 step of the whole grid (CFL dt + three fused stage
kernels) with the state resident in HBM.  The headline path is DSP-WENO
reconstruction (sp_weno_dl, the shipped 5-5-5-5-5 network) with FP32 flux and
reconstruction; the FP64 SP-WENO path is timed beside it.  Prints one JSON
line (metric, value in cell-updates/s, roofline of the stage kernel, CPU
baseline).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 2048]
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dsp-weno_amd"))

METRIC = "cell-updates/s, 2D Euler 2048² DSP-WENO mixed vs FP64 SP-WENO at equal L1 err"
NET_FILE = os.path.join(ROOT, "tests", "golden", "dsp_weno_net.txt")
HBM_PEAK_GBS = 8000.0
NVAR = 4
# untimed steps from the KH initial state before any path is timed (t ~ 0.03)
SPINUP_STEPS = 300


def kh_state(n, seed=2024):
    a = np.random.default_rng(seed).normal(size=8)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    inner = np.abs(Y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    u = np.where(inner, -0.5, 0.5)
    v = 0.01 * sum(a[k] * np.sin(2 * np.pi * (k + 1) * X) for k in range(8))
    p = np.full_like(rho, 2.5)
    return np.stack([rho, u, v, p]).reshape(4, -1), a


XAVIER_SEED = 6008  # the trainer's default seed (config.py)


def make_solver(tecno, n, precision, recon, mlp_dtype, weights="shipped"):
    """weights: "shipped" (the net_file) or "xavier" (the trainer's seeded
    initial network of the same 5-5-5-5-5 architecture: Xavier-uniform
    weights as PyTorch draws them under manual_seed(6008), tecno_init_net)"""
    cfg = tecno.make_cfg(n=(n, n, 2), dims=2, lo=(0, 0, 0), hi=(1, 1, 1), recon=recon,
                         flux="kepes_tecno4_roe", bc="periodic", cfl=0.5, precision=precision,
                         mlp_dtype=mlp_dtype, sign_fallback=1)
    s = tecno.Solver(cfg)
    if recon == "sp_weno_dl":
        if weights == "xavier":
            s.init_net([5, 5, 5, 5, 5], XAVIER_SEED)
        else:
            s.load_net(NET_FILE)
    return s


def time_steps(s, n, steps):
    """K steps bracketed by synchronisation; stage-kernel time from HIP events
    recorded on the solver's stream."""
    s.sync()
    fb0 = s.audit().n_sign_fallback
    s.timing(True)
    t0 = time.perf_counter()
    s.run_steps(steps, final_time=1.0, t_stamps=1)
    s.sync()
    el = time.perf_counter() - t0
    stage_ms, nst = s.timing_get()
    s.timing(False)
    a = s.audit()
    t, it = s.time()
    if a.n_nonfinite:
        raise RuntimeError("non-finite state")
    return dict(seconds=el, cell_updates_per_s=n * n * steps / el, stage_ms=stage_ms,
                n_stages=nst, ms_per_step=1e3 * el / steps, time=t, iters=it,
                sign_fallbacks=a.n_sign_fallback - fb0)


KH_PARAM = """grid
{{
   x_min   0.0
   x_max   1.0
   Nx_cell {{ {n} }}
   y_min   0.0
   y_max   1.0
   Ny_cell {{ {n} }}
   z_min   0.0
   z_max   1.0
   Nz_cell {{ 2 }}
}}
numeric
{{
   time_scheme  ssprk3
   cfl          0.5
   max_iter     50000
   final_time   1.0
   min_residue  1.0e-6
   reconstruct  {rec}
   net_file     {net}
}}
material
{{
   gamma       1.4
   gas_const   1.0
   viscosity
   {{
      model    constant
      mu_ref   0.0
   }}
   prandtl     0.72
   model       euler
   flux        kepes_tecno4_roe
   EC1_fix     no
}}
constants
{{
{consts}
}}
initial_condition
{{
   density    2.0*(abs(y-0.5)<0.25) + 1.0*(abs(y-0.5)>=0.25)
   xvelocity  -0.5*(abs(y-0.5)<0.25) + 0.5*(abs(y-0.5)>=0.25)
   yvelocity  0.01*({pert})
   zvelocity  0.0
   pressure   2.5
}}
exact_soln
{{
   available   no
}}
boundary_condition
{{
   xleft
   {{
      type periodic
   }}
   xright
   {{
      type periodic
   }}
   yleft
   {{
      type periodic
   }}
   yright
   {{
      type periodic
   }}
   zleft
   {{
      type periodic
   }}
   zright
   {{
      type periodic
   }}
}}
output
{{
   OOC_study   no
   problem_dim 2
   time_stamps 1
   output_path {out}
}}
"""


def _kh_param_file(tmp, n_sample):
    a = np.random.default_rng(2024).normal(size=8)
    consts = "\n".join(f"   a{k}   {float(a[k])!r}" for k in range(8))
    pert = " + ".join(f"a{k}*sin(2*pi*{k + 1}*x)" for k in range(8))
    p = os.path.join(tmp, "param.in")
    with open(p, "w") as f:
        f.write(KH_PARAM.format(n=n_sample, rec="sp_weno_dl", net=NET_FILE, consts=consts,
                                pert=pert, out=tmp))
    return p


def cpu_baseline(n_sample=1024, steps=1):
    """The reference solver (built from its sources into oracle/_ref by
    oracle/Makefile) on one host core, on a bounded sample of the workload:
    the KH problem at n_sample^2 (the reference's Nz = 2 layout), `steps`
    steps of sp_weno_dl (about 15 s).  Falls back to the oracle's C port when
    the reference build is absent."""
    refh = os.path.join(ROOT, "oracle", "_ref", "refh_sys")
    if os.path.exists(refh):
        with tempfile.TemporaryDirectory() as tmp:
            p = _kh_param_file(tmp, n_sample)
            r = subprocess.run(["taskset", "-c", "0", refh, "time", p, str(steps)], capture_output=True,
                               text=True, timeout=600)
            if r.returncode == 0:
                tok = r.stdout.split()
                sec = float(tok[3])
                return {"value": n_sample * n_sample * steps / sec, "unit": "cell-updates/s",
                        "cores": 1, "kind": "reference",
                        "sample": f"KH {n_sample}x{n_sample} (x2 z-layers as the reference stores 2D), "
                                  f"{steps} SSPRK3 step(s), sp_weno_dl FP64, reference solver built from "
                                  f"source, 1 host core, {sec:.1f} s"}
    return _oracle_baseline(n_sample, steps, 1)


def _oracle_baseline(n_sample, steps, threads):
    """The oracle (the C restatement of the reference, its arithmetic and Nz = 2
    layout) on `threads` host threads (OpenMP over face rows)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as orc
    import tecno
    cfg = tecno.make_cfg(n=(n_sample, n_sample, 2), dims=2, recon="sp_weno_dl", bc="periodic")
    s = orc.OracleSolver(cfg, orc.load_net(NET_FILE), faithful=True, threads=threads)
    prim, _ = kh_state(n_sample)
    full = np.zeros((5, 2, n_sample * n_sample))
    full[[0, 1, 2, 4]] = prim[:, None, :]
    s.set_state(full.reshape(5, -1))
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(1.0)
    sec = time.perf_counter() - t0
    return {"value": n_sample * n_sample * steps / sec, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"KH {n_sample}x{n_sample}x2, {steps} SSPRK3 step(s), sp_weno_dl FP64, oracle "
                      f"C port of the reference, {threads} host thread(s), {sec:.1f} s"}


def cpu_baseline_all_cores(n_sample=1024, steps=2):
    """The same workload sample on all host cores available to this process
    (at most 16): the oracle with OpenMP threads, since the reference solver
    is serial."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return _oracle_baseline(n_sample, steps, max(1, min(16, cores)))


def state_deviation(a, b, steps):
    """Deviation of state a from state b (nv x N primitive): per variable the
    largest |a - b| relative to max|b|, and the fraction of cells whose largest
    relative deviation over the variables is below 1e-8, 1e-6 and 1e-4."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.abs(b).max(axis=1, keepdims=True)
    d = np.abs(a - b) / scale
    per_cell = d.max(axis=0)
    return {"steps": int(steps), "max_rel_per_var": [float(x) for x in d.max(axis=1)],
            "vars": ["rho", "u", "v", "p"],
            "frac_cells_within": {t: float(np.mean(per_cell <= float(t)))
                                  for t in ("1e-8", "1e-6", "1e-4")}}


VALU_PEAK = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions/s: CUs x SIMDs x clock / 4 cycles


# ---------------------------------------------------------------- ensemble
ENS_1D_N = (100, 200, 400, 600, 800, 1000)
ENS_2D_N = (64, 96, 128, 192, 256, 384, 512, 768, 1024)
ENS_T1D, ENS_T2D = 1.5, 0.5


def burgers_sine_exact(x, t):
    """entropy solution of u_t + (u^2/2)_x = 0, u(x,0) = sin x on [-pi, pi]
    periodic, t > 1: stationary shocks at x = +-pi; inside, u = sin(xi) with
    xi + t sin(xi) = |x| on the increasing branch xi in [0, arccos(-1/t)]
    (bisection), odd in x"""
    ax = np.abs(x)
    hi_ = np.arccos(-1.0 / t) if t > 1 else np.pi
    lo = np.zeros_like(ax)
    hi = np.full_like(ax, hi_)
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        f = mid + t * np.sin(mid) - ax
        lo = np.where(f < 0, mid, lo)
        hi = np.where(f < 0, hi, mid)
    return np.sign(x) * np.sin(0.5 * (lo + hi))


def ens_ic2d(n, seed=4):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 4, size=(8, 2))
    ph = rng.uniform(0, 2 * np.pi, 8)
    amp = rng.normal(size=8)
    amp *= 0.05 / np.abs(amp).sum()

    def f(X, Y):
        u = np.sin(np.pi * X) * np.sin(np.pi * Y)
        for q in range(8):
            u = u + amp[q] * np.sin(np.pi * (k[q, 0] * X + k[q, 1] * Y) + ph[q])
        return u
    h = 2.0 / n
    x = -1 + (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(x, x)
    return f, X, Y


def ensemble_bench(tecno, use_graph=True, precs=None):
    """Config 4: 48 periodic scalar problems (6 x 1D Burgers sin x to t=1.5,
    9 x 2D Burgers and 9 x 2D linear advection (a = (1,1)) of
    sin(pi x) sin(pi y) + 8 seeded Fourier modes of total amplitude 0.05 to
    t = 0.5, each at FP64 and FP32; bur_tecno4 / lin_tecno4 + sp_weno, SSPRK3
    CFL 0.5) in one ensemble: autotuned tile shapes, hipGraph steps.  precs:
    the precisions of the members (default FP64 and FP32)."""
    if precs is None:
        precs = (tecno.FP64, tecno.FP32)
    label = {tecno.FP64: "fp64", tecno.FP32: "fp32", tecno.FP32_STATE64: "fp32_state64"}
    probs = []
    for prec in precs:
        for n in ENS_1D_N:
            probs.append(dict(kind="burgers1d", n=n, prec=prec, t=ENS_T1D,
                              cfg=tecno.make_cfg(system="scalar", flux="bur_tecno4", recon="sp_weno",
                                                 n=(n, 1, 1), dims=1, lo=(-np.pi, 0, 0),
                                                 hi=(np.pi, 1, 1), bc="periodic", precision=prec)))
        for kind, flux in (("burgers2d", "bur_tecno4"), ("linadv2d", "lin_tecno4")):
            for n in ENS_2D_N:
                probs.append(dict(kind=kind, n=n, prec=prec, t=ENS_T2D,
                                  cfg=tecno.make_cfg(system="scalar", flux=flux, recon="sp_weno",
                                                     n=(n, n, 1), dims=2, lo=(-1, -1, 0),
                                                     hi=(1, 1, 1), vel=(1, 1, 0), bc="periodic",
                                                     precision=prec)))
    ens = tecno.Ensemble([p["cfg"] for p in probs])

    def init():
        for m, p in zip(ens.members, probs):
            if p["kind"] == "burgers1d":
                h = 2 * np.pi / p["n"]
                m.set_state(np.sin(-np.pi + (np.arange(p["n"]) + 0.5) * h))
            else:
                f, X, Y = ens_ic2d(p["n"])
                m.set_state(f(X, Y).ravel())
        for m in ens.members:
            m.sync()
    init()
    ens.autotune(steps=3)
    # untimed warm-up pass (graph capture, code load), then the timed pass
    ens.advance([p["t"] for p in probs], use_graph=use_graph)
    init()
    t0 = time.perf_counter()
    cu = ens.advance([p["t"] for p in probs], use_graph=use_graph)
    el = time.perf_counter() - t0
    # errors: linear advection and 1D Burgers against exact solutions (cell
    # centre values, as compute_errors.cc); 2D Burgers against the 1024^2
    # solution averaged onto the coarser meshes that divide it
    res = []
    fine = {}
    for m, p in zip(ens.members, probs):
        if p["kind"] == "burgers2d" and p["n"] == 1024:
            fine[p["prec"]] = m.get_state().reshape(1024, 1024)
    for m, p in zip(ens.members, probs):
        t, it = m.time()
        n = p["n"]
        if p["kind"] == "burgers1d":
            h = 2 * np.pi / n
            x = -np.pi + (np.arange(n) + 0.5) * h
            err = float(m.l1_error(burgers_sine_exact(x, t))[0])
        elif p["kind"] == "linadv2d":
            f, X, Y = ens_ic2d(n)
            Xs = (X - t + 1) % 2 - 1
            Ys = (Y - t + 1) % 2 - 1
            err = float(m.l1_error(f(Xs, Ys).ravel())[0])
        elif 1024 % n == 0 and n < 1024:
            r = 1024 // n
            ref = fine[p["prec"]].reshape(n, r, n, r).mean(axis=(1, 3))
            err = float(m.l1_error(ref.ravel())[0])
        else:
            err = None
        res.append(dict(kind=p["kind"], n=n, precision=label[p["prec"]], steps=it, time=t, l1=err))
    orders = {}
    for kind in ("burgers1d", "linadv2d", "burgers2d"):
        for prec in (label[q] for q in precs):
            rows = [r for r in res if r["kind"] == kind and r["precision"] == prec and r["l1"]]
            orders[f"{kind}_{prec}"] = [
                round(float(np.log(a["l1"] / b["l1"]) / np.log(b["n"] / a["n"])), 3)
                for a, b in zip(rows[:-1], rows[1:])]
    groups = ens.groups()
    ens.close()
    return dict(cell_updates=cu, seconds=el, cell_updates_per_s=cu / el, problems=res,
                orders=orders, groups=[dict(g, tuned_ms={f"{k[0]}x{k[1]}": v for k, v in
                                                           g["tuned_ms"].items()}) for g in groups])


def vortex_state(n, t=0.0, strength=5.0, gamma=1.4):
    """Isentropic vortex on the periodic [-5,5]^2 (IsentropicVortex/param.in:
    strength 5, free stream (1, 0), rho = p = 1), translated to time t."""
    h = 10.0 / n
    x = -5 + (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(x, x)
    Xs = (X - t + 5) % 10 - 5
    r2 = Xs ** 2 + Y ** 2
    T = 1.0 - (gamma - 1) * strength ** 2 / (8 * gamma * np.pi ** 2) * np.exp(1 - r2)
    e = np.exp(0.5 * (1 - r2))
    return np.stack([T ** (1 / (gamma - 1)), 1.0 - strength / (2 * np.pi) * Y * e,
                     strength / (2 * np.pi) * Xs * e, T ** (gamma / (gamma - 1))]).reshape(4, -1)


def l1_gate(tecno, ns=(128, 256, 512), t_end=1.0, net_file=None):
    """Equal-error gate on the isentropic vortex (exact solution): L1(rho) at
    t = 1 on each mesh and the observed orders, for the headline path (mixed,
    DSP-WENO, bf16 network), the same method in FP64 (reference arithmetic),
    the FP64 SP-WENO path and SP-WENO with the headline's arithmetic split
    (FP32 jump-form reconstruction on the FP64 state)."""
    paths = {"headline": (tecno.FP32, "sp_weno_dl", tecno.MLP_BF16),
             "fp64_dsp": (tecno.FP64, "sp_weno_dl", tecno.MLP_FP64),
             "fp64_spweno": (tecno.FP64, "sp_weno", tecno.MLP_FP64),
             "mixed_spweno": (tecno.FP32, "sp_weno", tecno.MLP_FP64)}
    out = {"case": "isentropic vortex, strength 5, periodic [-5,5]^2, t=1, L1(rho)", "n": list(ns)}
    for name, (prec, rec, mlp) in paths.items():
        errs = []
        for n in ns:
            cfg = tecno.make_cfg(n=(n, n, 2), dims=2, lo=(-5, -5, 0), hi=(5, 5, 1), recon=rec,
                                 flux="kepes_tecno4_roe", bc="periodic", cfl=0.5, precision=prec,
                                 mlp_dtype=mlp, sign_fallback=1)
            s = tecno.Solver(cfg)
            if rec == "sp_weno_dl":
                s.load_net(net_file or NET_FILE)
            s.set_state(vortex_state(n))
            s.advance(t_end, 1)
            t, _ = s.time()
            errs.append(float(s.l1_error(vortex_state(n, t))[0]))
            s.close()
        out[name] = {"l1": errs,
                     "order": [round(float(np.log(a / b) / np.log(2)), 3)
                               for a, b in zip(errs[:-1], errs[1:])]}
    h, d, b = out["headline"]["l1"], out["fp64_dsp"]["l1"], out["fp64_spweno"]["l1"]
    out["l1_ratio_vs_fp64_dsp"] = [round(x / y, 4) for x, y in zip(h, d)]
    out["l1_ratio_vs_fp64_spweno"] = [round(x / y, 4) for x, y in zip(h, b)]

    def same(a, ref):
        # L1 within 2% on every mesh and the convergence orders within 0.05
        oa, orf = out[a]["order"], out[ref]["order"]
        return bool(all(abs(x / y - 1) <= 0.02 for x, y in zip(out[a]["l1"], out[ref]["l1"])) and
                    all(abs(x - y) <= 0.05 for x, y in zip(oa, orf)))
    out["pass_vs_fp64_spweno"] = same("headline", "fp64_spweno")
    out["pass_vs_fp64_dsp"] = same("headline", "fp64_dsp")
    out["mixed_spweno_l1_ratio_vs_fp64_spweno"] = [
        round(x / y, 4) for x, y in zip(out["mixed_spweno"]["l1"], b)]
    out["mixed_spweno_pass_vs_fp64_spweno"] = same("mixed_spweno", "fp64_spweno")
    out["rule"] = "L1 within 2% on every mesh and convergence orders within 0.05"
    return out


def tgv_state(n, mach=0.1, gamma=1.4):
    """Taylor-Green vortex on the periodic [0, 2 pi]^3: rho = 1,
    u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
    p = 1/(gamma M^2) + (cos 2x + cos 2y)(cos 2z + 2)/16."""
    h = 2 * np.pi / n
    c = (np.arange(n) + 0.5) * h
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    p = 1.0 / (gamma * mach * mach) + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2) / 16
    return np.stack([np.ones_like(u), u, v, np.zeros_like(u), p]).reshape(5, -1)


def tgv_bench(tecno, n, t_end, recon, net_file, every=0.25, fix_visc=0, cfl=0.5, fp32=False,
              mlp=None):
    """Navier-Stokes TGV (ns_es, kepes_tecno4_roe, Re 1600: mu = 1/1600,
    Pr 0.71, Mach 0.1) to t_end, FP64 or (fp32) FP32 reconstruction on the
    FP64 state with the network arithmetic `mlp`: kinetic energy and total
    entropy (FP64 audit sums per unit volume) every `every` time units, and
    the throughput of the steps between."""
    # fix_zmom: the reference's TeCNO fluxes subtract the z-momentum
    # dissipation from the y-momentum (kepes_tecno4_roe.cc:170), which leaves
    # the z momentum undamped; in 3D that grows w to O(1) within t < 0.4 and
    # the run fails, so the TGV uses the corrected dissipation
    cfg = tecno.make_cfg(n=(n, n, n), dims=3, lo=(0, 0, 0), hi=(2 * np.pi,) * 3, recon=recon,
                         flux="kepes_tecno4_roe", bc="periodic", cfl=cfl, model="ns_es",
                         viscosity=("constant", 1.0 / 1600.0), prandtl=0.71, fix_zmom=1,
                         fix_visc=fix_visc, precision=tecno.FP32 if fp32 else tecno.FP64,
                         mlp_dtype=mlp if mlp is not None else tecno.MLP_FP64,
                         sign_fallback=1 if fp32 else 0)
    s = tecno.Solver(cfg)
    if recon == "sp_weno_dl":
        s.load_net(net_file)
    s.set_state(tgv_state(n))
    vol = (2 * np.pi) ** 3
    a = s.audit()
    hist = [(0.0, a.kinetic_energy / vol, a.total_entropy / vol)]
    steps, secs = 0, 0.0
    every = min(every, t_end)
    t_next = every
    while t_next <= t_end + 1e-12:
        _, it0 = s.time()
        t0 = time.perf_counter()
        s.advance(t_next, 1)
        s.sync()
        secs += time.perf_counter() - t0
        t, it1 = s.time()
        steps += it1 - it0
        a = s.audit()
        if a.n_nonfinite:
            raise RuntimeError(f"TGV {recon}: non-finite state at t={t}")
        hist.append((t, a.kinetic_energy / vol, a.total_entropy / vol))
        print(f"# tgv {recon} n={n} t={t:.3f} steps={it1} KE={hist[-1][1]:.6e} "
              f"S={hist[-1][2]:.9e} ({n ** 3 * steps / secs:.3e} cell-updates/s)",
              file=sys.stderr, flush=True)
        t_next += every
    s.close()
    rises = [(hist[i][0], hist[i][2] - hist[i - 1][2]) for i in range(1, len(hist))
             if hist[i][2] > hist[i - 1][2]]
    return dict(recon=recon, n=n, t_end=t_end, steps=steps, seconds=secs, cfl=cfl,
                fix_visc=fix_visc, cell_updates_per_s=n ** 3 * steps / secs, history=hist,
                entropy_rises=len(rises), largest_rise=max([r[1] for r in rises], default=0.0))


def load_profile(kind, variant):
    p = os.path.join(ROOT, "profiles", f"{kind}_{variant}.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def load_traffic(variant):
    p = os.path.join(ROOT, "profiles", f"traffic_{variant}.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # each path starts from a fresh solver (allocation, state upload); its
    # first steps run while the GPU clocks come back up (stage 0.345 ms after
    # 5 untimed steps, 0.326 ms after 30), so 30 untimed steps precede the
    # timed ones
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="",
                    help="headline | fp64 | fp64dsp | mixedspw (profiling runs) | matrix (network "
                         "variants)")
    ap.add_argument("--ensemble", action="store_true",
                    help="run the 48-problem scalar ensemble (batched, hipGraph) and exit")
    ap.add_argument("--no-graph", action="store_true", help="ensemble without the hipGraph")
    ap.add_argument("--tgv", default="",
                    help="sp_weno | sp_weno_dl: 3D Navier-Stokes Taylor-Green vortex run and exit")
    ap.add_argument("--t-end", type=float, default=10.0, help="TGV end time")
    ap.add_argument("--net", default=NET_FILE, help="net_file for --tgv sp_weno_dl")
    ap.add_argument("--fix-visc", action="store_true",
                    help="--tgv: keep the K21/K31/K32 viscous blocks the reference overwrites")
    ap.add_argument("--cfl", type=float, default=0.5, help="--tgv: CFL number")
    ap.add_argument("--fp32", action="store_true",
                    help="--tgv: FP32 reconstruction on the FP64 state")
    ap.add_argument("--mlp", default="bf16", choices=["fp32", "bf16", "fp16"],
                    help="--tgv --fp32 sp_weno_dl: network arithmetic")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one plain 1 GiB device copy (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        a.fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print("copied 1073741824 bytes")
        return
    if args.gpus != 1:
        print(f"bench.py: one GPU per run; --gpus {args.gpus} is not supported (a run holds one "
              f"GPU's share of the grid; slab decomposition with halo exchange is not built)",
              file=sys.stderr)
        sys.exit(2)
    import tecno

    if args.tgv:
        n = args.n if args.n != 2048 else 256
        mlp = {"fp32": tecno.MLP_FP32, "bf16": tecno.MLP_BF16, "fp16": tecno.MLP_FP16}[args.mlp]
        r = tgv_bench(tecno, n, args.t_end, args.tgv, args.net, fix_visc=int(args.fix_visc),
                      cfl=args.cfl, fp32=args.fp32, mlp=mlp if args.fp32 else None)
        arith = "fp64"
        if args.fp32:
            arith = ("fp64 state, flux, eigensystem and viscous terms; fp32 reconstruction" +
                     (f", {args.mlp} network" if args.tgv == "sp_weno_dl" else ""))
        print(json.dumps({"metric": "cell-updates/s, 3D Navier-Stokes Taylor-Green vortex",
                          "value": r["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1,
                          "higher_is_better": True, "data": "synthetic (analytic initial state)",
                          "dtype": arith, "net_file": os.path.basename(args.net)
                          if args.tgv == "sp_weno_dl" else None,
                          "config": {"workload": f"TGV {n}^3 periodic [0,2pi]^3, Re 1600, Mach 0.1, "
                                                 f"ns_es, kepes_tecno4_roe (z-momentum dissipation "
                                                 f"fixed) + {args.tgv}, SSPRK3 CFL {args.cfl}, to "
                                                 f"t={args.t_end}" +
                                                 (", viscous blocks kept" if args.fix_visc else "")},
                          **r}), flush=True)
        return

    if args.ensemble:
        r = ensemble_bench(tecno, use_graph=not args.no_graph)
        # the same 24 problems with FP32 arithmetic on an FP64 state
        # (TECNO_FP32_STATE64): orders and throughput, outside the metric
        r64 = ensemble_bench(tecno, use_graph=not args.no_graph, precs=(tecno.FP32_STATE64,))
        r["fp32_state64"] = {k: r64[k] for k in ("cell_updates", "seconds", "cell_updates_per_s",
                                                 "orders", "groups")}
        print(json.dumps({"metric": "cell-updates/s, 48-problem scalar ensemble",
                          "value": r["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1,
                          "higher_is_better": True, "data": "synthetic (seeded initial states)",
                          "dtype": "fp64 and fp32 members", "hipgraph": not args.no_graph,
                          "config": {"workload": "6 x 1D Burgers sin x (t=1.5), 9 x 2D Burgers, 9 x 2D "
                                                 "linear advection (t=0.5), x {fp64, fp32}; "
                                                 "bur_tecno4 / lin_tecno4 + sp_weno, SSPRK3 CFL 0.5"},
                          **r}), flush=True)
        return

    n = args.n
    head_prec, head_rec, head_mlp = tecno.FP32, "sp_weno_dl", tecno.MLP_BF16
    res = {}
    # Every path is timed from the same developed state: the headline solver
    # advances the KH initial state by SPINUP_STEPS untimed steps (t ~ 0.03,
    # the shear layer has rolled up into resolved eddies), and each path then
    # starts from a copy of that state, takes `warmup` untimed steps and
    # `steps` timed ones.  The profiles under profiles/ are taken on the same
    # state (same spin-up), so their per-launch counters apply to this run.
    sh = make_solver(tecno, n, head_prec, head_rec, head_mlp)
    prim, _ = kh_state(n)
    sh.set_state(prim)
    sh.run_steps(SPINUP_STEPS, final_time=1.0, t_stamps=1)
    sh.sync()
    spun = sh.get_state()
    t_spun, _ = sh.time()
    sh.close()
    finals = {}
    # name: (precision, reconstruction, network arithmetic, weights)
    paths = {"headline": (head_prec, head_rec, head_mlp, "shipped"),
             "fp64_dsp": (tecno.FP64, "sp_weno_dl", tecno.MLP_FP64, "shipped"),
             "fp64_spweno": (tecno.FP64, "sp_weno", tecno.MLP_FP64, "shipped"),
             # config 3's comparison matrix: the network as an fp16 matrix-core
             # GEMM and as FP32 SIMT, and seeded Xavier weights
             "mixed_fp16_mfma": (tecno.FP32, "sp_weno_dl", tecno.MLP_FP16, "shipped"),
             "mixed_fp32_simt": (tecno.FP32, "sp_weno_dl", tecno.MLP_FP32, "shipped"),
             "mixed_bf16_mfma_xavier": (tecno.FP32, "sp_weno_dl", tecno.MLP_BF16, "xavier"),
             "fp64_dsp_xavier": (tecno.FP64, "sp_weno_dl", tecno.MLP_FP64, "xavier"),
             # SP-WENO with the headline's arithmetic split, no network
             "mixed_spweno": (tecno.FP32, "sp_weno", tecno.MLP_FP64, "shipped")}
    sel = {"": tuple(paths), "headline": ("headline",), "fp64": ("fp64_spweno",),
           "fp64dsp": ("fp64_dsp",), "mixedspw": ("mixed_spweno",),
           "matrix": tuple(paths)}[args.only]
    for name in sel:
        prec, rec, mlp, weights = paths[name]
        s = make_solver(tecno, n, prec, rec, mlp, weights)
        s.set_state(spun)
        if args.warmup:
            s.run_steps(args.warmup, final_time=1.0, t_stamps=1)
        res[name] = time_steps(s, n, args.steps)
        if args.only in ("", "matrix") and (rec == "sp_weno_dl" or
                                            name in ("fp64_spweno", "mixed_spweno")):
            finals[name] = s.get_state()
        s.close()
        print(f"# {name} from the spun-up state (t={t_spun:.4f}): {res[name]}", file=sys.stderr,
              flush=True)
    if args.only and args.only != "matrix":
        return
    h = res["headline"]
    b = res["fp64_spweno"]
    # algorithmic HBM bytes per stage launch: 8 state passes of NVAR values per
    # cell-update (read U^(k-1) [and U^n], write U^(k)), one third per stage;
    # the state is FP64 on both paths
    esize = 8
    bytes_per_launch = n * n * 8 * NVAR * esize / 3.0
    achieved = bytes_per_launch / (h["stage_ms"] * 1e-3) / 1e9
    traffic = load_traffic("headline")
    valu = load_profile("valu", "headline")
    mfma = load_profile("mfma", "headline")
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "kernel": "stage_kernel<Euler2Mixed<DSP,BF16>>", "stage_ms": h["stage_ms"],
            "bytes_per_launch_alg": bytes_per_launch}
    if valu:
        # wave64 VALU instructions per stage launch from the committed counter
        # profile of the same state, divided by this run's measured stage time:
        # the issue roof is 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles
        rate = valu["valu_insts_per_dispatch"] / (h["stage_ms"] * 1e-3)
        roof["valu"] = {"achieved": rate, "peak": VALU_PEAK, "unit": "wave64 VALU instructions/s",
                        "frac": rate / VALU_PEAK, "insts_per_launch": valu["valu_insts_per_dispatch"],
                        "source": "profiles/valu_headline.json"}
        if rate / VALU_PEAK > roof["frac"]:
            roof["bound"] = "valu"
    if mfma:
        # matrix-core use of the same stage from the committed counter profile
        # of the same state: busy fraction, and the bf16 FLOP rate at this
        # run's stage time against the dense bf16 peak
        roof["mfma"] = {k: mfma[k] for k in mfma if k != "command"}
        fl = mfma["bf16_flop_per_dispatch"] / (h["stage_ms"] * 1e-3)
        roof["mfma"].update({"achieved": fl / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                             "frac": fl / 2.5e15, "source": "profiles/mfma_headline.json"})
    out = {
        "metric": METRIC,
        "value": h["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 state + fp64 entropy-conservative flux, fp32 dissipation/reconstruction, "
                 "bf16 MLP (fp32 accumulation)",
        "data": "synthetic (seeded KH initial state; shipped DSP-WENO weights)",
        "config": {"workload": f"2D Euler Kelvin-Helmholtz {n}x{n} periodic, kepes_tecno4_roe + "
                               f"sp_weno_dl (DSP-WENO), SSPRK3 CFL 0.5",
                   "cells": n * n, "flux": "kepes_tecno4_roe", "reconstruct": "sp_weno_dl",
                   "precision": "fp64 state, cell records, entropy-conservative flux and "
                                "eigen-projection; fp32 jump-form reconstruction and dissipation; "
                                "DSP-WENO network as bf16 MFMA 32x32x16 with fp32 accumulation",
                   "spinup_steps": SPINUP_STEPS, "timed_from_time": t_spun,
                   "sign_fallback": 1},
        "fp64_spweno_cell_updates_per_s": b["cell_updates_per_s"],
        "fp64_spweno_stage_ms": b["stage_ms"],
        "speedup_vs_fp64_spweno": h["cell_updates_per_s"] / b["cell_updates_per_s"],
        "fp64_dsp_cell_updates_per_s": res["fp64_dsp"]["cell_updates_per_s"],
        "speedup_vs_fp64_dsp": h["cell_updates_per_s"] / res["fp64_dsp"]["cell_updates_per_s"],
        "sign_fallbacks_per_step": h["sign_fallbacks"] / max(args.steps, 1),
        "roofline": roof,
    }
    # error of the headline path against the FP64 DSP-WENO path (the same
    # method in the reference's arithmetic) on the KH state after the same
    # warmup + timed steps from the same state
    out["deviation_vs_fp64_dsp"] = state_deviation(finals["headline"], finals["fp64_dsp"],
                                                   args.warmup + args.steps)
    # every network variant against the FP64 path with the same weights, from
    # the same state after the same number of steps
    matrix = {}
    for name in ("headline", "mixed_fp16_mfma", "mixed_fp32_simt", "mixed_bf16_mfma_xavier",
                 "fp64_dsp", "fp64_dsp_xavier"):
        ref = "fp64_dsp_xavier" if name.endswith("xavier") else "fp64_dsp"
        row = {"network": {tecno.MLP_BF16: "bf16 MFMA", tecno.MLP_FP16: "fp16 MFMA",
                           tecno.MLP_FP32: "fp32 SIMT", tecno.MLP_FP64: "fp64 SIMT"}[paths[name][2]],
               "weights": paths[name][3], "stage_ms": res[name]["stage_ms"],
               "cell_updates_per_s": res[name]["cell_updates_per_s"],
               "sign_fallbacks_per_step": res[name]["sign_fallbacks"] / max(args.steps, 1)}
        if name != ref:
            row["deviation_vs_fp64"] = state_deviation(finals[name], finals[ref],
                                                       args.warmup + args.steps)
        matrix[name] = row
    out["network_matrix"] = matrix
    # SP-WENO with the headline's arithmetic split against FP64 SP-WENO: speed
    # and deviation on the KH state; its equal-error flag comes from the gate
    m = res["mixed_spweno"]
    out["mixed_spweno"] = {
        "stage_ms": m["stage_ms"], "cell_updates_per_s": m["cell_updates_per_s"],
        "speedup_vs_fp64_spweno": m["cell_updates_per_s"] / b["cell_updates_per_s"],
        "deviation_vs_fp64_spweno": state_deviation(finals["mixed_spweno"], finals["fp64_spweno"],
                                                    args.warmup + args.steps)}
    if args.only == "matrix":
        print(json.dumps(out["network_matrix"]), flush=True)
        return
    gate = l1_gate(tecno)
    out["l1_gate"] = gate
    # the metric's condition: L1 within 2% of the FP64 SP-WENO path on every
    # mesh of the vortex and the same convergence order (to 0.05)
    out["equal_l1_vs_fp64_spweno"] = gate["pass_vs_fp64_spweno"]
    out["equal_l1_vs_fp64_dsp"] = gate["pass_vs_fp64_dsp"]
    out["mixed_spweno"]["equal_l1_vs_fp64_spweno"] = gate["mixed_spweno_pass_vs_fp64_spweno"]
    print(f"# equal-error gate: {gate}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
