#!/usr/bin/env python3
"""Generate the golden fixtures under tests/golden/ from the reference build.

Runs only in a container that holds the reference tree: it drives
oracle/_ref/refh_sys and oracle/_ref/refh_scalar (the reference solvers
compiled from their own sources by oracle/Makefile, driven by
oracle/ref_harness.cc) on seeded inputs and stores inputs and outputs as
.npz.  The param.in files it writes under tests/golden/cases/ use the
param.in grammar of the reference (Parameter::read, enflo/src/parameter.cc).

    make -C oracle all ref && python3 tests/golden/make_fixtures.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFH_SYS = os.path.join(ROOT, "oracle", "_ref", "refh_sys")
REFH_SCA = os.path.join(ROOT, "oracle", "_ref", "refh_scalar")
REF_NET = "/root/reference/enflo/networks/dsp_weno.txt"
NET_FIXTURE = os.path.join(HERE, "dsp_weno_net.txt")
CASES = os.path.join(HERE, "cases")

BC_NAMES = ["xleft", "xright", "yleft", "yright", "zleft", "zright"]


def param_in(*, system, lo, hi, n, recon, flux, final_time, ic, bc="periodic",
             cfl=0.5, t_stamps=5, pdim=1, vel=(1, 0, 0), model=None,
             constants=(), net=NET_FIXTURE, ec1_fix="no", out=".",
             max_iter=50000, exact=None, ooc="no", euler_model="euler",
             viscosity=("constant", 0.0), prandtl=0.72, time_scheme="ssprk3"):
    """Text of a param.in file (sections in the order Parameter::read needs)."""
    L = ["grid", "{"]
    for d, ax in enumerate("xyz"):
        L += [f"   {ax}_min   {lo[d]!r}", f"   {ax}_max   {hi[d]!r}",
              f"   N{ax}_cell {{ {', '.join(str(v) for v in np.atleast_1d(n[d]))} }}"]
    L += ["}", "", "numeric", "{", f"   time_scheme  {time_scheme}", f"   cfl          {cfl!r}",
          f"   max_iter     {max_iter}", f"   final_time   {final_time!r}",
          "   min_residue  1.0e-6", f"   reconstruct  {recon}", f"   net_file     {net}", "}", ""]
    L += ["material", "{"]
    if system == "scalar":
        if model == "burgers":
            L += ["   model       burgers"]
        else:
            L += ["   model       linadv", f"   vel         {{ {vel[0]!r} {vel[1]!r} {vel[2]!r} }}"]
        L += [f"   flux        {flux}"]
    else:
        vm = viscosity[0]
        vl = [f"      model    {vm}", f"      mu_ref   {viscosity[1]!r}"]
        if vm == "sutherland":
            vl += [f"      T_ref    {viscosity[2]!r}", f"      T_0      {viscosity[3]!r}"]
        elif vm == "power":
            vl += [f"      T_ref    {viscosity[2]!r}", f"      omega    {viscosity[3]!r}"]
        L += ["   gamma       1.4", "   gas_const   1.0", "   viscosity", "   {"] + vl + [
              "   }", f"   prandtl     {prandtl!r}",
              f"   model       {euler_model}", f"   flux        {flux}", f"   EC1_fix     {ec1_fix}"]
    L += ["}", "", "constants", "{"]
    for k, v in constants:
        L += [f"   {k}   {v!r}"]
    L += ["}", "", "initial_condition", "{"]
    if system == "scalar":
        L += [f"   {ic}"]
    else:
        for name, e in zip(["density", "xvelocity", "yvelocity", "zvelocity", "pressure"], ic):
            L += [f"   {name}   {e}"]
    L += ["}", "", "exact_soln", "{"]
    if exact is None:
        L += ["   available   no"]
    elif system == "scalar":
        L += ["   available   yes", f"   {exact}"]
    else:
        L += ["   available   yes"]
        for name, e in zip(["density", "xvelocity", "yvelocity", "zvelocity", "pressure"], exact):
            L += [f"   {name}   {e}"]
    L += ["}", "", "boundary_condition", "{"]
    bcs = [bc] * 6 if isinstance(bc, str) else list(bc)
    for nm, b in zip(BC_NAMES, bcs):
        L += [f"   {nm}", "   {", f"      type {b}", "   }"]
    L += ["}", "", "output", "{", f"   OOC_study   {ooc}", f"   problem_dim {pdim}",
          f"   time_stamps {t_stamps}", f"   output_path {out}", "}", ""]
    return "\n".join(L)


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd} failed: {r.stderr[-2000:]}")


def read_run(path):
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(20), dtype=np.int32)
        nx, ny, nz, nvar, ns = (int(v) for v in hdr)
        N = nx * ny * nz
        nv = 1 if nvar == 1 else 5
        data = np.frombuffer(f.read(), dtype=np.float64)
    init = data[:nv * N].reshape(nv, nz, ny, nx)
    rest = data[nv * N:].reshape(ns, 2 + nv * N)
    times, dts = rest[:, 0].copy(), rest[:, 1].copy()
    states = rest[:, 2:].reshape(ns, nv, nz, ny, nx)
    return init.copy(), times, dts, states.copy()


def harness_run(binary, text, nsteps, tmp):
    p = os.path.join(tmp, "param.in")
    with open(p, "w") as f:
        f.write(text)
    out = os.path.join(tmp, "run.bin")
    run([binary, "run", p, str(nsteps), out])
    return read_run(out)


def harness_recon(binary, text, st, ng, tmp):
    p = os.path.join(tmp, "param.in")
    with open(p, "w") as f:
        f.write(text)
    fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    np.ascontiguousarray(st, dtype=np.float64).tofile(fin)
    run([binary, "recon", p, str(st.shape[0]), str(ng), fin, fout])
    return np.fromfile(fout, dtype=np.float64).reshape(-1, 2)


def harness_flux(binary, text, data, axis, nout, tmp):
    p = os.path.join(tmp, "param.in")
    with open(p, "w") as f:
        f.write(text)
    fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    np.ascontiguousarray(data, dtype=np.float64).tofile(fin)
    run([binary, "flux", p, str(data.shape[0]), str(axis), fin, fout])
    return np.fromfile(fout, dtype=np.float64).reshape(-1, nout)


def harness_ent(text, prim, tmp):
    p = os.path.join(tmp, "param.in")
    with open(p, "w") as f:
        f.write(text)
    fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    np.ascontiguousarray(prim, dtype=np.float64).tofile(fin)
    run([REFH_SYS, "ent", p, str(prim.shape[0]), fin, fout])
    return np.fromfile(fout, dtype=np.float64).reshape(-1, 5)


# ---------------------------------------------------------------- stencils
def recon_stencils(rng, n):
    """4-cell stencils covering every SP-WENO / DSP-WENO branch: smooth
    polynomial data, extrema, jumps, constant pairs, theta == 1, psi == -1 and
    values of both signs and magnitudes."""
    out = []
    # smooth: samples of random cubics on a grid
    x = np.arange(4.0)
    for _ in range(n // 4):
        c = rng.normal(size=4) * rng.choice([1e-3, 1.0, 10.0])
        h = rng.choice([1e-3, 1e-2, 0.1, 1.0])
        out.append(c[0] + c[1] * h * x + c[2] * (h * x) ** 2 + c[3] * (h * x) ** 3)
    # random: uniform and heavy tailed
    out.extend(rng.uniform(-2, 2, size=(n // 4, 4)))
    out.extend(rng.standard_cauchy(size=(n // 8, 4)))
    # structured: step functions, flat pairs, equal jumps (theta = 1)
    for _ in range(n // 8):
        a, b = rng.normal(size=2)
        k = rng.integers(0, 6)
        if k == 0:
            s = [a, a, b, b]
        elif k == 1:
            s = [a, b, b, a]
        elif k == 2:
            d = rng.normal()
            s = [a, a + d, a + 2 * d, a + 3 * d]  # theta = 1 both sides
        elif k == 3:
            d = rng.normal()
            s = [a, a + d, a + 2 * d, b]
        elif k == 4:
            s = [a, b, b + 1e-14, a]  # |del1| below the 1e-13 threshold
        else:
            d = rng.normal()
            s = [a - d, a, a + d, a]  # psi = -1 type symmetric extrema
        out.append(np.array(s, dtype=np.float64))
    # tiny jumps around the 1e-15 DSP skip threshold
    for _ in range(n // 16):
        a = rng.normal()
        e = rng.choice([5e-16, 2e-15, 1e-12])
        out.append(np.array([a - e * rng.normal(), a, a + e, a + 2 * e * rng.normal()]))
    st = np.array(out, dtype=np.float64)
    return st[rng.permutation(len(st))]


def euler_stencils(rng, n, ng, threeD):
    """n random face stencils of 2*ng primitive states (rho,u,v,w,p) with
    entropy variables computed by the reference's prim_to_ent."""
    rho = rng.uniform(0.3, 2.0, size=(n, 2 * ng))
    p = rng.uniform(0.2, 3.0, size=(n, 2 * ng))
    u = rng.uniform(-1.0, 1.0, size=(n, 2 * ng, 3))
    if not threeD:
        u[..., 2] = 0.0
    # a fraction of smooth stencils (near-equal states exercise the logavg
    # series branch) and exactly equal pairs
    sm = rng.random(n) < 0.4
    base = rng.uniform(0.5, 1.5, size=(n, 1))
    ramp = np.arange(2 * ng)[None, :] * rng.uniform(-1e-3, 1e-3, size=(n, 1))
    rho[sm] = (base + ramp)[sm]
    p[sm] = (base * 1.3 + 0.5 * ramp)[sm]
    eq = rng.random(n) < 0.05
    rho[eq, ng] = rho[eq, ng - 1]
    p[eq, ng] = p[eq, ng - 1]
    P = np.concatenate([rho[..., None], u, p[..., None]], axis=-1)  # n, 2ng, 5
    return P


# ---------------------------------------------------------------- cases
def scalar_cases(tmp):
    fx = {}
    base = dict(system="scalar", lo=(-np.pi, 0.0, 0.0), hi=(np.pi, 1.0, 1.0), final_time=0.5,
                ic="pow(sin(x),4)", pdim=1, out=tmp)
    # config-1 family at a small grid, every reconstruction
    for rec, tag in [("sp_weno", "spweno"), ("sp_wenoc", "spwenoc"), ("sp_weno_corr2", "spwenocorr2"),
                     ("eno 2", "eno2"), ("eno 3", "eno3"), ("sp_weno_dl", "dsp")]:
        text = param_in(n=(64, 2, 2), recon=rec, flux="lin_tecno4", **base)
        init, t, dt, st = harness_run(REFH_SCA, text, 40, tmp)
        ns = len(t)
        fx[f"sc1d_{tag}"] = dict(init=init, times=t, dts=dt, states=st[[0, 4, ns - 1]],
                                 steps=np.array([1, 5, ns]))
    # linear advection of sin^4 x at Nx=400 to t=0.5 (65 steps), final state
    text = param_in(n=(400, 2, 2), recon="sp_weno", flux="lin_tecno4", **base)
    with open(os.path.join(CASES, "linadv400_spweno.param.in"), "w") as f:
        f.write(param_in(n=(400, 2, 2), recon="sp_weno", flux="lin_tecno4", system="scalar",
                         lo=(-np.pi, 0.0, 0.0), hi=(np.pi, 1.0, 1.0), final_time=0.5,
                         ic="pow(sin(x),4)", pdim=1, out=".", exact="pow(sin(x-t),4)"))
    init, t, dt, st = harness_run(REFH_SCA, text, 1000, tmp)
    fx["linadv400_spweno"] = dict(init=init, times=t, dts=dt, states=st[[0, -1]], steps=np.array([1, len(t)]))
    # other fluxes in 1D (lin_ec4 / lin_tecno / bur_*)
    for flux, model, ic in [("lin_tecno", None, "pow(sin(x),4)"), ("lin_ec4", None, "pow(sin(x),4)"),
                            ("bur_tecno4", "burgers", "sin(x)"), ("bur_tecno", "burgers", "sin(x)"),
                            ("bur_ec4", "burgers", "sin(x)")]:
        b = dict(base)
        b["ic"] = ic
        text = param_in(n=(64, 2, 2), recon="sp_weno", flux=flux, model=model, **b)
        init, t, dt, st = harness_run(REFH_SCA, text, 20, tmp)
        fx[f"sc1d_{flux}"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    # 2D scalar: burgers and linadv with seeded Fourier noise, neumann/wall/periodic
    rng = np.random.default_rng(20240611)
    amps = rng.normal(size=8) * 0.05
    modes = " + ".join(f"a{k}*sin({k + 1}*pi*x + {k % 3}*pi*y)" for k in range(8))
    consts = [(f"a{k}", float(amps[k])) for k in range(8)]
    for tag, flux, model, vel, bc in [("bur2d", "bur_tecno4", "burgers", (1, 0, 0), "periodic"),
                                      ("lin2d", "lin_tecno4", None, (1.0, 0.5, 0.0), "periodic"),
                                      ("lin2d_wall", "lin_tecno4", None, (1.0, -0.7, 0.0),
                                       ["wall", "neumann", "neumann", "wall", "periodic", "periodic"])]:
        text = param_in(system="scalar", lo=(-1.0, -1.0, 0.0), hi=(1.0, 1.0, 1.0), n=(24, 20, 2),
                        recon="sp_weno", flux=flux, model=model, vel=vel, final_time=0.3,
                        ic=f"sin(pi*x)*sin(pi*y) + {modes}", constants=consts, pdim=2, bc=bc, out=tmp)
        init, t, dt, st = harness_run(REFH_SCA, text, 8, tmp)
        fx[f"sc2d_{tag}"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    # 3D scalar linadv
    text = param_in(system="scalar", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(10, 8, 6),
                    recon="sp_weno", flux="lin_tecno4", vel=(1.0, 0.5, -0.25), final_time=0.2,
                    ic="sin(2*pi*x)*cos(2*pi*y)*sin(2*pi*z) + 0.3*cos(2*pi*(x+z))", pdim=3, out=tmp)
    init, t, dt, st = harness_run(REFH_SCA, text, 4, tmp)
    fx["sc3d_lin"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    # the MUSCL-family reconstructions and the other time schemes, 1D
    for rec in ["first", "second", "minmod", "minabs"]:
        text = param_in(n=(64, 2, 2), recon=rec, flux="lin_tecno4", **base)
        init, t, dt, st = harness_run(REFH_SCA, text, 40, tmp)
        ns = len(t)
        fx[f"sc1d_rec_{rec}"] = dict(init=init, times=t, dts=dt, states=st[[0, 4, ns - 1]],
                                     steps=np.array([1, 5, ns]))
    for ts in ["rk1", "jameson_rk4", "rk4"]:
        text = param_in(n=(64, 2, 2), recon="sp_weno", flux="lin_tecno4", time_scheme=ts, **base)
        init, t, dt, st = harness_run(REFH_SCA, text, 40, tmp)
        ns = len(t)
        fx[f"sc1d_ts_{ts}"] = dict(init=init, times=t, dts=dt, states=st[[0, 4, ns - 1]],
                                   steps=np.array([1, 5, ns]))
    return fx


def quadrant_ic(rng):
    rho = rng.uniform(0.5, 1.5, 4)
    u = rng.uniform(-0.5, 0.5, 4)
    v = rng.uniform(-0.5, 0.5, 4)
    p = rng.uniform(0.3, 1.5, 4)
    consts = []
    for q in range(4):
        consts += [(f"r{q}", float(rho[q])), (f"u{q}", float(u[q])), (f"v{q}", float(v[q])),
                   (f"p{q}", float(p[q]))]
    # quadrants: 0 = (x>.5,y>.5), 1 = (x<=.5,y>.5), 2 = (x<=.5,y<=.5), 3 = (x>.5,y<=.5)
    sel = ["(x>0.5)*(y>0.5)", "(x<=0.5)*(y>0.5)", "(x<=0.5)*(y<=0.5)", "(x>0.5)*(y<=0.5)"]
    def e(var):
        return " + ".join(f"{var}{q}*{sel[q]}" for q in range(4))
    return consts, [e("r"), e("u"), e("v"), "0.0", e("p")]


def euler_cases(tmp):
    fx = {}
    rng = np.random.default_rng(7)
    consts, ic = quadrant_ic(rng)
    for rec, tag in [("sp_weno", "spweno"), ("sp_wenoc", "spwenoc"), ("eno 3", "eno3"),
                     ("sp_weno_dl", "dsp")]:
        text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(20, 16, 2),
                        recon=rec, flux="kepes_tecno4_roe", final_time=0.25, ic=ic,
                        constants=consts, bc="neumann", pdim=2, out=tmp)
        init, t, dt, st = harness_run(REFH_SYS, text, 6, tmp)
        fx[f"eu2d_riemann_{tag}"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    # Kelvin-Helmholtz (periodic), sp_weno_dl and sp_weno
    kh_consts = [("dL", 2.0), ("uL", -0.5), ("dR", 1.0), ("uR", 0.5), ("pp", 2.5)]
    a = np.random.default_rng(11).normal(size=8)
    kh_consts += [(f"a{k}", float(a[k])) for k in range(8)]
    pert = " + ".join(f"a{k}*sin(2*pi*{k + 1}*x)" for k in range(8))
    kh_ic = ["dL*(abs(y-0.5)<0.25) + dR*(abs(y-0.5)>=0.25)",
             "uL*(abs(y-0.5)<0.25) + uR*(abs(y-0.5)>=0.25)",
             f"0.01*({pert})", "0.0", "pp"]
    for rec, tag in [("sp_weno", "spweno"), ("sp_weno_dl", "dsp")]:
        text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(24, 24, 2),
                        recon=rec, flux="kepes_tecno4_roe", final_time=1.0, ic=kh_ic,
                        constants=kh_consts, bc="periodic", pdim=2, out=tmp)
        init, t, dt, st = harness_run(REFH_SYS, text, 6, tmp)
        fx[f"eu2d_kh_{tag}"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
        with open(os.path.join(CASES, f"kh_{tag}_24.param.in"), "w") as f:
            f.write(param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(24, 24, 2),
                             recon=rec, flux="kepes_tecno4_roe", final_time=1.0, ic=kh_ic,
                             constants=kh_consts, bc="periodic", pdim=2, out=".", max_iter=6))
    # isentropic vortex (smooth, periodic), every TeCNO flux variant
    vc = [("xc", 0.0), ("yc", 0.0), ("gam", 1.4), ("beta", 5.0), ("M", 1.0)]
    T = "(1.0 - (gam-1)*(pow(beta,2)/(8*gam*pi*pi))*exp((1-pow(x-xc,2)-pow(y-yc,2))))"
    v_ic = [f"pow({T},1/(gam-1))",
            "M - (beta/(2*pi))*(y-yc)*exp(0.5*(1-pow(x-xc,2)-pow(y-yc,2)))",
            "(beta/(2*pi))*(x-xc)*exp(0.5*(1-pow(x-xc,2)-pow(y-yc,2)))", "0.0",
            f"pow({T},gam/(gam-1))"]
    for flux in ["kepes_tecno4_roe", "kepes_tecno_roe", "kepes_tecno4_rusanov", "kepes_tecno_rusanov",
                 "roe_tecno4_roe", "roe_tecno_roe", "kep_tecno_roe"]:
        for ec1 in (["no", "yes"] if flux == "kepes_tecno4_roe" else ["no"]):
            text = param_in(system="euler", lo=(-5.0, -5.0, 0.0), hi=(5.0, 5.0, 1.0), n=(20, 20, 2),
                            recon="sp_weno", flux=flux, final_time=10.0, ic=v_ic, constants=vc,
                            bc="periodic", pdim=2, ec1_fix=ec1, out=tmp)
            init, t, dt, st = harness_run(REFH_SYS, text, 3, tmp)
            tag = flux + ("_ec1" if ec1 == "yes" else "")
            fx[f"eu2d_vortex_{tag}"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    with open(os.path.join(CASES, "vortex_spweno_ooc.param.in"), "w") as f:
        f.write(param_in(system="euler", lo=(-5.0, -5.0, 0.0), hi=(5.0, 5.0, 1.0),
                         n=([32, 64], [32, 64], [2, 2]), recon="sp_weno", flux="kepes_tecno4_roe",
                         final_time=1.0, ic=v_ic, constants=vc, bc="periodic", pdim=2, out=".",
                         exact=v_ic[:1] + [e.replace("(x-xc)", "(x-t-xc)").replace("pow(x-xc", "pow(x-t-xc")
                                           for e in v_ic[1:2]] + v_ic[2:], ooc="yes"))
    # 3D Taylor-Green-like flow with w != 0 (exercises the y/z momentum
    # dissipation line of the reference), periodic, and a wall/neumann box
    tg = ["1.0 + 0.2*sin(x)*sin(y)*cos(z)", "sin(x)*cos(y)*cos(z)", "-cos(x)*sin(y)*cos(z)",
          "0.3*sin(z)*cos(x)", "10.0 + 0.25*(cos(2*x)+cos(2*y))*(cos(2*z)+2)"]
    text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, n=(8, 6, 6),
                    recon="sp_weno", flux="kepes_tecno4_roe", final_time=1.0, ic=tg, pdim=3,
                    bc="periodic", out=tmp)
    init, t, dt, st = harness_run(REFH_SYS, text, 3, tmp)
    fx["eu3d_tg"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    # Navier-Stokes Taylor-Green vortex (Mach 0.1 pressure field, Re = 1600
    # at unit scales) on uneven periodic boxes: ES and KE viscous fluxes,
    # constant and Sutherland viscosity, sp_weno and sp_weno_dl.  Ny >= Nz on
    # every box: the reference's node-weight diagnostic indexes with Nz in
    # place of Ny (fv.cc:1147,1154) and leaves its array when Nz > Ny
    tgv = ["1.0", "sin(x)*cos(y)*cos(z)", "-cos(x)*sin(y)*cos(z)", "0.0",
           "71.42857142857143 + 0.0625*(cos(2*x)+cos(2*y))*(cos(2*z)+2)"]
    for tag, em, visc, rec, n in [("ns3d_es_tg", "ns_es", ("constant", 1.0 / 1600), "sp_weno", (10, 8, 6)),
                                  ("ns3d_kep_tg", "ns_kep", ("constant", 1.0 / 1600), "sp_weno", (8, 10, 6)),
                                  ("ns3d_es_suth", "ns_es", ("sutherland", 0.05, 71.0, 40.0), "sp_weno", (6, 10, 8)),
                                  ("ns3d_es_dsp", "ns_es", ("power", 0.02, 70.0, 0.7), "sp_weno_dl", (8, 8, 8))]:
        text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, n=n, recon=rec,
                        flux="kepes_tecno4_roe", final_time=1.0, ic=tgv, pdim=3, bc="periodic", out=tmp,
                        euler_model=em, viscosity=visc)
        init, t, dt, st = harness_run(REFH_SYS, text, 3, tmp)
        fx[tag] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(16, 12, 2),
                    recon="sp_weno", flux="kepes_tecno4_roe", final_time=0.25, ic=ic, constants=consts,
                    bc=["wall", "wall", "neumann", "wall", "periodic", "periodic"], pdim=2, out=tmp)
    init, t, dt, st = harness_run(REFH_SYS, text, 4, tmp)
    fx["eu2d_wall"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    # fluxes without dissipation on reconstructed jumps, on the vortex
    for flux in ["roe", "kep4", "keps", "keps4"]:
        text = param_in(system="euler", lo=(-5.0, -5.0, 0.0), hi=(5.0, 5.0, 1.0), n=(20, 20, 2),
                        recon="sp_weno", flux=flux, final_time=10.0, ic=v_ic, constants=vc,
                        bc="periodic", pdim=2, out=tmp)
        init, t, dt, st = harness_run(REFH_SYS, text, 3, tmp)
        fx[f"eu2d_vortex_{flux}"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    # MUSCL-family reconstructions on the quadrant Riemann problem
    for rec in ["first", "second", "minmod", "minabs"]:
        text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(20, 16, 2),
                        recon=rec, flux="kepes_tecno4_roe", final_time=0.25, ic=ic,
                        constants=consts, bc="neumann", pdim=2, out=tmp)
        init, t, dt, st = harness_run(REFH_SYS, text, 6, tmp)
        fx[f"eu2d_riemann_{rec}"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    # time schemes on the quadrant Riemann problem
    for ts in ["rk1", "jameson_rk4", "rk4"]:
        text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(20, 16, 2),
                        recon="sp_weno", flux="kepes_tecno4_roe", final_time=0.25, ic=ic,
                        constants=consts, bc="neumann", pdim=2, out=tmp, time_scheme=ts)
        init, t, dt, st = harness_run(REFH_SYS, text, 6, tmp)
        fx[f"eu2d_ts_{ts}"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    # 3D: the extra fluxes, reconstructions and time schemes on the
    # Taylor-Green-like flow, and classical RK4 on a viscous case
    for tag, flux, rec, ts in [("roe", "roe", "sp_weno", "ssprk3"), ("keps4", "keps4", "sp_weno", "ssprk3"),
                               ("kep4", "kep4", "sp_weno", "ssprk3"), ("keps", "keps", "sp_weno", "ssprk3"),
                               ("minmod", "kepes_tecno4_roe", "minmod", "ssprk3"),
                               ("second", "kepes_tecno4_roe", "second", "ssprk3"),
                               ("jameson_rk4", "kepes_tecno4_roe", "sp_weno", "jameson_rk4")]:
        text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, n=(8, 6, 6),
                        recon=rec, flux=flux, final_time=1.0, ic=tg, pdim=3, bc="periodic", out=tmp,
                        time_scheme=ts)
        init, t, dt, st = harness_run(REFH_SYS, text, 3, tmp)
        fx[f"eu3d_tg_{tag}"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, n=(10, 8, 6), recon="minabs",
                    flux="kepes_tecno4_roe", final_time=1.0, ic=tgv, pdim=3, bc="periodic", out=tmp,
                    euler_model="ns_es", viscosity=("constant", 1.0 / 1600), time_scheme="rk4")
    init, t, dt, st = harness_run(REFH_SYS, text, 3, tmp)
    fx["ns3d_es_rk4"] = dict(init=init, times=t, dts=dt, states=st[[0, len(t) - 1]], steps=np.array([1, len(t)]))
    return fx


def point_fixtures(tmp):
    fx = {}
    rng = np.random.default_rng(1234)
    st = recon_stencils(rng, 8000)
    fx["recon_stencils"] = dict(st=st)
    for rec, tag in [("sp_weno", "spweno"), ("sp_wenoc", "spwenoc"), ("sp_weno_corr2", "spwenocorr2"),
                     ("sp_weno_dl", "dsp")]:
        text = param_in(system="scalar", lo=(-1.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(8, 2, 2),
                        recon=rec, flux="lin_tecno4", final_time=0.1, ic="x", out=tmp)
        lr = harness_recon(REFH_SCA, text, st, 2, tmp)
        fx[f"recon_{tag}"] = dict(l=lr[:, 0], r=lr[:, 1])
    for rec in ["first", "second", "minmod", "minabs"]:
        text = param_in(system="scalar", lo=(-1.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(8, 2, 2),
                        recon=rec, flux="lin_tecno4", final_time=0.1, ic="x", out=tmp)
        lr = harness_recon(REFH_SCA, text, st, 2, tmp)
        fx[f"recon_{rec}"] = dict(l=lr[:, 0], r=lr[:, 1])
    # ENO (stencils of 2k values)
    for k in (2, 3):
        base4 = recon_stencils(rng, 4000)
        stk = np.concatenate([base4, base4[:, 3:] + rng.normal(size=(len(base4), 2))], axis=1)[:, :2 * k]
        text = param_in(system="scalar", lo=(-1.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(8, 2, 2),
                        recon=f"eno {k}", flux="lin_tecno4", final_time=0.1, ic="x", out=tmp)
        lr = harness_recon(REFH_SCA, text, stk, k, tmp)
        fx[f"recon_eno{k}"] = dict(st=stk, l=lr[:, 0], r=lr[:, 1])
    # scalar fluxes
    s1 = recon_stencils(rng, 2000)
    fx["sflux_stencils"] = dict(st=s1)
    for flux, model in [("lin_ec", None), ("lin_ec4", None), ("lin_tecno", None), ("lin_tecno4", None),
                        ("bur_ec", "burgers"), ("bur_ec4", "burgers"), ("bur_tecno", "burgers"),
                        ("bur_tecno4", "burgers")]:
        text = param_in(system="scalar", lo=(-1.0, -2.0, 0.0), hi=(1.0, 1.0, 0.5), n=(8, 6, 4),
                        recon="sp_weno", flux=flux, model=model, vel=(1.25, -0.5, 0.75),
                        final_time=0.1, ic="x", out=tmp)
        outs = [harness_flux(REFH_SCA, text, s1, ax, 1, tmp)[:, 0] for ax in range(3)]
        fx[f"sflux_{flux}"] = dict(F=np.stack(outs))
    for flux, model in [("lin_tecno4", None), ("bur_tecno", "burgers")]:
        for rec in ["second", "minmod", "minabs"]:
            text = param_in(system="scalar", lo=(-1.0, -2.0, 0.0), hi=(1.0, 1.0, 0.5), n=(8, 6, 4),
                            recon=rec, flux=flux, model=model, vel=(1.25, -0.5, 0.75),
                            final_time=0.1, ic="x", out=tmp)
            outs = [harness_flux(REFH_SCA, text, s1, ax, 1, tmp)[:, 0] for ax in range(3)]
            fx[f"sfluxrec_{flux}_{rec}"] = dict(F=np.stack(outs))
    # Euler fluxes on random stencils (entropy variables from prim_to_ent)
    ent_text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(4, 4, 4),
                        recon="sp_weno", flux="kepes_tecno4_roe", final_time=0.1,
                        ic=["1", "0", "0", "0", "1"], out=tmp)
    nf = 400
    for threeD in (False, True):
        P = euler_stencils(rng, nf, 2, threeD)
        V = harness_ent(ent_text, P.reshape(-1, 5), tmp).reshape(nf, 4, 5)
        fx[f"ent_{'3d' if threeD else '2d'}"] = dict(P=P, V=V)
        data = np.concatenate([P, V], axis=-1)
        for flux in ["kepes_tecno4_roe", "kepes_tecno_roe", "kepes_tecno4_rusanov", "kepes_tecno_rusanov",
                     "roe_tecno4_roe", "roe_tecno_roe", "kep_tecno_roe", "kepec", "kepec4", "roe_ec",
                     "roe_ec4", "kep", "roe", "kep4", "keps", "keps4"]:
            recs = [("sp_weno", "spweno")] if flux != "kepes_tecno4_roe" else \
                [("sp_weno", "spweno"), ("sp_wenoc", "spwenoc"), ("sp_weno_dl", "dsp"),
                 ("first", "first"), ("second", "second"), ("minmod", "minmod"), ("minabs", "minabs")]
            for rec, rtag in recs:
                for ec1 in (["no", "yes"] if flux == "kepes_tecno4_roe" and rec == "sp_weno" else ["no"]):
                    text = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 2.0, 0.5), n=(8, 6, 4),
                                    recon=rec, flux=flux, final_time=0.1, ec1_fix=ec1,
                                    ic=["1", "0", "0", "0", "1"], out=tmp)
                    outs = [harness_flux(REFH_SYS, text, data, ax, 5, tmp) for ax in range(3)]
                    key = f"eflux_{'3d' if threeD else '2d'}_{flux}_{rtag}" + ("_ec1" if ec1 == "yes" else "")
                    fx[key] = dict(F=np.stack(outs))
    return fx


def main():
    if not (os.path.exists(REFH_SYS) and os.path.exists(REFH_SCA)):
        sys.exit("build the reference harness first: make -C oracle ref")
    os.makedirs(CASES, exist_ok=True)
    # the shipped DSP-WENO weights (enflo/networks/dsp_weno.txt), rewritten in
    # the same net_file format: 4 layers of 5x5 weights and 5 biases
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    buf = ctypes.create_string_buffer(1 << 16)
    if lib.tor_net_load(REF_NET.encode(), buf) != 0 or lib.tor_net_save(NET_FIXTURE.encode(), buf) != 0:
        sys.exit("cannot load the shipped network")
    with tempfile.TemporaryDirectory() as tmp:
        for name, gen in [("fixtures_point.npz", point_fixtures), ("fixtures_scalar.npz", scalar_cases),
                          ("fixtures_euler.npz", euler_cases)]:
            fx = gen(tmp)
            flat = {}
            for case, d in fx.items():
                for k, v in d.items():
                    flat[f"{case}/{k}"] = np.asarray(v)
            np.savez_compressed(os.path.join(HERE, name), **flat)
            print(name, len(fx), "cases", os.path.getsize(os.path.join(HERE, name)) // 1024, "KiB")


if __name__ == "__main__":
    main()
