#!/usr/bin/env python3
"""Generate the diagnostics and output-file fixtures from the reference build.

Runs only in a container that holds the reference tree: it drives
oracle/_ref/refh_sys and oracle/_ref/refh_scalar (the reference solvers
compiled from their own sources by oracle/Makefile, driven by
oracle/ref_harness.cc) and stores what they compute and write:

  fixtures_globals.npz   per step the quantities of FV::compute_global
                         (compute_global.cc:10-89: time, total entropy,
                         kinetic energy, KE_diss_pr, KE_diss_stress,
                         enstrophy; scalar: time, total entropy) and the
                         nodal vorticity of find_vorticity (find_vort.cc:9-29)
                         before the first and after the last step, for the
                         golden cases of make_fixtures.py (same param.in text)
  cli/<case>/            a param.in (output_path .) and the files the
                         reference's own FV::run writes for it under
                         SOLN_FILES_<reconstruct>/ (output.cc:10-172,
                         compute_global.cc:80-88, compute_ooc.cc)

    make -C oracle all ref && python3 tests/golden/make_globals_fixtures.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_fixtures import (CASES, REFH_SCA, REFH_SYS, param_in, quadrant_ic,  # noqa: E402
                           run)

CLI = os.path.join(HERE, "cli")


def read_globals(path, system):
    with open(path, "rb") as f:
        nx, ny, nz, ns = (int(v) for v in np.frombuffer(f.read(16), dtype=np.int32))
        data = np.frombuffer(f.read(), dtype=np.float64)
    ncol = 2 if system == "scalar" else 6
    rows = data[:ns * ncol].reshape(ns, ncol).copy()
    out = dict(rows=rows, shape=np.array([nx, ny, nz]))
    if system != "scalar":
        nn = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        v = data[ns * ncol:]
        assert v.size == 2 * nn
        out["vort0"] = v[:nn].reshape(3, nz + 1, ny + 1, nx + 1).copy()
        out["vort1"] = v[nn:].reshape(3, nz + 1, ny + 1, nx + 1).copy()
    return out


def harness_globals(binary, text, nsteps, tmp, system):
    p = os.path.join(tmp, "param.in")
    with open(p, "w") as f:
        f.write(text)
    out = os.path.join(tmp, "glob.bin")
    run([binary, "globals", p, str(nsteps), out])
    return read_globals(out, system)


def kh_case(rec, out, max_iter=50000):
    kh_consts = [("dL", 2.0), ("uL", -0.5), ("dR", 1.0), ("uR", 0.5), ("pp", 2.5)]
    a = np.random.default_rng(11).normal(size=8)
    kh_consts += [(f"a{k}", float(a[k])) for k in range(8)]
    pert = " + ".join(f"a{k}*sin(2*pi*{k + 1}*x)" for k in range(8))
    kh_ic = ["dL*(abs(y-0.5)<0.25) + dR*(abs(y-0.5)>=0.25)",
             "uL*(abs(y-0.5)<0.25) + uR*(abs(y-0.5)>=0.25)", f"0.01*({pert})", "0.0", "pp"]
    return param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(24, 24, 2), recon=rec,
                    flux="kepes_tecno4_roe", final_time=1.0, ic=kh_ic, constants=kh_consts,
                    bc="periodic", pdim=2, out=out, max_iter=max_iter)


def riemann_case(rec, out, bc="neumann", n=(20, 16, 2), max_iter=50000):
    consts, ic = quadrant_ic(np.random.default_rng(7))
    return param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=n, recon=rec,
                    flux="kepes_tecno4_roe", final_time=0.25, ic=ic, constants=consts, bc=bc, pdim=2,
                    out=out, max_iter=max_iter)


TGV = ["1.0", "sin(x)*cos(y)*cos(z)", "-cos(x)*sin(y)*cos(z)", "0.0",
       "71.42857142857143 + 0.0625*(cos(2*x)+cos(2*y))*(cos(2*z)+2)"]
TG = ["1.0 + 0.2*sin(x)*sin(y)*cos(z)", "sin(x)*cos(y)*cos(z)", "-cos(x)*sin(y)*cos(z)",
      "0.3*sin(z)*cos(x)", "10.0 + 0.25*(cos(2*x)+cos(2*y))*(cos(2*z)+2)"]


def ns_case(model, n, out, max_iter=50000, rec="sp_weno"):
    return param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, n=n, recon=rec,
                    flux="kepes_tecno4_roe", final_time=1.0, ic=TGV, pdim=3, bc="periodic", out=out,
                    euler_model=model, viscosity=("constant", 1.0 / 1600), max_iter=max_iter)


def globals_cases(tmp):
    fx = {}
    fx["eu2d_kh_spweno"] = harness_globals(REFH_SYS, kh_case("sp_weno", tmp), 6, tmp, "euler")
    fx["eu2d_kh_dsp"] = harness_globals(REFH_SYS, kh_case("sp_weno_dl", tmp), 6, tmp, "euler")
    fx["eu2d_riemann_spweno"] = harness_globals(REFH_SYS, riemann_case("sp_weno", tmp), 6, tmp, "euler")
    consts, ic = quadrant_ic(np.random.default_rng(7))
    wall = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(16, 12, 2),
                    recon="sp_weno", flux="kepes_tecno4_roe", final_time=0.25, ic=ic, constants=consts,
                    bc=["wall", "wall", "neumann", "wall", "periodic", "periodic"], pdim=2, out=tmp)
    fx["eu2d_wall"] = harness_globals(REFH_SYS, wall, 4, tmp, "euler")
    fx["ns3d_es_tg"] = harness_globals(REFH_SYS, ns_case("ns_es", (10, 8, 6), tmp), 3, tmp, "euler")
    fx["ns3d_kep_tg"] = harness_globals(REFH_SYS, ns_case("ns_kep", (8, 10, 6), tmp), 3, tmp, "euler")
    eu3 = param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, n=(8, 6, 6),
                   recon="sp_weno", flux="kepes_tecno4_roe", final_time=1.0, ic=TG, pdim=3,
                   bc="periodic", out=tmp)
    fx["eu3d_tg"] = harness_globals(REFH_SYS, eu3, 3, tmp, "euler")
    sc1 = param_in(system="scalar", lo=(-np.pi, 0.0, 0.0), hi=(np.pi, 1.0, 1.0), n=(64, 2, 2),
                   recon="sp_weno", flux="lin_tecno4", final_time=0.5, ic="pow(sin(x),4)", pdim=1,
                   out=tmp)
    fx["sc1d_spweno"] = harness_globals(REFH_SCA, sc1, 10, tmp, "scalar")
    return fx


def sod_case(out):
    """1D Euler shock tube on an Nx x 2 x 2 grid (the reference's 1D examples
    are stored that way: problem_dim 1)"""
    ic = ["1.0*(x<0.5) + 0.125*(x>=0.5)", "0.0", "0.0", "0.0", "1.0*(x<0.5) + 0.1*(x>=0.5)"]
    return param_in(system="euler", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(64, 2, 2),
                    recon="sp_weno", flux="kepes_tecno4_roe", final_time=0.05, ic=ic,
                    bc=["neumann", "neumann", "periodic", "periodic", "periodic", "periodic"],
                    pdim=1, out=out, t_stamps=2)


def cli_cases():
    """case name -> param.in text with output_path ."""
    sc = param_in(system="scalar", lo=(-np.pi, 0.0, 0.0), hi=(np.pi, 1.0, 1.0), n=(64, 2, 2),
                  recon="sp_weno", flux="lin_tecno4", final_time=0.2, ic="pow(sin(x),4)", pdim=1,
                  out=".", t_stamps=2)
    with open(os.path.join(CASES, "vortex_spweno_ooc.param.in")) as f:
        ooc = f.read()
    return {"kh_spweno_24": kh_case("sp_weno", ".", max_iter=6),
            "riemann_spweno_20x16": riemann_case("sp_weno", ".", max_iter=6),
            "sod1d_spweno_64": sod_case("."),
            "ns3d_es_8": ns_case("ns_es", (8, 8, 8), ".", max_iter=3),
            "linadv1d_spweno_64": sc,
            "vortex_spweno_ooc": ooc}


def main():
    if not (os.path.exists(REFH_SYS) and os.path.exists(REFH_SCA)):
        sys.exit("build the reference harness first: make -C oracle ref")
    with tempfile.TemporaryDirectory() as tmp:
        fx = globals_cases(tmp)
    flat = {f"{c}/{k}": np.asarray(v) for c, d in fx.items() for k, v in d.items()}
    dst = os.path.join(HERE, "fixtures_globals.npz")
    np.savez_compressed(dst, **flat)
    print("fixtures_globals.npz", len(fx), "cases", os.path.getsize(dst) // 1024, "KiB")
    shutil.rmtree(CLI, ignore_errors=True)
    for name, text in cli_cases().items():
        d = os.path.join(CLI, name)
        os.makedirs(d)
        with open(os.path.join(d, "param.in"), "w") as f:
            f.write(text)
        binary = REFH_SCA if "model       linadv" in text or "model       burgers" in text else REFH_SYS
        r = subprocess.run([binary, "full", "param.in"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: {r.stderr[-2000:]}")
        sol = [e for e in os.listdir(d) if e.startswith("SOLN_FILES_")]
        print(name, sol, sorted(os.listdir(os.path.join(d, sol[0]))) if sol else None)


if __name__ == "__main__":
    main()
