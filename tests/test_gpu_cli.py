"""`tecno -i param.in` end to end on the GPU: the reference's OOC tables
(the linear-advection sin^4 figures of the reference run) and the solution / globals files."""
import math
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import line_extract as LE
from cases import euler_cfg, euler_prim_for, case
from tecno import TIME_SCHEME, Solver, make_cfg

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dsp-weno_amd", "tecno")
CASES = os.path.join(ROOT, "tests", "golden", "cases")


def run_cli(param, cwd):
    r = subprocess.run([CLI, "-i", param, "--quiet"], capture_output=True, text=True, timeout=300,
                       cwd=cwd)
    assert r.returncode == 0, r.stderr
    return r


def test_linadv_sin4_ooc_table():
    """lin_tecno4 + sp_weno, sin^4 on [-pi, pi], Nx = 100..800, t = 0.5: the
    reference reports L1 = 1.75268e-05 at Nx = 400 and orders 3.17, 3.21, 3.25"""
    text = open(os.path.join(CASES, "linadv400_spweno.param.in")).read()
    text = text.replace("Nx_cell { 400 }", "Nx_cell { 100 200 400 800 }")
    text = text.replace("Ny_cell { 2 }", "Ny_cell { 2 2 2 2 }").replace("Nz_cell { 2 }", "Nz_cell { 2 2 2 2 }")
    text = text.replace("OOC_study   no", "OOC_study   yes")
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "param.in")
        open(p, "w").write(text)
        run_cli(p, d)
        rows = open(os.path.join(d, "SOLN_FILES_sp_weno", "l1_err.dat")).read().splitlines()[3:]
    toks = [r.split() for r in rows if r.strip()]
    assert [t[0] for t in toks] == ["100", "200", "400", "800"]
    assert toks[2][3] == "1.75268e-05"
    orders = [float(t[4]) for t in toks[1:]]
    assert [round(o, 2) for o in orders] == [3.17, 3.21, 3.25], orders


def test_kh_files_and_state():
    """KH 24x24 (x2 layers) to t = 1 with 5 output stamps: sol0000..sol0005.vtk,
    one globals line per step, and the last file holds the state the library
    reaches for the same run"""
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "param.in")
        text = open(os.path.join(CASES, "kh_spweno_24.param.in")).read()
        open(p, "w").write(re.sub(r"max_iter\s+\d+", "max_iter 50000", text))
        run_cli(p, d)
        out = os.path.join(d, "SOLN_FILES_sp_weno")
        files = sorted(f for f in os.listdir(out) if f.startswith("sol"))
        assert files == [f"sol{k:04d}.vtk" for k in range(6)]
        glob = [l for l in open(os.path.join(out, "globals.dat")).read().splitlines() if l.strip()]
        text = open(os.path.join(out, "sol0005.vtk")).read()
        line, _ = LE.extract(os.path.join(out, "sol0005.vtk"), 0.0, 1.0, 24)
    assert "DIMENSIONS 25 25 3" in text and "CELL_DATA  1152" in text
    dens = text.split("SCALARS density float 1\nLOOKUP_TABLE default\n")[1].split("VECTORS")[0].split()
    rho_file = np.array([float(v) for v in dens]).reshape(2, 24, 24)
    assert np.array_equal(rho_file[0], rho_file[1])
    # line_extract along y = z = 0.5 reads the file's own header: the 6-digit
    # SPACING of the reference's writer (1/24 -> 0.0416667) puts the probe
    # points i/24 just below face i, in cell i - 1, as the probe filter does;
    # y = 0.5 and z = 0.5 fall in cell row 11 (or 12) and layer 0 (or 1)
    sp = [float(v) for v in text.split("SPACING")[1].split()[:3]]
    col = [min(max(int(np.floor((i / 24) / sp[0])), 0), 23) for i in range(24)]
    row = int(np.floor(0.5 / sp[1]))
    assert np.array_equal(line[:, 1], rho_file[int(np.floor(0.5 / sp[2])), row, col])
    cfg, final, stamps = euler_cfg("eu2d_kh_spweno")
    s = Solver(cfg)
    s.set_state(euler_prim_for(case("euler", "eu2d_kh_spweno")["init"], 2))
    it, t = s.advance(final, stamps)
    assert len(glob) == it
    rho = s.get_state()[0].reshape(24, 24)
    assert np.allclose(rho_file[0], rho, rtol=1e-5, atol=0)
    assert float(glob[0].split()[0]) == 0.0


@pytest.mark.parametrize("ts,rec", [("rk4", "minmod"), ("jameson_rk4", "second"), ("rk1", "minabs"),
                                    ("ssprk3", "first")])
def test_time_schemes_and_reconstructions_end_to_end(ts, rec):
    """`tecno -i param.in` with the other time_scheme / reconstruct names of
    parameter.cc (the sin^4 linear-advection file at Nx = 100): the last solution file holds
    the state the library reaches for the same configuration (6 digits, as
    the file is written)."""
    text = open(os.path.join(CASES, "linadv400_spweno.param.in")).read()
    text = text.replace("Nx_cell { 400 }", "Nx_cell { 100 }")
    text = re.sub(r"time_scheme\s+\S+", f"time_scheme  {ts}", text)
    text = re.sub(r"reconstruct\s+\S+", f"reconstruct  {rec}", text)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "param.in")
        open(p, "w").write(text)
        run_cli(p, d)
        out = os.path.join(d, f"SOLN_FILES_{rec}")
        last = sorted(f for f in os.listdir(out) if f.startswith("sol"))[-1]
        _, _, _, fields = LE.read_vtk(os.path.join(out, last))
    u_file = fields["solution"].reshape(2, 2, 100)[0, 0]
    cfg = make_cfg(system="scalar", n=(100, 2, 2), dims=1, lo=(-math.pi, 0, 0), hi=(math.pi, 1, 1),
                   recon=rec, flux="lin_tecno4", time_scheme=TIME_SCHEME[ts])
    s = Solver(cfg)
    x = -math.pi + (np.arange(100) + 0.5) * 2 * math.pi / 100
    s.set_state(np.sin(x) ** 4)
    s.advance(0.5, 5)
    assert np.allclose(u_file, s.get_state()[0], rtol=2e-5, atol=2e-6)
