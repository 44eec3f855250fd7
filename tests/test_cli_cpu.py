"""`tecno -i param.in` on the CPU (no GPU is touched with --dry-run): the
param.in reader and the expression evaluator against the reference's own
initial states (the golden fixtures hold them), the error paths, and
operator semantics."""
import glob
import os
import subprocess
import tempfile

import numpy as np
import pytest

from cases import fixtures

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dsp-weno_amd", "tecno")
CASES = os.path.join(ROOT, "tests", "golden", "cases")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="tecno CLI not built")


def dry(path):
    r = subprocess.run([CLI, "-i", path, "--dry-run"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = {}
    ic = {}
    for line in r.stdout.splitlines():
        k, *v = line.split()
        if k == "ic":
            ic[int(v[0])] = (float(v[1]), float(v[2]))
        else:
            out[k] = v
    return out, ic


@pytest.mark.parametrize("case,fx,key", [
    ("linadv400_spweno.param.in", "scalar", "linadv400_spweno"),
    ("kh_spweno_24.param.in", "euler", "eu2d_kh_spweno"),
    ("kh_dsp_24.param.in", "euler", "eu2d_kh_dsp")])
def test_initial_condition_matches_reference(case, fx, key):
    """sums and sums of squares of every IC field over the reference's cells
    (duplicate layers included) equal those of the reference's initial state"""
    out, ic = dry(os.path.join(CASES, case))
    init = np.asarray(fixtures(fx)[f"{key}/init"])
    assert len(ic) == init.shape[0]
    for v in range(init.shape[0]):
        s1, s2 = ic[v]
        r1, r2 = init[v].sum(), (init[v] ** 2).sum()
        assert abs(s1 - r1) <= 1e-12 * max(1.0, np.abs(init[v]).sum())
        assert abs(s2 - r2) <= 1e-12 * max(1.0, r2)


def test_parsed_fields():
    out, _ = dry(os.path.join(CASES, "kh_dsp_24.param.in"))
    assert out["system"] == ["euler"] and out["flux"] == ["101"] and out["recon"][0] == "2"
    assert out["reconstruction"] == ["sp_weno_dl"] and out["t_stamps"] == ["5"]
    out, _ = dry(os.path.join(CASES, "vortex_spweno_ooc.param.in"))
    assert out["levels"] == ["2"] and out["ooc"] == ["1"] and out["pdim"] == ["2"]
    out, _ = dry(os.path.join(CASES, "linadv400_spweno.param.in"))
    assert out["system"] == ["scalar"] and out["flux"] == ["3"] and out["exact"] == ["1"]


SCALAR = """grid
{{
   x_min   {xmin}
   x_max   {xmax}
   Nx_cell {{ 1 }}
   y_min   0.0
   y_max   1.0
   Ny_cell {{ 1 }}
   z_min   0.0
   z_max   1.0
   Nz_cell {{ 1 }}
}}
numeric
{{
   time_scheme  {ts}
   cfl          0.5
   max_iter     10
   final_time   0.1
   min_residue  1.0e-6
   reconstruct  {rec}
   net_file     none
}}
material
{{
   model       linadv
   vel         {{ 1 0 0 }}
   flux        {flux}
}}
constants
{{
   a   2.5   # a comment
   b   -1.0
}}
initial_condition
{{
   {expr}
}}
exact_soln
{{
   available   no
}}
boundary_condition
{{
   xleft {{ type {bc} }}
   xright {{ type {bc} }}
   yleft {{ type periodic }}
   yright {{ type periodic }}
   zleft {{ type periodic }}
   zright {{ type periodic }}
}}
output
{{
   OOC_study   no
   problem_dim 1
   time_stamps 1
   output_path .
}}
"""


def scalar_value(expr, x=0.5, **kw):
    args = dict(xmin=0.0, xmax=2 * x, ts="ssprk3", rec="sp_weno", flux="lin_tecno4", bc="periodic")
    args.update(kw)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "param.in")
        with open(p, "w") as f:
            f.write(SCALAR.format(expr=expr, **args))
        r = subprocess.run([CLI, "-i", p, "--dry-run"], capture_output=True, text=True, timeout=60)
    return r


@pytest.mark.parametrize("expr,value", [
    ("-2^2", -4.0), ("2^-1", 0.5), ("2^3^2", 512.0), ("7%3", 1.0), ("(3>2)+(1=1)+(2!=2)", 2.0),
    ("(x<1)&(x>0)", 1.0), ("(x>1)|(x<0)", 0.0), ("!0 + !3", 1.0), ("pow(a, 2)*b", -6.25),
    ("if(x>0.4, 10, 20)", 10.0), ("abs(b)+sqrt(16)+exp(0)+log(1)", 6.0), ("min(a,b)+max(a,b)", 1.5),
    ("pi*2/pi", 2.0), ("atan2(1,1)*4", np.pi), ("1e-2*100", 1.0), ("sin(x)^2+cos(x)^2", 1.0)])
def test_expression_semantics(expr, value):
    r = scalar_value(expr)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("ic 0")][0]
    assert abs(float(line.split()[2]) - value) <= 1e-14 * max(1.0, abs(value)), line


@pytest.mark.parametrize("kw,msg", [
    (dict(flux="keps4"), "unknown scalar flux"), (dict(ts="rk5"), "time_scheme"),
    (dict(rec="weno5"), "reconstruct weno5"), (dict(bc="dirichlet"), "dirichlet"),
    (dict(expr="sin(x"), "missing )"), (dict(expr="foo(x)"), "unknown function"),
    (dict(expr="x + q"), "unknown name q")])
def test_errors_are_reported(kw, msg):
    args = dict(expr="x")
    args.update(kw)
    expr = args.pop("expr")
    r = scalar_value(expr, **args)
    assert r.returncode != 0 and msg in r.stderr, (r.returncode, r.stderr)


@pytest.mark.parametrize("kw,field,value", [
    (dict(ts="rk1"), "time_scheme", "0"), (dict(ts="jameson_rk4"), "time_scheme", "2"),
    (dict(ts="rk4"), "time_scheme", "3"), (dict(rec="first"), "recon", "3 0 2"),
    (dict(rec="second"), "recon", "4 0 2"), (dict(rec="minmod"), "recon", "5 0 2"),
    (dict(rec="minabs"), "recon", "6 0 2")])
def test_numeric_section_names(kw, field, value):
    """Every time_scheme and reconstruct name of parameter.cc:111-212 maps
    to its tecno_cfg value."""
    r = scalar_value("x", **kw)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith(field + " ")][0]
    assert line == f"{field} {value}", line


def test_missing_file():
    r = subprocess.run([CLI, "-i", "/nonexistent/param.in", "--dry-run"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "cannot open" in r.stderr


@pytest.mark.parametrize("case", sorted(os.listdir(CASES)))
def test_library_param_reader_matches_command_line(case):
    """tecno_read_param (the library's Parameter::read) gives the fields the
    command line parses from the same file."""
    import tecno
    out, _ = dry(os.path.join(CASES, case))
    cfg, info = tecno.read_param(os.path.join(CASES, case))
    assert int(out["flux"][0]) == cfg.flux
    assert [int(v) for v in out["recon"]] == [cfg.recon, cfg.spweno_corr, cfg.eno_level]
    assert int(out["time_scheme"][0]) == cfg.time_scheme
    assert [int(v) for v in out["bc"]] == list(cfg.bc)
    assert float(out["cfl"][0]) == cfg.cfl
    assert int(out["levels"][0]) == info["levels"]
    assert float(out["final_time"][0]) == info["final_time"]
    assert int(out["t_stamps"][0]) == info["time_stamps"]
    assert (cfg.system == tecno.SYS_SCALAR) == (out["system"] == ["scalar"])
    assert cfg.dims == (int(out["pdim"][0]) if out["system"] == ["scalar"] else max(2, int(out["pdim"][0])))


def test_library_param_reader_errors(tmp_path):
    """a missing file, an unknown name and a null argument fail with the
    reader's message"""
    import tecno
    with pytest.raises(tecno.TecnoError, match="cannot open"):
        tecno.read_param(tmp_path / "missing.param.in")
    p = tmp_path / "bad.param.in"
    p.write_text(SCALAR.format(expr="x", xmin=0.0, xmax=1.0, ts="rk5", rec="sp_weno",
                               flux="lin_tecno4", bc="periodic"))
    with pytest.raises(tecno.TecnoError, match="time_scheme rk5"):
        tecno.read_param(p)
    with pytest.raises(tecno.TecnoError, match="mesh level"):
        tecno.read_param(os.path.join(CASES, "kh_spweno_24.param.in"), level=1)
    assert tecno.lib().tecno_read_param(str(p).encode(), 0, None, None, None, None, None, None, 0) == -1
    assert b"null argument" in tecno.lib().tecno_last_error()


def test_mesh_levels_without_ooc(tmp_path):
    """several meshes listed with OOC_study no: FV::run solves only the first
    (parameter.cc:636-639), so the reader reports one level, while every
    listed mesh can still be read"""
    import tecno
    src = open(os.path.join(CASES, "vortex_spweno_ooc.param.in")).read()
    assert "OOC_study   yes" in src
    p = tmp_path / "two_meshes.param.in"
    p.write_text(src.replace("OOC_study   yes", "OOC_study   no"))
    cfg, info = tecno.read_param(p)
    assert info["levels"] == 1
    cfg1, _ = tecno.read_param(p, level=1)
    assert (cfg.n[0], cfg1.n[0]) == (32, 64)
    out, _ = dry(str(p))
    assert int(out["levels"][0]) == 1 and int(out["listed"][0]) == 2
    _, info_ooc = tecno.read_param(os.path.join(CASES, "vortex_spweno_ooc.param.in"))
    assert info_ooc["levels"] == 2


REF_EXAMPLES = sorted(glob.glob("/root/reference/*/examples/**/param.in", recursive=True))


@pytest.mark.skipif(not REF_EXAMPLES, reason="reference examples not present on this machine")
@pytest.mark.parametrize("path", REF_EXAMPLES or ["none"])
def test_reference_examples_parse(path):
    """every param.in the reference ships under examples/ (enflo: 1D/2D Euler;
    enflo_scalar: Burgers, linear advection) is read by the library's reader
    and by the command line's dry run"""
    import tecno
    cfg, info = tecno.read_param(path)
    assert info["levels"] >= 1 and info["final_time"] > 0 and info["net_file"]
    out, ic = dry(path)
    assert int(out["flux"][0]) == cfg.flux and len(ic) in (1, 5)
