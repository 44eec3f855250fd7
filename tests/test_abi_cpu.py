"""CPU checks of the product boundary: libtecno.so loads and exports every
symbol include/tecno.h declares; host-side argument validation."""
import ctypes as C
import os
import re

import pytest

import tecno

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "tecno.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tecno_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = tecno.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_version_and_defaults():
    lib = tecno.lib()
    assert b"gfx950" in lib.tecno_version()
    cfg = tecno.Cfg()
    lib.tecno_cfg_default(C.byref(cfg))
    assert cfg.dims == 2 and cfg.flux == tecno.EULER_FLUX["kepes_tecno4_roe"]


@pytest.mark.parametrize("bad", [dict(dims=4), dict(bc="dirichlet"), dict(n=(1, 8, 2)), dict(cfl=-1.0)])
def test_create_rejects_bad_configs(bad):
    """Invalid configurations fail with TECNO_EINVAL before any device call."""
    kw = dict(n=(16, 16, 2), dims=2)
    kw.update(bad)
    cfg = tecno.make_cfg(**kw)
    h = C.c_void_p()
    rc = tecno.lib().tecno_create(C.byref(cfg), C.byref(h))
    assert rc == -1
    assert tecno.lib().tecno_last_error()


@pytest.mark.parametrize("kw,msg", [
    (dict(bc=["periodic", "wall", "periodic", "periodic", "periodic", "periodic"]), "paired"),
    (dict(time_scheme=7), "time scheme"),
    (dict(recon="minmod", precision=tecno.FP32), "FP32 Euler path"),
    (dict(flux="roe", precision=tecno.FP32), "FP32 Euler path"),
    (dict(model="ns_es", dims=2), "3D grids"),
    (dict(model="ns_es", dims=3, n=(8, 8, 1), viscosity=("constant", 1e-3)), "at least 2 cells"),
    (dict(precision=tecno.FP32_STATE64), "FP32_STATE64"),
    (dict(system="scalar", flux="lin_tecno4", recon="eno 2", precision=tecno.FP32_STATE64),
     "FP32_STATE64"),
    (dict(precision=3), "precision must be"),
    (dict(dims=3, n=(8, 8, 8), flux="roe", precision=tecno.FP32), "3D FP32 path"),
    (dict(dims=3, n=(8, 8, 8), recon="eno 2", precision=tecno.FP32), "3D FP32 path")])
def test_create_rejects_with_message(kw, msg):
    """The messages name the rule that failed (periodic pairing as
    parameter.cc:570-581, the supported precision combinations, ...)."""
    args = dict(n=(16, 16, 2), dims=2)
    args.update(kw)
    cfg = tecno.make_cfg(**args)
    h = C.c_void_p()
    assert tecno.lib().tecno_create(C.byref(cfg), C.byref(h)) == -1
    assert msg in tecno.lib().tecno_last_error().decode(), tecno.lib().tecno_last_error()


def test_cfg_struct_layout_matches_header():
    """The ctypes mirror of tecno_cfg has the size of the C struct
    (sizeof(tecno_cfg) == 272 with gcc/hipcc on x86-64)."""
    assert C.sizeof(tecno.Cfg) == 272
