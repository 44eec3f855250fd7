"""Configurations of the golden cases (tests/golden/make_fixtures.py writes
the matching param.in files; these are the same fields as tecno_cfg)."""
import math
import os

import numpy as np

from tecno import TIME_SCHEME, make_cfg

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

_cache = {}


def fixtures(name):
    if name not in _cache:
        _cache[name] = np.load(os.path.join(GOLDEN, f"fixtures_{name}.npz"))
    return _cache[name]


def case(fx, key):
    d = fixtures(fx)
    return {k.split("/", 1)[1]: d[k] for k in d.files if k.split("/", 1)[0] == key}


SC1D = dict(system="scalar", lo=(-math.pi, 0.0, 0.0), hi=(math.pi, 1.0, 1.0), dims=1, cfl=0.5)
FINAL_1D = 0.5


def scalar_cfg(key, **kw):
    """tecno_cfg of the scalar golden case `key` (fixtures_scalar.npz)."""
    recs = {"spweno": "sp_weno", "spwenoc": "sp_wenoc", "spwenocorr2": "sp_weno_corr2",
            "eno2": "eno 2", "eno3": "eno 3", "dsp": "sp_weno_dl"}
    if key.startswith("sc1d_") and key[5:] in recs:
        c = make_cfg(n=(64, 2, 2), recon=recs[key[5:]], flux="lin_tecno4", **SC1D, **kw)
        return c, 0.5, 5
    if key.startswith("sc1d_rec_"):
        return make_cfg(n=(64, 2, 2), recon=key[9:], flux="lin_tecno4", **SC1D, **kw), 0.5, 5
    if key.startswith("sc1d_ts_"):
        return make_cfg(n=(64, 2, 2), recon="sp_weno", flux="lin_tecno4",
                        time_scheme=TIME_SCHEME[key[8:]], **SC1D, **kw), 0.5, 5
    if key == "linadv400_spweno":
        return make_cfg(n=(400, 2, 2), recon="sp_weno", flux="lin_tecno4", **SC1D, **kw), 0.5, 5
    if key.startswith("sc1d_"):
        flux = key[5:]
        return make_cfg(n=(64, 2, 2), recon="sp_weno", flux=flux, **SC1D, **kw), 0.5, 5
    if key.startswith("sc2d_"):
        tag = key[5:]
        base = dict(system="scalar", lo=(-1.0, -1.0, 0.0), hi=(1.0, 1.0, 1.0), n=(24, 20, 2), dims=2,
                    recon="sp_weno", cfl=0.5)
        if tag == "bur2d":
            return make_cfg(flux="bur_tecno4", **base, **kw), 0.3, 5
        if tag == "lin2d":
            return make_cfg(flux="lin_tecno4", vel=(1.0, 0.5, 0.0), **base, **kw), 0.3, 5
        if tag == "lin2d_wall":
            return make_cfg(flux="lin_tecno4", vel=(1.0, -0.7, 0.0),
                            bc=["wall", "neumann", "neumann", "wall", "periodic", "periodic"],
                            **base, **kw), 0.3, 5
    if key == "sc3d_lin":
        return make_cfg(system="scalar", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(10, 8, 6), dims=3,
                        recon="sp_weno", flux="lin_tecno4", vel=(1.0, 0.5, -0.25), cfl=0.5, **kw), 0.2, 5
    raise KeyError(key)


def euler_cfg(key, **kw):
    """tecno_cfg of the Euler golden case `key` (fixtures_euler.npz)."""
    recs = {"spweno": "sp_weno", "spwenoc": "sp_wenoc", "eno3": "eno 3", "dsp": "sp_weno_dl",
            "first": "first", "second": "second", "minmod": "minmod", "minabs": "minabs"}
    if key.startswith("eu2d_ts_"):
        return make_cfg(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(20, 16, 2), dims=2,
                        recon="sp_weno", flux="kepes_tecno4_roe", bc="neumann",
                        time_scheme=TIME_SCHEME[key[8:]], **kw), 0.25, 5
    if key.startswith("eu2d_riemann_"):
        return make_cfg(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(20, 16, 2), dims=2,
                        recon=recs[key[13:]], flux="kepes_tecno4_roe", bc="neumann", **kw), 0.25, 5
    if key.startswith("eu2d_kh_"):
        return make_cfg(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(24, 24, 2), dims=2,
                        recon=recs[key[8:]], flux="kepes_tecno4_roe", bc="periodic", **kw), 1.0, 5
    if key.startswith("eu2d_vortex_"):
        tag = key[12:]
        ec1 = tag.endswith("_ec1")
        flux = tag[:-4] if ec1 else tag
        return make_cfg(lo=(-5.0, -5.0, 0.0), hi=(5.0, 5.0, 1.0), n=(20, 20, 2), dims=2, recon="sp_weno",
                        flux=flux, bc="periodic", ec1_fix=int(ec1), **kw), 10.0, 5
    if key == "eu3d_tg":
        return make_cfg(lo=(0.0, 0.0, 0.0), hi=(2 * math.pi,) * 3, n=(8, 6, 6), dims=3, recon="sp_weno",
                        flux="kepes_tecno4_roe", bc="periodic", **kw), 1.0, 5
    if key.startswith("eu3d_tg_"):
        tag = key[8:]
        flux, rec, ts = "kepes_tecno4_roe", "sp_weno", "ssprk3"
        if tag in ("roe", "kep4", "keps", "keps4"):
            flux = tag
        elif tag in TIME_SCHEME:
            ts = tag
        else:
            rec = tag
        return make_cfg(lo=(0.0, 0.0, 0.0), hi=(2 * math.pi,) * 3, n=(8, 6, 6), dims=3, recon=rec,
                        flux=flux, bc="periodic", time_scheme=TIME_SCHEME[ts], **kw), 1.0, 5
    if key == "ns3d_es_rk4":
        return make_cfg(lo=(0.0, 0.0, 0.0), hi=(2 * math.pi,) * 3, n=(10, 8, 6), dims=3, recon="minabs",
                        flux="kepes_tecno4_roe", bc="periodic", model="ns_es",
                        viscosity=("constant", 1.0 / 1600), time_scheme=TIME_SCHEME["rk4"], **kw), 1.0, 5
    if key.startswith("ns3d_"):
        spec = {"ns3d_es_tg": ("ns_es", ("constant", 1.0 / 1600), "sp_weno", (10, 8, 6)),
                "ns3d_kep_tg": ("ns_kep", ("constant", 1.0 / 1600), "sp_weno", (8, 10, 6)),
                "ns3d_es_suth": ("ns_es", ("sutherland", 0.05, 71.0, 40.0), "sp_weno", (6, 10, 8)),
                "ns3d_es_dsp": ("ns_es", ("power", 0.02, 70.0, 0.7), "sp_weno_dl", (8, 8, 8))}[key]
        return make_cfg(lo=(0.0, 0.0, 0.0), hi=(2 * math.pi,) * 3, n=spec[3], dims=3, recon=spec[2],
                        flux="kepes_tecno4_roe", bc="periodic", model=spec[0], viscosity=spec[1],
                        **kw), 1.0, 5
    if key == "eu2d_wall":
        return make_cfg(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), n=(16, 12, 2), dims=2, recon="sp_weno",
                        flux="kepes_tecno4_roe",
                        bc=["wall", "wall", "neumann", "wall", "periodic", "periodic"], **kw), 0.25, 5
    raise KeyError(key)


def collapse(state, dims):
    """Fixture state (nv, nz, ny, nx) of a run with duplicated dummy layers ->
    the layers of a `dims`-dimensional grid (takes layer 0 of collapsed axes
    after checking the duplicates agree)."""
    s = np.asarray(state)
    if dims <= 2:
        assert np.array_equal(s[:, 0], s[:, -1])
        s = s[:, :1]
    if dims <= 1:
        assert np.array_equal(s[:, :, 0], s[:, :, -1])
        s = s[:, :, :1]
    return s


def euler_prim_for(state, dims):
    """(5, nz, ny, nx) primitive -> the variables a `dims`-D context holds
    (2D drops w: rho, u, v, p)."""
    s = collapse(state, dims)
    if dims == 2:
        assert np.all(s[3] == 0.0)
        s = s[[0, 1, 2, 4]]
    return s.reshape(s.shape[0], -1)


# ---------------------------------------------------------------- NS boundaries
def ns_box_cfg(n, hi, bc, **kw):
    """3D Navier-Stokes (ns_es, mu = 1/1600) on [0, hi] with kepes_tecno4_roe +
    sp_weno and the z-momentum fix."""
    return make_cfg(lo=(0.0, 0.0, 0.0), hi=hi, n=n, dims=3, recon="sp_weno", flux="kepes_tecno4_roe",
                    bc=bc, model="ns_es", viscosity=("constant", 1.0 / 1600), fix_zmom=1, **kw)


def tgv_sub_box(m, kz, rng, z_uniform=False):
    """(5, kz, m, m) primitive state on the cell centres of [0, pi]^2 x [0, kz*h]
    (h = pi/m): Taylor-Green velocity and pressure plus seeded noise, so that no
    symmetry other than the mirror planes of the test holds."""
    h = math.pi / m
    x = (np.arange(m) + 0.5) * h
    z = (np.arange(kz) + 0.5) * h
    Z, Y, X = np.meshgrid(z, x, x, indexing="ij")
    cz = np.ones_like(Z) if z_uniform else np.cos(Z)
    u = np.sin(X) * np.cos(Y) * cz
    v = -np.cos(X) * np.sin(Y) * cz
    w = np.zeros_like(X) if z_uniform else 0.3 * np.sin(Z) * np.cos(X)
    p = 10.0 + 0.0625 * (np.cos(2 * X) + np.cos(2 * Y)) * ((np.ones_like(Z) if z_uniform else np.cos(2 * Z)) + 2)
    rho = 1.0 + 0.1 * np.sin(X + 2 * Y)
    s = np.stack([rho, u, v, w, p])
    noise = rng.normal(size=s.shape) * 0.01
    if z_uniform:
        noise[:] = noise[:, :1]
        noise[3] = 0.0
    return s + noise


def mirror_full(sub, axes):
    """Periodic state of twice the extent along `axes` (0 = x, 1 = y, 2 = z of
    a (5, nz, ny, nx) array) whose mirror planes at 0 and the sub-box edge
    hold `sub` and its images with the normal velocity negated."""
    s = np.asarray(sub)
    for d in axes:
        ax = 3 - d  # array axis of direction d
        img = np.flip(s, axis=ax).copy()
        img[1 + d] = -img[1 + d]
        s = np.concatenate([s, img], axis=ax)
    return s
