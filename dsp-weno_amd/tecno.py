"""Python host binding of libtecno.so (the MI355X TeCNO finite-volume library).

Mirrors the reference's operator seams (FV::num_flux, FV::reconstruct,
Neural_Network::forward, FV::run — enflo/src/fv.h:24-183) through the C ABI
declared in include/tecno.h.  The library is loaded from this directory; a
missing or unbuilt library raises at import time (there is no fallback).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtecno.so")

# ---- enums (include/tecno.h) ------------------------------------------------
SYS_SCALAR, SYS_EULER = 0, 1
MODEL_LINADV, MODEL_BURGERS, MODEL_EULER, MODEL_NS_KEP, MODEL_NS_ES = 0, 1, 10, 11, 12
SCALAR_FLUX = {"lin_ec": 0, "lin_ec4": 1, "lin_tecno": 2, "lin_tecno4": 3,
               "bur_ec": 4, "bur_ec4": 5, "bur_tecno": 6, "bur_tecno4": 7}
EULER_FLUX = {"kepes_tecno_roe": 100, "kepes_tecno4_roe": 101, "kepes_tecno_rusanov": 102,
              "kepes_tecno4_rusanov": 103, "roe_tecno_roe": 104, "roe_tecno4_roe": 105,
              "kep_tecno_roe": 106, "kepec": 120, "kepec4": 121, "roe_ec": 122,
              "roe_ec4": 123, "kep": 124, "roe": 125, "kep4": 126, "keps": 127, "keps4": 128}
REC_SP_WENO, REC_ENO, REC_SP_WENO_DL, REC_FIRST = 0, 1, 2, 3
REC_SECOND, REC_MINMOD, REC_MINABS = 4, 5, 6
BC = {"periodic": 0, "dirichlet": 1, "neumann": 2, "wall": 3}
TS_RK1, TS_SSPRK3, TS_JAMESON_RK4, TS_RK4 = 0, 1, 2, 3
TIME_SCHEME = {"rk1": TS_RK1, "ssprk3": TS_SSPRK3, "jameson_rk4": TS_JAMESON_RK4, "rk4": TS_RK4}
FP64, FP32, FP32_STATE64 = 0, 1, 2
MLP_FP64, MLP_FP32, MLP_BF16, MLP_FP16 = 0, 1, 2, 3


def recon_from_name(name):
    """Map a param.in `reconstruct` value to (recon, spweno_corr, eno_level)
    (enflo/src/parameter.cc:155-212)."""
    toks = name.split()
    if toks[0] == "sp_weno":
        return REC_SP_WENO, 0, 2
    if toks[0] == "sp_wenoc":
        return REC_SP_WENO, 1, 2
    if toks[0] == "sp_weno_corr2":
        return REC_SP_WENO, 2, 2
    if toks[0] == "sp_weno_dl":
        return REC_SP_WENO_DL, 0, 2
    if toks[0] == "eno":
        return REC_ENO, 0, int(toks[1])
    if toks[0] == "first":
        return REC_FIRST, 0, 2
    if toks[0] == "second":
        return REC_SECOND, 0, 2
    if toks[0] == "minmod":
        return REC_MINMOD, 0, 2
    if toks[0] == "minabs":
        return REC_MINABS, 0, 2
    raise ValueError(f"unknown reconstruction {name!r}")


class Cfg(C.Structure):
    _fields_ = [
        ("system", C.c_int32), ("model", C.c_int32), ("flux", C.c_int32), ("recon", C.c_int32),
        ("spweno_corr", C.c_int32), ("eno_level", C.c_int32), ("time_scheme", C.c_int32),
        ("precision", C.c_int32), ("mlp_dtype", C.c_int32), ("dims", C.c_int32),
        ("n", C.c_int32 * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("bc", C.c_int32 * 6), ("cfl", C.c_double), ("vel", C.c_double * 3),
        ("gamma", C.c_double), ("gas_const", C.c_double), ("prandtl", C.c_double),
        ("mu_model", C.c_int32), ("mu_ref", C.c_double), ("T_ref", C.c_double),
        ("T_0", C.c_double), ("omega", C.c_double), ("ec1_fix", C.c_int32),
        ("fix_zmom", C.c_int32), ("sign_fallback", C.c_int32), ("fix_edges", C.c_int32),
        ("fix_visc", C.c_int32), ("reserved", C.c_int32 * 6),
    ]


MU_CONSTANT, MU_SUTHERLAND, MU_POWER = 0, 1, 2


def make_cfg(*, system="euler", flux="kepes_tecno4_roe", recon="sp_weno", n=(32, 32, 2),
             lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), dims=2, bc="periodic", cfl=0.5,
             vel=(1.0, 0.0, 0.0), model=None, precision=FP64, mlp_dtype=MLP_FP64,
             gamma=1.4, gas_const=1.0, ec1_fix=0, fix_zmom=0, sign_fallback=0, fix_edges=0,
             time_scheme=TS_SSPRK3, viscosity=("constant", 0.0), prandtl=0.72, fix_visc=0):
    """viscosity: ("constant", mu_ref) | ("sutherland", mu_ref, T_ref, T_0) |
    ("power", mu_ref, T_ref, omega), as param.in's material/viscosity block"""
    c = Cfg()
    if system == "scalar":
        c.system = SYS_SCALAR
        c.flux = SCALAR_FLUX[flux]
        if model is None:
            model = "burgers" if flux.startswith("bur") else "linadv"
        c.model = MODEL_BURGERS if model == "burgers" else MODEL_LINADV
    else:
        c.system = SYS_EULER
        c.flux = EULER_FLUX[flux]
        c.model = {"euler": MODEL_EULER, "ns_kep": MODEL_NS_KEP, "ns_es": MODEL_NS_ES}[model or "euler"]
    c.recon, c.spweno_corr, c.eno_level = recon_from_name(recon)
    c.time_scheme = time_scheme
    c.precision = precision
    c.mlp_dtype = mlp_dtype
    c.dims = dims
    for d in range(3):
        c.n[d] = int(n[d])
        c.lo[d] = float(lo[d])
        c.hi[d] = float(hi[d])
        c.vel[d] = float(vel[d])
    bcs = [bc] * 6 if isinstance(bc, str) else list(bc)
    for i, b in enumerate(bcs):
        c.bc[i] = BC[b] if isinstance(b, str) else int(b)
    c.cfl = cfl
    c.gamma = gamma
    c.gas_const = gas_const
    c.prandtl = prandtl
    vm = viscosity[0]
    c.mu_model = {"constant": MU_CONSTANT, "sutherland": MU_SUTHERLAND, "power": MU_POWER}[vm]
    c.mu_ref = float(viscosity[1])
    if vm == "sutherland":
        c.T_ref, c.T_0 = float(viscosity[2]), float(viscosity[3])
    elif vm == "power":
        c.T_ref, c.omega = float(viscosity[2]), float(viscosity[3])
    c.ec1_fix = int(ec1_fix)
    c.fix_zmom = int(fix_zmom)
    c.sign_fallback = int(sign_fallback)
    c.fix_edges = int(fix_edges)
    c.fix_visc = int(fix_visc)
    return c


class Globals(C.Structure):
    _fields_ = [("time", C.c_double), ("total_entropy", C.c_double),
                ("kinetic_energy", C.c_double), ("ke_diss_pr", C.c_double),
                ("ke_diss_stress", C.c_double), ("enstrophy", C.c_double)]


class EnsGroup(C.Structure):
    _fields_ = [("precision", C.c_int), ("dims", C.c_int), ("n_members", C.c_int),
                ("n_tiles", C.c_int), ("grid", C.c_int), ("tile_x", C.c_int), ("tile_y", C.c_int),
                ("n_candidates", C.c_int), ("cand_x", C.c_int * 8), ("cand_y", C.c_int * 8),
                ("tuned_ms", C.c_double * 8)]


class Audit(C.Structure):
    _fields_ = [("total_entropy", C.c_double), ("kinetic_energy", C.c_double),
                ("mass", C.c_double), ("momentum", C.c_double * 3), ("energy", C.c_double),
                ("max_speed", C.c_double), ("n_nonfinite", C.c_int64),
                ("n_sign_fallback", C.c_int64)]


class Adam(C.Structure):
    _fields_ = [("lr", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double), ("weight_decay", C.c_double), ("step", C.c_int64)]


_lib = None


def lib():
    """Load libtecno.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with make -C dsp-weno_amd")
        L = C.CDLL(LIB_PATH)
        P, D, I64, I32 = C.c_void_p, C.c_double, C.c_int64, C.c_int
        dp = C.POINTER(C.c_double)
        L.tecno_create.argtypes = [C.POINTER(Cfg), C.POINTER(P)]
        L.tecno_read_param.argtypes = [C.c_char_p, I32, C.POINTER(Cfg), C.POINTER(C.c_int), dp,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, I32]
        L.tecno_destroy.argtypes = [P]
        L.tecno_destroy.restype = None
        L.tecno_load_net.argtypes = [P, C.c_char_p]
        L.tecno_set_net.argtypes = [P, I32, C.POINTER(C.c_int32), dp, dp]
        L.tecno_get_net.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int32), dp, dp]
        L.tecno_num_cells.argtypes = [P]
        L.tecno_num_cells.restype = I64
        L.tecno_num_vars.argtypes = [P]
        L.tecno_set_state.argtypes = [P, dp, I64]
        L.tecno_get_state.argtypes = [P, dp, I64]
        L.tecno_step.argtypes = [P, D, dp]
        L.tecno_advance.argtypes = [P, D, I32, I32, C.POINTER(C.c_int), dp]
        L.tecno_run_steps.argtypes = [P, I32, D, I32]
        L.tecno_get_time.argtypes = [P, dp, C.POINTER(C.c_int)]
        L.tecno_synchronize.argtypes = [P]
        L.tecno_audit_state.argtypes = [P, C.POINTER(Audit)]
        L.tecno_audit_state_noalloc.argtypes = [P, C.POINTER(Audit)]
        L.tecno_l1_error.argtypes = [P, dp, I64, dp]
        L.tecno_compute_globals.argtypes = [P, C.POINTER(Globals)]
        L.tecno_vorticity.argtypes = [P, dp, I64]
        L.tecno_reconstruct_batch.argtypes = [P, dp, I64, dp, dp]
        L.tecno_flux_batch.argtypes = [P, dp, dp, I64, I32, dp]
        L.tecno_mlp_forward_batch.argtypes = [P, dp, I64, dp, dp]
        L.tecno_train_init.argtypes = [P, C.POINTER(Adam)]
        L.tecno_train_step.argtypes = [P, dp, I64, I32, dp]
        L.tecno_train_grad.argtypes = [P, dp, I64, I32, dp, dp]
        L.tecno_make_train_rows.argtypes = [dp, dp, I64, dp]
        L.tecno_train_set_data.argtypes = [P, dp, I64]
        L.tecno_train_epoch.argtypes = [P, I64, C.c_uint64, I32, dp]
        L.tecno_save_net.argtypes = [P, C.c_char_p]
        L.tecno_init_net.argtypes = [P, I32, C.POINTER(C.c_int32), C.c_uint64]
        L.tecno_trainer_epoch_seeds.argtypes = [I32, C.POINTER(C.c_int32), C.c_uint64, I32,
                                                C.POINTER(C.c_uint64)]
        L.tecno_ens_create.argtypes = [C.POINTER(Cfg), I32, C.POINTER(P)]
        L.tecno_ens_destroy.argtypes = [P]
        L.tecno_ens_destroy.restype = None
        L.tecno_ens_size.argtypes = [P]
        L.tecno_ens_member.argtypes = [P, I32]
        L.tecno_ens_member.restype = P
        L.tecno_ens_advance.argtypes = [P, dp, I32, I32, I32, C.POINTER(C.c_int64)]
        L.tecno_ens_autotune.argtypes = [P, I32]
        L.tecno_ens_num_groups.argtypes = [P]
        L.tecno_ens_group_info.argtypes = [P, I32, C.POINTER(EnsGroup)]
        L.tecno_timing_enable.argtypes = [P, I32]
        L.tecno_timing_get.argtypes = [P, dp, C.POINTER(C.c_int64)]
        L.tecno_last_error.restype = C.c_char_p
        L.tecno_version.restype = C.c_char_p
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class TecnoError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise TecnoError(f"libtecno error {rc}: {lib().tecno_last_error().decode()}")


class Solver:
    """A context holding one grid in HBM (tecno_create ... tecno_destroy)."""

    def __init__(self, cfg, _handle=None):
        self.cfg = cfg
        self.h = C.c_void_p()
        self._owned = _handle is None
        if _handle is None:
            _check(lib().tecno_create(C.byref(cfg), C.byref(self.h)))
        else:
            self.h = C.c_void_p(_handle)

    def close(self):
        if self.h and self._owned:
            lib().tecno_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ncells(self):
        return lib().tecno_num_cells(self.h)

    @property
    def nvars(self):
        return lib().tecno_num_vars(self.h)

    def load_net(self, path):
        _check(lib().tecno_load_net(self.h, path.encode()))

    def set_net(self, widths, W, b):
        widths = np.ascontiguousarray(widths, dtype=np.int32)
        W = np.ascontiguousarray(W, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().tecno_set_net(self.h, len(widths) - 1,
                                   widths.ctypes.data_as(C.POINTER(C.c_int32)), _dp(W), _dp(b)))

    def get_net(self):
        nl = C.c_int()
        widths = np.zeros(9, dtype=np.int32)
        W = np.zeros(8 * 16 * 16)
        b = np.zeros(8 * 16)
        _check(lib().tecno_get_net(self.h, C.byref(nl), widths.ctypes.data_as(C.POINTER(C.c_int32)),
                                   _dp(W), _dp(b)))
        widths = widths[:nl.value + 1]
        nw = int(sum(widths[i] * widths[i + 1] for i in range(nl.value)))
        nb = int(sum(widths[1:]))
        return widths, W[:nw].copy(), b[:nb].copy()

    def set_state(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(-1)
        _check(lib().tecno_set_state(self.h, _dp(prim), self.ncells))

    def get_state(self):
        out = np.zeros(self.nvars * self.ncells)
        _check(lib().tecno_get_state(self.h, _dp(out), self.ncells))
        return out.reshape(self.nvars, -1)

    def step(self, t_stamp=1e300):
        dt = C.c_double()
        _check(lib().tecno_step(self.h, t_stamp, C.byref(dt)))
        return dt.value

    def advance(self, final_time, t_stamps=1, max_iter=2 ** 31 - 1):
        it = C.c_int()
        t = C.c_double()
        _check(lib().tecno_advance(self.h, final_time, t_stamps, max_iter, C.byref(it), C.byref(t)))
        return it.value, t.value

    def run_steps(self, n, final_time=1e300, t_stamps=1):
        _check(lib().tecno_run_steps(self.h, n, final_time, t_stamps))

    def time(self):
        t = C.c_double()
        it = C.c_int()
        _check(lib().tecno_get_time(self.h, C.byref(t), C.byref(it)))
        return t.value, it.value

    def sync(self):
        _check(lib().tecno_synchronize(self.h))

    def audit(self):
        a = Audit()
        _check(lib().tecno_audit_state(self.h, C.byref(a)))
        return a

    def audit_noalloc(self):
        """tecno_audit_state_noalloc: the audit with no allocation"""
        a = Audit()
        _check(lib().tecno_audit_state_noalloc(self.h, C.byref(a)))
        return a

    def globals(self):
        """FV::compute_global of the current state (tecno_compute_globals)"""
        g = Globals()
        _check(lib().tecno_compute_globals(self.h, C.byref(g)))
        return g

    def vorticity(self):
        """nodal vorticity (3, Nz+1, Ny+1, Nx+1) over the reference's nodes
        (a 2D grid: Nz = n[2] of its configuration)"""
        c = self.cfg
        nx, ny = c.n[0] + 1, (c.n[1] + 1) if c.dims >= 2 else 2
        nz = c.n[2] + 1
        out = np.zeros(3 * nx * ny * nz)
        _check(lib().tecno_vorticity(self.h, _dp(out), out.size // 3))
        return out.reshape(3, nz, ny, nx)

    def l1_error(self, exact):
        exact = np.ascontiguousarray(exact, dtype=np.float64).reshape(-1)
        out = np.zeros(self.nvars)
        _check(lib().tecno_l1_error(self.h, _dp(exact), self.ncells, _dp(out)))
        return out

    def reconstruct_batch(self, st):
        st = np.ascontiguousarray(st, dtype=np.float64)
        n = st.shape[0]
        l, r = np.zeros(n), np.zeros(n)
        _check(lib().tecno_reconstruct_batch(self.h, _dp(st), n, _dp(l), _dp(r)))
        return l, r

    def flux_batch(self, P, V, axis):
        P = np.ascontiguousarray(P, dtype=np.float64)
        n = P.shape[0]
        if V is None:
            V = np.zeros(1)
        V = np.ascontiguousarray(V, dtype=np.float64)
        nout = 1 if self.cfg.system == SYS_SCALAR else 5
        F = np.zeros((n, nout))
        _check(lib().tecno_flux_batch(self.h, _dp(P), _dp(V), n, axis, _dp(F)))
        return F

    def mlp_forward_batch(self, u4):
        u4 = np.ascontiguousarray(u4, dtype=np.float64)
        n = u4.shape[0]
        l, r = np.zeros(n), np.zeros(n)
        _check(lib().tecno_mlp_forward_batch(self.h, _dp(u4), n, _dp(l), _dp(r)))
        return l, r

    def train_init(self, lr=1e-3, beta1=0.5, beta2=0.9, eps=1e-8, weight_decay=1e-5):
        a = Adam(lr, beta1, beta2, eps, weight_decay, 0)
        _check(lib().tecno_train_init(self.h, C.byref(a)))

    def train_step(self, rows, use_bf16=True):
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        loss = C.c_double()
        _check(lib().tecno_train_step(self.h, _dp(rows), rows.shape[0], int(use_bf16), C.byref(loss)))
        return loss.value

    def train_grad(self, rows, use_bf16=False):
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        loss = C.c_double()
        widths, W, b = self.get_net()
        g = np.zeros(len(W) + len(b))
        _check(lib().tecno_train_grad(self.h, _dp(rows), rows.shape[0], int(use_bf16), C.byref(loss), _dp(g)))
        return loss.value, g

    def train_set_data(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        _check(lib().tecno_train_set_data(self.h, _dp(rows), rows.shape[0]))

    def train_epoch(self, batch, seed, use_bf16=True):
        loss = C.c_double()
        _check(lib().tecno_train_epoch(self.h, int(batch), int(seed), int(use_bf16), C.byref(loss)))
        return loss.value

    def save_net(self, path):
        _check(lib().tecno_save_net(self.h, str(path).encode()))

    def init_net(self, widths, seed):
        w = (C.c_int32 * len(widths))(*[int(x) for x in widths])
        _check(lib().tecno_init_net(self.h, len(widths) - 1, w, int(seed)))

    def timing(self, on=True):
        _check(lib().tecno_timing_enable(self.h, int(on)))

    def timing_get(self):
        ms = C.c_double()
        n = C.c_int64()
        _check(lib().tecno_timing_get(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


def read_param(path, level=0):
    """Parameter::read of a param.in file through the library
    (tecno_read_param): the tecno_cfg of mesh level `level` and a dict of
    levels, final_time, time_stamps, max_iter, net_file."""
    cfg = Cfg()
    nl, ts, mi = C.c_int(), C.c_int(), C.c_int()
    ft = C.c_double()
    buf = C.create_string_buffer(4096)
    _check(lib().tecno_read_param(str(path).encode(), level, C.byref(cfg), C.byref(nl), C.byref(ft),
                                  C.byref(ts), C.byref(mi), buf, len(buf)))
    return cfg, dict(levels=nl.value, final_time=ft.value, time_stamps=ts.value, max_iter=mi.value,
                     net_file=buf.value.decode())


def trainer_epoch_seeds(widths, seed, n_epochs):
    """Shuffle seeds of the trainer's first n_epochs epochs for seed_no = seed
    and these network widths (pass seeds[e] to Solver.train_epoch of epoch e
    to take the trainer's row order)."""
    w = (C.c_int32 * len(widths))(*[int(x) for x in widths])
    out = (C.c_uint64 * max(int(n_epochs), 1))()
    _check(lib().tecno_trainer_epoch_seeds(len(widths) - 1, w, int(seed), int(n_epochs), out))
    return [int(out[i]) for i in range(int(n_epochs))]


def make_train_rows(u4, target):
    u4 = np.ascontiguousarray(u4, dtype=np.float64)
    target = np.ascontiguousarray(target, dtype=np.float64)
    rows = np.zeros((u4.shape[0], 21))
    _check(lib().tecno_make_train_rows(_dp(u4), _dp(target), u4.shape[0], _dp(rows)))
    return rows


def harvest_stencils_1d(cells, face_values, periodic=True):
    """Training stencils from a 1D solver state: every interface i+1/2 with
    its 4-cell stencil (u[i-1], u[i], u[i+1], u[i+2]) and the target value on
    both sides (face_values[i] = value at x[i+1/2]).  Stencils whose centre
    jump is below 1e-12 of the stencil scale are dropped: the solver skips the
    network there (|u2 - u1| < 1e-15) and their ratios carry no information."""
    u = np.asarray(cells, dtype=np.float64)
    n = u.shape[0]
    idx = np.arange(n)
    if periodic:
        u4 = np.stack([u[(idx - 1) % n], u, u[(idx + 1) % n], u[(idx + 2) % n]], axis=1)
        f = np.asarray(face_values, dtype=np.float64)
    else:
        sel = idx[1:n - 2]
        u4 = np.stack([u[sel - 1], u[sel], u[sel + 1], u[sel + 2]], axis=1)
        f = np.asarray(face_values, dtype=np.float64)[sel]
    scale = np.maximum(1.0, np.abs(u4).max(axis=1))
    keep = np.abs(u4[:, 2] - u4[:, 1]) > 1e-12 * scale
    return u4[keep], np.stack([f[keep], f[keep]], axis=1)


class Ensemble:
    """Independent scalar problems advanced together (tecno_ens_*): one
    persistent launch per precision x dimension group and stage, one
    hipGraph per ensemble step."""

    def __init__(self, cfgs):
        arr = (Cfg * len(cfgs))(*cfgs)
        self.h = C.c_void_p()
        _check(lib().tecno_ens_create(arr, len(cfgs), C.byref(self.h)))
        self.members = [Solver(cfgs[i], _handle=lib().tecno_ens_member(self.h, i))
                        for i in range(len(cfgs))]

    def close(self):
        if self.h:
            for m in self.members:
                m.close()
            lib().tecno_ens_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.members)

    def advance(self, final_times, t_stamps=1, max_iter=2 ** 31 - 1, use_graph=True):
        ft = np.ascontiguousarray(final_times, dtype=np.float64)
        cu = C.c_int64()
        _check(lib().tecno_ens_advance(self.h, _dp(ft), t_stamps, max_iter, int(use_graph),
                                       C.byref(cu)))
        return cu.value

    def autotune(self, steps=3):
        _check(lib().tecno_ens_autotune(self.h, steps))

    def groups(self):
        out = []
        for g in range(lib().tecno_ens_num_groups(self.h)):
            info = EnsGroup()
            _check(lib().tecno_ens_group_info(self.h, g, C.byref(info)))
            label = {FP64: "fp64", FP32: "fp32", FP32_STATE64: "fp32_state64"}[info.precision]
            out.append(dict(precision=label, dims=info.dims,
                            members=info.n_members, tiles=info.n_tiles, grid=info.grid,
                            tile=(info.tile_x, info.tile_y),
                            tuned_ms={(info.cand_x[v], info.cand_y[v]): info.tuned_ms[v]
                                      for v in range(info.n_candidates)}))
        return out
