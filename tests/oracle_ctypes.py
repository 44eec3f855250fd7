"""ctypes binding of oracle/liboracle.so — the CPU FP64 restatement of the
reference hot path.  Test infrastructure: the checker, never the product."""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "dsp-weno_amd"))
from tecno import Cfg, make_cfg, SYS_SCALAR  # noqa: E402  (struct layout only)

LIB = os.path.join(ROOT, "oracle", "liboracle.so")
NET_FILE = os.path.join(ROOT, "tests", "golden", "dsp_weno_net.txt")

_lib = None


class Net(C.Structure):
    _fields_ = [("n_layers", C.c_int), ("width", C.c_int * 9),
                ("W", (C.c_double * 256) * 8), ("b", (C.c_double * 16) * 8)]


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        P = C.c_void_p
        L.tor_logavg.restype = C.c_double
        L.tor_logavg.argtypes = [C.c_double, C.c_double]
        L.tor_limited_slope.restype = C.c_double
        L.tor_limited_slope.argtypes = [C.c_double, C.c_double, C.c_int]
        L.tor_net_load.argtypes = [C.c_char_p, C.POINTER(Net)]
        L.tor_net_save.argtypes = [C.c_char_p, C.POINTER(Net)]
        L.tor_sp_weno_batch.argtypes = [C.c_int, dp, C.c_int64, dp, dp]
        L.tor_reconstruct_batch.argtypes = [C.POINTER(Cfg), C.POINTER(Net), dp, C.c_int, C.c_int64, dp, dp]
        L.tor_nn_forward_batch.argtypes = [C.POINTER(Net), dp, C.c_int64, dp, dp]
        L.tor_nn_features.argtypes = [dp, dp, dp, dp]
        L.tor_logavg_batch.argtypes = [dp, dp, C.c_int64, dp]
        L.tor_flux_batch.argtypes = [C.POINTER(Cfg), C.POINTER(Net), dp, dp, C.c_int, C.c_int64,
                                     C.c_int, dp, dp]
        L.tor_prim_to_ent_batch.argtypes = [C.POINTER(Cfg), dp, C.c_int64, dp]
        L.tor_create.restype = P
        L.tor_create.argtypes = [C.POINTER(Cfg), C.POINTER(Net), C.c_int]
        L.tor_destroy.argtypes = [P]
        L.tor_num_cells.restype = C.c_int64
        L.tor_num_cells.argtypes = [P]
        L.tor_num_vars.argtypes = [P]
        L.tor_set_prim.argtypes = [P, dp]
        L.tor_get_prim.argtypes = [P, dp]
        L.tor_step.argtypes = [P, C.c_double, dp]
        L.tor_advance.argtypes = [P, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), dp]
        L.tor_set_threads.argtypes = [C.c_int]
        L.tor_audit.argtypes = [P, dp]
        L.tor_es_node_flux.argtypes = [C.POINTER(Cfg), dp, dp, dp, dp, dp]
        L.tor_train_rows.argtypes = [dp, dp, C.c_int64, dp]
        L.tor_train_loss_grad.argtypes = [C.POINTER(Net), dp, C.c_int64, C.c_int, C.POINTER(Net)]
        L.tor_train_loss_grad.restype = C.c_double
        L.tor_adam_step.argtypes = [C.POINTER(Net), C.POINTER(Net), C.POINTER(Net), C.POINTER(Net),
                                    C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def load_net(path=NET_FILE):
    n = Net()
    rc = lib().tor_net_load(path.encode(), C.byref(n))
    if rc != 0:
        raise IOError(f"cannot load {path}: {rc}")
    return n


def net_from_arrays(widths, W, b):
    n = Net()
    n.n_layers = len(widths) - 1
    for i, w in enumerate(widths):
        n.width[i] = int(w)
    ow = ob = 0
    for l in range(n.n_layers):
        ni, no = widths[l], widths[l + 1]
        for o in range(no):
            for i in range(ni):
                n.W[l][o * 16 + i] = W[ow + o * ni + i]
            n.b[l][o] = b[ob + o]
        ow += ni * no
        ob += no
    return n


def net_arrays(net):
    widths = [net.width[i] for i in range(net.n_layers + 1)]
    W, b = [], []
    for l in range(net.n_layers):
        for o in range(widths[l + 1]):
            for i in range(widths[l]):
                W.append(net.W[l][o * 16 + i])
            b.append(net.b[l][o])
    return np.array(widths, dtype=np.int32), np.array(W), np.array(b)


def logavg(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros_like(a)
    lib().tor_logavg_batch(_dp(a), _dp(b), a.size, _dp(out))
    return out


def limited_slope(a, b, use_minabs):
    return lib().tor_limited_slope(float(a), float(b), int(bool(use_minabs)))


def sp_weno(st, corr=0):
    st = np.ascontiguousarray(st, dtype=np.float64)
    n = st.shape[0]
    l, r = np.zeros(n), np.zeros(n)
    lib().tor_sp_weno_batch(corr, _dp(st), n, _dp(l), _dp(r))
    return l, r


def reconstruct(cfg, st, net=None):
    st = np.ascontiguousarray(st, dtype=np.float64)
    n, s = st.shape
    l, r = np.zeros(n), np.zeros(n)
    lib().tor_reconstruct_batch(C.byref(cfg), C.byref(net) if net else None, _dp(st), s // 2, n,
                                _dp(l), _dp(r))
    return l, r


def nn_forward(net, u4):
    u4 = np.ascontiguousarray(u4, dtype=np.float64)
    n = u4.shape[0]
    l, r = np.zeros(n), np.zeros(n)
    lib().tor_nn_forward_batch(C.byref(net), _dp(u4), n, _dp(l), _dp(r))
    return l, r


def nn_features(u4):
    u4 = np.ascontiguousarray(u4, dtype=np.float64)
    f, x, y = np.zeros(5), np.zeros(5), np.zeros(5)
    lib().tor_nn_features(_dp(u4), _dp(f), _dp(x), _dp(y))
    return f, x, y


def flux(cfg, P, V, axis, h, net=None):
    P = np.ascontiguousarray(P, dtype=np.float64)
    n = P.shape[0]
    s = P.shape[1]
    Vv = np.ascontiguousarray(V if V is not None else np.zeros(1), dtype=np.float64)
    nout = 1 if cfg.system == SYS_SCALAR else 5
    F = np.zeros((n, nout))
    hh = np.ascontiguousarray(h, dtype=np.float64)
    lib().tor_flux_batch(C.byref(cfg), C.byref(net) if net else None, _dp(P), _dp(Vv), s // 2, n, axis,
                         _dp(hh), _dp(F))
    return F


def prim_to_ent(cfg, P):
    P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, 5)
    V = np.zeros_like(P)
    lib().tor_prim_to_ent_batch(C.byref(cfg), _dp(P), P.shape[0], _dp(V))
    return V


class OracleSolver:
    def __init__(self, cfg, net=None, faithful=False, threads=1):
        self.cfg = cfg
        lib().tor_set_threads(threads)
        self.h = lib().tor_create(C.byref(cfg), C.byref(net) if net else None, int(faithful))
        if not self.h:
            raise MemoryError("oracle allocation failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().tor_destroy(self.h)
            self.h = None

    @property
    def ncells(self):
        return lib().tor_num_cells(self.h)

    @property
    def nvars(self):
        return lib().tor_num_vars(self.h)

    def set_state(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(-1)
        assert prim.size == self.nvars * self.ncells
        lib().tor_set_prim(self.h, _dp(prim))

    def get_state(self):
        out = np.zeros(self.nvars * self.ncells)
        lib().tor_get_prim(self.h, _dp(out))
        return out.reshape(self.nvars, -1)

    def step(self, t_stamp=1e300):
        dt = C.c_double()
        lib().tor_step(self.h, t_stamp, C.byref(dt))
        return dt.value

    def advance(self, final_time, t_stamps=1, max_iter=2 ** 31 - 1):
        it = C.c_int()
        t = C.c_double()
        lib().tor_advance(self.h, final_time, t_stamps, max_iter, C.byref(it), C.byref(t))
        return it.value, t.value

    def audit(self):
        out = np.zeros(8)
        lib().tor_audit(self.h, _dp(out))
        return out


def es_node_flux(cfg, avg, Tx, Ty, Tz):
    """ES viscous node fluxes before the face areas: (3 fluxes, 4 components:
    momentum x, y, z, energy)"""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (avg, Tx, Ty, Tz)]
    out = np.zeros(12)
    lib().tor_es_node_flux(C.byref(cfg), *[_dp(v) for v in a], _dp(out))
    return out.reshape(3, 4)


# ---------------------------------------------------------------- trainer
def train_rows(u4, target):
    u4 = np.ascontiguousarray(u4, dtype=np.float64).reshape(-1, 4)
    target = np.ascontiguousarray(target, dtype=np.float64).reshape(-1, 2)
    rows = np.zeros((u4.shape[0], 21))
    lib().tor_train_rows(_dp(u4), _dp(target), u4.shape[0], _dp(rows))
    return rows


def grad_arrays(g):
    """flatten a gradient held in a Net: W (layer, out, in) then b"""
    _, W, b = net_arrays(g)
    return np.concatenate([W, b])


def train_loss_grad(net, rows, bf16=False):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    g = Net()
    loss = lib().tor_train_loss_grad(C.byref(net), _dp(rows), rows.shape[0], int(bf16), C.byref(g))
    return loss, grad_arrays(g)


def train_loss(net, rows, bf16=False):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    return lib().tor_train_loss_grad(C.byref(net), _dp(rows), rows.shape[0], int(bf16), None)


def adam_step(net, rows, m, v, t, lr=1e-3, beta1=0.5, beta2=0.9, eps=1e-8, wd=1e-5):
    """one trainer step on the oracle: gradient of rows, then Adam; m, v are
    Nets holding the moments (zero-initialised by the caller)"""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    g = Net()
    loss = lib().tor_train_loss_grad(C.byref(net), _dp(rows), rows.shape[0], 0, C.byref(g))
    lib().tor_adam_step(C.byref(net), C.byref(g), C.byref(m), C.byref(v), t, lr, beta1, beta2, eps, wd)
    return loss


def zero_like(net):
    z = Net()
    z.n_layers = net.n_layers
    for i in range(net.n_layers + 1):
        z.width[i] = net.width[i]
    return z
