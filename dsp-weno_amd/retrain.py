"""Retraining of the DSP-WENO network on stencils harvested from the solver's
own runs, on the GPU (libtecno's forward / backward / Adam kernels).

The reference trains its network offline in PyTorch
(DSP-WENO_trainer/DSP-WENO_network_trainer.py) from a data file of 4-cell
stencils with their exact interface values (rows of features, vertices,
cells and targets, trainer:281-346); the file itself is not part of the
reference.  Here the stencils come from runs of the solver:

* linear advection of sin^4 x (enflo_scalar/examples/LinAdvTest2) at several
  meshes and times: every interface's stencil (u[i-1], u[i], u[i+1], u[i+2])
  with the exact value sin^4(x_{i+1/2} - t) as both targets;
* the 2D Euler isentropic vortex (enflo/examples/2D/IsentropicVortex): at
  every x and y face the four cells' entropy variables projected on the
  face's scaled eigenvectors -- exactly the stencils the TeCNO flux
  reconstructs (kepes_tecno4_roe.cc:39-141) -- with the projection of the
  exact solution's entropy variables at the face centre as targets.

The loss, head and optimiser are the trainer's (MSE of both interface values,
softmax-vertex head, Adam with betas (0.5, 0.9), weight decay 1e-5); the
arithmetic is the device trainer's (bf16 activations, FP32 master weights).

    python retrain.py --out net.txt [--stencils 1e7] [--epochs 40] [--batch 65536]
    python retrain.py --out net.txt --data DSP-WENO_training_data.dat
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import tecno  # noqa: E402

GAMMA = 1.4


# ---------------------------------------------------------------- states
def vortex_prim(n, t=0.0, strength=5.0, gamma=GAMMA, xs=None, ys=None):
    """Isentropic vortex on the periodic [-5,5]^2 (free stream (1, 0),
    rho = p = 1 at infinity) at time t; point values at (xs, ys) (default:
    the cell centres of an n x n mesh).  Returns rho, u, v, p."""
    if xs is None:
        h = 10.0 / n
        c = -5 + (np.arange(n) + 0.5) * h
        xs, ys = np.meshgrid(c, c)
    X = (xs - t + 5) % 10 - 5
    r2 = X ** 2 + ys ** 2
    T = 1.0 - (gamma - 1) * strength ** 2 / (8 * gamma * math.pi ** 2) * np.exp(1 - r2)
    e = np.exp(0.5 * (1 - r2))
    return (T ** (1 / (gamma - 1)), 1.0 - strength / (2 * math.pi) * ys * e,
            strength / (2 * math.pi) * X * e, T ** (gamma / (gamma - 1)))


def entropy_vars(rho, u, v, p, gamma=GAMMA):
    """V = ((gamma - s)/(gamma - 1) - beta |u|^2, 2 beta u, 2 beta v, -2 beta),
    beta = rho / (2p), s = ln p - gamma ln rho (fv.cc:421-431, 557-560)."""
    beta = rho / (2 * p)
    s = np.log(p) - gamma * np.log(rho)
    return ((gamma - s) / (gamma - 1) - beta * (u * u + v * v), 2 * beta * u, 2 * beta * v,
            -2 * beta)


def logavg(a, b):
    """misc_func.h:14-32 (vectorised)."""
    xi = b / a
    f = (xi - 1) / (xi + 1)
    u = f * f
    with np.errstate(divide="ignore", invalid="ignore"):
        F = np.where(u < 1e-2, 1 + u / 3 + u * u / 5 + u * u * u / 7, np.log(xi) / 2 / f)
    return 0.5 * (a + b) / F


def face_eigvecs(rl, nl, tl, pl, rr, nr, tr, pr, gamma=GAMMA):
    """Scaled eigenvectors of the face between two cells in the face frame
    (normal velocity n, tangential t), as kepes_tecno4_roe.cc:39-89 builds
    them from the KEPEC means: rows (rho, n, t, E) of the waves u-a,
    entropy, shear, u+a; shape (4 waves, 4 rows, ...)."""
    bl, br = rl / (2 * pl), rr / (2 * pr)
    rho = logavg(rl, rr)
    beta = logavg(bl, br)
    p = 0.5 * (rl + rr) / (bl + br)
    vn, vt = 0.5 * (nl + nr), 0.5 * (tl + tr)
    a = np.sqrt(0.5 * gamma / beta)
    vsq = vn * vn + vt * vt
    H = a * a / (gamma - 1) + 0.5 * vsq
    c1 = np.sqrt(rho / gamma)
    S0, S1, S2 = math.sqrt(0.5) * c1, math.sqrt(gamma - 1) * c1, np.sqrt(p)
    z = np.zeros_like(vn)
    return np.array([[S0, (vn - a) * S0, vt * S0, (H - a * vn) * S0],
                     [S1, vn * S1, vt * S1, 0.5 * vsq * S1],
                     [z, z, -S2, -vt * S2],
                     [S0, (vn + a) * S0, vt * S0, (H + a * vn) * S0]])


def project(R, V):
    """Z_w = sum_r R[w, r] V_r (V in the face frame)."""
    return np.einsum("wr...,r...->w...", R, np.asarray(V))


# ---------------------------------------------------------------- harvests
def harvest_scalar_1d(cells, face_exact, drop_rel=1e-12):
    """Every periodic interface i+1/2 of a 1D state: stencil (u[i-1], u[i],
    u[i+1], u[i+2]) and its exact face value twice (left and right
    targets).  Stencils with |u[i+1] - u[i]| <= drop_rel * max(1, max|u|)
    are dropped: the solver skips the network there (reconstruct.cc:152)."""
    u = np.asarray(cells, dtype=np.float64)
    idx = np.arange(u.shape[0])
    n = u.shape[0]
    u4 = np.stack([u[(idx - 1) % n], u, u[(idx + 1) % n], u[(idx + 2) % n]], axis=1)
    f = np.asarray(face_exact, dtype=np.float64)
    keep = np.abs(u4[:, 2] - u4[:, 1]) > drop_rel * np.maximum(1.0, np.abs(u4).max(axis=1))
    return u4[keep], np.stack([f[keep], f[keep]], axis=1)


def harvest_euler_2d(prim, exact_face, drop_rel=1e-12, gamma=GAMMA):
    """Characteristic stencils of every x and y face of a periodic 2D Euler
    state prim = (rho, u, v, p), each (ny, nx).  exact_face(axis) returns
    the exact (rho, u, v, p) at the centres of the faces i+1/2 (axis 0) or
    j+1/2 (axis 1), shaped like the state.  For every face and wave the
    stencil is the four cells' entropy variables projected on the face's
    scaled eigenvectors; the target is the same projection of the exact
    face state's entropy variables."""
    rho, u, v, p = (np.asarray(q, dtype=np.float64) for q in prim)
    V = entropy_vars(rho, u, v, p, gamma)
    out_u4, out_t = [], []
    for axis in (0, 1):
        ax = 1 - axis  # numpy axis along which the face normal runs (x = columns)
        sh = lambda q, k: np.roll(q, -k, axis=ax)  # noqa: E731  q[i + k]
        vn_, vt_ = (u, v) if axis == 0 else (v, u)
        R = face_eigvecs(rho, vn_, vt_, p, sh(rho, 1), sh(vn_, 1), sh(vt_, 1), sh(p, 1), gamma)
        Vf = (V[0], V[1], V[2], V[3]) if axis == 0 else (V[0], V[2], V[1], V[3])
        Z = [project(R, [sh(q, k) for q in Vf]) for k in (-1, 0, 1, 2)]
        er, eu, ev, ep = exact_face(axis)
        Ve = entropy_vars(er, eu, ev, ep, gamma)
        Vef = (Ve[0], Ve[1], Ve[2], Ve[3]) if axis == 0 else (Ve[0], Ve[2], Ve[1], Ve[3])
        Zt = project(R, Vef)
        u4 = np.stack([z.reshape(4, -1) for z in Z], axis=-1).reshape(-1, 4)
        tg = Zt.reshape(-1)
        keep = np.abs(u4[:, 2] - u4[:, 1]) > drop_rel * np.maximum(1.0, np.abs(u4).max(axis=1))
        out_u4.append(u4[keep])
        out_t.append(np.stack([tg[keep], tg[keep]], axis=1))
    return np.concatenate(out_u4), np.concatenate(out_t)


# ---------------------------------------------------------------- data files
def read_trainer_data(path):
    """Rows (n x 21: 15 features, 4 cells, 2 targets, the layout of
    tecno_make_train_rows) from a data file of the reference trainer
    (DSP-WENO_network_trainer.py:281-333): one stencil per line, values
    separated by commas.  With tanh features (the trainer's default,
    config.py:14) a line has 23 values: the 4 cells, theta-, theta+,
    tanh(theta-), tanh(theta+), the 3 scaled jump magnitudes, 5 x-vertices,
    5 y-vertices, the left and right targets; without, 21 values (no tanh
    columns, the network inputs starting at theta-)."""
    rows = []
    with open(path) as f:
        for line in f:
            tok = [t for t in line.replace("\n", "").split(",") if t.strip()]
            if not tok:
                continue
            v = [float(t) for t in tok]
            if len(v) == 23:
                feat = v[6:21]
            elif len(v) == 21:
                feat = v[4:19]
            else:
                raise ValueError(f"{path}: a line has {len(v)} values (expected 21 or 23)")
            rows.append(feat + v[0:4] + v[-2:])
    return np.array(rows, dtype=np.float64).reshape(-1, 21)


def write_trainer_data(path, rows):
    """The inverse of read_trainer_data in the 23-value layout: cells,
    theta- = D2/D1, theta+ = D0/D1 of the cells, the 15 features, targets
    (%.17g, comma separated)."""
    rows = np.asarray(rows, dtype=np.float64)
    c = rows[:, 15:19]
    d0, d1, d2 = c[:, 1] - c[:, 0], c[:, 2] - c[:, 1], c[:, 3] - c[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        tm, tp = d2 / d1, d0 / d1
    out = np.column_stack([c, tm, tp, rows[:, 0:15], rows[:, 19:21]])
    with open(path, "w") as f:
        for r in out:
            f.write(",".join(f"{x:.17g}" for x in r) + "\n")


# ---------------------------------------------------------------- runs
def linadv_runs(meshes=(100, 200, 400, 800), times=(0.1, 0.2, 0.3, 0.4, 0.5)):
    """Config 1 (LinAdvTest2: sin^4 x on [-pi, pi], lin_tecno4 + sp_weno,
    CFL 0.5): stencils at each output time of each mesh."""
    parts = []
    for n in meshes:
        h = 2 * math.pi / n
        x = -math.pi + (np.arange(n) + 0.5) * h
        s = tecno.Solver(tecno.make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno",
                                        n=(n, 2, 2), dims=1, lo=(-math.pi, 0, 0), hi=(math.pi, 1, 1),
                                        bc="periodic", cfl=0.5))
        s.set_state(np.sin(x) ** 4)
        for t in times:
            s.advance(t, 1)
            tt, _ = s.time()
            parts.append(harvest_scalar_1d(s.get_state().reshape(-1), np.sin(x + 0.5 * h - tt) ** 4))
        s.close()
    return parts


def vortex_runs(meshes=(256, 512), times=(0.25, 0.5, 0.75, 1.0), recon="sp_weno"):
    """Config 2's isentropic vortex (kepes_tecno4_roe, CFL 0.5, periodic):
    characteristic stencils at each output time of each mesh."""
    parts = []
    for n in meshes:
        h = 10.0 / n
        c = -5 + (np.arange(n) + 0.5) * h
        X, Y = np.meshgrid(c, c)
        s = tecno.Solver(tecno.make_cfg(n=(n, n, 2), dims=2, lo=(-5, -5, 0), hi=(5, 5, 1),
                                        recon=recon, flux="kepes_tecno4_roe", bc="periodic",
                                        cfl=0.5))
        s.set_state(np.stack(vortex_prim(n)).reshape(4, -1))
        for t in times:
            s.advance(t, 1)
            tt, _ = s.time()
            st = s.get_state().reshape(4, n, n)

            def exact_face(axis, tt=tt):
                if axis == 0:
                    return vortex_prim(n, tt, xs=X + 0.5 * h, ys=Y)
                return vortex_prim(n, tt, xs=X, ys=Y + 0.5 * h)

            parts.append(harvest_euler_2d(st, exact_face))
        s.close()
    return parts


def informative(u4, min_jump=1e-3):
    """Stencils whose jump scale max|u[k+1] - u[k]| exceeds min_jump *
    max(1, max|u|): below that the cells' own discretisation error is
    comparable to the jumps (see normalize_rows)."""
    s = np.abs(np.diff(u4, axis=1)).max(axis=1)
    return s > min_jump * np.maximum(1.0, np.abs(u4).max(axis=1))


def build_dataset(target_rows=10_000_000, seed=7, vortex_meshes=(256, 512, 1024)):
    """Harvest, keep the informative stencils, then draw up to target_rows
    of them (vortex and linear advection pooled) in a seeded random order.
    Returns the stencils, targets and the harvested count."""
    parts = linadv_runs() + vortex_runs(meshes=vortex_meshes)
    u4 = np.concatenate([p[0] for p in parts])
    tg = np.concatenate([p[1] for p in parts])
    n_harvest = len(u4)
    keep = informative(u4)
    u4, tg = u4[keep], tg[keep]
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(u4))[:int(target_rows)]
    return u4[perm], tg[perm], n_harvest


def normalize_rows(rows, min_jump=1e-3):
    """Jump-normalised rows: cells and targets (columns 15-20) mapped by
    u -> (u - u1) / max|jump|, features and vertices (columns 0-14, computed
    from the raw stencil, as the solver sees it) unchanged.  The interface
    values of the head are affine-equivariant in the cells, so the loss of a
    row becomes its raw loss divided by the squared jump scale: smooth and
    steep stencils weigh alike, and the gradient is not swamped by the
    trainer's weight decay (1e-5) when jumps are O(h^3) small.  Rows whose
    jump scale is below min_jump * max(1, max|u|) are dropped: there the
    solver's own error in the cells (~1e-5 at these meshes) is comparable
    to the jumps and the target is not a function of the stencil."""
    c = rows[:, 15:19]
    s = np.abs(np.diff(c, axis=1)).max(axis=1)
    keep = s > min_jump * np.maximum(1.0, np.abs(c).max(axis=1))
    out = rows[keep].copy()
    out[:, 15:21] = (out[:, 15:21] - out[:, 16:17]) / s[keep][:, None]
    return out


def train(u4, tg, *, epochs=40, batch=65536, lr=1e-3, seed=6008, holdout=200_000,
          init_net=None, normalize=True, log=print):
    """Rows on the device, Xavier-initialised 5-5-5-5-5 network (or the
    given net_file), Adam as the trainer; returns the trained solver context
    and a summary (loss per epoch, held-out interface errors)."""
    rows = tecno.make_train_rows(u4, tg)
    if normalize:
        rows = normalize_rows(rows)
    te, tr = rows[:holdout], rows[holdout:]
    s = tecno.Solver(tecno.make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno_dl",
                                    n=(8, 2, 2), dims=1))
    if init_net:
        s.load_net(init_net)
    else:
        s.init_net([5, 5, 5, 5, 5], seed=seed)

    def err(sol):
        l, r = sol.mlp_forward_batch(te[:, 15:19])
        return float(np.mean(np.abs(l - te[:, 19])) + np.mean(np.abs(r - te[:, 20])))

    e0 = err(s)
    s.train_init(lr=lr, beta1=0.5, beta2=0.9, eps=1e-8, weight_decay=1e-5)
    s.train_set_data(tr)
    losses = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        losses.append(s.train_epoch(batch, seed=seed + ep, use_bf16=True))
        log(f"epoch {ep}: loss {losses[-1]:.4e}")
    t_train = time.perf_counter() - t0
    sp = tecno.Solver(tecno.make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno",
                                     n=(8, 2, 2), dims=1))
    l, r = sp.reconstruct_batch(te[:, 15:19])
    espw = float(np.mean(np.abs(l - te[:, 19])) + np.mean(np.abs(r - te[:, 20])))
    sp.close()
    summary = dict(rows=int(len(tr)), holdout=int(len(te)), batch=batch, epochs=epochs,
                   normalized=bool(normalize),
                   steps=int(epochs * math.ceil(len(tr) / batch)), train_seconds=t_train,
                   loss_first=losses[0], loss_last=losses[-1], heldout_err_untrained=e0,
                   heldout_err_trained=err(s), heldout_err_spweno=espw)
    return s, summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True, help="net_file to write")
    ap.add_argument("--stencils", type=float, default=1e7)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--init", default=None, help="start from this net_file instead of Xavier")
    ap.add_argument("--raw", action="store_true", help="train on raw (not jump-normalised) rows")
    ap.add_argument("--data", default=None,
                    help="train on a data file in the reference trainer's layout (its data_file) "
                         "instead of harvesting")
    ap.add_argument("--write-data", default=None,
                    help="also write the harvested stencils in the reference trainer's layout")
    a = ap.parse_args()
    t0 = time.perf_counter()
    if a.data:
        rows = read_trainer_data(a.data)
        u4, tg = rows[:, 15:19], rows[:, 19:21]
        print(f"read {len(u4)} stencils from {a.data}", flush=True)
    else:
        u4, tg, n_harvest = build_dataset(a.stencils)
        print(f"harvested {n_harvest} stencils, {len(u4)} informative kept, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        if a.write_data:
            write_trainer_data(a.write_data, tecno.make_train_rows(u4, tg))
    s, summary = train(u4, tg, epochs=a.epochs, batch=a.batch, init_net=a.init,
                       normalize=not a.raw, log=lambda m: print(m, flush=True))
    s.save_net(a.out)
    s.close()
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
