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
  exact solution's entropy variables at the face centre as targets;
* the 2D Euler four-quadrant Riemann problem (seeded quadrant states,
  neumann boundaries): a run on a 4x finer mesh, restricted to the coarse
  mesh by block means of the conserved variables; the coarse faces'
  characteristic stencils, with the fine mesh's SP-WENO face values of the
  same projection, averaged along the face, as the left and right targets.
  Here the stencils have large jumps, the vertex polygon of the network's
  head is large, and the network decides the interface value;
* 5% (of the harvest) random oscillatory stencils with SP-WENO's values as
  targets, which keep the network's dissipation on grid-scale oscillations.

evaluate_riemann measures what a network does inside the solver: config 2's
Riemann problem on seeds the harvest did not use, against a 4x finer run.

The loss, head and optimiser are the trainer's (MSE of both interface values,
softmax-vertex head, Adam with betas (0.5, 0.9), weight decay 1e-5); the
arithmetic is the device trainer's (bf16 activations, FP32 master weights).

    python retrain.py --out net.txt [--stencils 1e7] [--epochs 40] [--batch 65536]
                      [--compare shipped=dsp_weno.txt] [--eval-riemann 256]
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


def cons_of(rho, u, v, p, gamma=GAMMA):
    return rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)


def prim_of(rho, mx, my, E, gamma=GAMMA):
    u, v = mx / rho, my / rho
    return rho, u, v, (gamma - 1) * (E - 0.5 * rho * (u * u + v * v))


def restrict(prim, f, gamma=GAMMA):
    """The coarse state whose cells are f x f blocks of the fine mesh: block
    means of the conserved variables (the coarse cell averages of the fine
    solution), returned as rho, u, v, p."""
    rho = np.asarray(prim[0])
    ny, nx = rho.shape[0] // f, rho.shape[1] // f
    cons = [q.reshape(ny, f, nx, f).mean(axis=(1, 3)) for q in cons_of(*prim, gamma)]
    return prim_of(*cons, gamma)


def harvest_euler_2d_fine(prim_fine, f, recon, drop_rel=1e-12, gamma=GAMMA, chunk=1 << 20,
                          with_faces=False):
    """Characteristic stencils of the coarse state that restricts a fine 2D
    Euler state (f x f blocks, f >= 2), with targets from the fine state.

    Stencils: at every interior x and y face of the coarse mesh, the four
    coarse cells' entropy variables projected on the face's scaled
    eigenvectors (as harvest_euler_2d); faces whose stencil would leave the
    domain are dropped, so the boundaries may be of any kind.  Targets: the
    fine cells' entropy variables projected on the same eigenvectors along
    each of the f fine rows that make up the coarse face, reconstructed at
    the face by recon (recon(st) -> (left, right) for st of shape (m, 4):
    the fine mesh's SP-WENO) and averaged over the f rows -- the face mean
    that the transverse cell averages of the coarse stencil stand for.  The
    left target is the mean of the fine left values, the right target that
    of the fine right values, so a discontinuity on the face keeps its two
    sides.  Stencils whose centre jump is below drop_rel of their scale are
    dropped (the solver skips the network there).  with_faces: also return
    each row's face as (axis, wave, J, I), the face I + 1/2 of coarse row J
    along the axis (coarse indices in the axis' frame: x faces (y, x), y
    faces (x, y))."""
    rho, u, v, p = (np.asarray(q, dtype=np.float64) for q in prim_fine)
    assert f >= 2 and rho.shape[0] % f == 0 and rho.shape[1] % f == 0
    coarse = restrict((rho, u, v, p), f, gamma)
    Vc = entropy_vars(*coarse, gamma)
    Vfine = entropy_vars(rho, u, v, p, gamma)
    out_u4, out_t, out_f = [], [], []
    for axis in (0, 1):
        # the face normal runs along the last array axis: x faces as stored,
        # y faces on the transposed arrays with the velocity roles swapped
        T = (lambda q: q) if axis == 0 else (lambda q: np.ascontiguousarray(q.T))  # noqa: E731
        cr, cn, ct, cp = T(coarse[0]), T(coarse[1 + axis]), T(coarse[2 - axis]), T(coarse[3])
        vc = [T(Vc[0]), T(Vc[1 + axis]), T(Vc[2 - axis]), T(Vc[3])]
        vf = [T(Vfine[0]), T(Vfine[1 + axis]), T(Vfine[2 - axis]), T(Vfine[3])]
        ny, nx = cr.shape
        I = np.arange(1, nx - 2)  # faces I + 1/2 with cells I-1 .. I+2 inside
        R = face_eigvecs(cr[:, I], cn[:, I], ct[:, I], cp[:, I],
                         cr[:, I + 1], cn[:, I + 1], ct[:, I + 1], cp[:, I + 1], gamma)
        Z = np.stack([project(R, [q[:, I - 1 + k] for q in vc]) for k in range(4)])  # (k, w, J, I)
        u4 = Z.reshape(4, -1).T
        keep = np.abs(u4[:, 2] - u4[:, 1]) > drop_rel * np.maximum(1.0, np.abs(u4).max(axis=1))
        sel = np.nonzero(keep)[0]
        w, J, Ii = np.unravel_index(sel, (4, ny, len(I)))
        tg = np.zeros((len(sel), 2))
        for a in range(0, len(sel), chunk):
            b = min(len(sel), a + chunk)
            rows = J[a:b, None] * f + np.arange(f)[None, :]                    # (m, f)
            cols = (I[Ii[a:b]] + 1)[:, None] * f - 2 + np.arange(4)[None, :]   # (m, 4)
            Rw = R[w[a:b], :, J[a:b], Ii[a:b]]                                  # (m, 4 rows of R)
            zf = sum(Rw[:, c, None, None] * vf[c][rows[:, :, None], cols[:, None, :]]
                     for c in range(4))                                         # (m, f, 4)
            l, r = recon(zf.reshape(-1, 4))
            tg[a:b, 0] = l.reshape(-1, f).mean(axis=1)
            tg[a:b, 1] = r.reshape(-1, f).mean(axis=1)
        out_u4.append(u4[sel])
        out_t.append(tg)
        out_f.append(np.stack([np.full(len(sel), axis), w, J, I[Ii]], axis=1))
    if with_faces:
        return np.concatenate(out_u4), np.concatenate(out_t), np.concatenate(out_f)
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


def quadrant_prim(n, seed):
    """Config 2's four-quadrant Riemann problem on [0,1]^2 split at
    (0.5, 0.5): per-quadrant rho ~ U[0.5,1.5], u, v ~ U[-0.5,0.5],
    p ~ U[0.3,1.5] drawn from the seed, at the cell centres of an n x n
    mesh.  Returns rho, u, v, p, each (n, n)."""
    rng = np.random.default_rng(seed)
    q = np.stack([rng.uniform(0.5, 1.5, 4), rng.uniform(-0.5, 0.5, 4),
                  rng.uniform(-0.5, 0.5, 4), rng.uniform(0.3, 1.5, 4)])
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    idx = (X >= 0.5).astype(int) + 2 * (Y >= 0.5).astype(int)
    return tuple(q[k][idx] for k in range(4))


def riemann_runs(meshes=(256, 512), seeds=(7, 8, 9), times=(0.05, 0.1, 0.15, 0.2, 0.25),
                 factor=4):
    """Config 2's four-quadrant Riemann problem (kepes_tecno4_roe + sp_weno,
    FP64, CFL 0.5, neumann boundaries) run on a mesh `factor` times finer
    than each listed mesh; at each output time the stencils of the
    restricted state with the fine run's face values as targets
    (harvest_euler_2d_fine, fine-mesh SP-WENO)."""
    sp = tecno.Solver(tecno.make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno",
                                     n=(8, 2, 2), dims=1))
    parts = []
    for n in meshes:
        nf = n * factor
        for seed in seeds:
            s = tecno.Solver(tecno.make_cfg(n=(nf, nf, 2), dims=2, recon="sp_weno",
                                            flux="kepes_tecno4_roe", bc="neumann", cfl=0.5))
            s.set_state(np.stack(quadrant_prim(nf, seed)).reshape(4, -1))
            for t in times:
                s.advance(t, 1)
                st = s.get_state().reshape(4, nf, nf)
                parts.append(harvest_euler_2d_fine(st, factor, sp.reconstruct_batch))
            s.close()
    sp.close()
    return parts


def oscillatory_stencils(m, seed=3):
    """m random stencils that alternate (u[1]-u[0], u[2]-u[1], u[3]-u[2]
    change sign twice), at scales spread over several decades, with the
    SP-WENO interface values (limiter.h:103-252) as targets.  A harvested
    solution holds almost none of them and gives them no exact interface
    value; without such rows the network, never shown one, returns a small
    interface jump there, i.e. removes the TeCNO dissipation
    (kepes_tecno4_roe.cc:141-176) from grid-scale oscillations."""
    rng = np.random.default_rng(seed)
    out = np.zeros((0, 4))
    while len(out) < m:
        k = 4 * (m - len(out)) + 64
        u = rng.normal(size=(k, 4)) * rng.lognormal(0.0, 2.0, size=(k, 1)) + rng.normal(size=(k, 1))
        d = np.diff(u, axis=1)
        out = np.concatenate([out, u[(d[:, 0] * d[:, 1] < 0) & (d[:, 1] * d[:, 2] < 0)]])
    out = out[:m]
    sp = tecno.Solver(tecno.make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno",
                                     n=(8, 2, 2), dims=1))
    tg = np.stack(sp.reconstruct_batch(out), axis=1)
    sp.close()
    return out, tg


def jump_scale(u4):
    """max_k |u[k+1] - u[k]| of each stencil."""
    return np.abs(np.diff(u4, axis=1)).max(axis=1)


def informative(u4, min_jump=1e-3):
    """Stencils whose jump scale max|u[k+1] - u[k]| exceeds min_jump *
    max(1, max|u|): below that the cells' own discretisation error is
    comparable to the jumps (see normalize_rows)."""
    return jump_scale(u4) > min_jump * np.maximum(1.0, np.abs(u4).max(axis=1))


SOURCES = ("linadv", "vortex", "riemann", "oscillatory")


def build_dataset(target_rows=10_000_000, seed=7, linadv_meshes=(100, 200, 400, 800),
                  vortex_meshes=(256, 512, 1024), riemann_meshes=(256, 512),
                  riemann_seeds=(7, 8, 9), riemann_times=(0.05, 0.1, 0.15, 0.2, 0.25),
                  osc_frac=0.05, log=print):
    """Harvest configs 1 and 2 (linear advection, isentropic vortex,
    four-quadrant Riemann problem), keep the informative stencils, add
    osc_frac of their number of oscillatory stencils (oscillatory_stencils),
    then draw up to target_rows of them in a seeded random order.  Returns
    the stencils, targets, the source of each row (an index into SOURCES)
    and the counts per source."""
    runs = [linadv_runs(meshes=linadv_meshes), vortex_runs(meshes=vortex_meshes),
            riemann_runs(meshes=riemann_meshes, seeds=riemann_seeds, times=riemann_times)]
    n_inf = sum(int(informative(np.concatenate([p[0] for p in parts])).sum())
                for parts in runs if parts)
    m = int(osc_frac * n_inf)
    runs.append([oscillatory_stencils(m)] if m else [])
    u4s, tgs, srcs, counts = [], [], [], {}
    for k, parts in enumerate(runs):
        u4 = np.concatenate([p[0] for p in parts]) if parts else np.zeros((0, 4))
        tg = np.concatenate([p[1] for p in parts]) if parts else np.zeros((0, 2))
        keep = informative(u4)
        counts[SOURCES[k]] = dict(harvested=int(len(u4)), informative=int(keep.sum()))
        log(f"{SOURCES[k]}: {len(u4)} stencils, {int(keep.sum())} informative")
        u4s.append(u4[keep])
        tgs.append(tg[keep])
        srcs.append(np.full(int(keep.sum()), k, dtype=np.int8))
    u4, tg, src = np.concatenate(u4s), np.concatenate(tgs), np.concatenate(srcs)
    perm = np.random.default_rng(seed).permutation(len(u4))[:int(target_rows)]
    return u4[perm], tg[perm], src[perm], counts


def normalize_rows(rows, min_jump=1e-3):
    """Jump-normalised rows: cells and targets (columns 15-20) mapped by
    u -> (u - u1) / max|jump|, features and vertices (columns 0-14, computed
    from the raw stencil, as the solver sees it) unchanged.  The interface
    values of the head are affine-equivariant in the cells, so the loss of a
    row becomes its raw loss divided by the squared jump scale: smooth and
    steep stencils weigh alike.  Rows whose jump scale is below min_jump *
    max(1, max|u|) are dropped: there the solver's own error in the cells
    (~1e-5 at these meshes) is comparable to the jumps and the target is not
    a function of the stencil."""
    c = rows[:, 15:19]
    s = jump_scale(c)
    keep = s > min_jump * np.maximum(1.0, np.abs(c).max(axis=1))
    out = rows[keep].copy()
    out[:, 15:21] = (out[:, 15:21] - out[:, 16:17]) / s[keep][:, None]
    return out


def interface_error(l, r, u4, tg):
    """Per-stencil interface error in units of the stencil's jump scale:
    (|l - t_left| + |r - t_right|) / max|jump|."""
    return (np.abs(l - tg[:, 0]) + np.abs(r - tg[:, 1])) / jump_scale(u4)


def gradient_scale(sol, rows, weight_decay):
    """The loss gradient of the network at its current weights on rows
    (FP32 backward) next to the weight-decay term weight_decay * theta that
    Adam adds to it: root-mean-square of each and their ratio.  Below a
    ratio of about 1 the decay steers the update and the weights shrink."""
    loss, g = sol.train_grad(rows, use_bf16=False)
    _, W, b = sol.get_net()
    theta = np.concatenate([W, b])
    grms = float(np.sqrt(np.mean(g * g)))
    drms = float(weight_decay * np.sqrt(np.mean(theta * theta)))
    return dict(loss=float(loss), grad_rms=grms, decay_rms=drms, ratio=grms / drms)


def train(u4, tg, src=None, *, epochs=40, batch=65536, lr=1e-3, seed=6008, holdout=200_000,
          init_net=None, normalize=True, weight_decay=1e-5, compare=None, log=print):
    """Rows on the device, Xavier-initialised 5-5-5-5-5 network (or the
    given net_file), Adam as the trainer; returns the trained solver
    context and a summary.

    The first `holdout` stencils are held out.  On them each network (the
    untrained one, the trained one and every net_file in `compare`, a dict
    name -> path) and SP-WENO reconstruct the raw stencils as the solver
    does, and the summary reports the mean interface error in units of
    each stencil's jump scale, pooled and per source.  The summary also
    holds the loss gradient against the weight decay at the start, on raw
    and on normalised rows, pooled and per source (gradient_scale)."""
    if src is None:
        src = np.zeros(len(u4), dtype=np.int8)
    te_u4, te_tg, te_src = u4[:holdout], tg[:holdout], src[:holdout]
    raw = tecno.make_train_rows(u4[holdout:], tg[holdout:])
    tr_src = src[holdout:]
    if normalize:
        s_ = jump_scale(raw[:, 15:19])
        keep = s_ > 1e-3 * np.maximum(1.0, np.abs(raw[:, 15:19]).max(axis=1))
        raw, tr_src = raw[keep], tr_src[keep]
        tr = normalize_rows(raw)
    else:
        tr = raw
    s = tecno.Solver(tecno.make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno_dl",
                                    n=(8, 2, 2), dims=1))
    if init_net:
        s.load_net(init_net)
    else:
        s.init_net([5, 5, 5, 5, 5], seed=seed)

    def errors(l, r):
        e = interface_error(l, r, te_u4, te_tg)
        out = {"all": float(e.mean())}
        for k, name in enumerate(SOURCES):
            m = te_src == k
            if m.any():
                out[name] = float(e[m].mean())
        return out

    def net_err(sol):
        return errors(*sol.mlp_forward_batch(te_u4))

    held = {"untrained": net_err(s)}
    s.train_init(lr=lr, beta1=0.5, beta2=0.9, eps=1e-8, weight_decay=weight_decay)
    grad = {}
    sample = np.random.default_rng(seed).permutation(len(tr))[:batch]
    for tag, rows in (("raw", raw), ("normalized", tr)):
        grad[tag] = {"all": gradient_scale(s, rows[sample], weight_decay)}
        for k, name in enumerate(SOURCES):
            m = np.nonzero(tr_src[sample] == k)[0]
            if len(m):
                grad[tag][name] = gradient_scale(s, rows[sample[m]], weight_decay)
    s.train_init(lr=lr, beta1=0.5, beta2=0.9, eps=1e-8, weight_decay=weight_decay)
    s.train_set_data(tr)
    losses = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        losses.append(s.train_epoch(batch, seed=seed + ep, use_bf16=True))
        log(f"epoch {ep}: loss {losses[-1]:.4e}")
    t_train = time.perf_counter() - t0
    held["trained"] = net_err(s)
    for name, path in (compare or {}).items():
        c = tecno.Solver(tecno.make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno_dl",
                                        n=(8, 2, 2), dims=1))
        c.load_net(path)
        held[name] = net_err(c)
        c.close()
    sp = tecno.Solver(tecno.make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno",
                                     n=(8, 2, 2), dims=1))
    held["sp_weno"] = errors(*sp.reconstruct_batch(te_u4))
    sp.close()
    summary = dict(rows=int(len(tr)), holdout=int(len(te_u4)), batch=batch, epochs=epochs,
                   normalized=bool(normalize), weight_decay=weight_decay,
                   rows_per_source={n: int((tr_src == k).sum()) for k, n in enumerate(SOURCES)},
                   steps=int(epochs * math.ceil(len(tr) / batch)), train_seconds=t_train,
                   loss_first=losses[0] if losses else None,
                   loss_last=losses[-1] if losses else None,
                   heldout_error=held, gradient_vs_decay=grad)
    return s, summary


def evaluate_riemann(nets, n=256, seeds=(11, 12, 13, 14), t_end=0.25, factor=4, log=print):
    """Config 2's four-quadrant Riemann problem (kepes_tecno4_roe, CFL 0.5,
    neumann, FP64) to t_end on an n x n mesh with SP-WENO and with DSP-WENO
    under each net_file of `nets` (name -> path), against the same problem
    run with SP-WENO on a `factor` times finer mesh and restricted (block
    means of the conserved variables).  Seeds other than the harvest's give
    states the networks were not trained on.  Returns name -> {"rho": mean
    |rho - rho_ref|, "all": mean over the four primitive variables} averaged
    over the seeds, or None for a run that failed."""
    res = {}
    for seed in seeds:
        nf = n * factor
        f = tecno.Solver(tecno.make_cfg(n=(nf, nf, 2), dims=2, recon="sp_weno",
                                        flux="kepes_tecno4_roe", bc="neumann", cfl=0.5))
        f.set_state(np.stack(quadrant_prim(nf, seed)).reshape(4, -1))
        f.advance(t_end, 1)
        ref = np.stack(restrict(f.get_state().reshape(4, nf, nf), factor))
        f.close()
        for name, path in {"sp_weno": None, **nets}.items():
            s = tecno.Solver(tecno.make_cfg(n=(n, n, 2), dims=2,
                                            recon="sp_weno" if path is None else "sp_weno_dl",
                                            flux="kepes_tecno4_roe", bc="neumann", cfl=0.5))
            if path is not None:
                s.load_net(path)
            s.set_state(np.stack(quadrant_prim(n, seed)).reshape(4, -1))
            try:
                s.advance(t_end, 1)
                d = np.abs(s.get_state().reshape(4, n, n) - ref).mean(axis=(1, 2))
                e = np.array([d[0], d.mean()])
            except RuntimeError as exc:
                log(f"riemann seed {seed} {name}: {exc}")
                e = None
            s.close()
            prev = res.get(name, np.zeros(2))
            res[name] = None if (e is None or prev is None) else prev + e / len(seeds)
    return {k: (None if v is None else {"rho": float(v[0]), "all": float(v[1])})
            for k, v in res.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True, help="net_file to write")
    ap.add_argument("--stencils", type=float, default=1e7)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--weight-decay", type=float, default=1e-5)
    ap.add_argument("--init", default=None, help="start from this net_file instead of Xavier")
    ap.add_argument("--raw", action="store_true", help="train on raw (not jump-normalised) rows")
    ap.add_argument("--compare", action="append", default=[], metavar="NAME=NET_FILE",
                    help="also report this network's held-out error")
    ap.add_argument("--data", default=None,
                    help="train on a data file in the reference trainer's layout (its data_file) "
                         "instead of harvesting")
    ap.add_argument("--write-data", default=None,
                    help="also write the harvested stencils in the reference trainer's layout")
    ap.add_argument("--eval-riemann", type=int, default=0, metavar="N",
                    help="also solve config 2's Riemann problem on N x N with each network and "
                         "report its L1 distance from a 4x finer SP-WENO run (evaluate_riemann)")
    ap.add_argument("--eval-only", action="store_true",
                    help="no harvest or training: evaluate the --compare networks only")
    a = ap.parse_args()
    compare = dict(c.split("=", 1) for c in a.compare)
    if a.eval_only:
        print(json.dumps({"riemann_l1": evaluate_riemann(compare, n=a.eval_riemann or 256)}))
        return
    t0 = time.perf_counter()
    src = None
    if a.data:
        rows = read_trainer_data(a.data)
        u4, tg = rows[:, 15:19], rows[:, 19:21]
        print(f"read {len(u4)} stencils from {a.data}", flush=True)
    else:
        u4, tg, src, counts = build_dataset(a.stencils, log=lambda m: print(m, flush=True))
        print(f"harvest {json.dumps(counts)}, {len(u4)} drawn, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        if a.write_data:
            write_trainer_data(a.write_data, tecno.make_train_rows(u4, tg))
    s, summary = train(u4, tg, src, epochs=a.epochs, batch=a.batch, init_net=a.init,
                       normalize=not a.raw, weight_decay=a.weight_decay, compare=compare,
                       log=lambda m: print(m, flush=True))
    s.save_net(a.out)
    s.close()
    if a.eval_riemann:
        summary["riemann_l1"] = evaluate_riemann({"trained": a.out, **compare}, n=a.eval_riemann,
                                                 log=lambda m: print(m, flush=True))
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
