"""Host side of the retraining pipeline (dsp-weno_amd/retrain.py): the
characteristic stencils harvested from 2D Euler states and the row
normalisation.  The eigenvectors are pinned by Barth's identity, the targets
by interpolation of exact stencils, not by restating the formulas."""
import math

import numpy as np
import pytest

import retrain as R


def _vortex_harvest(n, t=0.0):
    h = 10.0 / n
    c = -5 + (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(c, c)

    def exact_face(axis):
        if axis == 0:
            return R.vortex_prim(n, t, xs=X + 0.5 * h, ys=Y)
        return R.vortex_prim(n, t, xs=X, ys=Y + 0.5 * h)

    return R.harvest_euler_2d(R.vortex_prim(n, t), exact_face)


def _jacobians(prim, gamma=1.4, e=1e-6):
    def U_of(q):
        r, a, b, p = q
        return np.array([r, r * a, r * b, p / (gamma - 1) + 0.5 * r * (a * a + b * b)])

    def V_of(q):
        return np.array(R.entropy_vars(*q, gamma=gamma))

    JU, JV = np.zeros((4, 4)), np.zeros((4, 4))
    for k in range(4):
        d = np.zeros(4)
        d[k] = e
        JU[:, k] = (U_of(prim + d) - U_of(prim - d)) / (2 * e)
        JV[:, k] = (V_of(prim + d) - V_of(prim - d)) / (2 * e)
    return JU @ np.linalg.inv(JV)


def test_face_eigenvectors_satisfy_barth_scaling():
    """R~ R~^T = dU/dV at equal left/right states (the scaled eigenvectors
    of kepes_tecno4_roe.cc:66-89 symmetrise the entropy Jacobian); a dropped
    or wrong scaling factor breaks the identity."""
    rng = np.random.default_rng(3)
    for _ in range(5):
        rho, un, ut, p = rng.uniform(0.5, 2.0), rng.uniform(-1, 1), rng.uniform(-1, 1), \
            rng.uniform(0.3, 3.0)
        a = [np.array([x]) for x in (rho, un, ut, p)]
        Rm = R.face_eigvecs(*a, *a)[..., 0]  # [wave, row]
        dUdV = _jacobians(np.array([rho, un, ut, p]))
        assert np.abs(Rm.T @ Rm - dUdV).max() <= 1e-7 * np.abs(dUdV).max()


def test_harvest_targets_converge_third_order():
    """On exact vortex states the target (projection of the exact face state)
    and the 4-point interpolation of the stencil differ by O(h^3): halving h
    divides the largest difference by about 8."""
    d = []
    for n in (64, 128):
        u4, tg = _vortex_harvest(n)
        interp = (-u4[:, 0] + 9 * u4[:, 1] + 9 * u4[:, 2] - u4[:, 3]) / 16
        d.append(np.abs(tg[:, 0] - interp).max())
        assert np.array_equal(tg[:, 0], tg[:, 1])
    assert 6.0 < d[0] / d[1] < 10.0


def test_harvest_counts_and_zero_jump_drop():
    """Every kept stencil has a centre jump above the drop threshold; a
    uniform state yields no stencil at all."""
    n = 32
    u4, tg = _vortex_harvest(n)
    assert len(u4) <= 2 * 4 * n * n and len(u4) == len(tg)
    scale = np.maximum(1.0, np.abs(u4).max(axis=1))
    assert np.all(np.abs(u4[:, 2] - u4[:, 1]) > 1e-12 * scale)
    ones = np.ones((n, n))
    u4u, _ = R.harvest_euler_2d((ones, 0.5 * ones, 0.0 * ones, ones),
                                lambda axis: (ones, 0.5 * ones, 0.0 * ones, ones))
    assert len(u4u) == 0


def test_scalar_harvest_periodic_stencils():
    n = 50
    h = 2 * math.pi / n
    x = -math.pi + (np.arange(n) + 0.5) * h
    u = np.sin(x)
    u4, tg = R.harvest_scalar_1d(u, np.sin(x + 0.5 * h))
    assert len(u4) == n
    assert np.allclose(u4[0], [u[-1], u[0], u[1], u[2]])
    assert np.allclose(tg[:, 0], np.sin(x + 0.5 * h))


def test_normalize_rows_is_affine_and_drops_noise():
    """Cells and targets map by (u - u1)/max|jump|; features untouched;
    rows whose jumps are below 1e-3 of the stencil magnitude are dropped."""
    rng = np.random.default_rng(5)
    rows = rng.normal(size=(6, 21))
    rows[2, 15:19] = [1.0, 1.0 + 1e-6, 1.0 + 2e-6, 1.0 + 3e-6]  # noise-level jumps
    out = R.normalize_rows(rows)
    assert len(out) == 5
    keep = np.array([0, 1, 3, 4, 5])
    c = rows[keep, 15:19]
    s = np.abs(np.diff(c, axis=1)).max(axis=1)
    assert np.allclose(out[:, :15], rows[keep, :15])
    assert np.allclose(out[:, 15:21], (rows[keep, 15:21] - c[:, 1:2]) / s[:, None])
    assert np.allclose(np.abs(np.diff(out[:, 15:19], axis=1)).max(axis=1), 1.0)


def test_trainer_data_file_round_trip(tmp_path):
    """A data file in the reference trainer's 23-column layout is read into
    device-trainer rows and written back value for value; the feature
    columns of a written file are the oracle's row builder's (the same
    features the solver computes, neural_network.cc:16-77)."""
    import oracle_ctypes as orc
    rng = np.random.default_rng(9)
    u4 = rng.normal(size=(50, 4))
    tg = rng.normal(size=(50, 2))
    rows = orc.train_rows(u4, tg)
    p = tmp_path / "data.dat"
    R.write_trainer_data(str(p), rows)
    back = R.read_trainer_data(str(p))
    assert np.array_equal(back, rows)
    line = open(p).readline().split(",")
    assert len(line) == 23
    v = [float(x) for x in line]
    assert np.allclose(v[0:4], u4[0]) and np.allclose(v[21:23], tg[0])
    assert v[4] == pytest.approx((u4[0, 3] - u4[0, 2]) / (u4[0, 2] - u4[0, 1]))
    # 21-column layout (no tanh features): inputs start at theta-
    p21 = tmp_path / "d21.dat"
    with open(p21, "w") as f:
        for r, c, t in zip(rows, u4, tg):
            f.write(",".join(f"{x:.17g}" for x in [*c, *r[0:15], *t]) + "\n")
    assert np.array_equal(R.read_trainer_data(str(p21)), rows)
