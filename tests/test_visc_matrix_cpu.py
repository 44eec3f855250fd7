"""The entropy-stable viscous flux (nodal_vlux_ES, vflux_ES.cc:15-151) as a
linear map K of the entropy-variable gradients, built column by column from
the oracle (tor_es_node_flux) at random node states, against closed forms:

  * with fix_visc (the K21, K31, K32 blocks kept), grad V^T K grad V equals
    the Navier-Stokes entropy production (tau : grad u / T + kappa |grad T|^2
    / T^2) / R for every gradient, and K is symmetric positive
    semi-definite;
  * the reference's flux (those blocks added, then overwritten at
    vflux_ES.cc:103-145) differs from it exactly in those blocks and is not
    symmetric; its symmetric part keeps the off-diagonal coupling at half
    strength and stays positive definite at the sampled states, so the
    reference's viscous entropy production still has the right sign (the
    overwrite alone does not make the total entropy rise)."""
import numpy as np
import pytest

import oracle_ctypes as orc
from tecno import make_cfg


def _cfg(fix, mu=0.02, pr=0.71, gamma=1.4, R=1.0):
    return make_cfg(n=(8, 8, 8), dims=3, model="ns_es", viscosity=("constant", mu), prandtl=pr,
                    gamma=gamma, gas_const=R, fix_visc=fix)


def _node(rng, gamma=1.4, R=1.0):
    rho, p = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
    u = rng.uniform(-1, 1, 3)
    beta = rho / (2 * p)
    s = np.log(p) - gamma * np.log(rho)
    return np.array([(gamma - s) / (gamma - 1) - beta * u @ u, *(2 * beta * u), -2 * beta])


def _matrix(cfg, avg):
    """K[i, j]: flux component i (x, y, z fluxes x 4) per unit gradient j
    (d/dx, d/dy, d/dz of V_u, V_v, V_w, V_l)"""
    K = np.zeros((12, 12))
    for j in range(12):
        T = np.zeros((3, 5))
        T[j // 4, 1 + j % 4] = 1.0
        K[:, j] = orc.es_node_flux(cfg, avg, T[0], T[1], T[2]).reshape(12)
    return K


def _ns_entropy_production(avg, G, mu, pr, gamma=1.4, R=1.0):
    """(tau : grad u / T + kappa |grad T|^2 / T^2) / R for node state avg and
    entropy-variable gradients G (3 x 5)"""
    Vu, Vl = avg[1:4], avg[4]
    u = -Vu / Vl
    T = -1.0 / (R * Vl)
    gu = np.array([[(-G[i, 1 + k] + Vu[k] * G[i, 4] / Vl) / Vl for k in range(3)] for i in range(3)])
    gT = G[:, 4] / (R * Vl * Vl)
    div = np.trace(gu)
    tau = mu * (gu + gu.T - (2.0 / 3.0) * div * np.eye(3))
    kappa = mu * gamma * R / (gamma - 1) / pr
    return (np.sum(tau * gu) / T + kappa * gT @ gT / T ** 2) / R


@pytest.mark.parametrize("seed", range(6))
def test_fixed_es_flux_is_ns_entropy_production(seed):
    rng = np.random.default_rng(seed)
    mu, pr = 0.02, 0.71
    cfg = _cfg(1, mu, pr)
    avg = _node(rng)
    K = _matrix(cfg, avg)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-13 * scale
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert ev.min() >= -1e-12 * ev.max(), ev
    for _ in range(20):
        G = rng.normal(size=(3, 5))
        g = G[:, 1:].reshape(12)
        q = g @ K @ g
        assert q == pytest.approx(_ns_entropy_production(avg, G, mu, pr), rel=1e-11)


def test_reference_es_flux_not_symmetric_but_dissipative():
    rng = np.random.default_rng(3)
    for _ in range(6):
        avg = _node(rng)
        K = _matrix(_cfg(0), avg)
        Kf = _matrix(_cfg(1), avg)
        assert np.abs(K - K.T).max() > 1e-3 * np.abs(K).max()
        # the dropped blocks are exactly the difference to the fixed matrix:
        # the y-flux rows per x-gradient and the z-flux rows per x, y gradient
        D = Kf - K
        assert np.abs(D[0:4, :]).max() == 0.0                 # x flux unchanged
        assert np.abs(D[4:8, 4:12]).max() == 0.0              # y flux: only V_x columns
        assert np.abs(D[8:12, 8:12]).max() == 0.0             # z flux: only V_x, V_y columns
        # the symmetric part of the reference's K is the fixed K with the
        # dropped blocks halved on both sides of the diagonal
        assert np.abs(0.5 * (K + K.T) - (Kf - 0.5 * (D + D.T))).max() <= 1e-15 * np.abs(K).max()
        assert np.linalg.eigvalsh(0.5 * (K + K.T)).min() > 0.0
