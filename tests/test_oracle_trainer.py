"""The oracle's trainer restatement (loss, analytic gradient, Adam) against
PyTorch golden vectors, finite differences and closed forms."""
import numpy as np
import pytest

import oracle_ctypes as orc
from cases import fixtures


def trainer_fx():
    return fixtures("trainer")


def net_of(fx, name, step):
    return orc.net_from_arrays(list(fx[f"{name}_widths"]), fx[f"{name}_W{step}"], fx[f"{name}_b{step}"])


@pytest.mark.parametrize("name", ["shipped", "w865"])
def test_loss_and_gradient_match_torch(name):
    """loss and autograd gradient of the PyTorch head/loss, both Adam steps"""
    fx = trainer_fx()
    rows = fx["rows"]
    for step in range(2):
        net = net_of(fx, name, step)
        loss, g = orc.train_loss_grad(net, rows)
        ref = np.concatenate([fx[f"{name}_gW{step}"], fx[f"{name}_gb{step}"]])
        assert abs(loss - fx[f"{name}_loss{step}"]) <= 1e-13 * abs(fx[f"{name}_loss{step}"])
        assert np.abs(g - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["shipped", "w865"])
def test_adam_matches_torch(name):
    """two torch.optim.Adam steps (betas 0.5/0.9, weight decay 1e-5)"""
    fx = trainer_fx()
    rows = fx["rows"]
    net = net_of(fx, name, 0)
    m, v = orc.zero_like(net), orc.zero_like(net)
    for t in (1, 2):
        orc.adam_step(net, rows, m, v, t)
        _, W, b = orc.net_arrays(net)
        assert np.abs(W - fx[f"{name}_W{t}"]).max() <= 1e-14
        assert np.abs(b - fx[f"{name}_b{t}"]).max() <= 1e-14


def test_gradient_central_differences():
    """analytic gradient vs central differences of the oracle's own loss"""
    fx = trainer_fx()
    rows = fx["rows"][:200]
    widths, W, b = orc.net_arrays(orc.load_net())
    _, g = orc.train_loss_grad(orc.net_from_arrays(list(widths), W, b), rows)
    p = np.concatenate([W, b])
    nW = W.size
    rng = np.random.default_rng(1)
    for q in rng.choice(p.size, 30, replace=False):
        h = 1e-6 * max(1.0, abs(p[q]))
        lp, lm = p.copy(), p.copy()
        lp[q] += h
        lm[q] -= h
        fp = orc.train_loss(orc.net_from_arrays(list(widths), lp[:nW], lp[nW:]), rows)
        fm = orc.train_loss(orc.net_from_arrays(list(widths), lm[:nW], lm[nW:]), rows)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1e-6, np.abs(g).max()), (q, fd, g[q])


def test_training_forward_is_the_solver_forward():
    """rows built from stencils, with the network's own reconstruction as the
    targets, give zero loss: the training forward equals Neural_Network::forward
    (pinned by the reference's fixtures elsewhere)"""
    fx = trainer_fx()
    u4 = fx["rows"][:, 15:19]
    net = orc.load_net()
    l, r = orc.nn_forward(net, u4)
    rows = orc.train_rows(u4, np.stack([l, r], axis=1))
    assert orc.train_loss(net, rows) <= 1e-28


def test_adam_first_step_closed_form():
    """with zero moments, step 1 moves every parameter by
    -lr * g' / (|g'| + eps), g' = g + wd * p (bias correction cancels)"""
    fx = trainer_fx()
    rows = fx["rows"]
    net = orc.load_net()
    widths, W0, b0 = orc.net_arrays(net)
    _, g = orc.train_loss_grad(net, rows)
    p0 = np.concatenate([W0, b0])
    gp = g + 1e-5 * p0
    m, v = orc.zero_like(net), orc.zero_like(net)
    orc.adam_step(net, rows, m, v, 1, lr=1e-3, beta1=0.5, beta2=0.9, eps=1e-8, wd=1e-5)
    _, W1, b1 = orc.net_arrays(net)
    dp = np.concatenate([W1, b1]) - p0
    assert np.allclose(dp, -1e-3 * gp / (np.abs(gp) + 1e-8), rtol=1e-10, atol=1e-18)


def test_bf16_loss_close():
    """bfloat16 weights and activations change the loss by a few per mille"""
    fx = trainer_fx()
    rows = fx["rows"]
    net = orc.load_net()
    l64 = orc.train_loss(net, rows)
    l16 = orc.train_loss(net, rows, bf16=True)
    assert abs(l16 - l64) <= 2e-2 * l64
