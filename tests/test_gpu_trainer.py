"""GPU trainer (forward, backward, Adam) against the oracle and PyTorch
golden vectors; training on harvested stencils and what it gains."""
import os
import tempfile

import numpy as np
import pytest

import oracle_ctypes as orc
from cases import fixtures
from tecno import (FP32, MLP_FP32, Solver, harvest_stencils_1d, make_cfg, make_train_rows,
                   trainer_epoch_seeds)

pytestmark = pytest.mark.gpu
NET_FILE = os.path.join(os.path.dirname(__file__), "golden", "dsp_weno_net.txt")


def net_solver(widths=None, W=None, b=None):
    s = Solver(make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno_dl", n=(8, 2, 2), dims=1,
                        precision=FP32, mlp_dtype=MLP_FP32))
    if widths is None:
        s.load_net(NET_FILE)
    else:
        s.set_net(widths, W, b)
    return s


def test_make_train_rows_matches_oracle():
    fx = fixtures("trainer")
    rows = fx["rows"]
    got = make_train_rows(rows[:, 15:19], rows[:, 19:21])
    ref = orc.train_rows(rows[:, 15:19], rows[:, 19:21])
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("name", ["shipped", "w865"])
def test_gradient_matches_oracle_and_torch(name):
    """FP32 forward/backward: gradient within 2e-6 of the largest component of
    the FP64 gradient (oracle and PyTorch autograd), loss within 1e-6"""
    fx = fixtures("trainer")
    rows = fx["rows"]
    widths = list(fx[f"{name}_widths"])
    s = net_solver(widths, fx[f"{name}_W0"], fx[f"{name}_b0"])
    s.train_init()
    loss, g = s.train_grad(rows, use_bf16=False)
    ref = np.concatenate([fx[f"{name}_gW0"], fx[f"{name}_gb0"]])
    lo, go = orc.train_loss_grad(orc.net_from_arrays(widths, fx[f"{name}_W0"], fx[f"{name}_b0"]), rows)
    scale = np.abs(ref).max()
    assert np.abs(g - ref).max() <= 2e-6 * scale
    assert np.abs(g - go).max() <= 2e-6 * scale
    assert abs(loss - fx[f"{name}_loss0"]) <= 1e-6 * fx[f"{name}_loss0"]
    assert abs(loss - lo) <= 1e-6 * lo


def test_bf16_gradient_matches_oracle_bf16():
    """bfloat16 weights/activations: GPU gradient within 1% of the oracle's
    bfloat16 gradient (a tie rounding differently moves one activation by one
    bfloat16 step), and within 5% of the FP64 gradient"""
    fx = fixtures("trainer")
    rows = fx["rows"]
    s = net_solver()
    s.train_init()
    _, g = s.train_grad(rows, use_bf16=True)
    net = orc.load_net()
    _, gb = orc.train_loss_grad(net, rows, bf16=True)
    _, g64 = orc.train_loss_grad(net, rows)
    scale = np.abs(g64).max()
    assert np.abs(g - gb).max() <= 1e-2 * scale
    assert np.abs(g - g64).max() <= 5e-2 * scale


@pytest.mark.parametrize("name", ["shipped", "w865"])
def test_adam_steps_match_torch(name):
    """two FP32 Adam steps land on PyTorch's weights: every update is within
    1e-6 of torch's (updates are ~lr = 1e-3; a parameter whose gradient sits at
    the FP32 noise floor may move by up to 2 lr differently)"""
    fx = fixtures("trainer")
    rows = fx["rows"]
    widths = list(fx[f"{name}_widths"])
    s = net_solver(widths, fx[f"{name}_W0"], fx[f"{name}_b0"])
    s.train_init(lr=1e-3, beta1=0.5, beta2=0.9, eps=1e-8, weight_decay=1e-5)
    for t in (1, 2):
        s.train_step(rows, use_bf16=False)
        _, W, b = s.get_net()
        got = np.concatenate([W, b])
        ref = np.concatenate([fx[f"{name}_W{t}"], fx[f"{name}_b{t}"]])
        d = np.abs(got - ref)
        g = np.abs(np.concatenate([fx[f"{name}_gW0"], fx[f"{name}_gb0"]]))
        solid = g > 1e-4 * g.max()
        assert d[solid].max() <= 1e-6, d[solid].max()
        assert d.max() <= 2.1e-3 * t


def test_save_and_reload_net():
    fx = fixtures("trainer")
    s = net_solver()
    s.train_init()
    s.train_step(fx["rows"], use_bf16=True)
    w0, W0, b0 = s.get_net()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "trained.txt")
        s.save_net(path)
        s2 = net_solver()
        s2.load_net(path)
        w1, W1, b1 = s2.get_net()
        text = open(path).read()
    assert text.startswith("Weights\nLayer 1\nBEGIN\n") and "\nBiases\n" in text
    assert list(w0) == list(w1)
    assert np.array_equal(W0, W1) and np.array_equal(b0, b1)


def smooth_stencils(rng, n):
    """cell averages of A sin(k x + phi) + c on unit cells and the exact value
    at the centre face, k h in [0.05, 1.2]"""
    kh = rng.uniform(0.05, 1.2, n)
    A = rng.uniform(0.1, 2.0, n) * rng.choice([-1, 1], n)
    phi = rng.uniform(0, 2 * np.pi, n)
    c = rng.uniform(-1, 1, n)
    j = np.arange(4)[None, :] - 1.0
    avg = (A[:, None] * (np.cos(kh[:, None] * (j - 0.5) + phi[:, None]) -
                         np.cos(kh[:, None] * (j + 0.5) + phi[:, None])) / kh[:, None] + c[:, None])
    face = A * np.sin(kh * 0.5 + phi) + c
    keep = np.abs(avg[:, 2] - avg[:, 1]) > 1e-9
    return avg[keep], np.stack([face[keep], face[keep]], axis=1)


def linadv_harvest(n, t):
    """stencils of the 1D linear-advection solver (sin^4 on [-pi, pi],
    lin_tecno4 + sp_weno) after time t, targets the exact face values"""
    h = 2 * np.pi / n
    x = -np.pi + (np.arange(n) + 0.5) * h
    s = Solver(make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno", n=(n, 2, 2), dims=1,
                        lo=(-np.pi, 0, 0), hi=(np.pi, 1, 1), bc="periodic", cfl=0.5, precision=FP32))
    s.set_state(np.sin(x) ** 4)
    s.advance(t, 1)
    u = s.get_state().reshape(-1)
    tt, _ = s.time()
    xf = x + 0.5 * h
    return harvest_stencils_1d(u, np.sin(xf - tt) ** 4)


def test_init_net_matches_pytorch_draws():
    """tecno_init_net through the C ABI: the weights and biases PyTorch draws
    for the trainer's initialisation under torch.manual_seed(seed)
    (DSP-WENO_network_trainer.py:54-60, 296), bit for bit, for the three
    golden width / seed cases."""
    fix = np.load(os.path.join(os.path.dirname(__file__), "golden", "fixtures_init.npz"))
    s = net_solver()
    for k in range(3):
        widths = [int(x) for x in fix[f"widths{k}"]]
        s.init_net(widths, int(fix[f"seed{k}"]))
        w, W, b = s.get_net()
        assert list(w) == widths
        assert np.array_equal(W, fix[f"W{k}"]) and np.array_equal(b, fix[f"b{k}"])
    s.close()


def test_epoch_takes_the_dataloader_order():
    """tecno_train_epoch with the trainer's epoch seeds (tecno_trainer_epoch_seeds)
    equals Adam steps on the batches of the DataLoader's row order recorded from
    PyTorch (golden epoch_orders: 1000 rows, batches of 64, the last one of 40;
    trainer:348, 384-392)."""
    fix = np.load(os.path.join(os.path.dirname(__file__), "golden", "fixtures_init.npz"))
    orders = fix["epoch_orders"]
    rng = np.random.default_rng(3)
    u4, tg = smooth_stencils(rng, 1200)
    rows = make_train_rows(u4[:1000], tg[:1000])
    seeds = trainer_epoch_seeds([5, 5, 5, 5, 5], 6008, 2)
    a, b = net_solver(), net_solver()
    for s in (a, b):
        s.init_net([5, 5, 5, 5, 5], 6008)
        s.train_init(lr=1e-3, beta1=0.5, beta2=0.9, eps=1e-8, weight_decay=1e-5)
    a.train_set_data(rows)
    for e in range(2):
        a.train_epoch(64, seed=seeds[e], use_bf16=True)
        for r0 in range(0, 1000, 64):
            b.train_step(rows[orders[e, r0:r0 + 64]], use_bf16=True)
    _, Wa, ba = a.get_net()
    _, Wb, bb = b.get_net()
    np.testing.assert_allclose(Wa, Wb, rtol=1e-12, atol=0)
    np.testing.assert_allclose(ba, bb, rtol=1e-12, atol=0)
    # a different epoch order gives different weights at this tolerance
    c = net_solver()
    c.init_net([5, 5, 5, 5, 5], 6008)
    c.train_init(lr=1e-3, beta1=0.5, beta2=0.9, eps=1e-8, weight_decay=1e-5)
    c.train_set_data(rows)
    for e in range(2):
        c.train_epoch(64, seed=seeds[e] + 1, use_bf16=True)
    _, Wc, _ = c.get_net()
    assert not np.allclose(Wa, Wc, rtol=1e-9, atol=0)
    for s in (a, b, c):
        s.close()


def test_training_on_harvested_stencils_beats_spweno():
    """Xavier-initialised 5-5-5-5-5 network trained on the GPU (bf16
    activations, fp32 master weights, Adam 1e-3 / (0.5, 0.9) / 1e-5) on stencils
    harvested from linear-advection runs plus smooth sine stencils: the loss
    falls, and on held-out stencils its interface error is below both the
    untrained network's and SP-WENO's (the method without the network)."""
    rng = np.random.default_rng(11)
    parts = [linadv_harvest(n, 0.3) for n in (40, 80, 160)] + [smooth_stencils(rng, 40000)]
    u4 = np.concatenate([p[0] for p in parts])
    tg = np.concatenate([p[1] for p in parts])
    perm = rng.permutation(len(u4))
    u4, tg = u4[perm], tg[perm]
    rows = make_train_rows(u4, tg)
    nte = 5000
    tr, te = rows[nte:], rows[:nte]
    s = net_solver()
    s.init_net([5, 5, 5, 5, 5], seed=6008)

    def err(sol):
        l, r = sol.mlp_forward_batch(te[:, 15:19])
        return np.mean(np.abs(l - te[:, 19])) + np.mean(np.abs(r - te[:, 20]))

    e0 = err(s)
    s.train_init(lr=1e-3, beta1=0.5, beta2=0.9, eps=1e-8, weight_decay=1e-5)
    s.train_set_data(tr)
    losses = [s.train_epoch(500, seed=ep, use_bf16=True) for ep in range(20)]
    e1 = err(s)
    sp = Solver(make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno", n=(8, 2, 2), dims=1))
    l, r = sp.reconstruct_batch(te[:, 15:19])
    espw = np.mean(np.abs(l - te[:, 19])) + np.mean(np.abs(r - te[:, 20]))
    print(f"held-out interface error: untrained {e0:.4e} trained {e1:.4e} sp_weno {espw:.4e}; "
          f"loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.5 * losses[0]
    assert e1 < e0 and e1 < espw
