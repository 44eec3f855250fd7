"""Golden vectors for the DSP-WENO trainer semantics, written with PyTorch.

A project-written float64 PyTorch restatement of the trainer's network head
and loss (DSP-WENO_trainer/DSP-WENO_network_trainer.py:63-106 forward and
inter_evaluation, 355-356 loss, 353 optimiser) produces, for a fixed batch:
the loss, its autograd gradient and the weights after two torch.optim.Adam
steps (lr 1e-3, betas (0.5, 0.9), weight_decay 1e-5, eps 1e-8).  The batch
rows (features, vertices, cells, targets) come from the oracle's row builder
on seeded stencils.  Run from the repository root:

    python tests/golden/make_trainer_fixtures.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
import oracle_ctypes as orc  # noqa: E402

torch.set_default_dtype(torch.float64)


def stencils(rng, n):
    """smooth and non-smooth 4-cell stencils, no zero centre jump"""
    x = rng.uniform(-1, 1, (n, 1)) + np.arange(4)[None, :] * rng.uniform(0.01, 0.3, (n, 1))
    f = np.where(rng.random((n, 1)) < 0.5, np.sin(3 * x), np.sign(x) * x ** 2)
    f += 0.05 * rng.standard_normal((n, 4)) * (rng.random((n, 1)) < 0.3)
    keep = np.abs(f[:, 2] - f[:, 1]) > 1e-6
    f = f[keep]
    xs = x[keep]
    h = xs[:, 1] - xs[:, 0]
    xf = xs[:, 1] + 0.5 * h
    t = np.stack([np.sin(3 * xf), np.sin(3 * xf)], axis=1)
    return f, t


class Head(torch.nn.Module):
    def __init__(self, widths, W, b):
        super().__init__()
        self.layers = torch.nn.ModuleList()
        ow = ob = 0
        for l in range(len(widths) - 1):
            ni, no = widths[l], widths[l + 1]
            lin = torch.nn.Linear(ni, no)
            with torch.no_grad():
                lin.weight.copy_(torch.tensor(W[ow:ow + ni * no]).reshape(no, ni))
                lin.bias.copy_(torch.tensor(b[ob:ob + no]))
            self.layers.append(lin)
            ow += ni * no
            ob += no

    def forward(self, rows):
        s = rows[:, 0:5]
        for i, lin in enumerate(self.layers):
            s = lin(s)
            if i < len(self.layers) - 1:
                s = torch.relu(s)
        s = torch.softmax(s, dim=1)
        C1 = torch.sum(s * rows[:, 5:10], 1)
        C2 = torch.sum(s * rows[:, 10:15], 1)
        w1 = 0.25 - 2 * C1
        tw0 = 0.25 - 2 * C2
        u0, u1, u2, u3 = rows[:, 15], rows[:, 16], rows[:, 17], rows[:, 18]
        left = (1 - w1) * (0.5 * u1 + 0.5 * u2) + w1 * (-0.5 * u0 + 1.5 * u1)
        right = tw0 * (1.5 * u2 - 0.5 * u3) + (1 - tw0) * (0.5 * u1 + 0.5 * u2)
        return left, right


def params(m):
    W = np.concatenate([l.weight.detach().numpy().ravel() for l in m.layers])
    b = np.concatenate([l.bias.detach().numpy().ravel() for l in m.layers])
    return W, b


def grads(m):
    W = np.concatenate([l.weight.grad.numpy().ravel() for l in m.layers])
    b = np.concatenate([l.bias.grad.numpy().ravel() for l in m.layers])
    return W, b


def case(name, widths, W, b, rows, out):
    m = Head(widths, W, b)
    opt = torch.optim.Adam(m.parameters(), lr=1e-3, betas=(0.5, 0.9), weight_decay=1e-5, eps=1e-8)
    R = torch.tensor(rows)
    mse = torch.nn.MSELoss()
    for step in range(2):
        left, right = m(R)
        loss = (mse(left, R[:, 19]) + mse(right, R[:, 20])) / 2
        opt.zero_grad()
        loss.backward()
        gW, gb = grads(m)
        out[f"{name}_loss{step}"] = np.array(loss.item())
        out[f"{name}_gW{step}"] = gW
        out[f"{name}_gb{step}"] = gb
        opt.step()
        pW, pb = params(m)
        out[f"{name}_W{step + 1}"] = pW
        out[f"{name}_b{step + 1}"] = pb
    out[f"{name}_widths"] = np.array(widths, dtype=np.int32)
    out[f"{name}_W0"] = W
    out[f"{name}_b0"] = b


def main():
    rng = np.random.default_rng(6008)
    u4, tgt = stencils(rng, 600)
    rows = orc.train_rows(u4, tgt)
    out = {"rows": rows}
    widths, W, b = orc.net_arrays(orc.load_net())
    case("shipped", list(widths), W, b, rows, out)
    # a 5-8-6-5 network with seeded weights
    w2 = [5, 8, 6, 5]
    nW = sum(w2[i] * w2[i + 1] for i in range(3))
    nb = sum(w2[1:])
    case("w865", w2, rng.uniform(-0.6, 0.6, nW), rng.uniform(-0.2, 0.2, nb), rows, out)
    np.savez_compressed(os.path.join(HERE, "fixtures_trainer.npz"), **out)
    print("rows", rows.shape, "written")


if __name__ == "__main__":
    main()
