"""Golden vectors for the seeded initial network (tecno_init_net), written
with PyTorch.

The trainer seeds PyTorch's CPU generator with torch.manual_seed(seed_no)
(DSP-WENO_trainer/DSP-WENO_network_trainer.py:296, seed_no 6008 by default,
config.py:32) with float64 as the default dtype (:254), then builds its
network as one nn.Linear per layer in order (:54-56) -- each one's default
initialisation draws its weights (kaiming_uniform_, a = sqrt(5)) and then
its bias U(+-1/sqrt(fan_in)) -- and finally redraws every weight with
nn.init.xavier_uniform_ (:58-60).  This script performs that call sequence
(torch.nn.Linear and torch.nn.init, not the trainer's file) for the shipped
5-5-5-5-5 architecture and two other widths / seeds, and stores the weights
(row-major out x in, layers concatenated) and biases.

It also records the training row order: torch.randperm under seeded
generators (the RandomSampler's permutation; whole permutations for small n,
the ends and a checksum for 3M rows), and the orders of the first three epochs
of a DataLoader(shuffle=True) over 1000 rows created, as in the trainer
(:348, 351, 384-392), around the seed-6008 5-5-5-5-5 network.  Run from the
repository root:

    python tests/golden/make_init_fixtures.py
"""
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = [([5, 5, 5, 5, 5], 6008), ([5, 8, 3, 5], 1234), ([5, 16, 16, 16, 16, 5], 2 ** 32 + 77)]


def init_net(widths, seed):
    torch.set_default_dtype(torch.float64)
    torch.manual_seed(seed)
    layers = [torch.nn.Linear(widths[i], widths[i + 1]) for i in range(len(widths) - 1)]
    for lay in layers:
        torch.nn.init.xavier_uniform_(lay.weight)
    W = np.concatenate([lay.weight.detach().numpy().ravel() for lay in layers])
    b = np.concatenate([lay.bias.detach().numpy().ravel() for lay in layers])
    return W, b


PERMS = [(10, 1), (1000, 2 ** 40 + 3), (5000, 987654321), (3000000, 6008)]


def checksum(p):
    """sum of i * p[i] modulo 2^64"""
    i = np.arange(len(p), dtype=np.uint64)
    return int(np.sum(i * p.astype(np.uint64), dtype=np.uint64))


def dataloader_orders(n=1000, batch=64, epochs=3):
    torch.set_default_dtype(torch.float64)
    torch.manual_seed(6008)
    loader = torch.utils.data.DataLoader(torch.arange(n, dtype=torch.float64),
                                         batch_size=batch, shuffle=True)
    widths = [5, 5, 5, 5, 5]
    layers = [torch.nn.Linear(widths[i], widths[i + 1]) for i in range(len(widths) - 1)]
    for lay in layers:
        torch.nn.init.xavier_uniform_(lay.weight)
    return np.stack([torch.cat(list(loader)).long().numpy().astype(np.int32)
                     for _ in range(epochs)])


def main():
    out = {}
    for k, (widths, seed) in enumerate(CASES):
        W, b = init_net(widths, seed)
        out[f"widths{k}"] = np.array(widths, dtype=np.int32)
        out[f"seed{k}"] = np.array(seed, dtype=np.uint64)
        out[f"W{k}"] = W
        out[f"b{k}"] = b
    for k, (n, seed) in enumerate(PERMS):
        g = torch.Generator()
        g.manual_seed(seed)
        p = torch.randperm(n, generator=g).numpy()
        out[f"perm_n{k}"] = np.array(n, dtype=np.int64)
        out[f"perm_seed{k}"] = np.array(seed, dtype=np.uint64)
        if n <= 5000:
            out[f"perm{k}"] = p.astype(np.int32)
        else:
            out[f"perm_head{k}"] = p[:64].astype(np.int32)
            out[f"perm_tail{k}"] = p[-64:].astype(np.int32)
            out[f"perm_sum{k}"] = np.array(checksum(p), dtype=np.uint64)
    out["epoch_orders"] = dataloader_orders()
    out["torch_version"] = np.array(torch.__version__)
    np.savez(os.path.join(HERE, "fixtures_init.npz"), **out)
    print("wrote", len(CASES), "cases with torch", torch.__version__)


if __name__ == "__main__":
    main()
