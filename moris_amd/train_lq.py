"""Train the learned cut-cell quadrature network on the GPU and write its weights.

Training set: seeded random cut cells (planes, spheres, gyroid patches) decomposed by the XTK
restatement; targets are the 27-point rules that reproduce the tri-quadratic moments of each
cell's material region.  The network (8 -> 64 -> 64 -> 27 SiLU) is trained with the library's
own HIP forward/backward kernels (BF16 MFMA, FP32 master weights, Adam, cosine learning rate).

    python -m moris_amd.train_lq [--samples 1000000] [--epochs 20] [--out moris_amd/data/lq_mlp.f32]

Writes the FP32 parameters (little endian, layout W1[8][64] b1 W2[64][64] b2 W3[64][27] b3) and a
JSON sidecar with the loss history and a held-out evaluation.
"""
import argparse
import json
import os
import time

import numpy as np

from . import learned as L
from .workloads import BASE_SEED

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "lq_mlp.f32")


def load(path=DEFAULT_OUT):
    """A network with the stored weights, or None if the file is absent."""
    if not os.path.exists(path):
        return None
    p = np.fromfile(path, dtype="<f4")
    net = L.MLP(seed=1)
    if p.size != net.n_params:
        raise ValueError("%s holds %d parameters, the network has %d" % (path, p.size, net.n_params))
    net.params = p
    return net


def evaluate(net, n=20000, seed=BASE_SEED + 41):
    """Held-out mean squared error of the weights / 8 and of the volume-scaled Gauss prior."""
    x, y, m = L.dataset(n, seed, with_moments=True)
    _, _, g = L.operators()
    yp = net.forward(x)
    prior = (g[None, :] * m[:, :1] / 8.0) / 8.0
    return {"mse_network": float(np.mean((yp - y) ** 2)), "mse_gauss_prior": float(np.mean((prior - y) ** 2)),
            "samples": n, "seed": seed}


def train(samples=1_000_000, epochs=20, batch=1024, lr=3e-3, seed=BASE_SEED + 4, net_seed=11, log=print):
    t0 = time.time()
    x, y, _ = L.dataset(samples, seed)
    t_data = time.time() - t0
    net = L.MLP(seed=net_seed)
    t0 = time.time()
    hist = net.train(x, y, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    t_train = time.time() - t0
    log("dataset %.1fs, training %.1fs, loss %.3e -> %.3e" % (t_data, t_train, hist[0], hist[-1]))
    info = {"samples": samples, "epochs": epochs, "batch": batch, "lr": lr, "data_seed": seed, "net_seed": net_seed,
            "loss_history": [float(v) for v in hist], "dataset_s": t_data, "train_s": t_train,
            "samples_per_s": samples * epochs / t_train}
    return net, info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    net, info = train(a.samples, a.epochs, a.batch, a.lr)
    info["held_out"] = evaluate(net)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.params.astype("<f4").tofile(a.out)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(info, f, indent=1)
    print(json.dumps({k: v for k, v in info.items() if k != "loss_history"}))


if __name__ == "__main__":
    main()
