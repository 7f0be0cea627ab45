"""TEST INFRASTRUCTURE ONLY -- CPU restatement (numpy) of the learned-quadrature network
8 -> 64 -> 64 -> 27 with SiLU, its mean-squared-error loss and its parameter gradients.

`bf16=True` rounds the GEMM operands to bfloat16 (round to nearest even) at the same points as the
HIP kernels (inputs, weights, hidden activations, dY, dZ), with products summed in float64; this
isolates the kernels' layout/indexing from the BF16 rounding they are built on.
"""
import numpy as np

NI, NH, NO = 8, 64, 27
OFF_W1 = 0
OFF_B1 = OFF_W1 + NI * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NO
N_PARAMS = OFF_B3 + NO


def bf16(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NI, NH), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(NH, NO), p[OFF_B3:N_PARAMS])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, bf=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    x = np.asarray(x, np.float64)
    z1 = r(x) @ r(W1) + b1
    h1 = r(_silu(z1))
    z2 = h1 @ r(W2) + b2
    h2 = r(_silu(z2))
    y = h2 @ r(W3) + b3
    return y, (x, z1, h1, z2, h2)


def loss_and_grad(p, x, t, bf=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, (x, z1, h1, z2, h2) = forward(p, x, bf)
    n = x.shape[0]
    e = y - np.asarray(t, np.float64)
    loss = np.sum(e * e) / (n * NO)
    dy = 2.0 * e / (n * NO)
    g = np.zeros(N_PARAMS)
    g[OFF_B3:N_PARAMS] = dy.sum(0)
    g[OFF_W3:OFF_B3] = (h2.T @ r(dy)).ravel()
    dz2 = (r(dy) @ r(W3).T) * _dsilu(z2)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (h1.T @ r(dz2)).ravel()
    dz1 = (r(dz2) @ r(W2).T) * _dsilu(z1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (r(x).T @ r(dz1)).ravel()
    return loss, g
