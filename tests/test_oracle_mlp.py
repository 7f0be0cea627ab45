"""Pins of the oracle's network restatement (oracle/mlp.py) that do not retype its formulas.

* SiLU and its derivative at closed-form points, through the identity z sigma(z) =
  z (1 + tanh(z / 2)) / 2 and a central difference of SiLU itself;
* the forward pass routes each layer through the intended operand (no transposed weight): with a
  sparse, non-symmetric choice of W1, W2, W3 the outputs are SiLU(SiLU(c_i x_i)) + b3 on the routed
  units, computed here with scalar math;
* loss_and_grad (FP64, bf=False) against central differences of its own loss, per parameter block,
  for every parameter, at <= 1e-6 relative (per block, in the max norm).
"""
import math

import numpy as np
import pytest

from oracle import mlp as R


def test_silu_closed_form_values():
    assert R._silu(np.array(0.0)) == 0.0
    assert abs(R._silu(np.array(1.0)) - 1.0 / (1.0 + math.exp(-1.0))) < 1e-16
    assert abs(R._silu(np.array(1.0)) - 0.7310585786300049) < 1e-16
    assert abs(R._silu(np.array(-1.0)) + 0.2689414213699951) < 1e-16
    z = np.linspace(-12.0, 12.0, 241)
    assert np.max(np.abs(R._silu(z) - z * (1.0 + np.tanh(0.5 * z)) / 2.0)) < 1e-14
    assert abs(R._dsilu(np.array(0.0)) - 0.5) < 1e-16
    h = 1e-5
    fd = (R._silu(z + h) - R._silu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - R._dsilu(z))) < 1e-9
    # limits: SiLU -> z for large z, -> 0 for very negative z; its minimum is at z ~ -1.2785
    assert abs(R._silu(np.array(40.0)) - 40.0) < 1e-12 and abs(R._silu(np.array(-40.0))) < 1e-15
    zmin = -1.2784645427610738
    assert abs(R._dsilu(np.array(zmin))) < 1e-12


def test_forward_routes_layers_through_the_intended_operands():
    p = np.zeros(R.N_PARAMS)
    W1 = np.zeros((R.NI, R.NH))
    W2 = np.zeros((R.NH, R.NH))
    W3 = np.zeros((R.NH, R.NO))
    c = np.array([0.5, -1.5, 2.0, 0.25, -0.75, 1.25, -2.5, 3.0])
    perm = lambda j: (j + 3) % R.NH   # noqa: E731
    for i in range(R.NI):
        W1[i, 2 * i + 1] = c[i]              # input i -> hidden unit 2i+1
    for j in range(R.NH):
        W2[j, perm(j)] = 1.0                 # hidden j -> hidden perm(j) (not symmetric)
    for i in range(R.NI):
        W3[perm(2 * i + 1), i] = 1.0         # routed unit -> output i
    b3 = np.linspace(-0.3, 0.3, R.NO)
    p[R.OFF_W1:R.OFF_B1] = W1.ravel()
    p[R.OFF_W2:R.OFF_B2] = W2.ravel()
    p[R.OFF_W3:R.OFF_B3] = W3.ravel()
    p[R.OFF_B3:] = b3
    x = np.random.default_rng(3).uniform(-1.0, 1.0, (5, R.NI))
    y, _ = R.forward(p, x)
    sil = lambda v: v / (1.0 + math.exp(-v))   # noqa: E731
    for n in range(5):
        for o in range(R.NO):
            expect = b3[o] + (sil(sil(c[o] * x[n, o])) if o < R.NI else 0.0)
            assert abs(y[n, o] - expect) < 1e-15, (n, o)


BLOCKS = [("W1", R.OFF_W1, R.OFF_B1), ("b1", R.OFF_B1, R.OFF_W2), ("W2", R.OFF_W2, R.OFF_B2),
          ("b2", R.OFF_B2, R.OFF_W3), ("W3", R.OFF_W3, R.OFF_B3), ("b3", R.OFF_B3, R.N_PARAMS)]


@pytest.mark.parametrize("seed", [1, 2])
def test_loss_and_grad_matches_central_differences(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.3, R.N_PARAMS)
    x = rng.uniform(-1.0, 1.0, (12, R.NI))
    t = rng.normal(0.0, 0.2, (12, R.NO))
    loss, g = R.loss_and_grad(p, x, t, bf=False)
    y, _ = R.forward(p, x)
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-15 * max(1.0, loss)
    h = 1e-5
    fd = np.zeros(R.N_PARAMS)
    for k in range(R.N_PARAMS):
        pp = p.copy()
        pp[k] += h
        lp, _ = R.loss_and_grad(pp, x, t, bf=False)
        pp[k] -= 2 * h
        lm, _ = R.loss_and_grad(pp, x, t, bf=False)
        fd[k] = (lp - lm) / (2 * h)
    for name, a, b in BLOCKS:
        scale = np.max(np.abs(fd[a:b]))
        assert scale > 0, name
        err = np.max(np.abs(g[a:b] - fd[a:b])) / scale
        assert err <= 1e-6, (name, err)
