"""The learned proposal's CPU oracle, pinned without the GPU.

- bf16 rounding: known answers (exact values, ties to even, carries);
- the float64 model's gradient equals central finite differences of its
  loss on sampled parameters of every layer of both towers (this pins the
  backward formulas the GPU kernels follow);
- with a vanishing logit scale the loss is log C (uniform softmax over the
  C keys) and the gradient vanishes.
"""
import numpy as np
import pytest

from oracle import lp_oracle as lo


def test_bf16_known_answers():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0, 3.0e38], np.float32)
    r = lo.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa wins (1.0);
    # 1 + 3*2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert r[0] == 1.0
    assert r[1] == 1.0
    assert r[2] == 1.0 + 2 ** -6
    assert r[3] == 1.0
    assert r[4] == -2.5 and r[5] == 0.0
    assert np.isfinite(r[6])


def _small_problem(seed=0, B=6, C=10):
    rng = np.random.default_rng(seed)
    p = np.zeros(2 * lo.TOWER)
    for t in range(2):
        for l in range(4):
            w, b = lo.LAYERS[l]
            base = t * lo.TOWER
            lim = np.sqrt(6.0 / lo.DIN[l])
            p[base + w:base + w + lo.DIN[l] * lo.DOUT[l]] = rng.uniform(-lim, lim, lo.DIN[l] * lo.DOUT[l])
            p[base + b:base + b + lo.DOUT[l]] = rng.normal(0, 0.05, lo.DOUT[l])
    q = rng.normal(size=(B, 240))
    k = rng.normal(size=(C, 240))
    return p, q, k


def test_oracle_gradient_matches_finite_differences():
    p, q, k = _small_problem()
    inv_tau = 3.0
    loss, g = lo.loss_grad(p, q, k, inv_tau, rnd=False)
    rng = np.random.default_rng(1)
    idx = []
    for t in range(2):
        for l in range(4):
            w, b = lo.LAYERS[l]
            base = t * lo.TOWER
            idx += list(base + w + rng.choice(lo.DIN[l] * lo.DOUT[l], 3, replace=False))
            idx += list(base + b + rng.choice(lo.DOUT[l], 1, replace=False))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (lo.loss_grad(pp, q, k, inv_tau, rnd=False)[0] - lo.loss_grad(pm, q, k, inv_tau, rnd=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_uninformative_logits_give_log_c():
    p, q, k = _small_problem(B=4, C=7)
    loss, g = lo.loss_grad(p, q, k, 1e-6, rnd=False)
    assert abs(loss - np.log(7)) < 1e-4
    # and the gradient vanishes with the logit scale
    assert np.max(np.abs(g)) < 1e-3


def test_features_layout():
    canon = np.zeros((80, 2), np.int32)
    canon[0] = (1, 0)
    canon[1] = (0, 1)
    canon[2] = (-1, 0)
    canon[3:] = [(9, i) for i in range(77)]
    offs = np.zeros((1, 80, 2), np.int32)
    vals = np.full((1, 80), np.nan, np.float32)
    offs[0, 0] = (0, 1)   # canonical slot 1
    vals[0, 0] = 5.0
    offs[0, 1] = (3, -2)  # off the stencil: dropped
    vals[0, 1] = 7.0
    f = lo.features_from_neighbours(offs, vals, canon, 1.0, 2.0).reshape(80, 3)
    assert np.allclose(f[1], [0, 1 / 16, 2.0]) and np.all(f[0] == 0) and np.all(f[2:] == 0)
    ti = np.arange(12, dtype=np.float32)  # 4 x 3 grid, dim0 fastest
    kf = lo.key_features(ti, (4, 3), canon, [0], 0.0, 1.0).reshape(80, 3)
    assert np.allclose(kf[0], [1 / 16, 0, 1.0]) and np.allclose(kf[1], [0, 1 / 16, 4.0])
    assert np.all(kf[2] == 0)  # (-1, 0) leaves the grid
