"""Pin the ranking network's CPU oracle: backward pass against central finite differences
(FP64), circular-shift equivariance, and the loss's closed form."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cnn_oracle as cn

RNG = np.random.default_rng(8)


def _batch(n: int, B: int = 64):
    e = RNG.normal(size=(n, B))
    e[:, :6] += 4.0  # a pulse, so the pooled maxima are well separated
    v = RNG.uniform(40.0, 60.0, (n, B))
    folds = np.stack([e, v], axis=1).astype(np.float32)
    feats = RNG.normal(size=(n, cn.NFEAT)).astype(np.float32)
    labels = (RNG.random(n) < 0.4).astype(np.float64)
    return folds, feats, labels


def test_parameter_layout():
    assert cn.NPARAMS == 32 * 5 + 32 + 2 * (32 * 32 * 5 + 32) + 32 * 40 + 32 + 32 + 1
    th = cn.init_params(1)
    assert th.shape == (cn.NPARAMS,) and th.dtype == np.float32
    P = cn.unpack(th)
    assert P["W2"].shape == (32, 32, 5) and P["bf2"].shape == (1,)


@pytest.mark.parametrize("B", [32, 64])
def test_backward_matches_finite_differences(B):
    folds, feats, labels = _batch(3, B)
    th = cn.init_params(3).astype(np.float64)
    th += RNG.normal(size=th.shape) * 0.01  # non-zero biases
    g, dx0, _ = cn.backward(th, folds, feats, labels, 1.7, 0.6)

    def f(t):
        return cn.loss(cn.forward(t, folds, feats), labels, 1.7, 0.6)

    o = 0
    for name, shape in cn.SHAPES:
        n = int(np.prod(shape))
        idx = o + RNG.choice(n, size=min(n, 12), replace=False)
        for i in idx:
            h = 1e-6 * max(1.0, abs(th[i]))
            tp, tm = th.copy(), th.copy()
            tp[i] += h
            tm[i] -= h
            fd = (f(tp) - f(tm)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, g[i])
        o += n
    # the input gradient (w.r.t. the standardised profile)
    x0 = cn.standardise(folds)
    for s, p in [(0, 3), (1, 40 % B), (2, B - 1)]:
        xp, xm = x0.copy(), x0.copy()
        xp[s, p] += 1e-6
        xm[s, p] -= 1e-6
        fd = (cn.loss(cn.forward(th, None, feats, x0=xp), labels, 1.7, 0.6)
              - cn.loss(cn.forward(th, None, feats, x0=xm), labels, 1.7, 0.6)) / 2e-6
        assert abs(fd - dx0[s, p]) <= 1e-6 * max(1.0, abs(fd))


def test_circular_shift_equivariance():
    folds, feats, _ = _batch(4)
    th = cn.init_params(5)
    base, c0 = cn.forward(th, folds, feats, cache=True)
    for r in (1, 17, 63):
        rolled = np.roll(folds, r, axis=2)
        lg, c = cn.forward(th, rolled, feats, cache=True)
        np.testing.assert_allclose(c["a3"], np.roll(c0["a3"], r, axis=2), rtol=1e-12,
                                   atol=1e-12)
        np.testing.assert_allclose(lg, base, rtol=1e-12, atol=1e-12)


def test_loss_closed_form():
    logit = np.array([-3.0, 0.0, 2.5, 40.0])
    y = np.array([1.0, 0.0, 1.0, 0.0])
    exp = 2.0 * np.log1p(np.exp(3.0)) + 0.5 * np.log(2.0) + 2.0 * np.log1p(np.exp(-2.5)) \
        + 0.5 * 40.0
    assert abs(cn.loss(logit, y, 2.0, 0.5) - exp) < 1e-9
