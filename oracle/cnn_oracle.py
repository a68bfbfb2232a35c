"""CPU oracle of the learned branch-ranking network -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this.
The network has no counterpart in the reference (it ranks EP survivors for a budgeted,
best-first search); this module is the definition the HIP kernels are checked against,
in FP64:

- input: a survivor's fold [2][B] (e, v) -> profile p = e / sqrt(v), standardised
  x = (p - mean p) / (std p + 1e-6) (population std over the B bins);
- conv1 (1 -> 32), conv2 (32 -> 32), conv3 (32 -> 32): kernel 5, circular padding
  (a profile is periodic in phase), y[c][p] = relu(b[c] + sum_{i,t} W[c][i][t] x[i][p+t-2]);
- h[c] = max_p a3[c][p] (the first maximum's position carries the gradient);
- z = (h, features[8]); u = relu(Wf1 z + bf1) (32 units); logit = Wf2 . u + bf2;
- loss: sum over samples of w_pos y softplus(-logit) + w_neg (1 - y) softplus(logit).

Parameters are one flat vector in the order W1[32][5], b1[32], W2[32][32][5], b2[32],
W3[32][32][5], b3[32], Wf1[32][40], bf1[32], Wf2[32], bf2[1] (``unpack``).

Pinning (tests/test_cnn_oracle.py): the backward pass against central finite differences of
the forward pass (FP64) for every parameter group and for the input; circular-shift
equivariance of the conv stack and invariance of the pooled features; the loss against
its closed form.
"""

from __future__ import annotations

import numpy as np

C = 32
K = 5
NFEAT = 8
NHID = 32
SHAPES = [("W1", (C, 1, K)), ("b1", (C,)), ("W2", (C, C, K)), ("b2", (C,)),
          ("W3", (C, C, K)), ("b3", (C,)), ("Wf1", (NHID, C + NFEAT)), ("bf1", (NHID,)),
          ("Wf2", (NHID,)), ("bf2", (1,))]
NPARAMS = int(sum(np.prod(s) for _, s in SHAPES))


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation (FP32 values)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("b"):
            parts.append(np.zeros(shape))
            continue
        fan_in = int(np.prod(shape[1:])) if len(shape) > 1 else NHID
        parts.append(rng.normal(0, np.sqrt(2.0 / fan_in), shape))
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def standardise(folds: np.ndarray) -> np.ndarray:
    p = folds[:, 0].astype(np.float64) / np.sqrt(folds[:, 1].astype(np.float64))
    mu = p.mean(axis=1, keepdims=True)
    sd = np.sqrt(((p - mu) ** 2).mean(axis=1, keepdims=True))
    return (p - mu) / (sd + 1e-6)


def _conv(x: np.ndarray, W: np.ndarray, b: np.ndarray) -> np.ndarray:
    """x [n][Cin][B], W [Cout][Cin][K] -> pre-activation [n][Cout][B] (circular)."""
    out = np.broadcast_to(b[None, :, None], (x.shape[0], W.shape[0], x.shape[2])).copy()
    for t in range(K):
        xs = np.roll(x, -(t - 2), axis=2)  # xs[..., p] = x[..., p + t - 2]
        out += np.einsum("ci,nip->ncp", W[:, :, t], xs)
    return out


def _conv_back(x: np.ndarray, W: np.ndarray, dy: np.ndarray):
    dW = np.zeros_like(W)
    dx = np.zeros_like(x)
    for t in range(K):
        xs = np.roll(x, -(t - 2), axis=2)
        dW[:, :, t] = np.einsum("ncp,nip->ci", dy, xs)
        dx += np.roll(np.einsum("ci,ncp->nip", W[:, :, t], dy), t - 2, axis=2)
    return dW, dy.sum(axis=(0, 2)), dx


def forward(theta: np.ndarray, folds: np.ndarray | None, feats: np.ndarray, cache: bool = False,
            x0: np.ndarray | None = None):
    """Logits [n]; ``x0`` (the standardised profiles [n][B]) replaces ``folds`` if given."""
    P = unpack(theta.astype(np.float64))
    x0 = (standardise(folds) if x0 is None else np.asarray(x0, dtype=np.float64))[:, None, :]
    z1 = _conv(x0, P["W1"], P["b1"])
    a1 = np.maximum(z1, 0)
    z2 = _conv(a1, P["W2"], P["b2"])
    a2 = np.maximum(z2, 0)
    z3 = _conv(a2, P["W3"], P["b3"])
    a3 = np.maximum(z3, 0)
    arg = np.argmax(a3, axis=2)
    h = np.take_along_axis(a3, arg[:, :, None], axis=2)[:, :, 0]
    zz = np.concatenate([h, feats.astype(np.float64)], axis=1)
    q = zz @ P["Wf1"].T + P["bf1"]
    u = np.maximum(q, 0)
    logit = u @ P["Wf2"] + P["bf2"][0]
    if not cache:
        return logit
    return logit, {"x0": x0, "a1": a1, "a2": a2, "a3": a3, "arg": arg, "zz": zz, "q": q,
                   "u": u, "P": P}


def loss(logit: np.ndarray, labels: np.ndarray, w_pos: float, w_neg: float) -> float:
    sp_neg = np.logaddexp(0, -logit)
    sp_pos = np.logaddexp(0, logit)
    return float(np.sum(w_pos * labels * sp_neg + w_neg * (1 - labels) * sp_pos))


def backward(theta: np.ndarray, folds: np.ndarray, feats: np.ndarray, labels: np.ndarray,
             w_pos: float, w_neg: float):
    """Gradient of ``loss`` w.r.t. the flat parameters (and the standardised input)."""
    logit, c = forward(theta, folds, feats, cache=True)
    P = c["P"]
    sig = 1.0 / (1.0 + np.exp(-logit))
    dlogit = w_pos * labels * (sig - 1.0) + w_neg * (1 - labels) * sig
    g = {"bf2": np.array([dlogit.sum()]), "Wf2": dlogit @ c["u"]}
    du = dlogit[:, None] * P["Wf2"][None, :]
    dq = du * (c["q"] > 0)
    g["Wf1"] = dq.T @ c["zz"]
    g["bf1"] = dq.sum(axis=0)
    dzz = dq @ P["Wf1"]
    dh = dzz[:, :C]
    da3 = np.zeros_like(c["a3"])
    np.put_along_axis(da3, c["arg"][:, :, None], dh[:, :, None], axis=2)
    dz3 = da3 * (c["a3"] > 0)
    g["W3"], g["b3"], da2 = _conv_back(c["a2"], P["W3"], dz3)
    dz2 = da2 * (c["a2"] > 0)
    g["W2"], g["b2"], da1 = _conv_back(c["a1"], P["W2"], dz2)
    dz1 = da1 * (c["a1"] > 0)
    g["W1"], g["b1"], dx0 = _conv_back(c["x0"], P["W1"], dz1)
    flat = np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])
    return flat, dx0[:, 0, :], logit
