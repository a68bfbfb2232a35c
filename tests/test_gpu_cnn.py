"""GPU parity of the ranking network's kernels with the FP64 oracle (forward logits,
backward gradient, Adam step), and that training through them learns a simple task.

Tolerances (FP32 kernels vs the FP64 oracle): logits within 2e-4 of max(1, |logit|);
each parameter group's gradient within 2e-3 of that group's largest |gradient| (samples
whose max-pool argmax is a near-tie are left out: FP32 may pick the other position)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import cnn_oracle as cn
from pyloki_amd import cnn
from pyloki_amd.device import ptr, stream_ptr

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _batch(n: int, B: int, seed: int, pulse: float = 4.0):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=(n, B))
    phase = rng.integers(0, B, n)
    for k in range(n):
        e[k, (phase[k] + np.arange(3)) % B] += pulse
    v = rng.uniform(40.0, 60.0, (n, B))
    folds = np.stack([e, v], axis=1).astype(np.float32)
    feats = rng.normal(size=(n, cn.NFEAT)).astype(np.float32)
    labels = (rng.random(n) < 0.3).astype(np.float32)
    return folds, feats, labels


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_cnn_parameter_count(gpu_device):
    r = cnn.BranchRanker(64)
    assert r.nparams == cn.NPARAMS
    np.testing.assert_array_equal(cnn.init_params(3), cn.init_params(3))


@pytest.mark.parametrize("B", [32, 64, 128])
def test_cnn_forward_matches_oracle(gpu_device, B):
    _case(f"cnn forward B={B}")
    folds, feats, _ = _batch(301, B, B)
    th = cn.init_params(7) + np.random.default_rng(1).normal(0, 0.02, cn.NPARAMS).astype(
        np.float32)
    r = cnn.BranchRanker(B, params=th)
    got = r.logits(_dev(folds), _dev(feats)).cpu().numpy()
    exp = cn.forward(th, folds, feats)
    err = np.abs(got - exp) / np.maximum(1.0, np.abs(exp))
    print(f"[result] logits max error {err.max():.2e}", flush=True)
    assert err.max() <= 2e-4


def test_cnn_forward_complex_folds(gpu_device):
    _case("cnn forward, complex folds")
    B = 64
    folds, feats, _ = _batch(130, B, 11)
    cf = np.fft.rfft(folds.astype(np.float64), axis=-1).astype(np.complex64)
    th = cn.init_params(2)
    r = cnn.BranchRanker(B, params=th)
    got = r.logits(_dev(cf), _dev(feats)).cpu().numpy()
    back = np.fft.irfft(cf.astype(np.complex128), n=B, axis=-1)
    exp = cn.forward(th, back, feats)
    err = np.abs(got - exp) / np.maximum(1.0, np.abs(exp))
    print(f"[result] complex logits max error {err.max():.2e}", flush=True)
    assert err.max() <= 2e-4


@pytest.mark.parametrize("B", [32, 64])
def test_cnn_backward_matches_oracle(gpu_device, B):
    _case(f"cnn backward B={B}")
    folds, feats, labels = _batch(90, B, 100 + B)
    th = cn.init_params(4) + np.random.default_rng(2).normal(0, 0.02, cn.NPARAMS).astype(
        np.float32)
    # keep the samples whose pooled maxima are unambiguous in FP32
    _, c = cn.forward(th, folds, feats, cache=True)
    a3 = np.sort(c["a3"], axis=2)
    gap = (a3[:, :, -1] - a3[:, :, -2]) / np.maximum(1.0, a3[:, :, -1])
    ok = np.all((gap > 1e-4) | (a3[:, :, -1] == 0), axis=1)
    folds, feats, labels = folds[ok], feats[ok], labels[ok]
    assert len(folds) > 60
    lib = cnn._lib.load()
    r = cnn.BranchRanker(B, params=th)
    n = len(folds)
    stride = lib.pyl_cnn_cache_floats(B)
    cache = torch.empty(n * stride, dtype=torch.float32, device="cuda")
    lg = r.logits(_dev(folds), _dev(feats), cache)
    y = _dev(labels)
    grad = torch.empty(cn.NPARAMS, dtype=torch.float32, device="cuda")
    work = torch.empty(cn.NPARAMS * 7, dtype=torch.float32, device="cuda")  # 7 blocks
    cnn._lib.check(lib.pyl_cnn_backward(ptr(r.params), ptr(cache), ptr(lg), ptr(y), n, B,
                                        1.5, 0.7, ptr(work), int(work.numel()), ptr(grad),
                                        stream_ptr()))
    g_ora, _, _ = cn.backward(th, folds, feats, labels.astype(np.float64), 1.5, 0.7)
    g = grad.cpu().numpy()
    o = 0
    worst = 0.0
    for name, shape in cn.SHAPES:
        k = int(np.prod(shape))
        ref = g_ora[o:o + k]
        scale = max(np.abs(ref).max(), 1e-12)
        e = np.abs(g[o:o + k] - ref).max() / scale
        worst = max(worst, e)
        assert e <= 2e-3, (name, e)
        o += k
    print(f"[result] gradient worst group error {worst:.2e}", flush=True)
    # the gradient does not depend on how many blocks share the work
    grad1 = torch.empty_like(grad)
    cnn._lib.check(lib.pyl_cnn_backward(ptr(r.params), ptr(cache), ptr(lg), ptr(y), n, B,
                                        1.5, 0.7, ptr(work), cn.NPARAMS, ptr(grad1),
                                        stream_ptr()))
    np.testing.assert_allclose(grad1.cpu().numpy(), g, rtol=1e-4, atol=1e-6 * np.abs(g).max())


def test_cnn_adam_step(gpu_device):
    _case("cnn adam")
    rng = np.random.default_rng(3)
    th = rng.normal(size=cn.NPARAMS).astype(np.float32)
    r = cnn.BranchRanker(64, params=th)
    g = rng.normal(size=cn.NPARAMS).astype(np.float32)
    m = np.zeros(cn.NPARAMS)
    v = np.zeros(cn.NPARAMS)
    p = th.astype(np.float64)
    lib = cnn._lib.load()
    gd = _dev(g)
    for step in (1, 2, 3):
        cnn._lib.check(lib.pyl_cnn_adam(ptr(r.params), ptr(gd), ptr(r.m), ptr(r.v),
                                        cn.NPARAMS, 1e-2, 0.9, 0.999, 1e-8, step, stream_ptr()))
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g.astype(np.float64) ** 2
        p = p - 1e-2 * (m / (1 - 0.9**step)) / (np.sqrt(v / (1 - 0.999**step)) + 1e-8)
    np.testing.assert_allclose(r.params.cpu().numpy(), p, rtol=1e-5, atol=1e-5)


def test_cnn_training_learns_pulses(gpu_device):
    """Pulse (label 1, matched-filter S/N 6) against noise (label 0) profiles; the
    features carry nothing."""
    _case("cnn training")
    B = 64
    rng = np.random.default_rng(9)
    n = 4096
    e = rng.normal(size=(n, B))
    y = (rng.random(n) < 0.5).astype(np.float32)
    ph = rng.integers(0, B, n)
    for k in np.where(y > 0)[0]:
        e[k, (ph[k] + np.arange(4)) % B] += 3.0
    folds = _dev(np.stack([e, np.ones((n, B))], axis=1).astype(np.float32))
    feats = torch.zeros((n, 8), dtype=torch.float32, device="cuda")
    yd = _dev(y)
    r = cnn.BranchRanker(B, seed=1)
    first = r.train_step(folds, feats, yd, lr=2e-3)
    for _ in range(200):
        last = r.train_step(folds, feats, yd, lr=2e-3)
    acc = float(((r.logits(folds, feats) > 0).float() == yd).float().mean().item())
    print(f"[result] training loss {first:.3f} -> {last:.3f}, accuracy {acc:.3f}", flush=True)
    assert last < 0.3 * first and acc > 0.97


def test_budgeted_search_pipeline(gpu_device):
    """Exhaustive vs budgeted circular EP on a smaller orbit workload: the budgeted
    search's candidates are exhaustive candidates (no overflow: the same thresholds), it
    scores fewer leaves, and the learned budget keeps most of the exhaustive candidates."""
    import bench

    _case("budgeted search pipeline")
    w = dict(bench.EP_CNN, nsamps=2**22, nseg=64, max_sugg=2**21)
    thr = np.linspace(0.5, 8.0, w["nseg"] - 1)
    res = cnn.budget_experiment(lambda s: bench.ep_circular_dyp(s, w), (1, 2), 9, thr,
                                w["max_sugg"], (2**13,), epochs=6)
    ex = res["exhaustive"]
    c, s = res["cnn_8192"], res["snr_8192"]
    print(f"[result] exhaustive {ex}; learned {c}; S/N {s}; train {res['train']}", flush=True)
    assert ex["n_final"] > 0
    assert c["precision"] == 1.0 and s["precision"] == 1.0
    assert c["trials"] < ex["trials"] and s["trials"] < ex["trials"]
    assert c["recall"] >= 0.8


def test_network_gain_over_the_snr_budget(gpu_device):
    """What the trained network adds, on held-out series of the benchmark's circular-orbit
    configuration (2^24 samples, 128 segments): trained on the exhaustive runs of seeds
    1-4, then 8 other series are searched as one forest under the same survivor budget
    keyed by S/N alone and by the hybrid key (the better of the network's and the S/N's
    rank). The hybrid budget recovers at least as many of the exhaustive candidates while
    scoring fewer leaves."""
    import bench

    _case("network gain over the S/N budget (held-out series)")
    w = bench.EP_CNN
    thr = np.linspace(w["thresholds"][0], w["thresholds"][1], w["nseg"] - 1)
    ranker, info = cnn.train_ranker(bench.ep_circular_dyp, (1, 2, 3, 4), thr,
                                    w["max_sugg"], epochs=w["epochs"])
    series = [bench.ep_circular_dyp(s) for s in range(9, 17)]
    b = w["survey_budget"]
    res = cnn.forest_throughput(series, thr, w["max_sugg"], {
        "exhaustive": None,
        "snr": lambda: cnn.SurvivorBudget(b),
        "hybrid": lambda: cnn.SurvivorBudget(b, ranker, hybrid=True)})
    s, h = res["snr"], res["hybrid"]
    print(f"[result] budget {b}: S/N recall {s['recall']:.4f} leaves {s['trials']}; "
          f"hybrid recall {h['recall']:.4f} leaves {h['trials']}; train {info}", flush=True)
    assert h["recall"] >= s["recall"]
    assert h["trials"] <= s["trials"]


@pytest.mark.parametrize("B,cplx", [(32, False), (64, False), (128, False), (64, True)])
def test_cnn_forward_bf16_ranks_like_fp32(gpu_device, B, cplx):
    """The MFMA (BF16 operands) inference pass against the FP64 oracle: logits within
    0.05 of max(1, |logit|) (bf16 weights and activations), and the ranking of the
    samples nearly unchanged (Spearman correlation with the FP32 logits >= 0.995)."""
    _case(f"cnn forward bf16 B={B} complex={cplx}")
    folds, feats, _ = _batch(777, B, 31 + B)
    th = cn.init_params(5) + np.random.default_rng(8).normal(0, 0.02, cn.NPARAMS).astype(
        np.float32)
    r = cnn.BranchRanker(B, params=th)
    fd = folds
    if cplx:
        fd = np.fft.rfft(folds.astype(np.float64), axis=-1).astype(np.complex64)
        folds = np.fft.irfft(fd.astype(np.complex128), n=B, axis=-1)
    got = r.logits(_dev(fd), _dev(feats), precision="bf16").cpu().numpy()
    f32 = r.logits(_dev(fd), _dev(feats)).cpu().numpy()
    exp = cn.forward(th, folds, feats)
    err = np.abs(got - exp) / np.maximum(1.0, np.abs(exp))
    rank = lambda x: np.argsort(np.argsort(x))  # noqa: E731
    rho = np.corrcoef(rank(got), rank(f32))[0, 1]
    print(f"[result] bf16 logits max error {err.max():.2e}, spearman {rho:.5f}", flush=True)
    assert err.max() <= 0.05
    assert rho >= 0.995
