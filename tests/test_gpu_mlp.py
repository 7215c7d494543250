"""The GPU-trained MLP (lespmv_mlp_*): gradients against central finite
differences of a float64 CPU forward pass, loss agreement, learning on a
separable problem, determinism, save / load.

The CPU forward below is test infrastructure: the network the library
documents (standardise, W1 ReLU, W2 ReLU, W3, per-head softmax cross-entropy
averaged over labelled samples, summed over heads) written in numpy float64.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

IN, H1, H2, HEADS = 12, 16, 16, [3, 4]


def unpack(p, i=IN, h1=H1, h2=H2, o=sum(HEADS)):
    k = 0
    out = []
    for shape in ((h1, i), (h1,), (h2, h1), (h2,), (o, h2), (o,)):
        n = int(np.prod(shape))
        out.append(p[k:k + n].reshape(shape).astype(np.float64))
        k += n
    assert k == len(p)
    return out


def cpu_loss(p, X, L, mean, std, heads=HEADS):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = (X.astype(np.float64) - mean) / std
    a1 = np.maximum(x @ W1.T + b1, 0)
    a2 = np.maximum(a1 @ W2.T + b2, 0)
    z = a2 @ W3.T + b3
    loss, off = 0.0, 0
    for h, sz in enumerate(heads):
        zz = z[:, off:off + sz]
        lab = L[:, h]
        ok = lab >= 0
        if ok.any():
            m = zz.max(axis=1, keepdims=True)
            lse = (m + np.log(np.exp(zz - m).sum(axis=1, keepdims=True)))[:, 0]
            loss += np.sum((lse - zz[np.arange(len(zz)), np.maximum(lab, 0)])[ok]) / ok.sum()
        off += sz
    return loss


def make_batch(rng, n=40):
    X = rng.normal(size=(n, IN)).astype(np.float32) * 3 + 1
    L = np.stack([rng.integers(0, 3, size=n), rng.integers(-1, 4, size=n)], axis=1).astype(np.int32)
    return X, L


def test_mlp_gradient_matches_finite_differences(gpu):
    lib = gpu
    rng = np.random.default_rng(0)
    mlp = lib.Mlp(IN, H1, H2, HEADS, seed=5)
    X, L = make_batch(rng)
    mean, std = X.mean(axis=0), X.std(axis=0)
    mlp.set_normalizer(mean, std)
    p = mlp.params()
    loss, g = mlp.loss_grad(X, L)
    ref = cpu_loss(p, X, L, mean.astype(np.float64), std.astype(np.float64))
    assert loss == pytest.approx(ref, rel=1e-5)
    # central differences on a sample of coordinates of every layer
    coords = np.concatenate([rng.choice(len(p), size=120, replace=False), [len(p) - 1, 0]])
    h = 1e-4
    bad = []
    for i in coords:
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[i] += h
        pm[i] -= h
        fd = (cpu_loss(pp, X, L, mean, std) - cpu_loss(pm, X, L, mean, std)) / (2 * h)
        if abs(fd - g[i]) > 2e-4 + 2e-3 * abs(fd):
            bad.append((int(i), float(fd), float(g[i])))
    assert len(bad) <= 1, bad  # a ReLU kink inside +-h can spoil one coordinate


def test_mlp_learns_and_is_deterministic(gpu, tmp_path):
    lib = gpu
    rng = np.random.default_rng(1)
    n = 2000
    X = rng.normal(size=(n, IN)).astype(np.float32)
    A = rng.normal(size=(IN, 3))
    B = rng.normal(size=(IN, 4))
    L = np.stack([np.argmax(X @ A, axis=1), np.argmax(X @ B, axis=1)], axis=1).astype(np.int32)
    L[::7, 1] = -1
    runs = []
    for _ in range(2):
        mlp = lib.Mlp(IN, 32, 32, HEADS, seed=3)
        mlp.set_normalizer(X.mean(axis=0), X.std(axis=0))
        hist = mlp.train(X, L, epochs=40, batch=128, lr=3e-3, seed=11)
        runs.append((mlp, hist))
    (m1, h1), (m2, h2) = runs
    assert h1[-1] < 0.3 * h1[0]
    np.testing.assert_array_equal(m1.params(), m2.params())  # fixed reduction order
    probs = m1.forward(X)
    acc0 = np.mean(np.argmax(probs[:, :3], axis=1) == L[:, 0])
    assert acc0 > 0.9
    path = str(tmp_path / "mlp.bin")
    m1.save(path)
    m3 = lib.Mlp(path=path)
    np.testing.assert_array_equal(m3.forward(X), probs)


def test_mlp_head_probabilities_sum_to_one(gpu):
    lib = gpu
    mlp = lib.Mlp(IN, H1, H2, HEADS, seed=2)
    X = np.random.default_rng(4).normal(size=(5, IN)).astype(np.float32)
    P = mlp.forward(X)
    np.testing.assert_allclose(P[:, :3].sum(axis=1), 1.0, rtol=1e-5)
    np.testing.assert_allclose(P[:, 3:].sum(axis=1), 1.0, rtol=1e-5)
