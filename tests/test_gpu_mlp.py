"""Closure-network training kernels on the GPU against the FP64 numpy model
(tests/mlp_model.py) that keeps the kernels' BF16 rounding points.

Tolerances: FP32 arithmetic and accumulation against FP64 (~1e-6 relative),
plus rare BF16 rounding flips where the FP32 and FP64 hidden activations
straddle a rounding midpoint (each moves y by ~1e-4)."""
import numpy as np
import pytest

import mlp_model as mm
from structured_amd import lib

pytestmark = pytest.mark.gpu


def data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, (n, 5)).astype(np.float32)
    t = (0.3 * np.tanh(1.5 * x[:, 0] - x[:, 2] + 0.5 * x[:, 3] * x[:, 4])).astype(np.float32)
    return x, t


def test_forward_matches_model():
    net = lib.Mlp(max_rows=4096, seed=11)
    p = net.params
    x, _ = data(10000, 1)  # more rows than max_rows: chunked
    y = net.forward(x)
    ym, _ = mm.forward(p.astype(np.float64), x, bf=True)
    err = np.abs(y - ym)
    print("forward: median %.2e max %.2e" % (np.median(err), err.max()))
    assert np.median(err) <= 2e-6
    assert err.max() <= 1e-3
    # the BF16 rounding points matter at this level: the unrounded network differs more
    y64, _ = mm.forward(p.astype(np.float64), x, bf=False)
    assert np.median(np.abs(y64 - ym)) > 5 * np.median(err)


def test_initial_parameters_glorot():
    net = lib.Mlp(max_rows=256, seed=5)
    w = mm.unpack(net.params)
    for k, fan in (("W1", 69), ("W2", 128), ("W3", 65)):
        lim = np.sqrt(6 / fan)
        assert np.abs(w[k]).max() <= lim and np.abs(w[k]).max() > 0.9 * lim
        assert abs(w[k].mean()) < 0.2 * lim
    for k in ("b1", "b2"):
        assert np.all(w[k] == 0)
    assert np.array_equal(net.params, lib.Mlp(max_rows=256, seed=5).params)
    assert not np.array_equal(net.params, lib.Mlp(max_rows=256, seed=6).params)


@pytest.mark.parametrize("n", [4096, 3001])
def test_loss_and_gradient_match_model(n):
    net = lib.Mlp(max_rows=4096, seed=3)
    rng = np.random.default_rng(9)
    p = (mm.glorot(seed=2) + 0.02 * rng.standard_normal(mm.NPARAM)).astype(np.float32)
    net.params = p
    x, t = data(n, 2)
    loss, g = net.loss_grad(x, t)
    lm, gm = mm.loss_grad(p.astype(np.float64), x, t, bf=True)
    assert abs(loss - lm) <= 1e-4 * lm
    for k, (o, shp) in mm.OFF.items():
        cnt = int(np.prod(shp)) if shp else 1
        a, b = g[o:o + cnt], gm[o:o + cnt]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print("grad %s rel %.2e" % (k, rel))
        assert rel <= 2e-3, k
    # BF16 operands change the gradient measurably (the model is not trivially FP64)
    _, g64 = mm.loss_grad(p.astype(np.float64), x, t, bf=False)
    o = mm.OFF["W2"][0]
    assert np.linalg.norm(g[o:o + 4096] - g64[o:o + 4096]) > np.linalg.norm(g[o:o + 4096] - gm[o:o + 4096])


def test_adam_step_matches_model():
    net = lib.Mlp(max_rows=2048, seed=8)
    p0 = net.params.astype(np.float64)
    x, t = data(2048, 3)
    _, g = net.loss_grad(x, t)
    net.train(x, t, epochs=1, batch=2048, lr=1e-3, seed=1)  # one full-batch step
    p1 = net.params
    pm, _, _ = mm.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g.astype(np.float64), 1, 1e-3)
    np.testing.assert_allclose(p1, pm, rtol=0, atol=2e-6)


def test_training_reduces_loss_and_is_deterministic():
    x, t = data(65536, 4)
    runs = []
    for _ in range(2):
        net = lib.Mlp(max_rows=8192, seed=1)
        hist = net.train(x, t, epochs=30, batch=8192, lr=3e-3, seed=7)
        runs.append((hist, net.params))
    hist, p = runs[0]
    print("loss per epoch", hist[[0, 5, 10, 20, 29]])
    assert hist[-1] < 0.05 * hist[0]
    assert np.array_equal(p, runs[1][1])
    assert np.array_equal(hist, runs[0][0])
    # the trained GPU network predicts held-out samples
    xv, tv = data(4096, 99)
    yv = lib.Mlp(max_rows=4096, seed=0)
    yv.params = p
    err = np.mean((yv.forward(xv) - tv) ** 2)
    assert err < 0.1 * np.mean(tv ** 2)


def test_rejects_oversized_batch():
    net = lib.Mlp(max_rows=128, seed=0)
    x, t = data(512, 5)
    with pytest.raises(lib.SgError):
        net.loss_grad(x, t)
    with pytest.raises(lib.SgError):
        net.train(x, t, epochs=1, batch=256)
