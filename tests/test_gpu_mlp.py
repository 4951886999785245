"""The MFMA GEMM and the MLP engine against float64 numpy references.

The network of the learned acoustic corrector runs its matrix products on
BF16 operands with FP32 accumulation; the references apply the same BF16
rounding to the operands (inputs, weights, hidden activations and their
gradients) and accumulate in float64.
"""
from __future__ import annotations

import numpy as np
import pytest

from dynamical_model_amd.lac import MLP, gemm_bf16, to_bf16

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K", [(130, 70, 50), (256, 256, 64), (33, 300, 2160), (512, 160, 512)])
@pytest.mark.parametrize("a_t,b_t", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_matches_numpy(M, N, K, a_t, b_t):
    rng = np.random.default_rng(M * 7 + N + K)
    a = rng.normal(size=(K, M) if a_t else (M, K)).astype(np.float32)
    b = rng.normal(size=(N, K) if b_t else (K, N)).astype(np.float32)
    print(f"gemm M={M} N={N} K={K} a_t={a_t} b_t={b_t}", flush=True)
    c = gemm_bf16(a, b, a_t, b_t)
    A = to_bf16(a).astype(np.float64)
    B = to_bf16(b).astype(np.float64)
    ref = (A.T if a_t else A) @ (B.T if b_t else B)
    err = np.abs(c - ref).max() / np.abs(ref).max()
    assert err < 1e-5, err


def _ref_grads(x, t, W, b):
    """float64 forward/backward with the engine's BF16 rounding points."""
    r = lambda a: to_bf16(a.astype(np.float32)).astype(np.float64)
    X = r(x)
    Wb = [r(w) for w in W]
    z1 = X @ Wb[0] + b[0]
    h1 = r(np.maximum(z1, 0))
    z2 = h1 @ Wb[1] + b[1]
    h2 = r(np.maximum(z2, 0))
    y = h2 @ Wb[2] + b[2]
    n = y.size
    loss = np.sum((y - t) ** 2) / n
    dy = 2 * (y - t) / n
    g3 = h2.T @ r(dy)
    gb3 = dy.sum(0)
    dh2 = r(dy) @ Wb[2].T
    dh2 = dh2 * (h2 > 0)
    g2 = h1.T @ r(dh2)
    gb2 = dh2.sum(0)
    dh1 = r(dh2) @ Wb[1].T
    dh1 = dh1 * (h1 > 0)
    g1 = X.T @ r(dh1)
    gb1 = dh1.sum(0)
    return loss, y, [(g1, gb1), (g2, gb2), (g3, gb3)]


@pytest.mark.parametrize("nin,nhid,nout,rows", [(108, 64, 16, 200), (2160, 512, 160, 512)],
                         ids=["small", "corrector-kz40"])
def test_mlp_forward_and_gradients_match_reference(nin, nhid, nout, rows):
    """The engine against float64 with the same BF16 rounding points.  What
    remains is FP32 accumulation, and the rare activation whose BF16 rounding
    it flips to the neighbouring value (2^-8 relative); the second case is the
    corrector's network at 512 x 512 x 40 (2160 -> 512 -> 512 -> 160)."""
    rng = np.random.default_rng(3)
    net = MLP(nin, nhid, nout, seed=5)
    W, B = [], []
    for l in range(3):
        w, b = net.get_weights(l)
        b = rng.normal(scale=0.1, size=b.shape).astype(np.float32)
        net.set_weights(l, w, b)
        W.append(w.astype(np.float64))
        B.append(b.astype(np.float64))
    x = rng.normal(size=(rows, nin)).astype(np.float32)
    t = rng.normal(size=(rows, nout)).astype(np.float32)
    loss_ref, y_ref, grads_ref = _ref_grads(x, t, W, B)
    y = net.forward(x)
    ey = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    eym = np.abs(y - y_ref).max() / np.abs(y_ref).max()
    print(f"forward: rel L2 {ey:.2e} max {eym:.2e}")
    assert ey < 1e-3 and eym < 4e-3
    loss = net.grad(x, t)
    print(f"loss rel err {abs(loss - loss_ref) / loss_ref:.2e}")
    assert abs(loss - loss_ref) / loss_ref < 1e-4
    for l in range(3):
        gw, gb = net.gradients(l)
        rw, rb = grads_ref[l]
        ew = np.linalg.norm(gw - rw) / np.linalg.norm(rw)
        eb = np.linalg.norm(gb - rb) / np.linalg.norm(rb)
        print(f"layer {l}: rel err W {ew:.2e} b {eb:.2e}")
        assert ew < 2e-3 and eb < 2e-3, l
    net.close()


def test_gradients_are_deterministic():
    """The gradient products at the corrector's shape split k into slices
    (partial sums added in slice order by a second kernel), so the same batch
    gives the same bits every time; the forward pass, which never splits,
    gives the same bits too.  kz = 60 shape: 3312 -> 512 -> 512 -> 240."""
    rng = np.random.default_rng(11)
    net = MLP(3312, 512, 240, seed=9)
    x = rng.normal(size=(1024, 3312)).astype(np.float32)
    t = rng.normal(size=(1024, 240)).astype(np.float32)
    runs = []
    for _ in range(3):
        loss = net.grad(x, t)
        runs.append((loss, [net.gradients(l) for l in range(3)], net.forward(x)))
    for loss, grads, y in runs[1:]:
        assert loss == runs[0][0]
        assert np.array_equal(y, runs[0][2])
        for (gw, gb), (rw, rb) in zip(grads, runs[0][1]):
            assert np.array_equal(gw, rw) and np.array_equal(gb, rb)
    assert all(np.isfinite(g).all() and np.abs(g).max() > 0 for g, _ in runs[0][1])
    net.close()


def test_adam_steps_match_numpy():
    """Two Adam steps (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected)
    against float64 numpy from the engine's own gradients: the moments, the
    bias corrections and the update.  At the first step the bias-corrected
    update is lr * sign(g); without the corrections it would be 3.16 lr."""
    rng = np.random.default_rng(4)
    nin, nhid, nout, rows = 64, 64, 32, 128
    net = MLP(nin, nhid, nout, seed=8)
    x = rng.normal(size=(rows, nin)).astype(np.float32)
    t = rng.normal(size=(rows, nout)).astype(np.float32)
    lr = 1e-3
    w = [np.concatenate([a.ravel() for a in net.get_weights(l)]).astype(np.float64) for l in range(3)]
    m = [np.zeros_like(a) for a in w]
    v = [np.zeros_like(a) for a in w]
    for step in (1, 2):
        net.grad(x, t)
        g = [np.concatenate([a.ravel() for a in net.gradients(l)]).astype(np.float64) for l in range(3)]
        net.adam(lr)
        for l in range(3):
            m[l] = 0.9 * m[l] + 0.1 * g[l]
            v[l] = 0.999 * v[l] + 0.001 * g[l] ** 2
            upd = lr * (m[l] / (1 - 0.9 ** step)) / (np.sqrt(v[l] / (1 - 0.999 ** step)) + 1e-8)
            got = np.concatenate([a.ravel() for a in net.get_weights(l)]).astype(np.float64)
            dw = w[l] - got
            live = np.abs(g[l]) > 1e-6
            err = np.abs(dw - upd)[live].max() / lr
            print(f"step {step} layer {l}: max |update error| / lr {err:.2e}")
            assert err < 1e-3, (step, l)
            if step == 1:
                big = np.abs(g[l]) > 1e-4  # eps = 1e-8 negligible against |g|
                np.testing.assert_allclose(dw[big], lr * np.sign(g[l][big]), rtol=1e-3)
            w[l] = got
    net.close()


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(9)
    nin, nhid, nout, rows = 96, 128, 24, 512
    A = rng.normal(size=(nin, nout)) / np.sqrt(nin)
    x = rng.normal(size=(rows, nin)).astype(np.float32)
    t = (np.tanh(x @ A)).astype(np.float32)
    net = MLP(nin, nhid, nout, seed=2)
    first = net.grad(x, t)
    for _ in range(300):
        net.grad(x, t)
        net.adam(3e-3)
    last = net.grad(x, t)
    print(f"loss {first:.4f} -> {last:.4f}")
    assert last < 0.1 * first
    net.close()
