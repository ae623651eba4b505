"""The cost model's own HIP kernels: forward, backward and Adam against a
float64 NumPy implementation of the same network (gradients checked), plus a
finite-difference check and a learning check."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def np_forward(params, x):
    acts = [x]
    h = x
    n = len(params) // 2
    for i in range(n):
        w, b = params[2 * i], params[2 * i + 1]
        z = h @ w.T + b
        h = np.maximum(z, 0) if i < n - 1 else z
        acts.append(h)
    return acts


def np_backward(params, acts, dy):
    grads = [None] * len(params)
    n = len(params) // 2
    for i in reversed(range(n)):
        w = params[2 * i]
        y = acts[i + 1]
        dz = dy * (y > 0) if i < n - 1 else dy
        grads[2 * i] = dz.T @ acts[i]
        grads[2 * i + 1] = dz.sum(0)
        dy = dz @ w
    return grads


def test_forward_backward_match_numpy(gpu_lib):
    import torch

    from numbagg_amd.tuning import CostModel

    m = CostModel(seed=3, n_out=2)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((257, 5)).astype(np.float32)
    t = rng.standard_normal((257, 2)).astype(np.float32)
    xt, tt = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    loss, grads = m.loss_and_grads(xt, tt)
    params = [p.cpu().numpy().astype(np.float64) for p in m.params]
    acts = np_forward(params, x.astype(np.float64))
    y = acts[-1]
    want_loss = np.mean((y - t) ** 2)
    assert abs(loss.item() - want_loss) <= 1e-4 * max(1.0, want_loss)
    dy = 2 * (y - t) / y.size
    want = np_backward(params, acts, dy)
    for g, w in zip(grads, want):
        g = g.cpu().numpy()
        np.testing.assert_allclose(g, w, rtol=1e-4, atol=1e-6 * max(1.0, np.abs(w).max()))


def test_finite_differences(gpu_lib):
    import torch

    from numbagg_amd.tuning import CostModel

    m = CostModel(seed=1, n_out=2)
    rng = np.random.default_rng(1)
    x = torch.from_numpy(rng.standard_normal((64, 5)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.standard_normal((64, 2)).astype(np.float32)).cuda()
    _, grads = m.loss_and_grads(x, t)
    params = [p.cpu().numpy().astype(np.float64) for p in m.params]

    def loss_np(ps):
        y = np_forward(ps, x.cpu().numpy().astype(np.float64))[-1]
        return np.mean((y - t.cpu().numpy()) ** 2)

    for li, idx in [(0, (3, 2)), (2, (10, 40)), (4, (1, 5)), (5, (0,))]:
        eps = 1e-4
        p_hi = [p.copy() for p in params]
        p_lo = [p.copy() for p in params]
        p_hi[li][idx] += eps
        p_lo[li][idx] -= eps
        fd = (loss_np(p_hi) - loss_np(p_lo)) / (2 * eps)
        g = grads[li].cpu().numpy()[idx]
        assert abs(g - fd) <= 1e-3 * max(1e-3, abs(fd)), (li, idx, g, fd)


def test_adam_step_matches_numpy(gpu_lib):
    import ctypes

    import torch

    from numbagg_amd import _device as dev
    from numbagg_amd import _lib

    rng = np.random.default_rng(2)
    p = rng.standard_normal(1000).astype(np.float32)
    g = rng.standard_normal(1000).astype(np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    tp, tg, tm, tv = (torch.from_numpy(a.copy()).cuda() for a in (p, g, m, v))
    for step in (1, 2, 3):
        _lib.call("nbg_adam_step", dev.ptr(tp), dev.ptr(tg), dev.ptr(tm), dev.ptr(tv), 1000, ctypes.c_float(1e-2),
                  ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8), step, dev.stream_handle())
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - 1e-2 * (m / (1 - 0.9**step)) / (np.sqrt(v / (1 - 0.999**step)) + 1e-8)
    np.testing.assert_allclose(tp.cpu().numpy(), p, rtol=1e-5, atol=1e-6)


def test_model_learns_a_strategy_boundary(gpu_lib):
    """Synthetic costs where LDS wins for few labels or skewed labels: the model
    recovers the boundary and beats always-global and the label heuristic."""
    from numbagg_amd.tuning import CostModel, features

    rng = np.random.default_rng(5)
    n = 600
    L = 2 ** rng.uniform(7, 20, n)
    K = np.round(2 ** rng.uniform(3, 10, n))
    skew = rng.uniform(0, 0.6, n)
    nan = rng.uniform(0, 0.5, n)
    X = features(L, K, skew, nan, 50)
    t_lds = 0.05 + 0.0004 * K / np.log2(L)
    t_glb = 0.08 + 0.3 * skew
    T = np.log(np.stack([t_lds, t_glb], 1))
    tr, te = slice(0, 450), slice(450, n)
    m = CostModel(seed=0, n_out=2)
    hist = m.fit(X[tr], T[tr], epochs=1500, lr=3e-3)
    assert hist[-1] < 0.05 * hist[0]
    choice = m.choose(X[te], K[te])
    cost = np.where(choice == 0, t_lds[te], t_glb[te]).sum()
    best = np.minimum(t_lds[te], t_glb[te]).sum()
    heur = np.where(K[te] <= 256, t_lds[te], t_glb[te]).sum()
    assert cost / best - 1 < 0.02
    assert cost < heur and cost < t_glb[te].sum()
