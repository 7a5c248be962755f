"""Surrogate network on the GPU: forward pass and gradients of the BF16 MFMA kernels against the
CPU oracle (oracle/mlp_oracle.py), and training."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def _net(max_batch=4096, seed=11, physics=False):
    from crystalscatter_amd.surrogate import MLP, physics_features
    net = MLP(device=0, max_batch=max_batch, seed=seed)
    if physics:
        net.set_features(*physics_features())
    return net


def _inputs(n, nsets=5, seed=0):
    import torch
    rng = np.random.default_rng(seed)
    par = rng.uniform(-1, 1, (nsets, 6)).astype(np.float32)
    sets = rng.integers(0, nsets, n).astype(np.int32)
    q = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    dev = lambda a: torch.from_numpy(a).cuda()
    return (par, sets, q, t), (dev(par), dev(sets), dev(q), dev(t))


@pytest.mark.parametrize("n,physics", [(1000, False), (4096, False), (1000, True)])
def test_forward_matches_oracle(n, physics):
    net = _net(physics=physics)
    (par, sets, q, _), (dpar, dsets, dq, _) = _inputs(n)
    y = net.forward(dpar, dsets, dq).cpu().numpy().astype(np.float64)
    p = net.get_params()
    X = mo.features(par, sets, q, *net.get_features())
    yb, _ = mo.forward(p.astype(np.float64), X, bf16=True)
    y64, _ = mo.forward(p.astype(np.float64), X, bf16=False)
    scale = np.abs(y64).max()
    # bf16-emulating oracle: differs by FP32 accumulation order and the fast sin/cos/exp
    assert np.abs(y - yb).max() <= 2e-2 * scale
    assert np.sqrt(np.mean((y - yb) ** 2)) <= 3e-3 * scale
    # float64 network: the bf16 storage of weights and activations (8-bit mantissa)
    assert np.sqrt(np.mean((y - y64) ** 2)) <= 2e-2 * scale


@pytest.mark.parametrize("n,physics", [(1, False), (1000, True), (4096, False), (70001, True)])
def test_fused_inference_equals_forward(n, physics):
    """csx_mlp_predict (one fused kernel, activations on chip) gives forward's y bit for bit,
    also beyond max_batch rows (forward then runs in chunks) and after a training step."""
    import torch
    net = _net(physics=physics)
    (_, _, _, _), (dpar, dsets, dq, dt) = _inputs(n, seed=3)
    for step in range(2):
        y = torch.empty(n, dtype=torch.float32, device="cuda")
        for a in range(0, n, net.max_batch):
            b = min(n, a + net.max_batch)
            net.forward(dpar, dsets[a:b].contiguous(), dq[a:b].contiguous(), out=y[a:b])
        yp = net.predict(dpar, dsets, dq)
        assert torch.equal(y, yp), (n, step, float((y - yp).abs().max()))
        m = min(n, net.max_batch)
        net.train_step(dpar, dsets[:m].contiguous(), dq[:m].contiguous(), dt[:m].contiguous(), 1e-3)


def test_gradient_matches_oracle():
    n = 1500
    net = _net(physics=True)
    (par, sets, q, t), (dpar, dsets, dq, dt) = _inputs(n, seed=4)
    loss = net.gradient(dpar, dsets, dq, dt)
    g = net.get_grads().astype(np.float64)
    p = net.get_params()
    X = mo.features(par, sets, q, *net.get_features())
    lb, gb = mo.loss_and_grad(p.astype(np.float64), X, t, bf16=True)
    l64, g64 = mo.loss_and_grad(p.astype(np.float64), X, t, bf16=False)
    assert abs(loss - lb) <= 1e-3 * lb
    lay, (wo, bo), _ = mo.offsets()
    blocks = [(w, w + mo.HID * mo.layer_in(l)) for l, (w, b) in enumerate(lay)] + \
             [(b, b + mo.HID) for (w, b) in lay] + [(wo, wo + mo.HID), (bo, bo + 1)]
    for a, b in blocks:
        rel = np.linalg.norm(g[a:b] - gb[a:b]) / np.linalg.norm(gb[a:b])
        assert rel <= 2e-2, (a, b, rel)
        rel64 = np.linalg.norm(g[a:b] - g64[a:b]) / np.linalg.norm(g64[a:b])
        assert rel64 <= 6e-2, (a, b, rel64)
    cos = g @ gb / (np.linalg.norm(g) * np.linalg.norm(gb))
    assert cos >= 0.9995
    # padding inputs: zero gradient
    W0 = g[lay[0][0]:lay[0][0] + mo.HID * mo.IN_PAD].reshape(mo.HID, mo.IN_PAD)
    assert np.all(W0[:, mo.IN:] == 0)


def test_adam_step_follows_the_update_rule():
    n = 512
    net = _net(seed=5)
    (_, _, _, _), (dpar, dsets, dq, dt) = _inputs(n, seed=5)
    p0 = net.get_params()
    net.gradient(dpar, dsets, dq, dt)
    g = net.get_grads()
    net._chk(net._L.csx_mlp_adam(net._h, 1e-3, 0.9, 0.999, 1e-8), "adam")
    p1 = net.get_params()
    # first Adam step: m^ = g, v^ = g^2 -> p1 = p0 - lr * g / (|g| + eps)
    exp = p0 - 1e-3 * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1, exp, rtol=1e-5, atol=1e-7)


def test_training_lowers_the_loss():
    import torch
    n = 4096
    net = _net(max_batch=n, seed=9)
    rng = np.random.default_rng(2)
    par = rng.uniform(-1, 1, (16, 6)).astype(np.float32)
    sets = rng.integers(0, 16, n).astype(np.int32)
    q = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    # a smooth target of parameters and q
    t = (np.sin(2 * q[:, 0] + par[sets, 0]) * np.cos(3 * q[:, 1] - par[sets, 1]) + 0.5 * par[sets, 2]).astype(np.float32)
    dev = lambda a: torch.from_numpy(a).cuda()
    dpar, dsets, dq, dt = dev(par), dev(sets), dev(q), dev(t)
    first = net.train_step(dpar, dsets, dq, dt, 1e-3, want_loss=True)
    for _ in range(300):
        net.train_step(dpar, dsets, dq, dt, 1e-3)
    last = net.gradient(dpar, dsets, dq, dt)
    assert last < 0.05 * first, (first, last)


def test_training_is_deterministic_with_fresh_batches():
    """Two runs of the same steps on freshly gathered batches give identical weights (the
    batches are torch tensors recycled by the caching allocator between steps)."""
    import torch
    from crystalscatter_amd import surrogate as sg
    rng = np.random.default_rng(8)
    n = 20000
    par = torch.from_numpy(rng.uniform(-1, 1, (64, 6)).astype(np.float32)).cuda()
    data = sg.TrainingSet(par, torch.from_numpy(rng.integers(0, 64, n).astype(np.int32)).cuda(),
                          torch.from_numpy(rng.uniform(-1, 1, (n, 2)).astype(np.float32)).cuda(),
                          torch.from_numpy(rng.normal(0, 1, n).astype(np.float32)).cuda(),
                          None, 0.0, 1.0, 64, n // 64)
    out = []
    for _ in range(2):
        net = _net(max_batch=8192, seed=4, physics=True)
        sg.train(net, data, 40, batch=8192, lr=1e-3)
        net.synchronize()
        out.append(net.get_params())
    assert np.array_equal(out[0], out[1])
