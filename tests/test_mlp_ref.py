"""Pins of the surrogate network's CPU restatement (oracle/mlp_ref.py)."""

import numpy as np
import torch

import mlp_ref


def test_bf16_rounding_equals_torch():
    rs = np.random.default_rng(0)
    x = np.concatenate([rs.normal(size=10000) * 10 ** rs.uniform(-6, 6, 10000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0])])
    want = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).double().numpy()
    assert np.array_equal(mlp_ref.bf16(x), want)


def _params(rs, F=16, H=24):
    return {"W1": rs.normal(size=(H, F)) / 4, "b1": rs.normal(size=H) / 10,
            "W2": rs.normal(size=(H, H)) / 5, "b2": rs.normal(size=H) / 10,
            "W3": rs.normal(size=(H, H)) / 5, "b3": rs.normal(size=H) / 10,
            "w4": rs.normal(size=H) / 5, "b4": np.float64(0.3)}


def test_float64_backward_equals_torch_autograd():
    rs = np.random.default_rng(1)
    W = _params(rs)
    X = rs.normal(size=(32, 16))
    t = rs.normal(size=32)
    y, cache = mlp_ref.forward(W, X, emulate=False)
    g = mlp_ref.backward(cache, t, emulate=False)
    tw = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in W.items()}
    h = torch.tensor(X)
    for l in (1, 2, 3):
        h = torch.nn.functional.gelu(h @ tw[f"W{l}"].T + tw[f"b{l}"])
    yt = h @ tw["w4"] + tw["b4"]
    loss = ((yt - torch.tensor(t)) ** 2).mean()
    loss.backward()
    assert np.allclose(y, yt.detach().numpy(), rtol=1e-13, atol=1e-13)
    for k in W:
        assert np.allclose(g[k], tw[k].grad.numpy(), rtol=1e-10, atol=1e-13), k


def test_adam_equals_torch_adam_without_epsilon():
    rs = np.random.default_rng(2)
    p0 = rs.normal(size=50)
    p = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-2, eps=0.0)
    q, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t in range(1, 4):
        g = rs.normal(size=50)
        p.grad = torch.tensor(g)
        opt.step()
        q, m, v = mlp_ref.adam(q, g, m, v, t, 1e-2, eps=0.0)
    assert np.allclose(q, p.detach().numpy(), rtol=1e-13, atol=1e-14)
