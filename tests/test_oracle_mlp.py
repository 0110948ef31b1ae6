"""Pins the FP64 oracle of the pose proposal network (oracle/mlp.py):

- bfloat16 rounding against torch's own float -> bfloat16 conversion (round to nearest even), incl. ties;
- GELU and its derivative against torch's exact-erf GELU and autograd;
- loss and every parameter gradient against torch autograd on the same FP64 network (an independent
  implementation of the same definition), and against central finite differences;
- Adam against torch.optim.Adam;
- a planted mistake (a GELU derivative without its second term) must fail the gradient comparison.
"""
import numpy as np
import pytest
import torch

from oracle import mlp


def _params(rng, fin, h, o, scale=0.3):
    p = {}
    dims = [(h, fin), (h, h), (h, h), (o, h)]
    for l, (a, b) in enumerate(dims, 1):
        p[f"W{l}"] = rng.normal(0, scale / np.sqrt(b), (a, b))
        p[f"b{l}"] = rng.normal(0, 0.1, a)
    return p


def _problem(seed=0, B=5, fin=12, h=16, K=2, C=6):
    rng = np.random.default_rng(seed)
    p = _params(rng, fin, h, K * C + 2)
    x = rng.normal(0, 1, (B, fin))
    t = rng.random((B, K * C)) ** 4
    t /= t.sum(axis=1, keepdims=True)
    s = rng.normal(0, 1, (B, 2))
    return p, x, t, s, K * C


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.normal(0, 1, 1000) * 10.0 ** rng.integers(-20, 20, 1000),
                        [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0, 65504.0]])
    want = torch.tensor(v, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mlp.to_bf16(v), want)


def test_gelu_matches_torch():
    a = np.linspace(-6, 6, 101)
    ta = torch.tensor(a, requires_grad=True)
    y = torch.nn.functional.gelu(ta)
    y.sum().backward()
    np.testing.assert_allclose(mlp.gelu(a), y.detach().numpy(), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(mlp.gelu_grad(a), ta.grad.numpy(), rtol=1e-12, atol=1e-14)


def _torch_loss(p, x, t, s, ncell, lam):
    tp = {k: torch.tensor(v, requires_grad=True) for k, v in p.items()}
    h = torch.tensor(x)
    for l in (1, 2, 3):
        h = torch.nn.functional.gelu(h @ tp[f"W{l}"].T + tp[f"b{l}"])
    z = h @ tp["W4"].T + tp["b4"]
    ce = torch.nn.functional.cross_entropy(z[:, :ncell], torch.tensor(t), reduction="mean")
    d = z[:, ncell:] - torch.tensor(s)
    loss = ce + 0.5 * lam * (d * d).sum(dim=1).mean()
    loss.backward()
    return float(loss.detach()), {k: v.grad.numpy() for k, v in tp.items()}


def test_gradients_match_autograd_and_finite_differences():
    p, x, t, s, ncell = _problem()
    lam = 0.7
    loss, g, _ = mlp.loss_and_grads(p, x, t, s, ncell, lam)
    tl, tg = _torch_loss(p, x, t, s, ncell, lam)
    assert abs(loss - tl) <= 1e-12 * abs(tl)
    for k in g:
        np.testing.assert_allclose(g[k], tg[k], rtol=1e-10, atol=1e-13, err_msg=k)
    # central differences on a few entries of every parameter
    rng = np.random.default_rng(2)
    for k in g:
        for _ in range(3):
            idx = tuple(int(rng.integers(n)) for n in p[k].shape)
            e = 1e-6
            pp = {kk: vv.copy() for kk, vv in p.items()}
            pp[k][idx] += e
            lp = mlp.loss_and_grads(pp, x, t, s, ncell, lam)[0]
            pp[k][idx] -= 2 * e
            lm = mlp.loss_and_grads(pp, x, t, s, ncell, lam)[0]
            assert abs((lp - lm) / (2 * e) - g[k][idx]) <= 1e-6 * max(1.0, abs(g[k][idx])), (k, idx)


def test_planted_gelu_derivative_mistake_fails():
    p, x, t, s, ncell = _problem(seed=3)
    good = mlp.gelu_grad
    try:
        mlp.gelu_grad = lambda a: 0.5 * (1.0 + mlp.erf(a * mlp.SQRT1_2))
        _, g, _ = mlp.loss_and_grads(p, x, t, s, ncell, 0.7)
    finally:
        mlp.gelu_grad = good
    _, tg = _torch_loss(p, x, t, s, ncell, 0.7)
    assert not np.allclose(g["W1"], tg["W1"], rtol=1e-6, atol=1e-9)


def test_bf16_mode_rounds_product_operands():
    p, x, t, s, ncell = _problem(seed=4)
    z64, _ = mlp.forward(p, x)
    z16, c16 = mlp.forward(p, x, bf16=True)
    assert np.array_equal(c16["h1"], mlp.to_bf16(c16["h1"]))
    rel = np.linalg.norm(z16 - z64) / np.linalg.norm(z64)
    assert 1e-5 < rel < 3e-2  # BF16 operands: a few 2^-8 relative, not zero


def test_adam_matches_torch():
    p, x, t, s, ncell = _problem(seed=5)
    state = {}
    tp = {k: torch.tensor(v, requires_grad=True) for k, v in p.items()}
    opt = torch.optim.Adam(list(tp.values()), lr=1e-2)
    for _ in range(3):
        _, g, _ = mlp.loss_and_grads(p, x, t, s, ncell, 0.7)
        for k, v in tp.items():
            v.grad = torch.tensor(g[k])
        opt.step()
        mlp.adam_step(p, g, state, 1e-2)
    for k, v in tp.items():
        np.testing.assert_allclose(p[k], v.detach().numpy(), rtol=1e-12, atol=1e-14, err_msg=k)
