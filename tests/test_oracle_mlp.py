"""Pins the proposal-network oracle (oracle/mlp_oracle.py) on the CPU:
bf16 rounding against torch's float->bfloat16 conversion, the real-valued
gradient against torch autograd in FP64 and against central finite
differences, Adam against torch.optim.Adam."""
import numpy as np
import pytest
import torch

import mlp_oracle as mo


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-30, 30, 10000),
                        [0.0, -0.0, 1.0, -1.0, 3.0e38, -3.0e38, 1e-40, np.inf, -np.inf],
                        # exact ties between two bf16 values: even mantissa wins
                        np.array([0x3F808000, 0x3F818000, 0xBF808000], np.uint32)
                        .view(np.float32).astype(np.float64)])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    got = mo.bf16_round(x)
    assert np.array_equal(got, ref)
    assert np.array_equal(mo.from_bf16_bits(mo.bf16_bits(x)), ref)


def _batch(seed, B):
    rng = np.random.default_rng(seed)
    P = mo.init_params(seed).astype(np.float64)
    P[mo.OFFSETS["b1"][0]:mo.N_PARAMS] += rng.standard_normal(mo.N_PARAMS - mo.OFFSETS["b1"][0]) * 0.1
    X = rng.standard_normal((B, mo.IN))
    t = rng.standard_normal(B)
    return P, X, t


def test_gradient_matches_torch_autograd():
    P, X, t = _batch(1, 32)
    loss, g, _ = mo.loss_and_grad(P, X, t, emulate=False)
    pt = torch.tensor(P, dtype=torch.float64, requires_grad=True)
    parts = {}
    for name, (o, shape) in mo.OFFSETS.items():
        parts[name] = pt[o:o + int(np.prod(shape))].reshape(shape)
    h = torch.tensor(X)
    for i in (1, 2, 3):
        h = torch.relu(h @ parts[f"W{i}"] + parts[f"b{i}"])
    y = h @ parts["w4"] + parts["b4"]
    lt = ((y - torch.tensor(t)) ** 2).mean()
    lt.backward()
    assert abs(lt.item() - loss) <= 1e-12 * abs(loss)
    ref = pt.grad.numpy()
    assert np.max(np.abs(g - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_gradient_matches_finite_differences():
    P, X, t = _batch(2, 8)
    _, g, _ = mo.loss_and_grad(P, X, t, emulate=False)
    rng = np.random.default_rng(3)
    checked = 0
    for name, (o, shape) in mo.OFFSETS.items():
        size = int(np.prod(shape))
        for i in rng.choice(size, min(size, 12), replace=False):
            k = o + int(i)
            fd = []
            for eps in (1e-5, 5e-6):
                Pp, Pm = P.copy(), P.copy()
                Pp[k] += eps
                Pm[k] -= eps
                fd.append((mo.loss_and_grad(Pp, X, t, False)[0] -
                           mo.loss_and_grad(Pm, X, t, False)[0]) / (2 * eps))
            if abs(fd[0] - fd[1]) > 1e-6 * max(1.0, abs(fd[0])):
                continue  # a ReLU kink inside the stencil
            assert abs(g[k] - fd[1]) <= 1e-6 * max(1e-3, abs(fd[1])), (name, i, g[k], fd)
            checked += 1
    assert checked >= 80


def test_emulated_gradient_close_to_real():
    """The bf16 rounding points move the gradient by a few percent (on this
    random network, rounding X alone moves dW1 by ~4% in norm: the batch sum
    cancels strongly), never more than 10%; the output-layer gradients, one
    rounding away from the loss, stay within 1%."""
    P, X, t = _batch(4, 64)
    _, g0, _ = mo.loss_and_grad(P, X, t, emulate=False)
    _, g1, _ = mo.loss_and_grad(P, X, t, emulate=True)
    for name, (o, shape) in mo.OFFSETS.items():
        a, b = g0[o:o + int(np.prod(shape))], g1[o:o + int(np.prod(shape))]
        tol = 0.01 if name in ("w4", "b4") else 0.1
        assert np.linalg.norm(a - b) <= tol * np.linalg.norm(a) + 1e-12, name


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(1000)
    pt = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for step in range(1, 6):
        g = rng.standard_normal(1000)
        pt.grad = torch.tensor(g)
        opt.step()
        p, m, v = mo.adam(p, g, m, v, step, 1e-3)
    assert np.max(np.abs(p - pt.detach().numpy())) <= 1e-12


@pytest.mark.parametrize("mistake", ["no_relu_mask", "transposed_W"])
def test_planted_mistakes_fail_autograd(mistake):
    """The autograd check catches a dropped ReLU mask or a transposed weight
    in the backward pass (the mistake is planted in a copy of the oracle)."""
    P, X, t = _batch(6, 16)
    Pd = mo.unpack(P)
    y, c = mo.forward(P, X, emulate=False)
    H, W = c["H"], c["W"]
    dy = 2.0 * (y - t) / len(t)
    G = np.where(H[3] > 0, np.outer(dy, Pd["w4"]), 0.0)
    G = G @ W[2].T
    if mistake == "no_relu_mask":
        bad = G
    else:
        bad = np.where(H[3] > 0, np.outer(dy, Pd["w4"]), 0.0) @ W[2]
        bad = np.where(H[2] > 0, bad, 0.0)
    gW2_bad = H[1].T @ bad
    _, g, _ = mo.loss_and_grad(P, X, t, emulate=False)
    o, shape = mo.OFFSETS["W2"]
    good = g[o:o + int(np.prod(shape))].reshape(shape)
    assert np.linalg.norm(gW2_bad - good) > 1e-3 * np.linalg.norm(good)
