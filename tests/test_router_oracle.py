"""Pins the router oracle (CPU): gradients against finite differences, planted mistakes caught,
bfloat16 rounding against torch and hand-worked ties, closed forms of the loss and Adam."""
import numpy as np
import pytest

import router_oracle as ro


def small_problem(seed=1, dim=7, H=9, nb=6, n=11, k=3):
    rng = np.random.default_rng(seed)
    params = ro.init_params(dim, H, nb, seed)
    lay, _ = ro.layout(dim, H, nb)
    for b in ("b1", "b2", "b3"):
        o, s = lay[b]
        params[o:o + s[0]] = rng.uniform(-0.3, 0.3, size=s[0])
    x = rng.normal(size=(dim, n)).astype(np.float32)
    tgt = rng.integers(0, nb, size=(k, n)).astype(np.int32)
    return params.astype(np.float64), x, tgt, H, nb


def fd_grad(params, x, tgt, H, nb, idx, h=1e-6):
    out = []
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        out.append((ro.loss_only(p1, x, tgt, H, nb) - ro.loss_only(p2, x, tgt, H, nb)) / (2 * h))
    return np.array(out)


def max_fd_error(plant=None, seed=1):
    params, x, tgt, H, nb = small_problem(seed)
    _, g = ro.loss_grad(params, x, tgt, H, nb, rnd=False, plant=plant)
    idx = np.arange(params.size)
    fd = fd_grad(params, x, tgt, H, nb, idx)
    return np.max(np.abs(fd - g[idx])) / (np.max(np.abs(fd)) + 1e-30)


def test_gradient_matches_finite_differences():
    for seed in (1, 2, 3):
        assert max_fd_error(seed=seed) < 1e-6


@pytest.mark.parametrize("plant", ["no_mask", "transpose", "bias"])
def test_planted_gradient_mistakes_are_caught(plant):
    # each mistake must move some gradient entry well beyond the finite-difference agreement
    assert max_fd_error(plant=plant) > 1e-3


def test_bf16_rounding_ties_and_torch():
    torch = pytest.importorskip("torch")
    one = np.float32(1.0)
    cases = {
        1.0: 1.0,
        1.0 + 2.0 ** -8: 1.0,                    # tie, even mantissa kept
        1.0 + 3 * 2.0 ** -8: 1.0 + 2.0 ** -6,     # tie, rounds up to even
        1.0 + 2.0 ** -8 + 2.0 ** -20: 1.0 + 2.0 ** -7,
        -2.5: -2.5,
    }
    for a, b in cases.items():
        assert ro.bf16_round(np.array([a], dtype=np.float32))[0] == np.float32(b), a
    assert np.isnan(ro.bf16_round(np.array([np.nan], dtype=np.float32))[0])
    assert ro.bf16_round(np.array([np.inf], dtype=np.float32))[0] == np.inf
    rng = np.random.default_rng(5)
    v = (rng.normal(size=100000) * np.exp(rng.uniform(-20, 20, size=100000))).astype(np.float32)
    t = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ro.bf16_round(v), t)
    assert one == 1.0


def test_loss_closed_form_at_zero_logits():
    dim, H, nb, n, k = 5, 4, 10, 6, 2
    params = np.zeros(ro.layout(dim, H, nb)[1])
    x = np.ones((dim, n), dtype=np.float32)
    tgt = np.zeros((k, n), dtype=np.int32)
    loss, g = ro.loss_grad(params, x, tgt, H, nb, rnd=False)
    assert abs(loss - np.log(nb)) < 1e-12
    lay, _ = ro.layout(dim, H, nb)
    o, s = lay["b3"]
    expect = np.full(nb, 1.0 / nb)
    expect[0] -= 1.0
    assert np.allclose(g[o:o + nb], expect, atol=1e-12)  # mean over samples of p - t


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    p0 = rng.normal(size=50)
    g = rng.normal(size=50)
    lr, eps = 1e-2, 1e-8
    p1, m, v = ro.adam(p0, g, np.zeros(50), np.zeros(50), 1, lr, eps=eps)
    assert np.allclose(p1, p0 - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-12)
    # second step with the same gradient moves by the same amount again
    p2, _, _ = ro.adam(p1, g, m, v, 2, lr, eps=eps)
    assert np.allclose(p2, p1 - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-12)
