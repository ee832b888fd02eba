"""TEST INFRASTRUCTURE ONLY -- CPU restatement of the learned block router.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and
only as the checker.  The library never calls it.

The reference (code/knncuda.cu, code/test.cpp) has no router: this module restates the router the
library specifies (knn-cuda_amd/csrc/knn_router.hip) so that its GPU kernels can be checked:

* MLP dim -> H -> H -> nb, ReLU after both hidden layers, weights stored out x in;
* every matrix-product operand rounded to bfloat16 (round to nearest even) where it is read,
  products accumulated exactly here (float64) and in FP32 on the GPU;
* loss: softmax cross-entropy against t_b = (number of the query's k neighbours in block b) / k,
  averaged over the batch;
* Adam with bias correction (Kingma & Ba 2015, Algorithm 1).

How it is pinned (tests/test_router_oracle.py): the analytic gradient (rounding off) against
central finite differences in float64 for every parameter group, with planted mistakes that the
check must catch; bfloat16 rounding against torch's own conversion and hand-worked ties; the loss
against log(nb) at zero logits; Adam's first step against its closed form -lr g / (|g| + eps).
"""
from __future__ import annotations

import numpy as np


def bf16_round(a) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32.  NaN stays NaN."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def layout(dim: int, H: int, nb: int):
    """Offsets of W1 [H][dim], b1 [H], W2 [H][H], b2 [H], W3 [nb][H], b3 [nb] in the vector."""
    shapes = [("W1", (H, dim)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (nb, H)),
              ("b3", (nb,))]
    out, off = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = (off, shp)
        off += n
    return out, off


def unpack(params, dim, H, nb):
    lay, total = layout(dim, H, nb)
    assert params.size == total
    return {k: params[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in lay.items()}


def init_params(dim, H, nb, seed=0):
    """He-uniform weights, zero biases (a convenient start for tests; any vector is valid)."""
    rng = np.random.default_rng(seed)
    p = unpack(np.zeros(layout(dim, H, nb)[1], dtype=np.float32), dim, H, nb)
    for w, fan in (("W1", dim), ("W2", H), ("W3", H)):
        a = np.sqrt(6.0 / fan)
        p[w][...] = rng.uniform(-a, a, size=p[w].shape)
    return np.concatenate([p[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")])


def _r(a, rnd):
    return bf16_round(a).astype(np.float64) if rnd else np.asarray(a, dtype=np.float64)


def forward(params, x, H, nb, rnd=True):
    """x: dim x n (dim-major).  Returns h1, h2 ([n][H]) and z3 ([n][nb]).  With rnd the
    parameters are taken as the FP32 values the GPU holds and h1, h2 come back as FP32 (as
    stored); without it everything stays float64 (for finite differences)."""
    dim = x.shape[0]
    p = unpack(np.asarray(params, dtype=np.float32 if rnd else np.float64), dim, H, nb)
    X = np.asarray(x, dtype=np.float32).T
    h1 = np.maximum(_r(X, rnd) @ _r(p["W1"], rnd).T + p["b1"].astype(np.float64), 0.0)
    h1 = h1.astype(np.float32) if rnd else h1
    h2 = np.maximum(_r(h1, rnd) @ _r(p["W2"], rnd).T + p["b2"].astype(np.float64), 0.0)
    h2 = h2.astype(np.float32) if rnd else h2
    z3 = _r(h2, rnd) @ _r(p["W3"], rnd).T + p["b3"].astype(np.float64)
    return h1, h2, z3


def targets(tgt, nb):
    """k x n block ids -> n x nb target distribution."""
    k, n = tgt.shape
    t = np.zeros((n, nb))
    for j in range(k):
        np.add.at(t, (np.arange(n), tgt[j]), 1.0 / k)
    return t


def loss_grad(params, x, tgt, H, nb, rnd=True, plant=None):
    """Mean loss and its gradient (float64, the parameter vector's layout).

    plant: a deliberate mistake for the tests that check the checks ("no_mask": ReLU mask
    dropped in the backward pass; "transpose": W2 used untransposed backwards; "bias": b2
    gradient summed over the wrong axis's length).
    """
    dim, n = x.shape
    p = unpack(np.asarray(params, dtype=np.float32 if rnd else np.float64), dim, H, nb)
    X = np.asarray(x, dtype=np.float32).T
    h1, h2, z3 = forward(params, x, H, nb, rnd)
    t = targets(np.asarray(tgt), nb)
    zm = z3 - z3.max(axis=1, keepdims=True)
    lse = np.log(np.exp(zm).sum(axis=1, keepdims=True))
    logp = zm - lse
    loss = float(-(t * logp).sum() / n)
    dz3 = (np.exp(logp) - t) / n
    dz3_f = dz3.astype(np.float32) if rnd else dz3
    gW3 = _r(dz3_f, rnd).T @ _r(h2, rnd)
    gb3 = dz3_f.astype(np.float64).sum(axis=0)
    d2 = _r(dz3_f, rnd) @ _r(p["W3"], rnd)
    if plant != "no_mask":
        d2 = d2 * (np.asarray(h2) > 0)
    d2 = d2.astype(np.float32) if rnd else d2
    gW2 = _r(d2, rnd).T @ _r(h1, rnd)
    gb2 = d2.astype(np.float64).sum(axis=0)
    if plant == "bias":
        gb2 = gb2 * (n / max(n - 1, 1))
    W2 = _r(p["W2"], rnd)
    d1 = _r(d2, rnd) @ (W2.T if plant == "transpose" else W2)
    d1 = d1 * (np.asarray(h1) > 0)
    d1 = d1.astype(np.float32) if rnd else d1
    gW1 = _r(d1, rnd).T @ _r(X, rnd)
    gb1 = d1.astype(np.float64).sum(axis=0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def loss_only(params, x, tgt, H, nb, rnd=False):
    n = x.shape[1]
    _, _, z3 = forward(params, x, H, nb, rnd)
    t = targets(np.asarray(tgt), nb)
    zm = z3 - z3.max(axis=1, keepdims=True)
    logp = zm - np.log(np.exp(zm).sum(axis=1, keepdims=True))
    return float(-(t * logp).sum() / n)


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (float64); returns (params, m, v)."""
    g = np.asarray(grad, dtype=np.float64)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return np.asarray(params, dtype=np.float64) - lr * mh / (np.sqrt(vh) + eps), m, v
