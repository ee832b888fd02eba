"""CPU check of the residual form's error bound (knn_scan_resid.hip).

The residual scan rounds e = fl32(x - c) to bfloat16 (nearest even) for both the query and the
reference and forms ||e_q~ - e_r~||^2.  The certificate relies on

    | ||e_q~ - e_r~|| - ||q - r|| |  <=  delta (||q - c|| + ||r - c||),   delta = 2^-8 + 2^-23.

This emulates the two roundings with torch's float32 -> bfloat16 conversion (round to nearest
even) and checks the bound in float64 on random clustered data and on adversarial coordinates
that sit on bfloat16 rounding midpoints, where the rounding error reaches its maximum.  A planted
mistake (delta = 2^-9, half the unit) must fail on the adversarial data.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

DELTA = 2.0 ** -8 + 2.0 ** -23


def bf16(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(torch.bfloat16).float().numpy()


def max_ratio(q, r, c):
    """Largest | ||eq~ - er~|| - ||q - r|| | / (||q - c|| + ||r - c||) over all (q, r) pairs."""
    eq = bf16((q - c).astype(np.float32)).astype(np.float64)
    er = bf16((r - c).astype(np.float32)).astype(np.float64)
    approx = np.linalg.norm(eq[:, None, :] - er[None, :, :], axis=2)
    q64, r64, c64 = q.astype(np.float64), r.astype(np.float64), c.astype(np.float64)
    exact = np.linalg.norm(q64[:, None, :] - r64[None, :, :], axis=2)
    scale = np.linalg.norm(q64 - c64, axis=1)[:, None] + np.linalg.norm(r64 - c64, axis=1)[None, :]
    return float(np.max(np.abs(approx - exact) / scale))


@pytest.mark.parametrize("dim", [7, 96, 256])
def test_bound_holds_on_clustered_data(dim):
    rng = np.random.default_rng(dim)
    c0 = rng.uniform(-1, 1, dim).astype(np.float32)
    q = (c0 + 0.05 * rng.standard_normal((64, dim))).astype(np.float32)
    r = (c0 + 0.05 * rng.standard_normal((256, dim))).astype(np.float32)
    c = r.astype(np.float64).mean(axis=0).astype(np.float32)
    assert max_ratio(q, r, c) <= DELTA


def adversarial(dim, rng):
    # residuals 1 + 2^-8 (a bfloat16 midpoint: rounds to 1, error 2^-8 relative) with signs chosen
    # so that the query's and the reference's errors add up
    c = np.zeros(dim, dtype=np.float32)
    sign = np.where(rng.random(dim) < 0.5, -1.0, 1.0).astype(np.float32)
    q = (sign * np.float32(1 + 2.0 ** -8))[None, :]
    r = (-sign * np.float32(1 + 2.0 ** -8))[None, :]
    return q, r, c


@pytest.mark.parametrize("dim", [8, 128])
def test_bound_is_reached_on_rounding_midpoints(dim):
    q, r, c = adversarial(dim, np.random.default_rng(1))
    ratio = max_ratio(q, r, c)
    assert ratio <= DELTA
    # the planted mistake: half the unit roundoff is not a bound
    assert ratio > 2.0 ** -9
