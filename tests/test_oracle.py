"""CPU tests of the oracle: pinned against the reference's own outputs and against closed forms.

The oracle restates code/test.cpp:129-175 (knn_c).  It is pinned by
  * golden vectors written by the reference's knn_c compiled from its sources (tests/golden);
  * the compiled reference itself (oracle/_ref) on random inputs, when built;
  * mathematics: FP64 brute force on well-separated data, points on a line, duplicates,
    k = 1 and k = ref_nb;
and each check is shown to catch a planted mistake (wrong tie-break, FMA contraction, transposed
layout, FP64 accumulation) applied to a numpy copy of the oracle's output path.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLDEN, "golden_index.json")) as f:
    INDEX = json.load(f)

ARRAY_CASES = sorted(n for n, v in INDEX.items() if v["kind"] == "arrays")
GEN_CASES = sorted(n for n, v in INDEX.items() if v["kind"] == "generated")


def load_case(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    meta = INDEX[name]
    if meta["kind"] == "arrays":
        ref, qry = z["ref"], z["query"]
    else:
        if "mixture" in meta:
            ncomp, sigma, cseed = meta["mixture"]
            ref = oracle.gen_mixture(meta["R"], meta["D"], ncomp, sigma, cseed, meta["seed_ref"])
            qry = oracle.gen_mixture(meta["Q"], meta["D"], ncomp, sigma, cseed, meta["seed_qry"])
        else:
            ref = oracle.gen_uniform(meta["R"], meta["D"], meta["seed_ref"], meta["scale"])
            qry = oracle.gen_uniform(meta["Q"], meta["D"], meta["seed_qry"], meta["scale"])
        assert np.array_equal(ref[:, :8], z["ref_head"]), "generator no longer reproduces the fixture"
        assert np.array_equal(qry[:, :8], z["query_head"]), "generator no longer reproduces the fixture"
    return ref, qry, int(z["k"]), z["dist"], z["index"]


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("name", ARRAY_CASES + GEN_CASES)
def test_oracle_matches_reference_golden(name):
    ref, qry, k, gd, gi = load_case(name)
    d, i = oracle.knn_c(ref, qry, k)
    assert np.array_equal(i, gi)
    assert np.array_equal(bits(d), bits(gd))


def test_single_thread_equals_threaded():
    ref, qry, k, gd, gi = load_case("uniform10_small")
    d1, i1 = oracle.knn_c(ref, qry, k, threads=False)
    d2, i2 = oracle.knn_c(ref, qry, k, threads=True)
    assert np.array_equal(i1, i2) and np.array_equal(bits(d1), bits(d2))


def test_query_subsample_equals_full():
    ref, qry, k, gd, gi = load_case("uniform10_small")
    d, i = oracle.knn_c_queries(ref, qry, k, 5, 17)
    assert np.array_equal(i[:, 5:17], gi[:, 5:17])
    assert np.array_equal(bits(d[:, 5:17]), bits(gd[:, 5:17]))


@pytest.mark.skipif(oracle.reference_lib() is None, reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_compiled_reference(seed):
    rng = np.random.default_rng(100 + seed)
    R, Q, D = [(700, 40, 33), (257, 19, 128), (90, 7, 2), (1200, 12, 96), (64, 64, 5), (333, 11, 1)][seed]
    k = [9, 16, 90, 1, 64, 40][seed]
    ref = (10 * rng.random((D, R))).astype(np.float32)
    qry = (10 * rng.random((D, Q))).astype(np.float32)
    if D <= 2:
        ref, qry = np.round(ref), np.round(qry)  # many exact ties
    d1, i1 = oracle.reference_knn_c(ref, qry, k)
    d2, i2 = oracle.knn_c(ref, qry, k)
    assert np.array_equal(i1, i2) and np.array_equal(bits(d1), bits(d2))


# --------------------------------------------------------------------------------------------
# Closed forms
# --------------------------------------------------------------------------------------------
def test_fp64_agrees_on_well_separated_data():
    # Integer lattice points with distinct squared distances: FP32 and FP64 orders coincide and
    # the FP32 key is sqrtf of an exactly representable integer.
    rng = np.random.default_rng(7)
    D, R, Q, k = 4, 300, 20, 10
    ref = rng.integers(0, 40, size=(D, R)).astype(np.float32)
    qry = rng.integers(0, 40, size=(D, Q)).astype(np.float32) + 0.25
    d32, i32 = oracle.knn_c(ref, qry, k)
    d64, i64 = oracle.knn_f64(ref, qry, k)
    # every FP32 key is the correctly rounded true distance (sums of squares of x.25 / x.75 are exact)
    assert np.array_equal(d32, d64.astype(np.float32))
    assert np.array_equal(i32, i64)


def test_line_known_order():
    # Points on a line at integer positions: order by |x - p|, ties to the lower index.
    rng = np.random.default_rng(3)
    pos = rng.integers(0, 30, size=200).astype(np.float32)
    qs = np.array([0.0, 4.0, 15.0, 29.0, 7.5], dtype=np.float32)
    k = 25
    d, i = oracle.knn_c(pos[None, :], qs[None, :], k)
    for c, x in enumerate(qs):
        keys = np.abs(pos - x)
        order = sorted(range(len(pos)), key=lambda j: (keys[j], j))[:k]
        assert list(i[:, c]) == order
        assert np.array_equal(d[:, c], keys[order])


def test_duplicates_lower_index_first():
    base = np.array([[0.0, 1.0, 2.0]], dtype=np.float32)
    ref = np.concatenate([base, base, base], axis=1)  # indexes j, j+3, j+6 coincide
    d, i = oracle.knn_c(ref, np.array([[1.0]], dtype=np.float32), 9)
    assert list(i[:, 0]) == [1, 4, 7, 0, 2, 3, 5, 6, 8]
    assert list(d[:, 0]) == [0, 0, 0, 1, 1, 1, 1, 1, 1]


def test_k1_is_argmin_and_k_equals_n_is_full_sort():
    rng = np.random.default_rng(5)
    ref = rng.random((6, 120), dtype=np.float32)
    qry = rng.random((6, 9), dtype=np.float32)
    full_d, full_i = oracle.knn_c(ref, qry, 120)
    one_d, one_i = oracle.knn_c(ref, qry, 1)
    assert np.array_equal(one_i[0], full_i[0]) and np.array_equal(one_d[0], full_d[0])
    for c in range(9):
        assert sorted(full_i[:, c]) == list(range(120))  # a permutation
        assert np.all(np.diff(full_d[:, c]) >= 0)  # non-decreasing
        keys = np.array([np.sqrt(oracle.canonical_sq(ref, qry, j, c)) for j in range(120)], dtype=np.float32)
        assert np.array_equal(full_d[:, c], keys[full_i[:, c]])


def test_reference_check_semantics():
    ref, qry, k, gd, gi = load_case("uniform10_small")
    assert oracle.check(gd, gi, gd, gi) == (1.0, 1.0)
    bad = gd.copy()
    bad.ravel()[:10] += 2e-3
    fd, fi = oracle.check(bad, gi, gd, gi)
    assert fd == pytest.approx(1 - 10 / gd.size) and fi == 1.0


def test_uniform_generator_range_and_determinism():
    a = oracle.gen_uniform(1000, 7, 99)
    b = oracle.gen_uniform(1000, 7, 99)
    assert np.array_equal(a, b)
    assert a.min() >= 0.0 and a.max() < 1.0 and abs(a.mean() - 0.5) < 0.02
    assert not np.array_equal(a, oracle.gen_uniform(1000, 7, 100))


def test_mixture_generator_clusters():
    m = oracle.gen_mixture(2000, 16, 4, 0.05, 1, 2)
    # four tight clusters: every point within ~6 sigma*sqrt(dim) of one of 4 centres
    pts = m.T
    from scipy.cluster.vq import kmeans2

    cent, lab = kmeans2(pts.astype(np.float64), 4, seed=1, minit="++")
    spread = np.linalg.norm(pts - cent[lab], axis=1)
    assert np.median(spread) < 0.05 * np.sqrt(16) * 1.5


# --------------------------------------------------------------------------------------------
# Planted mistakes: each plausible bug in a k-NN ground truth fails one of the pins above.
# --------------------------------------------------------------------------------------------
def _numpy_knn(ref, qry, k, *, fma=False, f64=False, transposed=False, tie_high=False):
    D, R = ref.shape
    if transposed:
        ref = ref.reshape(R, D).T  # reads the buffer as n x dim
    acc = np.zeros((R, qry.shape[1]), dtype=np.float64 if f64 else np.float32)
    for d in range(D):
        t = (ref[d][:, None].astype(acc.dtype) - qry[d][None, :].astype(acc.dtype))
        if fma:
            acc = (acc.astype(np.float64) + t.astype(np.float64) * t.astype(np.float64)).astype(np.float32)
        else:
            acc = acc + t * t
    key = np.sqrt(acc).astype(np.float32)
    out_i = np.zeros((k, qry.shape[1]), dtype=np.int32)
    for c in range(qry.shape[1]):
        idx = np.arange(R)
        order = np.lexsort((-idx if tie_high else idx, key[:, c]))[:k]
        out_i[:, c] = order
    return np.take_along_axis(key, out_i, axis=0), out_i


def test_numpy_restatement_agrees_then_planted_mistakes_fail():
    ref, qry, k, gd, gi = load_case("uniform10_small")
    d, i = _numpy_knn(ref, qry, k)
    assert np.array_equal(i, gi) and np.array_equal(bits(d), bits(gd))

    dd, di = _numpy_knn(*load_case("duplicates")[:3], tie_high=True)
    assert not np.array_equal(di, load_case("duplicates")[4])  # wrong tie-break caught

    d2, i2 = _numpy_knn(ref, qry, k, fma=True)
    assert not np.array_equal(bits(d2), bits(gd))  # contraction into FMA caught (bitwise)

    d3, i3 = _numpy_knn(ref, qry, k, f64=True)
    assert not np.array_equal(bits(d3), bits(gd))  # FP64 accumulation caught (bitwise)

    d4, i4 = _numpy_knn(ref, qry, k, transposed=True)
    assert not np.array_equal(i4, gi)  # transposed layout caught
