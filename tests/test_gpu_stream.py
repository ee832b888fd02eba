"""Streamed host searches: a device memory budget smaller than the problem.

The reference set goes through the device in chunks and the queries in
groups; the per-chunk exact results are merged on the device by (distance,
index).  Bar: bit-equal to the reference ground truth (code/test.cpp:129-175)
through every host entry point, with the chunking read back from the plan.
"""
import numpy as np
import pytest

import knnhip
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def say(msg):
    print(msg, flush=True)


@pytest.fixture(autouse=True)
def default_budget():
    yield
    knnhip.set_device_budget(0)


def find_budget(R, Q, dim, k, want, largest=False):
    """Smallest (or largest) budget in 1 MiB steps whose chunking satisfies want(ref_chunks, query_chunks)."""
    for mib in (range(1023, 7, -1) if largest else range(8, 1024)):
        try:
            rc, qc = knnhip.stream_chunks(R, Q, dim, k, mib << 20)
        except knnhip.KnnError:
            continue
        nr, nq = -(-R // rc), -(-Q // qc)
        if want(nr, nq):
            return mib << 20, nr, nq
    raise AssertionError(f"no budget gives the wanted chunking for {(R, Q, dim, k)}")


def check_entry_points(ref, qry, k, nr, nq):
    gd, gi = oracle.knn_c(ref, qry, k)
    for entry in ("knn_cuda_global", "knn_cuda_texture", "knn_cublas"):
        d, i = getattr(knnhip, entry)(ref, qry, k)
        plan = knnhip.last_plan()
        assert plan.get("streamed") and plan["ref_chunks"] == nr and plan["query_chunks"] == nq, plan
        assert np.array_equal(i, gi), f"{entry}: {np.count_nonzero(i != gi)} index mismatches"
        assert np.array_equal(bits(d), bits(gd)), f"{entry}: distance mismatches"
    kth = knnhip.kth_distance(ref, qry, k)
    assert np.array_equal(bits(kth), bits(gd[k - 1])), "kth_distance"


@pytest.mark.parametrize("largest", [False, True], ids=["tight", "loose"])
@pytest.mark.parametrize("R,Q,dim,k", [(30000, 1500, 24, 16), (20011, 2100, 7, 5), (24000, 900, 130, 33)])
def test_reference_chunks_and_query_groups(R, Q, dim, k, largest):
    rng = np.random.default_rng(R + Q + dim)
    ref = rng.random((dim, R), dtype=np.float32)
    qry = rng.random((dim, Q), dtype=np.float32)
    want = (lambda nr, nq: nr >= 3) if largest else (lambda nr, nq: nr >= 3 and nq >= 2)
    budget, nr, nq = find_budget(R, Q, dim, k, want, largest)
    say(f"R={R} Q={Q} dim={dim} k={k}: budget {budget >> 20} MiB, {nr} reference chunks, {nq} query groups")
    knnhip.set_device_budget(budget)
    check_entry_points(ref, qry, k, nr, nq)


def test_query_groups_with_resident_references():
    # the legacy split: every reference fits, the queries go in groups
    R, Q, dim, k = 9000, 14000, 8, 12
    rng = np.random.default_rng(7)
    ref = rng.random((dim, R), dtype=np.float32)
    qry = rng.random((dim, Q), dtype=np.float32)
    budget, nr, nq = find_budget(R, Q, dim, k, lambda nr, nq: nr == 1 and nq >= 2)
    say(f"budget {budget >> 20} MiB, {nq} query groups")
    knnhip.set_device_budget(budget)
    check_entry_points(ref, qry, k, nr, nq)


def test_ties_across_chunk_boundaries():
    # lattice points repeat all over the set: equal distances fall in
    # different chunks, and the lower index must win every tie
    R, Q, dim, k = 26000, 1200, 4, 40
    rng = np.random.default_rng(11)
    ref = (np.floor(rng.random((dim, R)) * 3) / 3).astype(np.float32)
    qry = (np.floor(rng.random((dim, Q)) * 3) / 3).astype(np.float32)
    budget, nr, nq = find_budget(R, Q, dim, k, lambda nr, nq: nr >= 3)
    say(f"budget {budget >> 20} MiB, {nr} reference chunks, {nq} query groups")
    knnhip.set_device_budget(budget)
    check_entry_points(ref, qry, k, nr, nq)


def test_back_to_back_calls_with_different_data():
    # the chunk and query buffers keep the previous call's contents: a search that ran ahead of
    # its own copies would return the previous data's neighbours
    R, Q, dim, k = 40000, 3000, 16, 10
    budget, nr, nq = find_budget(R, Q, dim, k, lambda nr, nq: nr >= 3 and nq >= 2, largest=True)
    say(f"budget {budget >> 20} MiB, {nr} reference chunks, {nq} query groups")
    knnhip.set_device_budget(budget)
    rng = np.random.default_rng(5)
    for rep in range(3):
        ref = rng.random((dim, R), dtype=np.float32)
        qry = rng.random((dim, Q), dtype=np.float32)
        gd, gi = oracle.knn_c(ref, qry, k)
        d, i = knnhip.knn_cuda_global(ref, qry, k)
        plan = knnhip.last_plan()
        assert plan["ref_chunks"] == nr and plan["query_chunks"] == nq
        assert np.array_equal(i, gi), f"call {rep}: {np.count_nonzero(i != gi)} index mismatches"
        assert np.array_equal(bits(d), bits(gd)), f"call {rep}: distances"


def test_budget_too_small_is_an_error():
    ref = np.zeros((16, 5000), dtype=np.float32)
    qry = np.zeros((16, 10), dtype=np.float32)
    knnhip.set_device_budget(1 << 20)
    with pytest.raises(knnhip.KnnError, match="budget"):
        knnhip.knn_cuda_global(ref, qry, 4)
    with pytest.raises(knnhip.KnnError):
        knnhip.set_device_budget(-1)
    knnhip.set_device_budget(0)
    d, i = knnhip.knn_cuda_global(ref, qry, 4)
    assert "streamed" not in knnhip.last_plan()
    assert np.array_equal(i[:, 0], np.arange(4))


def test_chunks_never_smaller_than_k():
    R, Q, dim, k = 20000, 500, 32, 600
    seen = 0
    for mib in range(8, 1024, 8):
        try:
            rc, qc = knnhip.stream_chunks(R, Q, dim, k, mib << 20)
        except knnhip.KnnError as e:
            assert "budget" in str(e)
            continue
        nr = -(-R // rc)
        assert R // nr >= k and 1 <= qc <= Q
        seen += 1
    assert seen > 0
