"""GPU tests of the batched form (one persistent scan launch over every class's tasks): every
problem of a batch gets exactly the oracle's result (and exactly what a search of that problem
alone returns), whatever the class mix, tile shape or tuning."""
import json
import os

import numpy as np
import pytest

import knnhip
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))


def say(m):
    print(m, flush=True)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def make_items(shapes, seed):
    items = []
    for j, (R, Q, D, k) in enumerate(shapes):
        ref = knnhip.generate_uniform(R, D, seed=seed + 2 * j)
        qry = knnhip.generate_uniform(Q, D, seed=seed + 2 * j + 1)
        items.append((ref, qry, k))
    return items


# mixed classes: shared (dim, k) classes of several problems, ragged sizes, a dimension that
# takes the chunked scan (512), k = 128 (kp 192), k equal to R, one-query problems
SHAPES = [
    (1500, 40, 32, 8), (3000, 7, 32, 8), (1024, 100, 32, 8),
    (2500, 33, 64, 1), (1100, 1, 64, 1),
    (4096, 64, 128, 16), (5000, 65, 128, 16),
    (1200, 20, 256, 64), (2048, 31, 512, 16), (3100, 17, 512, 16),
    (1300, 50, 96, 128), (130, 9, 24, 130),
]


def test_mixed_batch_equals_oracle():
    items = make_items(SHAPES, 900)
    outs = knnhip.search_batched(items)
    torch.cuda.synchronize()
    plan = knnhip.last_plan()
    say(f"plan {plan}; certificate failures {knnhip.batched_last_failures()}")
    assert plan["scan_launches"] == 1  # every class's tasks in one persistent scan launch
    for j, ((ref, qry, k), (d, i)) in enumerate(zip(items, outs)):
        gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
        assert np.array_equal(i.cpu().numpy(), gi), f"problem {j} {SHAPES[j]}: indexes"
        assert np.array_equal(bits(d.cpu().numpy()), bits(gd)), f"problem {j} {SHAPES[j]}: distances"


def test_every_tile_shape_equals_single_searches():
    items = make_items(SHAPES[:10], 1900)
    single = [knnhip.search_device(r, q, k) for r, q, k in items]
    classes = sorted({(r.shape[0], k) for r, q, k in items})
    for L in (2048, 8192, 32768, 131072):
        for dim, k in classes:
            knnhip.batched_set_slice(dim, k, L)
        outs = knnhip.search_batched(items)
        torch.cuda.synchronize()
        for j, ((d, i), (sd, si)) in enumerate(zip(outs, single)):
            assert torch.equal(i, si) and torch.equal(d, sd), f"slice {L}, problem {j}"
    for dim, k in classes:
        knnhip.batched_set_slice(dim, k, 0)
    assert knnhip.batched_table() == []


def test_autotune_then_search_c4_sample():
    # 24 problems of the committed batched workload, sizes capped for the oracle-free check
    P = json.load(open(os.path.join(HERE, "golden", "c4_problems.json")))["problems"][:24]
    items = []
    for p in P:
        R, Q = min(p["ref_nb"], 20000), min(p["query_nb"], 300)
        items.append((knnhip.generate_uniform(R, p["dim"], seed=p["ref_seed"]),
                      knnhip.generate_uniform(Q, p["dim"], seed=p["query_seed"]), p["k"]))
    single = [knnhip.search_device(r, q, k) for r, q, k in items]
    table = knnhip.batched_autotune(items)
    say(f"tuned table {table}")
    assert len(table) == len({(r.shape[0], k) for r, q, k in items})
    outs = knnhip.search_batched(items)
    torch.cuda.synchronize()
    plan = knnhip.last_plan()
    assert all(c["tuned"] for c in plan["classes"])
    for j, ((d, i), (sd, si)) in enumerate(zip(outs, single)):
        assert torch.equal(i, si) and torch.equal(d, sd), f"problem {j}"
    for c in table:
        knnhip.batched_set_slice(c["dim"], c["k"], 0)


def test_bad_problem_is_rejected():
    ref = knnhip.generate_uniform(100, 8, seed=1)
    qry = knnhip.generate_uniform(10, 8, seed=2)
    with pytest.raises(knnhip.KnnError):
        knnhip.search_batched([(ref, qry, 5), (ref, qry, 101)])  # k > ref_nb


def test_certificate_failures_across_classes_in_flight():
    # lattice data (many equal distances) with a one-candidate margin: certificates fail in
    # several classes that are in flight together, and each failed query is searched again
    rng = np.random.default_rng(44)
    shapes = [(900, 40, 16, 10), (1300, 70, 16, 10), (800, 33, 24, 6), (1100, 50, 24, 6),
              (700, 20, 40, 12), (1500, 64, 40, 12), (600, 30, 8, 4), (1000, 45, 8, 4), (950, 25, 32, 9)]
    items = []
    for R, Q, D, k in shapes:
        ref = torch.from_numpy((np.floor(rng.random((D, R)) * 3) / 3).astype(np.float32)).cuda()
        qry = torch.from_numpy((np.floor(rng.random((D, Q)) * 3) / 3).astype(np.float32)).cuda()
        items.append((ref, qry, k))
    outs = knnhip.search_batched(items, margin=1)
    torch.cuda.synchronize()
    fails = knnhip.batched_last_failures()
    say(f"classes {len({(r.shape[0], k) for r, q, k in items})}, certificate failures {fails}")
    assert fails > 0
    for j, ((ref, qry, k), (d, i)) in enumerate(zip(items, outs)):
        gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
        assert np.array_equal(i.cpu().numpy(), gi), f"problem {j}: indexes"
        assert np.array_equal(bits(d.cpu().numpy()), bits(gd)), f"problem {j}: distances"
