"""GPU parity tests of the block index (router form): every search equals the oracle bit for bit,
whatever the probe count, ranking or data (clustered, uniform, ties)."""
import numpy as np
import pytest

import knnhip
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def check(d, i, gd, gi, what):
    d, i = d.cpu().numpy(), i.cpu().numpy()
    assert np.array_equal(i, gi), f"{what}: {int((i != gi).sum())} index mismatches"
    assert np.array_equal(bits(d), bits(gd)), f"{what}: distance mismatches"


def say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def clustered():
    R, Q, D, k = 20000, 600, 32, 8
    ref = knnhip.generate_mixture(R, D, 64, 0.05, 31, 32)
    qry = knnhip.generate_mixture(Q, D, 64, 0.05, 31, 33)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    return ref, qry, k, gd, gi


@pytest.mark.parametrize("probes", [1, 2, 4, 16])
def test_clustered_probes(clustered, probes):
    ref, qry, k, gd, gi = clustered
    say(f"index clustered probes={probes}")
    ix = knnhip.BlockIndex(ref, nblocks=64, iters=6)
    d, i = ix.search(qry, k, probes=probes)
    torch.cuda.synchronize()
    say(f"fallbacks {knnhip.last_fallback_count()} / {qry.shape[1]}")
    check(d, i, gd, gi, f"probes={probes}")
    ix.close()


def test_clustered_pruning_is_effective(clustered):
    ref, qry, k, gd, gi = clustered
    ix = knnhip.BlockIndex(ref, nblocks=64, iters=6)
    d, i = ix.search(qry, k, probes=4)
    nfb = knnhip.last_fallback_count()
    check(d, i, gd, gi, "probes=4")
    assert nfb <= qry.shape[1] // 10, f"too many exact fallbacks: {nfb}"


def test_uniform_data_stays_exact():
    R, Q, D, k = 8000, 200, 16, 10
    ref = knnhip.generate_uniform(R, D, seed=41)
    qry = knnhip.generate_uniform(Q, D, seed=42)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    ix = knnhip.BlockIndex(ref, nblocks=32, iters=4)
    d, i = ix.search(qry, k, probes=3)
    check(d, i, gd, gi, "uniform")


def test_arbitrary_scores_stay_exact(clustered):
    ref, qry, k, gd, gi = clustered
    ix = knnhip.BlockIndex(ref, nblocks=64, iters=6)
    scores = torch.rand((64, qry.shape[1]), device="cuda")
    d, i = ix.search(qry, k, probes=3, scores=scores)
    check(d, i, gd, gi, "random scores")


def test_duplicate_points_and_forced_fallback():
    rng = np.random.default_rng(5)
    base = np.round(3 * rng.random((8, 300))).astype(np.float32)
    ref = torch.from_numpy(np.concatenate([base, base, base], axis=1)).cuda()
    qry = torch.from_numpy(np.round(3 * rng.random((8, 40))).astype(np.float32)).cuda()
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), 12)
    ix = knnhip.BlockIndex(ref, nblocks=16, iters=3)
    check(*ix.search(qry, 12, probes=2), gd, gi, "duplicates")
    check(*ix.search(qry, 12, probes=2, force_fallback=True), gd, gi, "forced fallback")
    assert knnhip.last_fallback_count() == 40


def test_index_matches_plain_path_at_scale():
    # 262,144 refs (512 clusters, dim 96), 1024 queries, k 16: index == plain search
    R, Q, D, k = 1 << 18, 1024, 96, 16
    say("index vs plain at 262144 x 1024 x 96")
    ref = knnhip.generate_mixture(R, D, 512, 0.1, 51, 52)
    qry = knnhip.generate_mixture(Q, D, 512, 0.1, 51, 53)
    pd, pi = knnhip.search_device(ref, qry, k)
    ix = knnhip.BlockIndex(ref, nblocks=256, iters=8)
    d, i = ix.search(qry, k, probes=4)
    torch.cuda.synchronize()
    nfb = knnhip.last_fallback_count()
    say(f"index info {ix.info()}, fallbacks {nfb}")
    assert torch.equal(i, pi) and torch.equal(d, pd)
    sample = [0, 5, 500, 1023]
    gd, gi = oracle.knn_c(ref.cpu().numpy(), np.ascontiguousarray(qry.cpu().numpy()[:, sample]), k)
    check(d[:, sample], i[:, sample], gd, gi, "oracle sample")


def test_layout_is_a_partition(clustered):
    # every reference sits in exactly one block; blocks start on 32-point tiles and do not overlap
    ref, qry, k, gd, gi = clustered
    R = ref.shape[1]
    ix = knnhip.BlockIndex(ref, nblocks=64, iters=6)
    start, count, rb = ix.layout()
    rb = rb.cpu().numpy()
    assert int(count.sum()) == R
    assert np.all(start % 32 == 0)
    assert np.all(start[1:] >= start[:-1] + count[:-1])
    assert rb.min() >= 0 and rb.max() < 64
    assert np.array_equal(np.bincount(rb, minlength=64), count)
    ix.close()


def test_blocks_balanced_on_separated_clusters():
    # 256 well separated clusters of equal weight into 64 blocks: k-means++ seeding covers the
    # clusters, so no block swallows many of them (a block of 4 clusters holds about R/64 points)
    R, Q, D, k = 1 << 16, 256, 48, 8
    ref = knnhip.generate_mixture(R, D, 256, 0.05, 61, 62)
    qry = knnhip.generate_mixture(Q, D, 256, 0.05, 61, 63)
    ix = knnhip.BlockIndex(ref, nblocks=64, iters=8)
    start, count, rb = ix.layout()
    say(f"block sizes: min {count.min()} max {count.max()} mean {count.mean():.0f}")
    assert count.max() <= 4 * R // 64
    pd, pi = knnhip.search_device(ref, qry, k)
    d, i = ix.search(qry, k, probes=1)
    torch.cuda.synchronize()
    nfb = knnhip.last_fallback_count()
    say(f"fallbacks at one probe: {nfb}")
    assert torch.equal(i, pi) and torch.equal(d, pd)
    # neighbours lie in the query's own cluster, whose sub-balls are tight: one probe certifies most
    assert nfb <= Q // 10
    ix.close()


@pytest.mark.parametrize("ranking", ["bounds", "random"])
def test_many_queries_both_rankings(ranking):
    # 2000 queries: the block choice runs 64 queries per workgroup (the few-query form runs one
    # query per workgroup and is covered above); random scores force many certificate retries
    R, Q, D, k, nb = 20000, 2000, 32, 8, 64
    ref = knnhip.generate_mixture(R, D, 64, 0.05, 71, 72)
    qry = knnhip.generate_mixture(Q, D, 64, 0.05, 71, 73)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    ix = knnhip.BlockIndex(ref, nblocks=nb, iters=6)
    scores = None if ranking == "bounds" else torch.rand((nb, Q), device="cuda")
    d, i = ix.search(qry, k, probes=2, scores=scores)
    torch.cuda.synchronize()
    say(f"{ranking}: first-pass certificate failures {knnhip.last_fallback_count()} / {Q}")
    check(d, i, gd, gi, ranking)
    ix.close()


# precision scan in the index: every task-form instance (2, 4, 6, 8 chunks of 16 dims), ragged
# dims, one and several probes, uniform and clustered data
BF3_CASES = [
    (6000, 300, 24, 8, "uniform"),
    (9000, 257, 64, 16, "mixture"),
    (12000, 400, 96, 16, "mixture"),
    (7000, 129, 128, 32, "uniform"),
    (5000, 70, 100, 1, "mixture"),
]


@pytest.mark.parametrize("probes", [1, 3])
@pytest.mark.parametrize("case", BF3_CASES)
def test_precision_scan_equals_oracle(case, probes):
    R, Q, D, k, dist = case
    say(f"index bf16x3 {case} probes={probes}")
    if dist == "uniform":
        ref = knnhip.generate_uniform(R, D, seed=R + D)
        qry = knnhip.generate_uniform(Q, D, seed=R + D + 1)
    else:
        ref = knnhip.generate_mixture(R, D, 48, 0.1, R + D, 1)
        qry = knnhip.generate_mixture(Q, D, 48, 0.1, R + D, 2)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    ix = knnhip.BlockIndex(ref, nblocks=48, iters=4)
    d, i = ix.search(qry, k, probes=probes, precision="bf16x3")
    torch.cuda.synchronize()
    assert knnhip.last_plan()["precision"] == "bf16x3"
    say(f"certificate failures {knnhip.last_fallback_count()}")
    check(d, i, gd, gi, f"{case} probes={probes}")
    # the FP32 products give the same result
    d2, i2 = ix.search(qry, k, probes=probes)
    assert knnhip.last_plan()["precision"] == "fp32"
    assert torch.equal(d, d2) and torch.equal(i, i2)
    ix.close()


def test_precision_scan_with_scores_and_ties():
    # coarse lattice (many equal distances) and duplicated points, ranked by random scores
    rng = np.random.default_rng(17)
    base = (np.floor(rng.random((40, 2000)) * 4) / 4).astype(np.float32)
    ref = torch.from_numpy(np.concatenate([base, base[:, :500]], axis=1)).cuda()
    qry = torch.from_numpy((np.floor(rng.random((40, 150)) * 4) / 4).astype(np.float32)).cuda()
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), 12)
    ix = knnhip.BlockIndex(ref, nblocks=16, iters=3)
    scores = torch.rand((16, 150), device="cuda")
    d, i = ix.search(qry, 12, probes=2, scores=scores, precision="bf16x3")
    check(d, i, gd, gi, "lattice bf16x3")


def test_precision_scan_keeps_fp32_above_128_dims():
    ref = knnhip.generate_uniform(3000, 160, seed=3)
    qry = knnhip.generate_uniform(40, 160, seed=4)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), 5)
    ix = knnhip.BlockIndex(ref, nblocks=8, iters=2)
    d, i = ix.search(qry, 5, probes=2, precision="bf16x3")
    assert knnhip.last_plan()["precision"] == "fp32"
    check(d, i, gd, gi, "dim 160")


@pytest.mark.parametrize("k", [8, 100])
def test_wide_dims_staged_and_chunked_operands(k):
    # dim 768: the index's tiles are padded to 128 dims, so a search whose heaps leave no room
    # for the staged query rows (k = 100) takes the chunked operands over the same tiles
    R, Q, D = 6000, 90, 768
    say(f"index dim {D} k={k}")
    ref = knnhip.generate_mixture(R, D, 20, 0.2, 61, 62)
    qry = knnhip.generate_mixture(Q, D, 20, 0.2, 61, 63)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    ix = knnhip.BlockIndex(ref, nblocks=24, iters=3)
    d, i = ix.search(qry, k, probes=3)
    check(d, i, gd, gi, f"dim {D} k={k}")
    ix.close()


@pytest.mark.parametrize("dim,prec", [(24, "fp32"), (24, "bf16x3"), (96, "bf16x3"), (128, "bf16x3"),
                                      (100, "fp32"), (7, "bf16x3"), (200, "bf16x3")])
def test_block_bounds_never_exceed_the_true_distance(dim, prec):
    # what pruning rests on: every block bound is at most the float64 distance from the query to
    # every point of the block (FP32 products, or bfloat16 three-term products with their own
    # error allowance); far-away sigma keeps many bounds positive so the check has teeth
    R, Q, nb = 40000, 257, 64
    say(f"block bounds dim={dim} {prec}")
    ref = knnhip.generate_mixture(R, dim, 128, 0.03, 41 + dim, 42)
    qry = knnhip.generate_mixture(Q, dim, 128, 0.03, 41 + dim, 43)
    ix = knnhip.BlockIndex(ref, nblocks=nb, iters=6)
    lb = ix.bounds(qry, precision=prec).t().double()  # Q x nb
    _, _, ref_block = ix.layout()
    d64 = torch.cdist(qry.t().double(), ref.t().double(), compute_mode="donot_use_mm_for_euclid_dist")
    true = torch.full((Q, nb), float("inf"), dtype=torch.float64, device="cuda")
    true = true.scatter_reduce(1, ref_block.long().unsqueeze(0).expand(Q, R), d64, reduce="amin")
    bad = lb > true
    assert not bool(bad.any()), f"{int(bad.sum())} bounds above the true distance"
    live = torch.isfinite(true)
    positive = (lb[live] > 0).float().mean().item()
    tight = (lb[live] / true[live]).median().item()
    say(f"  positive bounds {positive:.3f}, median bound / distance {tight:.3f}")
    assert positive > 0.5
    if prec == "bf16x3":
        fp = ix.bounds(qry, precision="fp32").t().double()
        fin = torch.isfinite(fp)
        assert bool((fp[fin] - lb[fin]).abs().max() <= 1e-2 * (1 + fp[fin].abs().max()))
    ix.close()


@pytest.mark.parametrize("dim,prec,nb,Q", [(96, "bf16x3", 2048, 16384), (96, "fp32", 256, 1000),
                                           (24, "bf16x3", 8, 77), (24, "fp32", 5, 33), (128, "bf16x3", 512, 300),
                                           (7, "bf16x3", 64, 257), (200, "bf16x3", 64, 129)])
def test_slot_bounds_equal_the_per_ball_fold(dim, prec, nb, Q):
    # the searches' bounds kernel (blocks in 4/8/16/32-slot groups, minima over lanes) against the
    # per-ball fold (bounds ball by ball, folded block by block): the same per-ball bounds, so the
    # same minima, bit for bit; few blocks give blocks of many balls (8- to 32-slot groups), and
    # query counts that are not multiples of 32 or 64 test the padded columns
    R = 200000 if nb == 2048 else 30000
    say(f"slot bounds dim={dim} {prec} nb={nb} Q={Q}")
    ref = knnhip.generate_mixture(R, dim, 256, 0.05, 51 + dim, 52)
    qry = knnhip.generate_mixture(Q, dim, 256, 0.05, 51 + dim, 53)
    ix = knnhip.BlockIndex(ref, nblocks=nb, iters=4)
    a = ix.bounds(qry, precision=prec)
    b = ix.bounds(qry, precision=prec, per_ball=True)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.int32), b.view(torch.int32)), \
        f"{int((a != b).sum())} bounds differ between the slot form and the per-ball fold"
    assert bool(torch.isfinite(a).any())
    ix.close()
