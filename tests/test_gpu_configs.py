"""The benchmark configurations at their stated sizes, against the oracle.

Each test runs one workload at full size on the device and checks
  * size-independent properties on every query: distances ascending, equal distances in
    ascending index order (the reference's tie order), indexes distinct and in [0, R);
  * bit-exact equality with the oracle (the restatement of the reference ground truth
    code/test.cpp:129-175) on a fixed sample of queries spread over the whole set.

Workloads (the benchmark's own seeds):
  plain FP32 at scale  1,048,576 x 16,384, dim 128, k 32, uniform [0,1)
  precision form       1,048,576 x 8,192, dim 256, k 64, 1,024 Gaussians (sigma 0.05)
  batched form         the 512 problems of tests/golden/c4_problems.json
  learned-router form  4,194,304 x 16,384, dim 96, k 16, 4,096 Gaussians (sigma 0.1),
                       2,048 blocks
"""
import json
import os

import numpy as np
import pytest

import knnhip
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def check_properties(d, i, R, what):
    d = d.cpu().numpy() if hasattr(d, "cpu") else d
    i = i.cpu().numpy() if hasattr(i, "cpu") else i
    assert np.all(np.isfinite(d)), f"{what}: non-finite distance"
    step = np.diff(d, axis=0)
    assert np.all(step >= 0), f"{what}: distances not ascending"
    tie = step == 0
    assert np.all(np.diff(i, axis=0)[tie] > 0), f"{what}: a tie not in index order"
    assert np.all((i >= 0) & (i < R)), f"{what}: index out of range"
    assert np.all(np.diff(np.sort(i, axis=0), axis=0) > 0), f"{what}: repeated index"
    return d, i


def check_sample(ref_h, qry_h, k, d, i, sample, what):
    qs = np.ascontiguousarray(qry_h[:, sample])
    gd, gi = oracle.knn_c(ref_h, qs, k)
    bad = np.argwhere(i[:, sample] != gi)
    assert bad.size == 0, f"{what}: {len(bad)} index mismatches vs the oracle, first {bad[:3].tolist()}"
    badd = np.argwhere(bits(d[:, sample]) != bits(gd))
    assert badd.size == 0, f"{what}: {len(badd)} distance mismatches vs the oracle"


def spread(Q, n):
    return np.unique(np.linspace(0, Q - 1, n).astype(np.int64))


def test_plain_fp32_at_scale_full_size():
    R, Q, D, k = 1 << 20, 16384, 128, 32
    say(f"plain FP32 {R}x{Q}x{D} k={k}")
    ref = knnhip.generate_uniform(R, D, seed=0x6B4E4E02)
    qry = knnhip.generate_uniform(Q, D, seed=0x6B4E4E12)
    d, i = knnhip.search_device(ref, qry, k)
    torch.cuda.synchronize()
    plan = knnhip.last_plan()
    assert plan["query_blocks"] == Q // 32 and plan["precision"] == "fp32"
    assert knnhip.last_fallback_count() == 0
    d, i = check_properties(d, i, R, "plain FP32")
    check_sample(ref.cpu().numpy(), qry.cpu().numpy(), k, d, i, spread(Q, 48), "plain FP32")
    # the precision request on the same data: bfloat16 three-term products, equal output
    d3, i3 = knnhip.search_device(ref, qry, k, precision="bf16x3")
    torch.cuda.synchronize()
    assert knnhip.last_plan()["precision"] == "bf16x3"
    assert np.array_equal(i3.cpu().numpy(), i) and np.array_equal(bits(d3.cpu().numpy()), bits(d))


@pytest.fixture(scope="module")
def dense_clusters():
    R, Q, D, k = 1 << 20, 8192, 256, 64
    say(f"precision-form data {R}x{Q}x{D} k={k}, 1,024 Gaussians")
    ref = knnhip.generate_mixture(R, D, 1024, 0.05, 0x6B4E4E03, 0x6B4E4E13)
    qry = knnhip.generate_mixture(Q, D, 1024, 0.05, 0x6B4E4E03, 0x6B4E4E23)
    torch.cuda.synchronize()
    d, i = knnhip.search_device(ref, qry, k)
    torch.cuda.synchronize()
    d, i = check_properties(d, i, R, "dense clusters FP32")
    check_sample(ref.cpu().numpy(), qry.cpu().numpy(), k, d, i, spread(Q, 48), "dense clusters FP32")
    yield ref, qry, k, d, i
    del ref, qry
    torch.cuda.empty_cache()


def test_precision_form_full_size_plain(dense_clusters):
    ref, qry, k, d, i = dense_clusters
    d2, i2 = knnhip.search_device(ref, qry, k, precision="bf16x3")
    torch.cuda.synchronize()
    assert knnhip.last_fallback_count() == 0
    assert np.array_equal(i2.cpu().numpy(), i) and np.array_equal(bits(d2.cpu().numpy()), bits(d))


def test_precision_form_full_size_block_residuals(dense_clusters):
    # the precision form as the workload names it: bfloat16 products (single pass, on block-centroid
    # residuals) and the k + 64 candidate list re-ranked in FP32, through a block index
    ref, qry, k, d, i = dense_clusters
    ix = knnhip.BlockIndex(ref, nblocks=1024, iters=8)
    for probes in (1, 2):
        d2, i2 = ix.search(qry, k, probes=probes, precision="bf16")
        torch.cuda.synchronize()
        plan = knnhip.last_plan()
        assert plan["precision"] == "bf16-residual"
        say(f"residual P={probes}: certificate failures {knnhip.last_fallback_count()}")
        assert np.array_equal(i2.cpu().numpy(), i) and np.array_equal(bits(d2.cpu().numpy()), bits(d))
    ix.close()


def test_batched_form_all_512_problems():
    with open(os.path.join(GOLDEN, "c4_problems.json")) as f:
        P = json.load(f)["problems"]
    assert len(P) == 512
    say("batched form: 512 problems")
    items = [(knnhip.generate_uniform(p["ref_nb"], p["dim"], seed=p["ref_seed"]),
              knnhip.generate_uniform(p["query_nb"], p["dim"], seed=p["query_seed"]), p["k"]) for p in P]
    torch.cuda.synchronize()
    outs = knnhip.search_batched(items)
    torch.cuda.synchronize()
    say(f"batched certificate failures: {knnhip.batched_last_failures()}")
    for j, ((ref, qry, k), (d, i)) in enumerate(zip(items, outs)):
        R, Q = ref.shape[1], qry.shape[1]
        dh, ih = check_properties(d, i, R, f"problem {j}")
        sample = spread(Q, 3)
        check_sample(ref.cpu().numpy(), qry[:, sample].cpu().numpy(), k, dh[:, sample], ih[:, sample],
                     np.arange(len(sample)), f"problem {j} (R={R} Q={Q} dim={ref.shape[0]} k={k})")


def test_learned_router_form_full_size():
    R, Q, D, k, NB = 4 << 20, 16384, 96, 16, 2048
    say(f"learned-router form {R}x{Q}x{D} k={k}, {NB} blocks")
    seed_c = 0x6B4E4E05
    ref = knnhip.generate_mixture(R, D, 4096, 0.1, seed_c, 0x6B4E4E06)
    qry = knnhip.generate_mixture(Q, D, 4096, 0.1, seed_c, 0x6B4E4E16)
    torch.cuda.synchronize()
    d, i = knnhip.search_device(ref, qry, k)
    torch.cuda.synchronize()
    dh, ih = check_properties(d, i, R, "plain FP32")
    check_sample(ref.cpu().numpy(), qry.cpu().numpy(), k, dh, ih, spread(Q, 32), "plain FP32")
    ix = knnhip.BlockIndex(ref, nblocks=NB, iters=8)
    trq = knnhip.generate_mixture(16384, D, 4096, 0.1, seed_c, 0x6B4E4E26)
    mlp = knnhip.Router(D, 256, NB, seed=0x6B4E4E07)
    mlp.train(ix, trq, k, epochs=4, batch=1024, lr=1e-3)
    for what, kw in (("index fp32 P=5", dict(probes=5)),
                     ("index bf16x3 P=5", dict(probes=5, precision="bf16x3")),
                     ("index bf16x3 P=5 routed", dict(probes=5, precision="bf16x3", router=mlp)),
                     ("index bf16 residual P=5", dict(probes=5, precision="bf16")),
                     ("index bf16 residual P=5 routed", dict(probes=5, precision="bf16", router=mlp))):
        d2, i2 = ix.search(qry, k, **kw)
        torch.cuda.synchronize()
        say(f"{what}: certificate failures {knnhip.last_fallback_count()}")
        assert np.array_equal(i2.cpu().numpy(), ih), what
        assert np.array_equal(bits(d2.cpu().numpy()), bits(dh)), what
    mlp.close()
    ix.close()
