"""GPU tests of the residual precision form of the block index (options.precision = 2).

Single-pass bfloat16 products of (q - c) and (r - c), c the scanned block's centroid, with the
exact FP32 re-rank and a certificate widened by the residual error bound.  Bar: indexes equal
and distances bit-identical to the oracle (the reference ground truth, code/test.cpp:129-175),
and to the FP32 index search, on every case: clustered and uniform data (uniform blocks are wide,
their bound large, and most queries take the retry), tile tails (dims 7 .. 256), k = 1 and
large k, duplicated points and lattice ties, and router scores.
"""
import numpy as np
import pytest

import knnhip
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def check(d, i, gd, gi, what):
    d, i = d.cpu().numpy(), i.cpu().numpy()
    bad = np.argwhere(i != gi)
    assert bad.size == 0, f"{what}: {len(bad)} index mismatches, first {bad[:3].tolist()}"
    assert np.array_equal(bits(d), bits(gd)), f"{what}: distance mismatches"


CASES = [
    # R, Q, dim, k, data, blocks
    (8000, 200, 7, 5, "mixture", 32),
    (9000, 257, 24, 16, "mixture", 48),
    (12000, 300, 96, 16, "mixture", 48),
    (7000, 129, 100, 1, "mixture", 40),
    (10000, 190, 128, 32, "mixture", 48),
    (9000, 150, 200, 8, "mixture", 32),
    (16000, 100, 256, 64, "mixture", 64),
    (6000, 120, 64, 16, "uniform", 24),
]


@pytest.mark.parametrize("probes", [1, 3])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "R%d_Q%d_D%d_k%d_%s" % c[:5])
def test_residual_scan_equals_oracle(case, probes):
    R, Q, D, k, data, nb = case
    say(f"residual {case} probes={probes}")
    if data == "uniform":
        ref = knnhip.generate_uniform(R, D, seed=R + D)
        qry = knnhip.generate_uniform(Q, D, seed=R + D + 1)
    else:
        ref = knnhip.generate_mixture(R, D, nb, 0.05, R + D, 1)
        qry = knnhip.generate_mixture(Q, D, nb, 0.05, R + D, 2)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    ix = knnhip.BlockIndex(ref, nblocks=nb, iters=4)
    d, i = ix.search(qry, k, probes=probes, precision="bf16")
    torch.cuda.synchronize()
    assert knnhip.last_plan()["precision"] == "bf16-residual"
    say(f"certificate failures {knnhip.last_fallback_count()}")
    check(d, i, gd, gi, f"{case} probes={probes}")
    d2, i2 = ix.search(qry, k, probes=probes)
    assert torch.equal(d, d2) and torch.equal(i, i2)
    ix.close()


def test_residual_certificate_passes_on_tight_clusters():
    # separated clusters, one per block: the residual bound is far below the neighbour gaps, so
    # (almost) every query certifies at level 0
    R, Q, D, k, nb = 32768, 512, 128, 16, 32
    ref = knnhip.generate_mixture(R, D, nb, 0.05, 71, 72)
    qry = knnhip.generate_mixture(Q, D, nb, 0.05, 71, 73)
    ix = knnhip.BlockIndex(ref, nblocks=nb, iters=6)
    d, i = ix.search(qry, k, probes=2, precision="bf16")
    torch.cuda.synchronize()
    fails = knnhip.last_fallback_count()
    say(f"tight clusters: certificate failures {fails} of {Q}")
    assert fails <= Q // 50
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    check(d, i, gd, gi, "tight clusters")


def test_residual_scan_with_ties_duplicates_and_scores():
    rng = np.random.default_rng(23)
    base = (np.floor(rng.random((48, 3000)) * 4) / 4).astype(np.float32)
    ref = torch.from_numpy(np.concatenate([base, base[:, :700]], axis=1)).cuda()
    qry = torch.from_numpy((np.floor(rng.random((48, 160)) * 4) / 4).astype(np.float32)).cuda()
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), 12)
    ix = knnhip.BlockIndex(ref, nblocks=16, iters=3)
    for scores in (None, torch.rand((16, 160), device="cuda")):
        d, i = ix.search(qry, 12, probes=3, scores=scores, precision="bf16")
        check(d, i, gd, gi, "lattice residual")
    ix.close()


def test_residual_form_wider_than_256_dims_takes_fp32():
    ref = knnhip.generate_mixture(3000, 300, 8, 0.1, 5, 6)
    qry = knnhip.generate_mixture(40, 300, 8, 0.1, 5, 7)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), 5)
    ix = knnhip.BlockIndex(ref, nblocks=8, iters=2)
    d, i = ix.search(qry, 5, probes=2, precision="bf16")
    assert knnhip.last_plan()["precision"] == "fp32"
    check(d, i, gd, gi, "dim 300")
    ix.close()


def test_residual_form_after_load(tmp_path):
    # the residual tiles of a loaded index come from the reference set it is loaded over
    R, Q, D, k = 20000, 200, 48, 10
    ref = knnhip.generate_mixture(R, D, 40, 0.05, 31, 32)
    qry = knnhip.generate_mixture(Q, D, 40, 0.05, 31, 33)
    ix = knnhip.BlockIndex(ref, nblocks=40)
    path = str(tmp_path / "ix.bin")
    ix.save(path)
    ix.close()
    ix2 = knnhip.BlockIndex.load(path, ref)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    d, i = ix2.search(qry, k, probes=2, precision="bf16")
    check(d, i, gd, gi, "loaded index")
    ix2.close()
