"""GPU tests of the saved formats (block index, router) and the k-th-distance entry point."""
import os

import numpy as np
import pytest

import knnhip
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_kth_distance_equals_row_k_minus_1():
    R, Q, D, k = 5000, 333, 37, 11
    rng = np.random.default_rng(3)
    ref = rng.random((D, R), dtype=np.float32)
    qry = rng.random((D, Q), dtype=np.float32)
    gd, gi = oracle.knn_c(ref, qry, k)
    kth = knnhip.kth_distance(ref, qry, k)
    assert np.array_equal(kth.view(np.uint32), gd[k - 1].view(np.uint32))
    with pytest.raises(knnhip.KnnError):
        knnhip.kth_distance(ref, qry, R + 1)


def test_index_save_load_search_equal(tmp_path):
    R, Q, D, k = 30000, 500, 48, 8
    ref = knnhip.generate_mixture(R, D, 100, 0.05, 101, 102)
    qry = knnhip.generate_mixture(Q, D, 100, 0.05, 101, 103)
    ix = knnhip.BlockIndex(ref, nblocks=64, iters=5)
    d0, i0 = ix.search(qry, k, probes=2)
    path = str(tmp_path / "index.knnhip")
    ix.save(path)
    ix2 = knnhip.BlockIndex.load(path, ref)
    assert ix2.info()["nblocks"] == 64
    s0, c0, rb0 = ix.layout()
    s1, c1, rb1 = ix2.layout()
    assert np.array_equal(s0, s1) and np.array_equal(c0, c1) and torch.equal(rb0, rb1)
    d1, i1 = ix2.search(qry, k, probes=2)
    torch.cuda.synchronize()
    assert torch.equal(i0, i1) and torch.equal(d0, d1)
    # wrong reference set shape, wrong magic, truncation
    with pytest.raises(knnhip.KnnError):
        knnhip.BlockIndex.load(path, ref[:, : R - 1].contiguous())
    bad = tmp_path / "bad.knnhip"
    raw = open(path, "rb").read()
    bad.write_bytes(b"X" + raw[1:])
    with pytest.raises(knnhip.KnnError):
        knnhip.BlockIndex.load(str(bad), ref)
    bad.write_bytes(raw[:-5])
    with pytest.raises(knnhip.KnnError):
        knnhip.BlockIndex.load(str(bad), ref)
    bad.write_bytes(raw + b"\0")
    with pytest.raises(knnhip.KnnError):
        knnhip.BlockIndex.load(str(bad), ref)


def test_router_save_load_scores_equal(tmp_path):
    r = knnhip.Router(24, 32, 40, seed=9)
    x = torch.rand((24, 70), device="cuda")
    s0 = r.scores(x)
    path = str(tmp_path / "router.knnhip")
    r.save(path)
    r2 = knnhip.Router.load(path)
    assert (r2.dim, r2.hidden, r2.nblocks) == (24, 32, 40)
    assert np.array_equal(r.get_params(), r2.get_params())
    assert torch.equal(s0, r2.scores(x))
    raw = open(path, "rb").read()
    (tmp_path / "bad").write_bytes(raw[:-4])
    with pytest.raises(knnhip.KnnError):
        knnhip.Router.load(str(tmp_path / "bad"))
