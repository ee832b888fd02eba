"""GPU tests of ball probes in block-index searches: the scan visits only the balls of its probed
blocks whose lower bound lies below the reach of the nearest balls holding k + margin members (one
pass), or the nearest ball of each query's first block and then the balls whose lower bound lies
below the threshold that first scan left (two passes).  Every search equals the oracle bit for
bit, and equals the whole-block search; index files of version 2 keep the balls' position ranges, version-1 files
still load (whole-block searches)."""
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
    R, Q, D, k = 40000, 700, 48, 10
    ref = knnhip.generate_mixture(R, D, 80, 0.1, 61, 62)
    qry = knnhip.generate_mixture(Q, D, 80, 0.1, 61, 63)
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    ix = knnhip.BlockIndex(ref, nblocks=64, iters=6)
    yield ref, qry, k, gd, gi, ix
    ix.close()


@pytest.mark.parametrize("units", ["balls", "balls-threshold"])
@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16"])
@pytest.mark.parametrize("probes", [1, 3, 8])
def test_ball_probes_equal_oracle(clustered, precision, probes, units):
    ref, qry, k, gd, gi, ix = clustered
    say(f"{units} precision={precision} probes={probes}")
    d, i = ix.search(qry, k, probes=probes, precision=precision, timing=True, units=units)
    torch.cuda.synchronize()
    u = ix.last_units()
    say(f"units {u} fallbacks {knnhip.last_fallback_count()} work {ix.last_work()}")
    assert u["units"] == "balls" and u["balls"] > 64
    check(d, i, gd, gi, f"{units} {precision} P={probes}")
    d2, i2 = ix.search(qry, k, probes=probes, precision=precision, units="blocks")
    torch.cuda.synchronize()
    assert ix.last_units()["units"] == "blocks"
    check(d2, i2, gd, gi, f"blocks {precision} P={probes}")


def test_ball_probes_scan_fewer_pairs(clustered):
    ref, qry, k, gd, gi, ix = clustered
    ix.search(qry, k, probes=4, timing=True)
    wb = ix.last_work()
    ix.search(qry, k, probes=4, timing=True, units="blocks")
    wk = ix.last_work()
    say(f"balls {wb} blocks {wk}")
    assert wb["pairs"] < wk["pairs"]


def test_ball_probes_with_router_scores_and_ties():
    # duplicated points (exact ties across balls) and random first-probe scores
    rng = np.random.default_rng(8)
    base = np.round(4 * rng.random((12, 900))).astype(np.float32)
    ref = torch.from_numpy(np.concatenate([base, base], axis=1)).cuda()
    qry = torch.from_numpy(np.round(4 * rng.random((12, 150))).astype(np.float32)).cuda()
    k = 9
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    ix = knnhip.BlockIndex(ref, nblocks=16, iters=3)
    scores = torch.rand((16, qry.shape[1]), device="cuda")
    for prec in ("fp32", "bf16"):
        for units in ("balls", "balls-threshold"):
            check(*ix.search(qry, k, probes=3, scores=scores, precision=prec, units=units), gd, gi,
                  f"scores {prec} {units}")
            check(*ix.search(qry, k, probes=2, precision=prec, units=units), gd, gi, f"ties {prec} {units}")
    with pytest.raises(knnhip.KnnError):
        ix.search(qry, k, units="rows")
    ix.close()


def _strip_to_v1(raw):
    """The version-1 form of a version-2 index file: version word 1, ball ranges dropped."""
    hdr = np.frombuffer(bytes(raw[12:36]), dtype=np.int32)
    nf = int(hdr[3])
    assert np.frombuffer(bytes(raw[8:12]), dtype=np.uint32)[0] == 2
    v1 = bytearray(raw[: len(raw) - 8 * nf])
    v1[8:12] = np.array([1], dtype=np.uint32).tobytes()
    return bytes(v1)


def test_index_file_versions(clustered, tmp_path):
    ref, qry, k, gd, gi, ix = clustered
    path = tmp_path / "ix2.bin"
    ix.save(str(path))
    raw = path.read_bytes()
    ix2 = knnhip.BlockIndex.load(str(path), ref)
    check(*ix2.search(qry, k, probes=3, precision="bf16"), gd, gi, "v2 loaded")
    assert ix2.last_units()["units"] == "balls"
    # version 1: no ball ranges, whole-block searches, radii refitted to the nearest ball
    p1 = tmp_path / "ix1.bin"
    p1.write_bytes(_strip_to_v1(raw))
    ix1 = knnhip.BlockIndex.load(str(p1), ref)
    check(*ix1.search(qry, k, probes=3), gd, gi, "v1 loaded")
    u = ix1.last_units()
    assert u["units"] == "blocks" and u["balls"] == 0
    # a v1 index saves as version 1
    p1b = tmp_path / "ix1b.bin"
    ix1.save(str(p1b))
    assert p1b.read_bytes()[8:12] == np.array([1], dtype=np.uint32).tobytes()
    ix1.close()
    ix2.close()


def test_inconsistent_ball_ranges_are_refused(clustered, tmp_path):
    ref, qry, k, gd, gi, ix = clustered
    path = tmp_path / "ix.bin"
    ix.save(str(path))
    raw = bytearray(path.read_bytes())
    nf = int(np.frombuffer(bytes(raw[12:36]), dtype=np.int32)[3])
    cnt_off = len(raw) - 4 * nf
    cnt = np.frombuffer(bytes(raw[cnt_off:]), dtype=np.int32).copy()
    j = int(np.argmax(cnt))
    cnt[j] -= 1  # the block's ranges no longer cover it
    raw[cnt_off:] = cnt.tobytes()
    path.write_bytes(bytes(raw))
    with pytest.raises(knnhip.KnnError, match="inconsistent index"):
        knnhip.BlockIndex.load(str(path), ref)


def test_ball_probes_at_scale_equal_plain():
    # 262,144 refs (512 clusters, dim 96), 2048 queries, k 16: ball searches == plain search
    R, Q, D, k = 1 << 18, 2048, 96, 16
    say("balls vs plain at 262144 x 2048 x 96")
    ref = knnhip.generate_mixture(R, D, 512, 0.1, 71, 72)
    qry = knnhip.generate_mixture(Q, D, 512, 0.1, 71, 73)
    pd, pi = knnhip.search_device(ref, qry, k)
    ix = knnhip.BlockIndex(ref, nblocks=128, iters=8)
    for prec in ("fp32", "bf16x3", "bf16"):
        for units in ("balls", "balls-threshold"):
            d, i = ix.search(qry, k, probes=5, precision=prec, timing=True, units=units)
            torch.cuda.synchronize()
            say(f"{prec} {units}: units {ix.last_units()} fallbacks {knnhip.last_fallback_count()} "
                f"work {ix.last_work()} stages {ix.stage_ms()}")
            assert torch.equal(i, pi) and torch.equal(d, pd), (prec, units)
    sample = [0, 7, 1000, 2047]
    gd, gi = oracle.knn_c(ref.cpu().numpy(), np.ascontiguousarray(qry.cpu().numpy()[:, sample]), k)
    check(d[:, sample], i[:, sample], gd, gi, "oracle sample")
    ix.close()


def test_scan_only_timing_reads_the_scan_stage_alone():
    # timing="scan": the scan stage's time (positive), every other stage -1; the output is the
    # same as with full timing or none
    R, Q, D, k = 1 << 16, 512, 96, 16
    ref = knnhip.generate_mixture(R, D, 128, 0.1, 81, 82)
    qry = knnhip.generate_mixture(Q, D, 128, 0.1, 81, 83)
    ix = knnhip.BlockIndex(ref, nblocks=64, iters=8)
    d0, i0 = ix.search(qry, k, probes=5, precision="bf16x3")
    for mode in ("scan", True):
        d, i = ix.search(qry, k, probes=5, precision="bf16x3", timing=mode)
        torch.cuda.synchronize()
        st = ix.stage_ms()
        assert st["scan"] > 0, st
        others = [v for s, v in st.items() if s != "scan"]
        if mode == "scan":
            assert all(v == -1 for v in others), st
        else:
            assert all(v >= 0 for v in others), st
        assert torch.equal(i, i0) and torch.equal(d, d0), mode
    ix.close()


@pytest.mark.parametrize("D", [1, 5, 13, 130])
def test_ball_probes_odd_dims_equal_plain(D):
    # the ball choice's FP64 centre distances and the bounds over odd dimension counts: ball
    # searches == plain search
    R, Q, k = 20000, 300, 8
    ref = knnhip.generate_mixture(R, D, 40, 0.1, 91 + D, 92)
    qry = knnhip.generate_mixture(Q, D, 40, 0.1, 91 + D, 93)
    pd, pi = knnhip.search_device(ref, qry, k)
    ix = knnhip.BlockIndex(ref, nblocks=32, iters=6)
    for prec in ("fp32", "bf16x3"):
        d, i = ix.search(qry, k, probes=4, precision=prec, units="balls")
        torch.cuda.synchronize()
        say(f"dim {D} {prec}: units {ix.last_units()} fallbacks {knnhip.last_fallback_count()}")
        assert torch.equal(i, pi) and torch.equal(d, pd), (D, prec)
    ix.close()
