"""GPU test of knnhip_search_device_noalloc, the entry point that allocates no device memory.

Through the C ABI it must return exactly what knnhip_search_device returns (and the oracle), on
two calls in a row, the second with a smaller input (the scratch keeps the first call's
contents); it must leave the device's free memory unchanged while doing so, and refuse a search
whose buffers do not fit its scratch with an error naming the size.
"""
import numpy as np
import pytest

import knnhip
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module", autouse=True)
def scratch_in_place():
    # the scratch array is the library's static device memory; the runtime places it on the
    # device when the entry point first takes its address (once per device), so one small call
    # comes first and the free-memory checks below see only what the calls themselves do
    ref = knnhip.generate_uniform(300, 8, seed=1)
    qry = knnhip.generate_uniform(10, 8, seed=2)
    d = torch.empty((3, 10), dtype=torch.float32, device="cuda")
    i = torch.empty((3, 10), dtype=torch.int32, device="cuda")
    knnhip.search_device_noalloc(ref, qry, 3, d, i)
    torch.cuda.synchronize()


def run_both(R, Q, D, k, seed, **kw):
    ref = knnhip.generate_uniform(R, D, seed=seed)
    qry = knnhip.generate_uniform(Q, D, seed=seed + 1)
    d0, i0 = knnhip.search_device(ref, qry, k, **kw)
    d1 = torch.full((k, Q), -1.0, dtype=torch.float32, device=qry.device)
    i1 = torch.full((k, Q), -7, dtype=torch.int32, device=qry.device)
    torch.cuda.synchronize()
    free_before = torch.cuda.mem_get_info()[0]
    knnhip.search_device_noalloc(ref, qry, k, d1, i1, **kw)
    torch.cuda.synchronize()
    free_after = torch.cuda.mem_get_info()[0]
    return ref, qry, (d0, i0), (d1, i1), free_before, free_after


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_noalloc_equals_search_device_on_two_calls(precision):
    # first call, then a smaller one: the second must not read the first call's leftovers
    for (R, Q, D, k, seed) in ((30000, 1500, 64, 16, 901), (5000, 333, 24, 8, 903)):
        ref, qry, (d0, i0), (d1, i1), fb, fa = run_both(R, Q, D, k, seed, precision=precision)
        assert torch.equal(i0, i1), (R, Q, D, k)
        assert torch.equal(d0, d1), (R, Q, D, k)
        assert fa == fb, f"free device memory changed by {fb - fa} bytes"
        gd, gi = oracle.knn_c(ref.cpu().numpy(), qry[:, :64].cpu().numpy(), k)
        assert np.array_equal(i1[:, :64].cpu().numpy(), gi)
        assert np.array_equal(bits(d1[:, :64].cpu().numpy()), bits(gd))


def test_noalloc_certificate_failures_take_the_exact_scan():
    # duplicated lattice points: ties make the certificate fail; the exact scan answers them
    rng = np.random.default_rng(5)
    base = (np.floor(rng.random((16, 3000)) * 3) / 3).astype(np.float32)
    ref = torch.from_numpy(np.concatenate([base, base], axis=1)).cuda()
    qry = torch.from_numpy((np.floor(rng.random((16, 200)) * 3) / 3).astype(np.float32)).cuda()
    k = 12
    d = torch.empty((k, 200), dtype=torch.float32, device="cuda")
    i = torch.empty((k, 200), dtype=torch.int32, device="cuda")
    knnhip.search_device_noalloc(ref, qry, k, d, i, margin=1)
    torch.cuda.synchronize()
    gd, gi = oracle.knn_c(ref.cpu().numpy(), qry.cpu().numpy(), k)
    assert np.array_equal(i.cpu().numpy(), gi)
    assert np.array_equal(bits(d.cpu().numpy()), bits(gd))


def test_noalloc_refuses_searches_beyond_its_scratch():
    R, Q, D, k = 200000, 100, 128, 8  # reference layouts alone are about 200 MB
    ref = knnhip.generate_uniform(R, D, seed=11)
    qry = knnhip.generate_uniform(Q, D, seed=12)
    d = torch.empty((k, Q), dtype=torch.float32, device="cuda")
    i = torch.empty((k, Q), dtype=torch.int32, device="cuda")
    with pytest.raises(knnhip.KnnError, match="bytes of scratch"):
        knnhip.search_device_noalloc(ref, qry, k, d, i)
