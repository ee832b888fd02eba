"""CPU tests of the Python binding's argument checks: every malformed tensor is rejected before
any device pointer reaches the library (a narrower query, or outputs of the wrong shape, would
otherwise make the kernels read or write past an allocation)."""
import pytest
import torch

import knnhip


def pts(dim, n, dtype=torch.float32):
    return torch.zeros((dim, n), dtype=dtype)


@pytest.mark.parametrize("case,match", [
    ("query_narrower", "expected 8"),
    ("query_wider", "expected 8"),
    ("one_d", "2-D"),
    ("float64", "float32"),
    ("k_zero", "k must"),
    ("k_too_large", "k must"),
    ("rows_not_contiguous", "contiguous rows"),
    ("dist_shape", "dist must be a tensor of shape"),
    ("dist_dtype", "dist must be torch.float32"),
    ("index_dtype", "index must be torch.int32"),
    ("index_transposed", "index must be"),
    ("cpu_tensors", "device"),
])
def test_search_device_rejects(case, match):
    ref, qry, k = pts(8, 50), pts(8, 7), 4
    kw = {}
    if case == "query_narrower":
        qry = pts(7, 7)
    elif case == "query_wider":
        qry = pts(9, 7)
    elif case == "one_d":
        qry = torch.zeros(8)
    elif case == "float64":
        ref = pts(8, 50, torch.float64)
    elif case == "k_zero":
        k = 0
    elif case == "k_too_large":
        k = 51
    elif case == "rows_not_contiguous":
        ref = pts(50, 8).t()
    elif case == "dist_shape":
        kw["dist"] = torch.zeros((k, 6), dtype=torch.float32)
    elif case == "dist_dtype":
        kw["dist"] = torch.zeros((k, 7), dtype=torch.float64)
    elif case == "index_dtype":
        kw["index"] = torch.zeros((k, 7), dtype=torch.int64)
    elif case == "index_transposed":
        kw["index"] = torch.zeros((7, k), dtype=torch.int32).t()
    with pytest.raises(knnhip.KnnError, match=match):
        knnhip.search_device(ref, qry, k, **kw)


def test_batched_problem_with_mismatched_dims_is_rejected():
    with pytest.raises(knnhip.KnnError, match="problem 1: query has dim 5, expected 6"):
        knnhip._problems([(pts(4, 10), pts(4, 3), 2), (pts(6, 10), pts(5, 3), 2)])


def test_output_checks_accept_exact_tensors():
    # the checks themselves accept what search_device allocates
    knnhip._check_out("dist", torch.zeros((3, 5)), (3, 5), torch.float32, torch.device("cpu"))
    knnhip._check_points("ref", pts(4, 9))
    knnhip._check_points("query", pts(4, 2), dim=4)
