"""Python binding of libknnhip.so (the MI355X exact k-NN drop-in library).

Mirrors the reference interface (code/knncuda.h:14-21, :37-44, :61-68):
``knn_cuda_global``, ``knn_cuda_texture`` and ``knn_cublas`` take dim x n
row-major point sets and return k x query_nb distances and indexes, with the
reference's failure behaviour (an "ERROR:" line from the library and a false
result, raised here as ``KnnError``).

``search_device`` is the device-resident form used by the benchmark: torch
tensors already in GPU memory, queued on the current torch stream.

The library is loaded from this directory only; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libknnhip.so")

# Every symbol the public header declares (include/knncuda.h).
CXX_SYMBOLS = (
    "_Z15knn_cuda_globalPKfiS0_iiiPfPi",
    "_Z16knn_cuda_texturePKfiS0_iiiPfPi",
    "_Z10knn_cublasPKfiS0_iiiPfPi",
)
C_SYMBOLS = (
    "knnhip_global",
    "knnhip_texture",
    "knnhip_cublas",
    "knnhip_search_device",
    "knnhip_last_fallback_count",
    "knnhip_last_stage_ms",
    "knnhip_last_error",
    "knnhip_last_plan",
    "knnhip_generate_uniform",
    "knnhip_generate_mixture",
    "knnhip_index_build",
    "knnhip_index_search",
    "knnhip_index_free",
    "knnhip_index_info",
    "knnhip_index_stage_ms",
    "knnhip_index_layout",
    "knnhip_router_create",
    "knnhip_router_free",
    "knnhip_router_num_params",
    "knnhip_router_params",
    "knnhip_router_loss_grad",
    "knnhip_router_step",
    "knnhip_router_train",
    "knnhip_router_scores",
    "knnhip_search_batched",
    "knnhip_kth_distance",
    "knnhip_index_save",
    "knnhip_index_load",
    "knnhip_router_save",
    "knnhip_router_load",
    "knnhip_batched_last_failures",
    "knnhip_batched_autotune",
    "knnhip_batched_table",
    "knnhip_batched_set_slice",
    "knnhip_router_top1",
    "knnhip_index_bounds",
    "knnhip_index_search_routed",
    "knnhip_set_device_budget",
    "knnhip_device_budget",
    "knnhip_stream_chunks",
    "knnhip_version",
    "knnhip_index_last_work",
    "knnhip_index_last_units",
    "knnhip_search_device_noalloc",
)

VARIANTS = {"global": 0, "texture": 1, "cublas": 2}
# knnhip_options.precision: FP32 products; bfloat16 three-term products; single-pass bfloat16
# products on block-centroid residuals (block-index searches; plain searches take the three-term
# form).  Every form returns the same exact result.
PRECISIONS = {"fp32": 0, "bf16x3": 1, "bf16": 2}
# what an index search scans in its probed blocks: the balls that may hold a neighbour (chosen
# in one pass), whole blocks, or balls chosen in two passes (below the first ball's threshold)
UNITS = {"balls": 0, "blocks": 1, "balls-threshold": 2}


class KnnError(RuntimeError):
    pass


class Options(ctypes.Structure):
    """knnhip_options (include/knncuda.h)."""

    _fields_ = [
        ("variant", ctypes.c_int),
        ("margin", ctypes.c_int),
        ("slices", ctypes.c_int),
        ("force_fallback", ctypes.c_int),
        ("timing", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("block_queries", ctypes.c_int),
        ("score_all", ctypes.c_int),
        ("probe_units", ctypes.c_int),
        ("reserved", ctypes.c_int * 4),
    ]


_lib = None


def load() -> ctypes.CDLL:
    """Load libknnhip.so from this directory (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KnnError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    for name in ("knnhip_global", "knnhip_texture", "knnhip_cublas"):
        fn = getattr(lib, name)
        fn.argtypes = [fp, ctypes.c_int, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, ip]
        fn.restype = ctypes.c_int
    lib.knnhip_search_device.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.POINTER(Options), ctypes.c_void_p,
    ]
    lib.knnhip_search_device.restype = ctypes.c_int
    lib.knnhip_search_device_noalloc.argtypes = lib.knnhip_search_device.argtypes
    lib.knnhip_search_device_noalloc.restype = ctypes.c_int
    lib.knnhip_last_fallback_count.restype = ctypes.c_int
    lib.knnhip_last_stage_ms.argtypes = [ctypes.c_int]
    lib.knnhip_last_stage_ms.restype = ctypes.c_float
    lib.knnhip_last_error.restype = ctypes.c_char_p
    lib.knnhip_last_plan.restype = ctypes.c_char_p
    lib.knnhip_version.restype = ctypes.c_char_p
    lib.knnhip_generate_uniform.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_float,
        ctypes.c_void_p,
    ]
    lib.knnhip_generate_uniform.restype = ctypes.c_int
    lib.knnhip_generate_mixture.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
        ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_void_p,
    ]
    lib.knnhip_generate_mixture.restype = ctypes.c_int
    lib.knnhip_index_build.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
    ]
    lib.knnhip_index_build.restype = ctypes.c_int
    lib.knnhip_index_search.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
        ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Options),
        ctypes.c_void_p,
    ]
    lib.knnhip_index_search.restype = ctypes.c_int
    lib.knnhip_index_free.argtypes = [ctypes.c_void_p]
    lib.knnhip_index_free.restype = None
    lib.knnhip_index_info.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
        ctypes.POINTER(ctypes.c_float),
    ]
    lib.knnhip_index_info.restype = ctypes.c_int
    lib.knnhip_index_stage_ms.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    lib.knnhip_index_stage_ms.restype = ctypes.c_int
    lib.knnhip_index_last_work.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 4 + [
        ctypes.POINTER(ctypes.c_int)]
    lib.knnhip_index_last_work.restype = ctypes.c_int
    lib.knnhip_index_last_units.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int)] * 3
    lib.knnhip_index_last_units.restype = ctypes.c_int
    lib.knnhip_index_layout.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]
    lib.knnhip_index_layout.restype = ctypes.c_int
    vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float
    lib.knnhip_router_create.argtypes = [i32, i32, i32, ctypes.c_ulonglong, ctypes.POINTER(vp)]
    lib.knnhip_router_create.restype = i32
    lib.knnhip_router_free.argtypes = [vp]
    lib.knnhip_router_free.restype = None
    lib.knnhip_router_num_params.argtypes = [vp]
    lib.knnhip_router_num_params.restype = i64
    lib.knnhip_router_params.argtypes = [vp, vp, i64, i32]
    lib.knnhip_router_params.restype = i32
    lib.knnhip_router_loss_grad.argtypes = [vp, vp, i32, i64, vp, i64, i32, vp,
                                            ctypes.POINTER(f32), vp]
    lib.knnhip_router_loss_grad.restype = i32
    lib.knnhip_router_step.argtypes = [vp, vp, i32, i64, vp, i64, i32, f32, vp]
    lib.knnhip_router_step.restype = i32
    lib.knnhip_router_train.argtypes = [vp, vp, vp, i32, i64, i32, i32, i32, f32,
                                        ctypes.c_ulonglong, ctypes.POINTER(f32), vp]
    lib.knnhip_router_train.restype = i32
    lib.knnhip_router_scores.argtypes = [vp, vp, i32, i64, vp, vp]
    lib.knnhip_router_scores.restype = i32
    lib.knnhip_index_bounds.argtypes = [vp, vp, i32, i64, i32, vp, vp]
    lib.knnhip_index_bounds.restype = i32
    lib.knnhip_router_top1.argtypes = [vp, vp, i32, i64, vp, vp]
    lib.knnhip_router_top1.restype = i32
    lib.knnhip_index_search_routed.argtypes = [vp, vp, vp, i32, i64, i32, i32, vp, vp,
                                               ctypes.POINTER(Options), vp]
    lib.knnhip_index_search_routed.restype = i32
    lib.knnhip_kth_distance.argtypes = [vp, i32, vp, i32, i32, i32, vp]
    lib.knnhip_kth_distance.restype = i32
    lib.knnhip_index_save.argtypes = [vp, ctypes.c_char_p]
    lib.knnhip_index_save.restype = i32
    lib.knnhip_index_load.argtypes = [ctypes.c_char_p, vp, i32, i64, i32, ctypes.POINTER(vp), vp]
    lib.knnhip_index_load.restype = i32
    lib.knnhip_router_save.argtypes = [vp, ctypes.c_char_p]
    lib.knnhip_router_save.restype = i32
    lib.knnhip_router_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    lib.knnhip_router_load.restype = i32
    lib.knnhip_search_batched.argtypes = [vp, i32, ctypes.POINTER(Options), vp]
    lib.knnhip_search_batched.restype = i32
    lib.knnhip_batched_last_failures.argtypes = []
    lib.knnhip_batched_last_failures.restype = i64
    lib.knnhip_batched_autotune.argtypes = [vp, i32, i32, vp]
    lib.knnhip_batched_autotune.restype = i32
    lib.knnhip_batched_table.argtypes = [ctypes.c_char_p, i32]
    lib.knnhip_batched_table.restype = i32
    lib.knnhip_batched_set_slice.argtypes = [i32, i32, i32]
    lib.knnhip_batched_set_slice.restype = i32
    lib.knnhip_set_device_budget.argtypes = [i64]
    lib.knnhip_set_device_budget.restype = i32
    lib.knnhip_device_budget.argtypes = []
    lib.knnhip_device_budget.restype = i64
    lib.knnhip_stream_chunks.argtypes = [i32, i32, i32, i32, i64, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    lib.knnhip_stream_chunks.restype = i32
    _lib = lib
    return lib


def last_error() -> str:
    return load().knnhip_last_error().decode()


def _host_call(name: str, ref: np.ndarray, query: np.ndarray, k: int) -> Tuple[np.ndarray, np.ndarray]:
    lib = load()
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    query = np.ascontiguousarray(query, dtype=np.float32)
    if ref.ndim != 2 or query.ndim != 2 or ref.shape[0] != query.shape[0]:
        raise KnnError("ref and query must be dim x n arrays with the same dim")
    dim, ref_nb = ref.shape
    query_nb = query.shape[1]
    dist = np.empty((k, query_nb), dtype=np.float32)
    idx = np.empty((k, query_nb), dtype=np.int32)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    ok = getattr(lib, name)(
        ref.ctypes.data_as(fp), ref_nb, query.ctypes.data_as(fp), query_nb, dim, k,
        dist.ctypes.data_as(fp), idx.ctypes.data_as(ip),
    )
    if not ok:
        raise KnnError(last_error())
    return dist, idx


def knn_cuda_global(ref: np.ndarray, query: np.ndarray, k: int):
    """Drop-in for code/knncuda.h:14-21 on host arrays; returns (dist, index), each k x query_nb."""
    return _host_call("knnhip_global", ref, query, k)


def knn_cuda_texture(ref: np.ndarray, query: np.ndarray, k: int):
    """Drop-in for code/knncuda.h:37-44."""
    return _host_call("knnhip_texture", ref, query, k)


def knn_cublas(ref: np.ndarray, query: np.ndarray, k: int):
    """Drop-in for code/knncuda.h:61-68."""
    return _host_call("knnhip_cublas", ref, query, k)


def kth_distance(ref, query, k: int):
    """Distance to each query's k-th nearest neighbour (host numpy arrays, dim x n), as the
    reference's legacy programs without indexes return it: row k-1 of knn_cuda_global."""
    import numpy as np

    ref = np.ascontiguousarray(ref, dtype=np.float32)
    query = np.ascontiguousarray(query, dtype=np.float32)
    dim, ref_nb = ref.shape
    out = np.empty(query.shape[1], dtype=np.float32)
    if not load().knnhip_kth_distance(ref.ctypes.data, ref_nb, query.ctypes.data, query.shape[1],
                                      dim, k, out.ctypes.data):
        raise KnnError(last_error())
    return out


def _check_points(name: str, t, dim: Optional[int] = None):
    """A dim x n float32 point set with contiguous rows (shape and dtype first, device last, so
    that shape errors are reported the same way with or without a GPU)."""
    import torch

    if not isinstance(t, torch.Tensor) or t.ndim != 2:
        raise KnnError(f"{name} must be a 2-D (dim x n) tensor")
    if t.dtype != torch.float32:
        raise KnnError(f"{name} must be float32")
    if dim is not None and t.shape[0] != dim:
        raise KnnError(f"{name} has dim {t.shape[0]}, expected {dim} (ref and query must share dim)")
    if t.shape[1] < 1 or t.shape[0] < 1:
        raise KnnError(f"{name} must hold at least one point of at least one dimension")
    if t.stride(1) != 1 or t.stride(0) < t.shape[1]:
        raise KnnError(f"{name} must have contiguous rows (row stride >= n)")


def _check_out(name: str, t, shape, dtype, device):
    """A caller-supplied output: exactly `shape`, contiguous, of `dtype`, on `device`."""
    import torch

    if not isinstance(t, torch.Tensor) or tuple(t.shape) != tuple(shape):
        raise KnnError(f"{name} must be a tensor of shape {tuple(shape)}")
    if t.dtype != dtype:
        raise KnnError(f"{name} must be {dtype}")
    if not t.is_contiguous():
        raise KnnError(f"{name} must be contiguous")
    if t.device != device:
        raise KnnError(f"{name} must be on {device}")


def _check_device(*ts):
    for t in ts:
        if not t.is_cuda:
            raise KnnError("tensors must be device (GPU) tensors")
    if any(t.device != ts[0].device for t in ts):
        raise KnnError("tensors must be on one device")


def search_device(ref, query, k: int, dist=None, index=None, *, variant: str = "global",
                  margin: int = 0, slices: int = 0, force_fallback: bool = False,
                  precision: str = "fp32", block_queries: int = 0,
                  timing: bool = False, stream: Optional[int] = None):
    """Exact k-NN on device-resident torch tensors (dim x n, float32, row stride = n or more).

    Returns (dist, index) torch tensors of shape (k, query_nb).
    """
    import torch

    _check_points("ref", ref)
    _check_points("query", query, dim=ref.shape[0])
    dim, ref_nb = ref.shape
    query_nb = query.shape[1]
    if not 1 <= k <= ref_nb:
        raise KnnError(f"k must satisfy 1 <= k <= ref_nb (k={k} ref_nb={ref_nb})")
    if dist is not None:
        _check_out("dist", dist, (k, query_nb), torch.float32, query.device)
    if index is not None:
        _check_out("index", index, (k, query_nb), torch.int32, query.device)
    _check_device(ref, query)
    lib = load()
    if dist is None:
        dist = torch.empty((k, query_nb), dtype=torch.float32, device=query.device)
    if index is None:
        index = torch.empty((k, query_nb), dtype=torch.int32, device=query.device)
    opt = Options()
    opt.variant = VARIANTS[variant]
    opt.margin = margin
    opt.slices = slices
    opt.force_fallback = 1 if force_fallback else 0
    opt.timing = 1 if timing else 0
    if precision not in PRECISIONS:
        raise KnnError(f"precision must be one of {sorted(PRECISIONS)}")
    opt.precision = PRECISIONS[precision]
    opt.block_queries = block_queries
    if stream is None:
        stream = torch.cuda.current_stream(query.device).cuda_stream
    ok = lib.knnhip_search_device(
        ref.data_ptr(), ref_nb, ref.stride(0), query.data_ptr(), query_nb, query.stride(0), dim, k,
        dist.data_ptr(), index.data_ptr(), ctypes.byref(opt), ctypes.c_void_p(stream),
    )
    if not ok:
        raise KnnError(last_error())
    return dist, index


def search_device_noalloc(ref, query, k: int, dist, index, *, margin: int = 0, precision: str = "fp32",
                          stream: Optional[int] = None):
    """``search_device`` through the entry point that allocates no device memory (its buffers live
    in the library's static scratch array): the caller passes the (k, query_nb) outputs."""
    import torch

    _check_points("ref", ref)
    _check_points("query", query, dim=ref.shape[0])
    dim, ref_nb = ref.shape
    query_nb = query.shape[1]
    if not 1 <= k <= ref_nb:
        raise KnnError(f"k must satisfy 1 <= k <= ref_nb (k={k} ref_nb={ref_nb})")
    _check_out("dist", dist, (k, query_nb), torch.float32, query.device)
    _check_out("index", index, (k, query_nb), torch.int32, query.device)
    _check_device(ref, query)
    if precision not in PRECISIONS:
        raise KnnError(f"precision must be one of {sorted(PRECISIONS)}")
    opt = Options()
    opt.margin = margin
    opt.precision = PRECISIONS[precision]
    if stream is None:
        stream = torch.cuda.current_stream(query.device).cuda_stream
    ok = load().knnhip_search_device_noalloc(
        ref.data_ptr(), ref_nb, ref.stride(0), query.data_ptr(), query_nb, query.stride(0), dim, k,
        dist.data_ptr(), index.data_ptr(), ctypes.byref(opt), ctypes.c_void_p(stream),
    )
    if not ok:
        raise KnnError(last_error())
    return dist, index


def last_plan() -> dict:
    """Launch plan of the last search (slices, heap size, blocks per CU, ...)."""
    import json

    text = load().knnhip_last_plan().decode()
    return json.loads(text) if text else {}


def last_fallback_count() -> int:
    return load().knnhip_last_fallback_count()


def last_stage_ms(stage: int) -> float:
    return float(load().knnhip_last_stage_ms(stage))


def generate_uniform(n: int, dim: int, seed: int, scale: float = 1.0, device: str = "cuda"):
    """Seeded uniform point set generated on the device (dim x n float32 tensor)."""
    import torch

    out = torch.empty((dim, n), dtype=torch.float32, device=device)
    stream = torch.cuda.current_stream(out.device).cuda_stream
    if not load().knnhip_generate_uniform(out.data_ptr(), n, dim, seed, scale, ctypes.c_void_p(stream)):
        raise KnnError(last_error())
    return out


def generate_mixture(n: int, dim: int, ncomp: int, sigma: float, centre_seed: int, point_seed: int,
                     device: str = "cuda"):
    """Seeded Gaussian-mixture point set generated on the device (dim x n float32 tensor)."""
    import torch

    out = torch.empty((dim, n), dtype=torch.float32, device=device)
    stream = torch.cuda.current_stream(out.device).cuda_stream
    if not load().knnhip_generate_mixture(out.data_ptr(), n, dim, ncomp, sigma, centre_seed,
                                          point_seed, ctypes.c_void_p(stream)):
        raise KnnError(last_error())
    return out


class BlockIndex:
    """k-means block index over a device-resident reference set (router form).

    Searches return exactly what ``search_device`` returns on the same data: blocks are skipped
    only where the triangle inequality proves they hold no neighbour.  The index reads ``ref``
    during exact fallbacks, so ``ref`` is kept alive here and must not be modified.
    """

    def __init__(self, ref, nblocks: int, iters: int = 8, stream: Optional[int] = None):
        import torch

        lib = load()
        if ref.dtype != torch.float32 or not ref.is_cuda or ref.stride(1) != 1:
            raise KnnError("ref must be a float32 device tensor with contiguous rows")
        self.ref = ref
        self.dim, self.ref_nb = ref.shape
        self.nblocks = nblocks
        handle = ctypes.c_void_p()
        if stream is None:
            stream = torch.cuda.current_stream(ref.device).cuda_stream
        if not lib.knnhip_index_build(ref.data_ptr(), self.ref_nb, ref.stride(0), self.dim, nblocks,
                                      iters, ctypes.byref(handle), ctypes.c_void_p(stream)):
            raise KnnError(last_error())
        self._h = handle

    def info(self) -> dict:
        nb = ctypes.c_int()
        npos = ctypes.c_int()
        ms = ctypes.c_float()
        load().knnhip_index_info(self._h, ctypes.byref(nb), ctypes.byref(npos), ctypes.byref(ms))
        return {"nblocks": nb.value, "positions": npos.value, "build_ms": ms.value}

    STAGES = ("norms", "bounds", "pick", "tasks", "scan", "select", "fallback")

    def last_work(self) -> dict:
        """Scan work of the last search (executed and held (reference, query) pairs, tile bytes
        its tasks read, bytes of the distinct tiles they cover, products used)."""
        ex, pr, tb, ub = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        prec = ctypes.c_int()
        load().knnhip_index_last_work(self._h, ctypes.byref(ex), ctypes.byref(pr), ctypes.byref(tb),
                                      ctypes.byref(ub), ctypes.byref(prec))
        return {"exec_pairs": ex.value, "pairs": pr.value, "tile_bytes": tb.value, "unit_bytes": ub.value,
                "precision": {0: "fp32", 1: "bf16x3", 2: "bf16-residual"}.get(prec.value, "?")}

    def last_units(self) -> dict:
        """Scan units of the last search: balls or whole blocks, heap rows per query, and the
        index's ball count (0 when it does not know its balls' position ranges)."""
        b, rows, nballs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        load().knnhip_index_last_units(self._h, ctypes.byref(b), ctypes.byref(rows), ctypes.byref(nballs))
        return {"units": "balls" if b.value else "blocks", "rows_per_query": rows.value,
                "balls": nballs.value}

    def stage_ms(self) -> dict:
        buf = (ctypes.c_float * 7)()
        load().knnhip_index_stage_ms(self._h, buf, 7)
        return dict(zip(self.STAGES, list(buf)))

    def save(self, path: str):
        if not load().knnhip_index_save(self._h, path.encode()):
            raise KnnError(last_error())

    @classmethod
    def load(cls, path: str, ref, stream: Optional[int] = None) -> "BlockIndex":
        """Index saved by ``save``, rebuilt over the same (device) reference set."""
        import torch

        if ref.dtype != torch.float32 or not ref.is_cuda or ref.stride(1) != 1:
            raise KnnError("ref must be a float32 device tensor with contiguous rows")
        h = ctypes.c_void_p()
        st = stream if stream is not None else torch.cuda.current_stream(ref.device).cuda_stream
        if not load().knnhip_index_load(path.encode(), ref.data_ptr(), ref.shape[1], ref.stride(0),
                                        ref.shape[0], ctypes.byref(h), st):
            raise KnnError(last_error())
        self = cls.__new__(cls)
        self.ref = ref
        self.dim, self.ref_nb = ref.shape
        self._h = h
        self.nblocks = self.info()["nblocks"]
        return self

    def layout(self, stream: Optional[int] = None):
        """(block_start, block_count) as host int32 arrays and the block of every reference
        point as a device int32 tensor of ref_nb entries."""
        import numpy as np
        import torch

        nb = self.info()["nblocks"]
        start = np.zeros(nb, dtype=np.int32)
        count = np.zeros(nb, dtype=np.int32)
        ref_block = torch.empty(self.ref.shape[1], dtype=torch.int32, device=self.ref.device)
        st = stream if stream is not None else torch.cuda.current_stream(self.ref.device).cuda_stream
        ok = load().knnhip_index_layout(self._h, start.ctypes.data, count.ctypes.data,
                                        ref_block.data_ptr(), st)
        if not ok:
            raise KnnError(last_error())
        torch.cuda.synchronize(self.ref.device)
        return start, count, ref_block

    def bounds(self, query, precision: str = "fp32", per_ball: bool = False, stream: Optional[int] = None):
        """The block lower bounds searches rank and prune by: (nblocks, query_nb) float32.
        ``per_ball``: the per-ball fold form of the same bounds (searches use the slot form)."""
        import torch

        _check_points("query", query, dim=self.dim)
        _check_device(query, self.ref)
        out = torch.empty((self.nblocks, query.shape[1]), dtype=torch.float32, device=query.device)
        flags = (1 if precision == "bf16x3" else 0) | (2 if per_ball else 0)
        ok = load().knnhip_index_bounds(self._h, query.data_ptr(), query.shape[1], query.stride(0),
                                        flags, out.data_ptr(), _stream_of(query, stream))
        if not ok:
            raise KnnError(last_error())
        return out

    def search(self, query, k: int, probes: int = 4, scores=None, dist=None, index=None, *,
               margin: int = 0, force_fallback: bool = False, timing=False,
               precision: str = "fp32", score_all: bool = False, router=None,
               units: str = "balls", stream: Optional[int] = None):
        """Exact k-NN through the index.  ``scores`` (nblocks x query_nb) or ``router`` (a Router
        of this index's shape, its first probe taken from the fused inference kernel) choose each
        query's first probe; the lower bounds rank the rest.  ``units``: what the scan visits in
        the probed blocks (``UNITS``; indexes that do not know their balls scan whole blocks).
        ``timing``: True records every stage's device time (``stage_ms``); "scan" the scan
        stage's alone (two events, the other stages read -1)."""
        import torch

        lib = load()
        _check_points("query", query, dim=self.dim)
        dim, query_nb = query.shape
        if not 1 <= k <= self.ref_nb:
            raise KnnError(f"k must satisfy 1 <= k <= ref_nb (k={k} ref_nb={self.ref_nb})")
        if dist is not None:
            _check_out("dist", dist, (k, query_nb), torch.float32, query.device)
        if index is not None:
            _check_out("index", index, (k, query_nb), torch.int32, query.device)
        _check_device(query, self.ref)
        if dist is None:
            dist = torch.empty((k, query_nb), dtype=torch.float32, device=query.device)
        if index is None:
            index = torch.empty((k, query_nb), dtype=torch.int32, device=query.device)
        sp = None
        if scores is not None:
            if scores.shape != (self.nblocks, query_nb) or scores.dtype != torch.float32:
                raise KnnError("scores must be float32 (nblocks, query_nb): one column per query")
            scores = scores.contiguous()
            sp = scores.data_ptr()
        if precision not in PRECISIONS:
            raise KnnError(f"precision must be one of {sorted(PRECISIONS)}")
        opt = Options()
        opt.margin = margin
        opt.force_fallback = 1 if force_fallback else 0
        opt.timing = 2 if timing == "scan" else 1 if timing else 0
        opt.precision = PRECISIONS[precision]
        opt.score_all = 1 if score_all else 0
        if units not in UNITS:
            raise KnnError(f"units must be one of {sorted(UNITS)}")
        opt.probe_units = UNITS[units]
        if stream is None:
            stream = torch.cuda.current_stream(query.device).cuda_stream
        if router is not None:
            if scores is not None:
                raise KnnError("pass scores or a router, not both")
            ok = lib.knnhip_index_search_routed(self._h, router._h, query.data_ptr(), query_nb,
                                                query.stride(0), k, probes, dist.data_ptr(),
                                                index.data_ptr(), ctypes.byref(opt), ctypes.c_void_p(stream))
        else:
            ok = lib.knnhip_index_search(self._h, query.data_ptr(), query_nb, query.stride(0), k, probes,
                                         sp, dist.data_ptr(), index.data_ptr(), ctypes.byref(opt),
                                         ctypes.c_void_p(stream))
        if not ok:
            raise KnnError(last_error())
        return dist, index

    def close(self):
        if getattr(self, "_h", None):
            load().knnhip_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_of(t, stream):
    import torch

    return stream if stream is not None else torch.cuda.current_stream(t.device).cuda_stream


class Router:
    """Learned block router: MLP dim -> hidden -> hidden -> nblocks, trained on the GPU.

    Inputs are dim-major device tensors (dim x n), like the point sets; targets are k x n int32
    block ids (the block of each exact neighbour).  ``scores`` returns an (nblocks, n) tensor for
    ``BlockIndex.search(scores=...)``.
    """

    def __init__(self, dim: int, hidden: int, nblocks: int, seed: int = 0x6B4E4E07):
        h = ctypes.c_void_p()
        if not load().knnhip_router_create(dim, hidden, nblocks, seed, ctypes.byref(h)):
            raise KnnError(last_error())
        self._h = h
        self.dim, self.hidden, self.nblocks = dim, hidden, nblocks

    def save(self, path: str):
        if not load().knnhip_router_save(self._h, path.encode()):
            raise KnnError(last_error())

    @classmethod
    def load(cls, path: str) -> "Router":
        h = ctypes.c_void_p()
        if not load().knnhip_router_load(path.encode(), ctypes.byref(h)):
            raise KnnError(last_error())
        self = cls.__new__(cls)
        self._h = h
        n = self.num_params
        # dims from the file header (dim, hidden, nblocks follow the magic and version)
        with open(path, "rb") as f:
            head = f.read(24)
        import struct

        self.dim, self.hidden, self.nblocks = struct.unpack_from("<3i", head, 12)
        assert n > 0
        return self

    @property
    def num_params(self) -> int:
        return int(load().knnhip_router_num_params(self._h))

    def get_params(self):
        import numpy as np

        out = np.empty(self.num_params, dtype=np.float32)
        if not load().knnhip_router_params(self._h, out.ctypes.data, out.size, 0):
            raise KnnError(last_error())
        return out

    def set_params(self, params):
        import numpy as np

        p = np.ascontiguousarray(params, dtype=np.float32)
        if not load().knnhip_router_params(self._h, p.ctypes.data, p.size, 1):
            raise KnnError(last_error())

    def _check(self, x, tgt):
        import torch

        _check_points("x", x, dim=self.dim)
        if not isinstance(tgt, torch.Tensor) or tgt.ndim != 2 or tgt.dtype != torch.int32 or \
                tgt.stride(1) != 1 or tgt.shape[1] != x.shape[1] or tgt.shape[0] < 1:
            raise KnnError("tgt must be an int32 (k x n) tensor with contiguous rows, n as in x")
        _check_device(x, tgt)

    def loss_grad(self, x, tgt, stream: Optional[int] = None):
        import torch

        self._check(x, tgt)
        grad = torch.empty(self.num_params, dtype=torch.float32, device=x.device)
        loss = ctypes.c_float()
        ok = load().knnhip_router_loss_grad(self._h, x.data_ptr(), x.shape[1], x.stride(0),
                                            tgt.data_ptr(), tgt.stride(0), tgt.shape[0],
                                            grad.data_ptr(), ctypes.byref(loss),
                                            _stream_of(x, stream))
        if not ok:
            raise KnnError(last_error())
        return grad, loss.value

    def step(self, x, tgt, lr: float, stream: Optional[int] = None):
        self._check(x, tgt)
        ok = load().knnhip_router_step(self._h, x.data_ptr(), x.shape[1], x.stride(0),
                                       tgt.data_ptr(), tgt.stride(0), tgt.shape[0], lr,
                                       _stream_of(x, stream))
        if not ok:
            raise KnnError(last_error())

    def train(self, index: "BlockIndex", train_query, k: int, epochs: int = 10, batch: int = 1024,
              lr: float = 1e-3, seed: int = 0x6B4E4E17, stream: Optional[int] = None) -> float:
        _check_points("train_query", train_query, dim=self.dim)
        _check_device(train_query)
        loss = ctypes.c_float()
        ok = load().knnhip_router_train(self._h, index._h, train_query.data_ptr(),
                                        train_query.shape[1], train_query.stride(0), k, epochs,
                                        batch, lr, seed, ctypes.byref(loss),
                                        _stream_of(train_query, stream))
        if not ok:
            raise KnnError(last_error())
        return loss.value

    def scores(self, query, out=None, stream: Optional[int] = None):
        import torch

        _check_points("query", query, dim=self.dim)
        if out is not None:
            _check_out("out", out, (self.nblocks, query.shape[1]), torch.float32, query.device)
        _check_device(query)
        if out is None:
            out = torch.empty((self.nblocks, query.shape[1]), dtype=torch.float32,
                              device=query.device)
        ok = load().knnhip_router_scores(self._h, query.data_ptr(), query.shape[1],
                                         query.stride(0), out.data_ptr(), _stream_of(query, stream))
        if not ok:
            raise KnnError(last_error())
        return out

    def top1(self, query, out=None, stream: Optional[int] = None):
        """Each query's highest-scoring block (int32, query_nb) from the fused inference kernel."""
        import torch

        _check_points("query", query, dim=self.dim)
        if out is not None:
            _check_out("out", out, (query.shape[1],), torch.int32, query.device)
        _check_device(query)
        if out is None:
            out = torch.empty(query.shape[1], dtype=torch.int32, device=query.device)
        ok = load().knnhip_router_top1(self._h, query.data_ptr(), query.shape[1], query.stride(0),
                                       out.data_ptr(), _stream_of(query, stream))
        if not ok:
            raise KnnError(last_error())
        return out

    def close(self):
        if getattr(self, "_h", None):
            load().knnhip_router_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Problem(ctypes.Structure):
    """One search of a batch (device pointers; the layout of ``knnhip_problem``)."""

    _fields_ = [
        ("ref", ctypes.c_void_p), ("ref_nb", ctypes.c_int), ("ref_ld", ctypes.c_longlong),
        ("query", ctypes.c_void_p), ("query_nb", ctypes.c_int), ("query_ld", ctypes.c_longlong),
        ("dim", ctypes.c_int), ("k", ctypes.c_int),
        ("knn_dist", ctypes.c_void_p), ("knn_index", ctypes.c_void_p),
    ]


def _problems(items):
    """items: (ref, query, k) device tensors (dim x n).  Returns (ctypes array, outputs)."""
    import torch

    arr = (Problem * len(items))()
    outs = []
    for check_device in (False, True):
        for j, (ref, query, k) in enumerate(items):
            try:
                if check_device:
                    _check_device(ref, query)
                    continue
                _check_points("ref", ref)
                _check_points("query", query, dim=ref.shape[0])
                if not 1 <= k <= ref.shape[1]:
                    raise KnnError(f"k must satisfy 1 <= k <= ref_nb (k={k} ref_nb={ref.shape[1]})")
            except KnnError as e:
                raise KnnError(f"problem {j}: {e}") from None
    for j, (ref, query, k) in enumerate(items):
        d = torch.empty((k, query.shape[1]), dtype=torch.float32, device=query.device)
        i = torch.empty((k, query.shape[1]), dtype=torch.int32, device=query.device)
        arr[j] = Problem(ref.data_ptr(), ref.shape[1], ref.stride(0), query.data_ptr(),
                         query.shape[1], query.stride(0), ref.shape[0], k, d.data_ptr(),
                         i.data_ptr())
        outs.append((d, i))
    return arr, outs


def search_batched(items, *, margin: int = 0, stream: Optional[int] = None):
    """Exact k-NN of many independent problems in one call.

    items: sequence of (ref, query, k) with dim x n float32 device tensors.  Returns a list of
    (dist, index) pairs, each exactly what ``search_device`` returns for that problem.
    """
    import torch

    arr, outs = _problems(items)
    opt = Options()
    opt.margin = margin
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    if not load().knnhip_search_batched(arr, len(items), ctypes.byref(opt), st):
        raise KnnError(last_error())
    return outs


def batched_autotune(items, reps: int = 1, stream: Optional[int] = None) -> list:
    """Time the tile shapes of every (dim, k) class of these problems; keep the fastest."""
    import json
    import torch

    arr, _ = _problems(items)
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    if not load().knnhip_batched_autotune(arr, len(items), reps, st):
        raise KnnError(last_error())
    return batched_table()


def batched_table() -> list:
    import json

    n = load().knnhip_batched_table(None, 0)
    buf = ctypes.create_string_buffer(n)
    load().knnhip_batched_table(buf, n)
    return json.loads(buf.value.decode())


def batched_set_slice(dim: int, k: int, slice_len: int):
    if not load().knnhip_batched_set_slice(dim, k, slice_len):
        raise KnnError(last_error())


def batched_last_failures() -> int:
    return int(load().knnhip_batched_last_failures())


def set_device_budget(nbytes: int):
    """Device memory the host-pointer entry points may use (0: 90 % of free memory).

    Larger searches stream the reference set through it in chunks, with the same exact result."""
    if not load().knnhip_set_device_budget(int(nbytes)):
        raise KnnError(last_error())


def device_budget() -> int:
    return int(load().knnhip_device_budget())


def stream_chunks(ref_nb: int, query_nb: int, dim: int, k: int, device_bytes: int) -> tuple:
    """(reference chunk, query group) sizes of a streamed search within device_bytes."""
    rc, qc = ctypes.c_int32(0), ctypes.c_int32(0)
    if not load().knnhip_stream_chunks(ref_nb, query_nb, dim, k, int(device_bytes), ctypes.byref(rc),
                                       ctypes.byref(qc)):
        raise KnnError(last_error())
    return rc.value, qc.value
