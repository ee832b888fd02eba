"""ctypes wrapper of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by the smoke check of __graft_entry__.py and by the
cpu_baseline leg of bench.py — as the checker, never as the product.  The
library under knn-cuda_amd/ never imports this module.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(_HERE, "_ref", "libknn_ref.so")

_lib = None
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
    lib = ctypes.CDLL(LIB_PATH)
    sig = [_fp, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _ip]
    for name in ("oracle_knn_c", "oracle_knn_c_omp"):
        getattr(lib, name).argtypes = sig
        getattr(lib, name).restype = ctypes.c_int
    lib.oracle_knn_c_queries.argtypes = sig + [ctypes.c_int, ctypes.c_int]
    lib.oracle_knn_c_queries.restype = ctypes.c_int
    lib.oracle_knn_f64.argtypes = [_fp, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, _dp, _ip]
    lib.oracle_knn_f64.restype = ctypes.c_int
    lib.oracle_canonical_sq.argtypes = [_fp, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int]
    lib.oracle_canonical_sq.restype = ctypes.c_float
    lib.oracle_check.argtypes = [_fp, _ip, _fp, _ip, ctypes.c_long, ctypes.c_float, _dp, _dp]
    lib.oracle_check.restype = None
    lib.oracle_gen_uniform.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.c_float]
    lib.oracle_gen_uniform.restype = None
    lib.oracle_gen_mixture.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_float, ctypes.c_uint64, ctypes.c_uint64]
    lib.oracle_gen_mixture.restype = None
    lib.oracle_rng_u64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
    lib.oracle_rng_u64.restype = ctypes.c_uint64
    _lib = lib
    return lib


def _prep(ref, query):
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    query = np.ascontiguousarray(query, dtype=np.float32)
    assert ref.ndim == 2 and query.ndim == 2 and ref.shape[0] == query.shape[0]
    return ref, query


def knn_c(ref, query, k: int, threads: bool = True) -> Tuple[np.ndarray, np.ndarray]:
    """Restated reference ground truth (code/test.cpp:129-175): (dist, index), each k x query_nb."""
    ref, query = _prep(ref, query)
    dim, rn = ref.shape
    qn = query.shape[1]
    dist = np.zeros((k, qn), dtype=np.float32)
    idx = np.zeros((k, qn), dtype=np.int32)
    fn = load().oracle_knn_c_omp if threads else load().oracle_knn_c
    if not fn(ref.ctypes.data_as(_fp), rn, query.ctypes.data_as(_fp), qn, dim, k,
              dist.ctypes.data_as(_fp), idx.ctypes.data_as(_ip)):
        raise ValueError("oracle rejected the arguments")
    return dist, idx


def knn_c_queries(ref, query, k: int, q_begin: int, q_end: int):
    """Single-threaded ground truth for queries [q_begin, q_end) only (CPU baseline sample)."""
    ref, query = _prep(ref, query)
    dim, rn = ref.shape
    qn = query.shape[1]
    dist = np.zeros((k, qn), dtype=np.float32)
    idx = np.zeros((k, qn), dtype=np.int32)
    if not load().oracle_knn_c_queries(ref.ctypes.data_as(_fp), rn, query.ctypes.data_as(_fp), qn,
                                       dim, k, dist.ctypes.data_as(_fp), idx.ctypes.data_as(_ip),
                                       q_begin, q_end):
        raise ValueError("oracle rejected the arguments")
    return dist, idx


def knn_f64(ref, query, k: int):
    """FP64 brute force: (dist float64, index), each k x query_nb."""
    ref, query = _prep(ref, query)
    dim, rn = ref.shape
    qn = query.shape[1]
    dist = np.zeros((k, qn), dtype=np.float64)
    idx = np.zeros((k, qn), dtype=np.int32)
    if not load().oracle_knn_f64(ref.ctypes.data_as(_fp), rn, query.ctypes.data_as(_fp), qn, dim,
                                 k, dist.ctypes.data_as(_dp), idx.ctypes.data_as(_ip)):
        raise ValueError("oracle rejected the arguments")
    return dist, idx


def canonical_sq(ref, query, r: int, q: int) -> np.float32:
    ref, query = _prep(ref, query)
    return np.float32(load().oracle_canonical_sq(ref.ctypes.data_as(_fp), ref.shape[1],
                                                 query.ctypes.data_as(_fp), query.shape[1],
                                                 ref.shape[0], r, q))


def check(test_dist, test_idx, gt_dist, gt_idx, precision: float = 1e-3):
    """The reference acceptance check (code/test.cpp:256-272): (dist_frac, index_frac)."""
    td = np.ascontiguousarray(test_dist, dtype=np.float32).ravel()
    ti = np.ascontiguousarray(test_idx, dtype=np.int32).ravel()
    gd = np.ascontiguousarray(gt_dist, dtype=np.float32).ravel()
    gi = np.ascontiguousarray(gt_idx, dtype=np.int32).ravel()
    a = ctypes.c_double()
    b = ctypes.c_double()
    load().oracle_check(td.ctypes.data_as(_fp), ti.ctypes.data_as(_ip), gd.ctypes.data_as(_fp),
                        gi.ctypes.data_as(_ip), td.size, precision, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def gen_uniform(n: int, dim: int, seed: int, scale: float = 1.0) -> np.ndarray:
    out = np.empty((dim, n), dtype=np.float32)
    load().oracle_gen_uniform(out.ctypes.data_as(_fp), n, dim, seed, scale)
    return out


def gen_mixture(n: int, dim: int, ncomp: int, sigma: float, centre_seed: int,
                point_seed: int) -> np.ndarray:
    out = np.empty((dim, n), dtype=np.float32)
    load().oracle_gen_mixture(out.ctypes.data_as(_fp), n, dim, ncomp, sigma, centre_seed,
                              point_seed)
    return out


_ref_lib = None


def reference_lib() -> Optional[ctypes.CDLL]:
    """The reference's own knn_c compiled from its sources (oracle/_ref), or None if not built."""
    global _ref_lib
    if _ref_lib is not None:
        return _ref_lib
    if not os.path.exists(REF_LIB_PATH):
        return None
    lib = ctypes.CDLL(REF_LIB_PATH)
    fn = getattr(lib, "_Z5knn_cPKfiS0_iiiPfPi")
    fn.argtypes = [_fp, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _ip]
    fn.restype = ctypes.c_bool
    _ref_lib = lib
    return lib


def reference_knn_c(ref, query, k: int):
    """Run the reference's knn_c (code/test.cpp:129-175) as compiled from its sources."""
    lib = reference_lib()
    if lib is None:
        raise RuntimeError("oracle/_ref/libknn_ref.so not built")
    ref, query = _prep(ref, query)
    dim, rn = ref.shape
    qn = query.shape[1]
    dist = np.zeros((k, qn), dtype=np.float32)
    idx = np.zeros((k, qn), dtype=np.int32)
    if not lib._Z5knn_cPKfiS0_iiiPfPi(ref.ctypes.data_as(_fp), rn, query.ctypes.data_as(_fp), qn,
                                      dim, k, dist.ctypes.data_as(_fp), idx.ctypes.data_as(_ip)):
        raise RuntimeError("reference knn_c failed")
    return dist, idx
