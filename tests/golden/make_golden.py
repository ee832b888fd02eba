"""Write golden k-NN vectors with the reference's own ground truth.

Runs ``knn_c`` of the reference (code/test.cpp:129-175), compiled from the
reference sources by ``make -C oracle ref`` into oracle/_ref/libknn_ref.so,
on seeded inputs, and stores inputs and outputs as .npz fixtures in this
directory.  Only the data is committed; the reference never travels.  Large
inputs are stored as generator parameters (the oracle's counter-based
generator re-creates them bit for bit) instead of raw arrays.

Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import zlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402


def case_inputs(name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name == "uniform10_small":
        # test.cpp:32,37 data: 10 * rand()/RAND_MAX
        return (10 * rng.random((32, 512))).astype(np.float32), (10 * rng.random((32, 64))).astype(np.float32), 16
    if name == "tail_dims":
        return rng.random((17, 300), dtype=np.float32), rng.random((17, 37), dtype=np.float32), 5
    if name == "duplicates":
        base = np.round(4 * rng.random((8, 50))).astype(np.float32)
        ref = np.concatenate([base, base, base[:, ::-1], base], axis=1)  # every point 4 times
        qry = np.round(4 * rng.random((8, 20))).astype(np.float32)
        return ref, qry, 12
    if name == "line_1d":
        ref = np.round(50 * rng.random((1, 100))).astype(np.float32)
        qry = np.round(50 * rng.random((1, 10))).astype(np.float32)
        return ref, qry, 10
    if name == "k_equals_ref_nb":
        return rng.random((5, 40), dtype=np.float32), rng.random((5, 8), dtype=np.float32), 40
    if name == "k1_dim64":
        return rng.random((64, 500), dtype=np.float32), rng.random((64, 40), dtype=np.float32), 1
    if name == "query_in_ref":
        ref = rng.random((24, 400), dtype=np.float32)
        qry = np.ascontiguousarray(ref[:, [0, 7, 399, 123, 7]])
        return ref, qry, 8
    if name == "mixed_sign_dim96":
        ref = (rng.standard_normal((96, 400)) * 3).astype(np.float32)
        qry = (rng.standard_normal((96, 33)) * 3).astype(np.float32)
        return ref, qry, 20
    raise KeyError(name)


SMALL_CASES = ["uniform10_small", "tail_dims", "duplicates", "line_1d", "k_equals_ref_nb",
               "k1_dim64", "query_in_ref", "mixed_sign_dim96"]

# Larger cases: data from the oracle generator, only parameters stored.
GEN_CASES = {
    # repo default shape (code/test.cpp:295-298) on a query subset, data in [0,1)
    "c1_uniform_q48": dict(R=16384, Q=48, D=128, k=16, seed_ref=0x6B4E4E01, seed_qry=0x6B4E4E11, scale=1.0),
    # same shape with the reference's [0,10] scale
    "c1_uniform10_q32": dict(R=16384, Q=32, D=128, k=16, seed_ref=0x6B4E4E02, seed_qry=0x6B4E4E12, scale=10.0),
    # clustered data (near-ties), precision-form shape at reduced size
    "mixture_d256_k64": dict(R=8192, Q=24, D=256, k=64, mixture=(64, 0.05, 0x6B4E4E03), seed_ref=0x6B4E4E04,
                             seed_qry=0x6B4E4E14),
}


def gen_inputs(p):
    if "mixture" in p:
        ncomp, sigma, cseed = p["mixture"]
        ref = oracle.gen_mixture(p["R"], p["D"], ncomp, sigma, cseed, p["seed_ref"])
        qry = oracle.gen_mixture(p["Q"], p["D"], ncomp, sigma, cseed, p["seed_qry"])
    else:
        ref = oracle.gen_uniform(p["R"], p["D"], p["seed_ref"], p["scale"])
        qry = oracle.gen_uniform(p["Q"], p["D"], p["seed_qry"], p["scale"])
    return ref, qry


def main():
    if oracle.reference_lib() is None:
        sys.exit("build the reference first: make -C oracle ref")
    index = {}
    for name in SMALL_CASES:
        ref, qry, k = case_inputs(name)
        dist, idx = oracle.reference_knn_c(ref, qry, k)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), ref=ref, query=qry, k=np.int32(k),
                            dist=dist, index=idx)
        index[name] = {"kind": "arrays", "ref_nb": int(ref.shape[1]), "query_nb": int(qry.shape[1]),
                       "dim": int(ref.shape[0]), "k": int(k)}
    for name, p in GEN_CASES.items():
        ref, qry = gen_inputs(p)
        dist, idx = oracle.reference_knn_c(ref, qry, p["k"])
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), k=np.int32(p["k"]), dist=dist, index=idx,
                            ref_head=ref[:, :8].copy(), query_head=qry[:, :8].copy())
        index[name] = {"kind": "generated", **{kk: (list(v) if isinstance(v, tuple) else v) for kk, v in p.items()}}
    with open(os.path.join(HERE, "golden_index.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)
    print("wrote", len(index), "fixtures")


if __name__ == "__main__":
    main()
