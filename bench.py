"""Benchmark: exact k-NN queries/s on one MI355X.

Workload (one "step"): exact brute-force k-NN of 16,384 queries against
1,048,576 reference points, dim 128, k = 32, float32 uniform [0,1), dim x n
row-major, all inputs resident in HBM (generated on the device by the
library's seeded generator).  The step is the full path behind the drop-in
API: squared norms, fused MFMA scan + top-k, exact FP32 re-rank with
certificate, fallback launch, outputs in k x query_nb layout.

Prints one JSON line.  `roofline` is measured live for the dominant kernel
(the scan) from HIP events the library records on the stream it launches on;
`traffic` comes from the committed rocprofv3 counter summary of that kernel
when present.  `cpu_baseline` times the oracle's single-threaded restatement
of the reference ground truth on a bounded query sample of the same data.

`precision_form` runs the bfloat16 three-term scan on the same data (equality
with the FP32 result checked) and the dense-cluster workload (dim 256, k 64).
`router_form` measures the second half of the metric on the clustered
workload (4,194,304 refs, 16,384 queries, dim 96, k 16, 2,048 blocks): the
plain FP32 search against the block-index search on the same data, both exact
(the outputs are compared for equality), and the learned MLP router.
`batched_form` times the 512-problem batched workload in one call against
the same problems searched one call at a time (outputs compared for equality).
`reference_form` runs the reference's own benchmark: its three host-pointer
entry points on its test configuration, per call, beside its published times.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "knn-cuda_amd"))

METRIC = "exact-kNN queries/s on 1 MI355X; BF16+re-rank+router speedup vs plain FP32"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "scan_f32_1Mx16k_d128_k32_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ref-nb", type=int, default=1 << 20)
    ap.add_argument("--query-nb", type=int, default=16384)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the cpu_baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-precision-form", action="store_true",
                    help="skip the precision-form measurement (bfloat16 three-term scan)")
    ap.add_argument("--no-router-form", action="store_true",
                    help="skip the router-form measurement (4M clustered refs, block index)")
    ap.add_argument("--router-reps", type=int, default=5)
    ap.add_argument("--no-reference-form", action="store_true",
                    help="skip the reference's own 16K x 4K host-pointer benchmark")
    ap.add_argument("--no-batched-form", action="store_true",
                    help="skip the batched-form measurement (512 problems, tests/golden/c4_problems.json)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device copy of this many MiB before the timed region (a "
                         "known-size kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def precision_form(args, knnhip, torch, ref, qry, k, dist, idx, fp32_ms, fp32_scan_ms):
    """Precision form (bfloat16 three-term products, exact re-rank) on the plain workload's own
    data, against the FP32 result (compared for equality); and the dense-cluster workload
    (1,048,576 x 8,192, dim 256, k 64, mixture of 1,024 Gaussians, sigma 0.05), where the
    three-term bound is too wide for the neighbour gaps and the search runs FP32 products with
    the k+64 candidate list."""
    import time

    R, D = ref.shape[1], ref.shape[0]
    Q = qry.shape[1]
    d16 = torch.empty_like(dist)
    i16 = torch.empty_like(idx)
    stream = torch.cuda.current_stream()
    knnhip.search_device(ref, qry, k, d16, i16, precision="bf16x3", stream=stream.cuda_stream)
    torch.cuda.synchronize()
    scan = []
    t = time.perf_counter()
    for _ in range(args.steps):
        knnhip.search_device(ref, qry, k, d16, i16, precision="bf16x3", timing=True,
                             stream=stream.cuda_stream)
        scan.append(knnhip.last_stage_ms(1))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / args.steps
    plan = knnhip.last_plan()
    equal = bool(torch.equal(d16, dist) and torch.equal(i16, idx))
    scan_ms = sum(scan) / len(scan)
    flops = 2.0 * R * Q * D
    out = {"workload": f"{R} x {Q}, dim {D}, k {k}, uniform [0,1) (the plain workload's data)",
           "ms_per_search": ms, "queries_per_s": Q / ms * 1e3, "speedup_vs_fp32": fp32_ms / ms,
           "scan_ms": scan_ms, "scan_speedup_vs_fp32": fp32_scan_ms / scan_ms,
           "bf16_tflops_executed": 3.0 * flops / scan_ms / 1e9,
           "bf16_peak_frac": 3.0 * flops / scan_ms / 1e9 / 2500.0,
           "fp32_equivalent_tflops": flops / scan_ms / 1e9,
           "exact_scan_queries": knnhip.last_fallback_count(), "equal_to_fp32": equal,
           "precision": plan.get("precision")}
    # dense clusters (k 64)
    R3, Q3, D3, k3 = 1 << 20, 8192, 256, 64
    ref3 = knnhip.generate_mixture(R3, D3, 1024, 0.05, 0x6B4E4E03, 0x6B4E4E13)
    qry3 = knnhip.generate_mixture(Q3, D3, 1024, 0.05, 0x6B4E4E03, 0x6B4E4E23)
    knnhip.search_device(ref3, qry3, k3, precision="bf16x3")
    torch.cuda.synchronize()
    t = time.perf_counter()
    reps = 3
    for _ in range(reps):
        knnhip.search_device(ref3, qry3, k3, precision="bf16x3")
    torch.cuda.synchronize()
    ms3 = (time.perf_counter() - t) * 1e3 / reps
    p3 = knnhip.last_plan()
    out["dense_clusters"] = {"workload": f"{R3} x {Q3}, dim {D3}, k {k3}, mixture of 1,024 Gaussians "
                                         "(sigma 0.05)",
                             "ms_per_search": ms3, "queries_per_s": Q3 / ms3 * 1e3,
                             "precision": p3.get("precision"), "kp": p3.get("kp"),
                             "exact_scan_queries": knnhip.last_fallback_count(),
                             "fp32_scan_tflops": 2.0 * R3 * Q3 * D3 / ms3 / 1e9}
    del ref3, qry3
    return out


def router_form(args, knnhip, torch):
    """Router form on the clustered workload: 4,194,304 refs, 16,384 queries, dim 96, k 16,
    mixture of 4,096 Gaussians (sigma 0.1, centres U[-1,1]^96); 2,048 blocks.  Times the plain FP32
    search and the block-index search on the same device-resident data (both exact; the results
    are compared for equality), plus the learned MLP router trained on 65,536 mixture queries."""
    import time

    R, Q, D, k, NB = 4 << 20, 16384, 96, 16, 2048
    seed_c, seed_r, seed_q, seed_t = 0x6B4E4E05, 0x6B4E4E06, 0x6B4E4E16, 0x6B4E4E26
    ref = knnhip.generate_mixture(R, D, 4096, 0.1, seed_c, seed_r)
    qry = knnhip.generate_mixture(Q, D, 4096, 0.1, seed_c, seed_q)
    torch.cuda.synchronize()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            res = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps * 1e3, res

    plain_ms, (pd, pi) = timed(lambda: knnhip.search_device(ref, qry, k), max(2, args.router_reps // 2))
    t = time.perf_counter()
    ix = knnhip.BlockIndex(ref, nblocks=NB, iters=8)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t
    rows = []
    for prec in ("fp32", "bf16x3"):
        for P in (4, 5, 6, 8):
            ms, (d, i) = timed(lambda: ix.search(qry, k, probes=P, precision=prec), args.router_reps)
            equal = bool(torch.equal(i, pi) and torch.equal(d, pd))
            ix.search(qry, k, probes=P, precision=prec, timing=True)
            rows.append({"precision": prec, "probes": P, "ms": ms, "queries_per_s": Q / ms * 1e3,
                         "speedup_vs_plain": plain_ms / ms,
                         "certificate_failures": knnhip.last_fallback_count(), "equal_to_plain": equal,
                         "stages_ms": {a: round(b, 4) for a, b in ix.stage_ms().items()}})
    best = min(rows, key=lambda r: r["ms"])
    # learned router: trained on the GPU, then used as the block ranking with either product
    trq = knnhip.generate_mixture(65536, D, 4096, 0.1, seed_c, seed_t)
    mlp = knnhip.Router(D, 256, NB, seed=0x6B4E4E07)
    t = time.perf_counter()
    loss = mlp.train(ix, trq, k, epochs=8, batch=1024, lr=1e-3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t
    # the router picks each query's first probe (scanned first) and the bounds rank the rest;
    # "all": the router ranks every probe
    # "first (fused)": the router handle, its first probe from the fused inference kernel (the
    # three layers and the argmax in one launch, no score matrix)
    mlp_rows = []
    for prec, mode in (("fp32", "first"), ("bf16x3", "first"), ("bf16x3", "all"), ("fp32", "first (fused)"),
                       ("bf16x3", "first (fused)")):
        P = min((r for r in rows if r["precision"] == prec), key=lambda r: r["ms"])["probes"]
        if mode == "first (fused)":
            fn = lambda: ix.search(qry, k, probes=P, router=mlp, precision=prec)
        else:
            fn = lambda: ix.search(qry, k, probes=P, scores=mlp.scores(qry), precision=prec,
                                   score_all=mode == "all")
        ms_m, (d, i) = timed(fn, args.router_reps)
        mlp_rows.append({"precision": prec, "scores_rank": mode, "probes": P,
                         "ms": ms_m, "queries_per_s": Q / ms_m * 1e3, "speedup_vs_plain": plain_ms / ms_m,
                         "certificate_failures": knnhip.last_fallback_count(),
                         "equal_to_plain": bool(torch.equal(i, pi) and torch.equal(d, pd))})
    # the router's own inference: full score matrix, against the fused first-probe kernel
    scores_ms, _ = timed(lambda: mlp.scores(qry), args.router_reps)
    top1_ms, _ = timed(lambda: mlp.top1(qry), args.router_reps)
    mlp.close()
    ix.close()
    bf = min((r for r in mlp_rows if r["precision"] == "bf16x3"), key=lambda r: r["ms"])
    return {"workload": "4,194,304 refs x 16,384 queries, dim 96, k 16, mixture of 4,096 Gaussians "
                        "(sigma 0.1), 2,048 blocks",
            "plain_fp32_ms": plain_ms, "plain_queries_per_s": Q / plain_ms * 1e3,
            "index_build_s": build_s, "ranking": "certificate lower bounds",
            "best": {"precision": best["precision"], "probes": best["probes"], "ms": best["ms"],
                     "speedup_vs_plain": best["speedup_vs_plain"]},
            # the form the target names: bfloat16 three-term products, exact FP32 re-rank, learned router
            "bf16_rerank_router": {"probes": bf["probes"], "scores_rank": bf["scores_rank"], "ms": bf["ms"],
                                   "queries_per_s": bf["queries_per_s"],
                                   "speedup_vs_plain": bf["speedup_vs_plain"],
                                   "equal_to_plain": bf["equal_to_plain"]},
            "probes": rows,
            "mlp_router": {"train_s": train_s, "train_queries": 65536, "epochs": 8, "final_loss": loss,
                           "inference_ms": {"score_matrix": scores_ms, "fused_first_probe": top1_ms},
                           "layers": [D, 256, 256, NB], "dtype": "bf16 products, fp32 accumulation",
                           "searches": mlp_rows}}


def reference_form(knnhip, np):
    """The reference's own benchmark (code/test.cpp:295-298, 344-347; README.md:61-85): 16,384 refs,
    4,096 queries, dim 128, k 16, data 10*U[0,1) as host arrays, each of the three drop-in entry
    points timed per call host to host (H2D, search, D2H), next to the published per-call times
    (NVIDIA Titan X, README.md:76-78).  All three outputs are compared for equality."""
    import time

    R, Q, D, k = 16384, 4096, 128, 16
    rng = np.random.default_rng(0x6B4E4E01)
    ref = (10.0 * rng.random((D, R), dtype=np.float32)).astype(np.float32)
    qry = (10.0 * rng.random((D, Q), dtype=np.float32)).astype(np.float32)
    published = {"knn_cuda_global": 41.66, "knn_cuda_texture": 62.96, "knn_cublas": 31.12}
    calls = {"knn_cuda_global": knnhip.knn_cuda_global, "knn_cuda_texture": knnhip.knn_cuda_texture,
             "knn_cublas": knnhip.knn_cublas}
    rows, outs = {}, []
    for name, fn in calls.items():
        fn(ref, qry, k)
        n = 20
        t = time.perf_counter()
        for _ in range(n):
            d, i = fn(ref, qry, k)
        ms = (time.perf_counter() - t) / n * 1e3
        outs.append((d, i))
        rows[name] = {"ms_per_call": ms, "queries_per_s": Q / ms * 1e3,
                      "published_ms_per_call": published[name],
                      "speedup_vs_published": published[name] / ms}
    equal = all(np.array_equal(o[0].view(np.uint32), outs[0][0].view(np.uint32)) and
                np.array_equal(o[1], outs[0][1]) for o in outs)
    # the same entry point at the plain workload's size from host arrays: the references go up in
    # chunks, each scanned as it lands; against one copy of them followed by the device search
    import torch

    R2, Q2, D2, k2 = 1 << 20, 16384, 128, 32
    ref2 = rng.random((D2, R2), dtype=np.float32)
    qry2 = rng.random((D2, Q2), dtype=np.float32)
    knnhip.knn_cuda_global(ref2, qry2, k2)
    t = time.perf_counter()
    for _ in range(3):
        d2, i2 = knnhip.knn_cuda_global(ref2, qry2, k2)
    host_ms = (time.perf_counter() - t) / 3 * 1e3
    chunks = knnhip.last_plan().get("upload_chunks", 1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    ref_d = torch.from_numpy(ref2).cuda()
    qry_d = torch.from_numpy(qry2).cuda()
    torch.cuda.synchronize()
    copy_ms = (time.perf_counter() - t) * 1e3
    knnhip.search_device(ref_d, qry_d, k2)
    torch.cuda.synchronize()
    t = time.perf_counter()
    dd, di = knnhip.search_device(ref_d, qry_d, k2)
    torch.cuda.synchronize()
    dev_ms = (time.perf_counter() - t) * 1e3
    same = bool(np.array_equal(i2, di.cpu().numpy()) and np.array_equal(d2.view(np.uint32),
                                                                        dd.cpu().numpy().view(np.uint32)))
    del ref_d, qry_d, dd, di
    torch.cuda.empty_cache()
    # the same call under device memory budgets below the resident footprint: the references
    # stream through two chunk buffers, each chunk's copy overlapping the previous chunk's search
    streamed = []
    for mib in (700, 380):
        knnhip.set_device_budget(mib << 20)
        try:
            knnhip.knn_cuda_global(ref2, qry2, k2)
            t = time.perf_counter()
            for _ in range(2):
                d3, i3 = knnhip.knn_cuda_global(ref2, qry2, k2)
            ms = (time.perf_counter() - t) / 2 * 1e3
            plan = knnhip.last_plan()
        finally:
            knnhip.set_device_budget(0)
        streamed.append({"device_budget_mib": mib, "ref_chunks": plan.get("ref_chunks"),
                         "query_chunks": plan.get("query_chunks"), "ms_per_call": ms,
                         "equal_to_resident": bool(np.array_equal(i3, i2) and
                                                   np.array_equal(d3.view(np.uint32), d2.view(np.uint32)))})
    return {"workload": "16384 refs x 4096 queries, dim 128, k 16, 10*U[0,1) host arrays (the reference's "
                        "test.cpp configuration), per call host to host",
            "published_hardware": "NVIDIA Titan X (README.md:76-78)", "calls": rows,
            "outputs_equal": bool(equal),
            "host_arrays_at_scale": {"workload": "1048576 x 16384, dim 128, k 32, U[0,1) host arrays",
                                     "knn_cuda_global_ms": host_ms, "upload_chunks": chunks,
                                     "h2d_copy_ms": copy_ms, "device_search_ms": dev_ms,
                                     "copy_then_search_ms": copy_ms + dev_ms,
                                     "equal_to_device_search": same, "streamed": streamed}}


def batched_form(knnhip, torch):
    """Batched form: the 512 problems of tests/golden/c4_problems.json (R log-uniform in [1K, 256K],
    Q in [16, 4096], dim in {32..512}, k in {1..128}, uniform data from the listed seeds), all
    device-resident.  Times one call of the batched search (untuned, then after the autotuner has
    picked each class's tile shape) against the same problems searched one call at a time; all
    outputs are compared for equality."""
    import time

    with open(os.path.join(ROOT, "tests", "golden", "c4_problems.json")) as f:
        P = json.load(f)["problems"]
    items = [(knnhip.generate_uniform(p["ref_nb"], p["dim"], seed=p["ref_seed"]),
              knnhip.generate_uniform(p["query_nb"], p["dim"], seed=p["query_seed"]), p["k"]) for p in P]
    torch.cuda.synchronize()
    flops = sum(2.0 * r.shape[1] * q.shape[1] * r.shape[0] for r, q, _ in items)
    nq = sum(q.shape[1] for _, q, _ in items)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3, out

    seq_ms, seq = timed(lambda: [knnhip.search_device(r, q, k) for r, q, k in items])
    for c in knnhip.batched_table():
        knnhip.batched_set_slice(c["dim"], c["k"], 0)
    unt_ms, out = timed(lambda: knnhip.search_batched(items))
    equal = all(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]) for a, b in zip(out, seq))
    t = time.perf_counter()
    knnhip.batched_autotune(items)
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - t
    tun_ms, out = timed(lambda: knnhip.search_batched(items))
    equal = equal and all(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]) for a, b in zip(out, seq))
    return {"workload": "512 problems (tests/golden/c4_problems.json), uniform [0,1)",
            "flops": flops, "queries": nq, "sequential_ms": seq_ms, "batched_untuned_ms": unt_ms,
            "batched_tuned_ms": tun_ms, "autotune_s": tune_s,
            "queries_per_s": nq / tun_ms * 1e3, "tflops": flops / tun_ms / 1e9,
            "speedup_vs_sequential": seq_ms / tun_ms, "certificate_failures": knnhip.batched_last_failures(),
            "equal_to_single_searches": equal, "classes": len(knnhip.batched_table())}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the reference-set "
              "shard across GPUs is described in DESIGN.md and is not run here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import knnhip

    R, Q, D, k = args.ref_nb, args.query_nb, args.dim, args.k
    dev = torch.device("cuda:0")
    seed_r, seed_q = 0x6B4E4E02, 0x6B4E4E12
    ref = knnhip.generate_uniform(R, D, seed=seed_r, device=dev)
    qry = knnhip.generate_uniform(Q, D, seed=seed_q, device=dev)
    dist = torch.empty((k, Q), dtype=torch.float32, device=dev)
    idx = torch.empty((k, Q), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize()
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for _ in range(args.warmup):
        knnhip.search_device(ref, qry, k, dist, idx, stream=stream.cuda_stream)
    torch.cuda.synchronize()

    scan_ms = []
    stage_ms = [[] for _ in range(4)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        knnhip.search_device(ref, qry, k, dist, idx, timing=True, stream=stream.cuda_stream)
        for s in range(4):
            stage_ms[s].append(knnhip.last_stage_ms(s))
        scan_ms.append(knnhip.last_stage_ms(1))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fallbacks = knnhip.last_fallback_count()

    ms_per_step = 1000.0 * elapsed / args.steps
    qps = Q / (elapsed / args.steps)
    flops = 2.0 * R * Q * D
    scan_avg = sum(scan_ms) / len(scan_ms)
    achieved = flops / (scan_avg * 1e-3) / 1e12

    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("workload") == [R, Q, D, k]:
            traffic = t.get("hbm_bytes_per_launch")

    # Correctness gate + CPU baseline on the same data.
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    ref_h = ref.cpu().numpy()
    qry_h = qry.cpu().numpy()
    d_h = dist.cpu().numpy()
    i_h = idx.cpu().numpy()
    cpu = None
    gate_q = 0
    if not args.no_cpu_baseline:
        # the reference's own knn_c (code/test.cpp:129-175) compiled from its sources when that
        # build is present, else the oracle's restatement of it; single-threaded either way
        use_ref = oracle.reference_lib() is not None

        def cpu_knn(n):
            if use_ref:
                return oracle.reference_knn_c(ref_h, np.ascontiguousarray(qry_h[:, :n]), k)
            return oracle.knn_c_queries(ref_h, qry_h, k, 0, n)

        t1 = time.perf_counter()
        cpu_knn(1)
        one = time.perf_counter() - t1
        n = int(max(2, min(64, args.cpu_seconds / max(one, 1e-6))))
        t1 = time.perf_counter()
        gd, gi = cpu_knn(n)
        cpu_s = time.perf_counter() - t1
        gate_q = n
        ok = np.array_equal(gi[:, :n], i_h[:, :n]) and np.array_equal(
            gd[:, :n].view(np.uint32), d_h[:, :n].view(np.uint32))
        if not ok:
            print(json.dumps({"error": "parity gate failed", "queries_checked": n}), flush=True)
            sys.exit(1)
        what = ("the reference's own knn_c (code/test.cpp:129-175) compiled from its sources" if use_ref
                else "the oracle's restatement of knn_c")
        cpu = {"value": n / cpu_s, "unit": "queries/s", "cores": 1, "kind": "reference" if use_ref else "port",
               "sample": f"first {n} of {Q} queries against all {R} refs (dim {D}, k {k}), single-threaded "
                         f"{what}, {cpu_s:.1f} s"}

    plan = knnhip.last_plan()
    # selection / re-rank stage: HBM-bound gathers.  Algorithmic bytes per launch = the candidates'
    # point-major rows (4 D bytes each, kp per query) + the slice heaps (8 bytes per entry) + the
    # outputs (8 k bytes per query)
    sel_ms = sum(stage_ms[2]) / len(stage_ms[2])
    kp, slices = plan.get("kp", k), plan.get("slices", 1)
    sel_bytes = 4.0 * Q * kp * D + 8.0 * slices * Q * kp + 8.0 * k * Q
    select_stage = {"kernel": "select_rerank_kernel", "ms": sel_ms, "algorithmic_bytes": sel_bytes,
                    "achieved_gb_per_s": sel_bytes / (sel_ms * 1e-3) / 1e9,
                    "hbm_peak_frac": sel_bytes / (sel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    precision = None if args.no_precision_form else precision_form(args, knnhip, torch, ref, qry, k,
                                                                      dist, idx, ms_per_step, scan_avg)
    router = None if args.no_router_form else router_form(args, knnhip, torch)
    batched = None if args.no_batched_form else batched_form(knnhip, torch)
    reference = None if args.no_reference_form else reference_form(knnhip, np)

    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform [0,1), generated on device)",
        "config": {"workload": "plain FP32 exact kNN at scale", "ref_nb": R, "query_nb": Q,
                   "dim": D, "k": k, "layout": "dim x n row-major", "parallelism": "1 GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "scan_f32_kernel", "kernel_ms": scan_avg,
                     "flops_per_launch": flops},
        "cpu_baseline": cpu,
        "select_stage": select_stage,
        "stages_ms": {name: sum(v) / len(v) for name, v in
                      zip(("norms", "scan", "select_rerank", "fallback"), stage_ms)},
        "fallback_queries": fallbacks,
        "plan": plan,
        "parity_gate_queries": gate_q,
        "precision_form": precision,
        "router_form": router,
        "batched_form": batched,
        "reference_form": reference,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
