"""Benchmark: exact k-NN queries/s on one MI355X.

Headline workload (one "step"): the learned-router form of the clustered workload, one exact
search of 16,384 queries against 4,194,304 references, dim 96, k 16, float32 drawn from a mixture
of 4,096 isotropic Gaussians (sigma 0.1, centres uniform in [-1,1]^96), generated on the device by
the library's seeded generator and resident in HBM.  The search goes through the block index
(2,048 k-means blocks) with bfloat16 MFMA distance products, the MLP router's first probe
(96 -> 256 -> 256 -> 2048, trained on the GPU on the method's own exact neighbours), the
triangle-inequality certificate and the exact FP32 re-rank; its output is compared with the
plain FP32 search, which is also timed (the metric's speedup).  Index build and router training
are timed separately and reported beside the speedup (they are paid once per reference set).

Prints one JSON line.  `roofline` is measured live for the dominant kernel (the block scan) from
HIP events the library records on the stream it launches on, against the bfloat16 MFMA peak with
the MFMA work the scan's tasks executed.  `cpu_baseline` times the reference's own knn_c
(code/test.cpp:129-175), compiled from its sources, single-threaded on a bounded query sample
of the same data, next to the oracle's restatement on all host cores; the GPU results of the
sample must equal it bit for bit.

Further objects, each on its own workload: `plain_fp32` (1,048,576 x 16,384, dim 128, k 32,
uniform: the plain fused FP32 scan and its roofline), `precision_form` (the bfloat16 scans on that
data, and the 1,048,576 x 8,192, dim 256, k 64 dense-cluster workload through the block-residual
form, with the gap-against-bound tables of bfloat16 and FP8 residuals), `router_form` (every
probe count and product of the headline workload), `batched_form` (512 problems) and
`reference_form` (the reference's own benchmark configuration through its host entry points).
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
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
PLAIN_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "scan_f32_1Mx16k_d128_k32_traffic.json")
PLAIN_LDS_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "scan_lds_f32_1Mx16k_d128_k32_traffic.json")
ROUTER_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "router_scan_4Mx16k_d96_k16_traffic.json")
HEADLINE_FORM = ("balls", "bf16x3", 8)

# the learned-router workload (configs: 4,194,304 x 16,384, dim 96, k 16, 4,096 Gaussians)
C5 = dict(R=4 << 20, Q=16384, D=96, k=16, ncomp=4096, sigma=0.1, nblocks=2048,
          seed_c=0x6B4E4E05, seed_r=0x6B4E4E06, seed_q=0x6B4E4E16, seed_t=0x6B4E4E26)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the single-threaded cpu_baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain-form", action="store_true",
                    help="skip the plain FP32 workload (1M x 16K x 128)")
    ap.add_argument("--no-precision-form", action="store_true",
                    help="skip the precision-form measurements")
    ap.add_argument("--no-router-table", action="store_true",
                    help="skip the table of probe counts and products of the headline workload")
    ap.add_argument("--router-reps", type=int, default=5)
    ap.add_argument("--no-reference-form", action="store_true",
                    help="skip the reference's own 16K x 4K host-pointer benchmark")
    ap.add_argument("--no-batched-form", action="store_true",
                    help="skip the batched-form measurement (512 problems, tests/golden/c4_problems.json)")
    ap.add_argument("--form", default="",
                    help="pin the headline's search form as units:precision:probes (e.g. balls:bf16x3:6) "
                         "and time only it; by default every form is timed and the headline is "
                         "balls:bf16x3:8")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device copy of this many MiB before the timed region (a "
                         "known-size kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def timed(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        res = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps * 1e3, res


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = int(omp) if omp and omp.isdigit() else usable
    return {"model": model, "nproc": os.cpu_count(), "usable_cpus": usable, "omp_threads": min(threads, usable)}


def load_traffic(path, workload, form=None):
    """HBM bytes per launch of a kernel, from a profile summary of the same workload (and search
    form, where given), or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("workload") != workload:
        return None
    if form is None:
        return t.get("hbm_bytes_per_launch")
    for e in t.get("forms", []):
        if e.get("form") == form:
            return e.get("hbm_bytes_per_launch")
    return None


def headline(args, knnhip, torch):
    """The learned-router workload: plain FP32 search, index build, router training, the table of
    forms (unless skipped), then the timed steps of the fastest exact BF16 + router form."""
    R, Q, D, k, NB = C5["R"], C5["Q"], C5["D"], C5["k"], C5["nblocks"]
    ref = knnhip.generate_mixture(R, D, C5["ncomp"], C5["sigma"], C5["seed_c"], C5["seed_r"])
    qry = knnhip.generate_mixture(Q, D, C5["ncomp"], C5["sigma"], C5["seed_c"], C5["seed_q"])
    torch.cuda.synchronize()
    plain_ms, (pd, pi) = timed(torch, lambda: knnhip.search_device(ref, qry, k), 3)
    t = time.perf_counter()
    ix = knnhip.BlockIndex(ref, nblocks=NB, iters=8)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t
    trq = knnhip.generate_mixture(65536, D, C5["ncomp"], C5["sigma"], C5["seed_c"], C5["seed_t"])
    mlp = knnhip.Router(D, 256, NB, seed=0x6B4E4E07)
    t = time.perf_counter()
    loss = mlp.train(ix, trq, k, epochs=8, batch=1024, lr=1e-3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t
    del trq

    def same(d, i):
        return bool(torch.equal(i, pi) and torch.equal(d, pd))

    # candidate forms: bfloat16 products (three-term, or single-pass on centre residuals) with the
    # router's first probe, the scan visiting balls (one pass, the default; or two passes) or whole
    # blocks, each timed
    forms = []
    pinned = args.form.split(":") if args.form else None
    for units in ((pinned[0],) if pinned else ("balls", "balls-threshold", "blocks")):
        for prec in ((pinned[1],) if pinned else ("bf16x3", "bf16")):
            for P in ((int(pinned[2]),) if pinned else (4, 6, 8, 12) if units != "blocks" else (4, 5, 6)):
                ms, (d, i) = timed(torch, lambda: ix.search(qry, k, probes=P, router=mlp, precision=prec,
                                                            units=units), args.router_reps)
                forms.append({"units": units, "precision": prec, "probes": P,
                              "ms": ms, "equal_to_plain": same(d, i),
                              "certificate_failures": knnhip.last_fallback_count(),
                              "rows_per_query": ix.last_units()["rows_per_query"]})
    # the headline form: ball units, three-term products, 8 probes (the form whose scan traffic
    # is measured with rocprofv3), unless pinned; the fastest exact form if it were not exact
    want = (pinned[0], pinned[1], int(pinned[2])) if pinned else HEADLINE_FORM
    best = next((f for f in forms if (f["units"], f["precision"], f["probes"]) == want and f["equal_to_plain"]),
                None) or min((f for f in forms if f["equal_to_plain"]), key=lambda f: f["ms"])
    prec, P, units = best["precision"], best["probes"], best["units"]

    dist = torch.empty((k, Q), dtype=torch.float32, device=qry.device)
    idx = torch.empty((k, Q), dtype=torch.int32, device=qry.device)
    stream = torch.cuda.current_stream(qry.device)
    for _ in range(args.warmup):
        ix.search(qry, k, probes=P, router=mlp, precision=prec, dist=dist, index=idx, units=units,
                  stream=stream.cuda_stream)
    torch.cuda.synchronize()
    scan_ms, stages, works = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        # (HIP events around the scan launch only: each recorded event leaves a few us between kernels)
        ix.search(qry, k, probes=P, router=mlp, precision=prec, dist=dist, index=idx, timing="scan",
                  units=units, stream=stream.cuda_stream)
        scan_ms.append(ix.stage_ms()["scan"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fails = knnhip.last_fallback_count()
    work = ix.last_work()
    equal = same(dist, idx)
    # the stage breakdown, from as many more searches with every stage's events (not timed)
    for _ in range(args.steps):
        ix.search(qry, k, probes=P, router=mlp, precision=prec, dist=dist, index=idx, timing=True,
                  units=units, stream=stream.cuda_stream)
        stages.append(ix.stage_ms())
    equal = equal and same(dist, idx)
    ms_per_step = elapsed * 1e3 / args.steps
    scan_avg = sum(scan_ms) / len(scan_ms)
    # executed MFMA work of the scan: 2 D flops per executed (reference, query) pair of its
    # tiles, three bfloat16 products per pair for the three-term form
    mult = {"bf16x3": 3.0, "bf16-residual": 1.0, "fp32": 1.0}[work["precision"]]
    exec_flops = 2.0 * D * work["exec_pairs"] * mult
    peak = FP32_PEAK_TFLOPS if work["precision"] == "fp32" else BF16_PEAK_TFLOPS
    achieved = exec_flops / (scan_avg * 1e-3) / 1e12
    stage_avg = {s: sum(x[s] for x in stages) / len(stages) for s in stages[0]}
    table = None
    if not args.no_router_table:
        table = router_table(args, knnhip, torch, ix, mlp, qry, k, pd, pi, plain_ms)
    lu = ix.last_units()
    launches = 1 if units == "balls" else 2  # one scan phase with one-pass ball probes
    form = {"units": units, "precision": prec, "probes": P}
    # The scan is bound by neither MFMA (a few % of peak) nor HBM: its tasks hold 2-6 queries and
    # per-task candidate insertion dominates.  Its roofline is taken against HBM with the bytes of
    # the distinct reference tiles it scans (each ball once: what a scan reading every ball once
    # reads); the MFMA figures follow.
    out = {
        "qps": Q / (elapsed / args.steps), "ms_per_step": ms_per_step, "precision": prec, "probes": P,
        "units": units, "rows_per_query": lu["rows_per_query"], "balls": lu["balls"],
        "roofline": {"bound": "hbm", "achieved": work["unit_bytes"] / (scan_avg * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": work["unit_bytes"] / (scan_avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": load_traffic(ROUTER_TRAFFIC_FILE, [R, Q, D, k], form),
                     "kernel": "scan_res_task_kernel" if work["precision"] == "bf16-residual"
                     else "scan_bf3_task_kernel" if work["precision"] == "bf16x3" else "scan_f32_kernel",
                     "kernel_ms": scan_avg, "launches_per_step": launches,
                     "algorithmic_bytes_per_step": work["unit_bytes"],
                     "bytes_note": "distinct reference tiles of the scanned units (bytes per reference: "
                                   "384 three-term, 192 residual, 4 x padded dim FP32); kernel_ms is the scan "
                                   "stage (its launches and their task building) from HIP events",
                     "tile_bytes_read_per_step": work["tile_bytes"],
                     "mfma": {"achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                              "flops_per_step": exec_flops,
                              "flops_note": "executed MFMA work of the scan's tasks (whole 128-reference steps "
                                            "x 32 query columns, 2 x 96 flops per pair, x3 for the three-term "
                                            "products)"},
                     "pairs_held": work["pairs"], "pairs_executed": work["exec_pairs"]},
        "stages_ms": stage_avg, "certificate_failures": fails, "equal_to_plain_fp32": equal,
        "plain_fp32_ms": plain_ms, "plain_fp32_queries_per_s": Q / plain_ms * 1e3,
        "speedup_vs_plain_fp32": plain_ms / ms_per_step,
        "index_build_s": build_s, "router_train_s": train_s, "router_final_loss": loss,
        "setup_in_plain_searches": (build_s + train_s) * 1e3 / plain_ms,
        "forms": forms, "router_table": table,
    }
    return out, ref, qry, dist, idx, ix, mlp


def router_table(args, knnhip, torch, ix, mlp, qry, k, pd, pi, plain_ms):
    """Every probe count and product of the headline workload, with the lower bounds alone, the
    router's scores, and the fused router (its first probe)."""
    Q = qry.shape[1]
    rows = []
    for units in ("balls", "balls-threshold", "blocks"):
        for prec in ("fp32", "bf16x3", "bf16"):
            for P in (4, 6, 8):
                ms, (d, i) = timed(torch, lambda: ix.search(qry, k, probes=P, precision=prec, units=units),
                                   args.router_reps)
                rows.append({"ranking": "bounds", "units": units, "precision": prec,
                             "probes": P, "ms": ms, "speedup_vs_plain": plain_ms / ms,
                             "certificate_failures": knnhip.last_fallback_count(),
                             "equal_to_plain": bool(torch.equal(i, pi) and torch.equal(d, pd))})
    P = 6
    for prec, mode in (("bf16x3", "first"), ("bf16x3", "all"), ("bf16", "first")):
        ms, (d, i) = timed(torch, lambda: ix.search(qry, k, probes=P, scores=mlp.scores(qry), precision=prec,
                                                    score_all=mode == "all"), args.router_reps)
        rows.append({"ranking": f"router scores ({mode})", "units": "balls", "precision": prec, "probes": P,
                     "ms": ms,
                     "speedup_vs_plain": plain_ms / ms, "certificate_failures": knnhip.last_fallback_count(),
                     "equal_to_plain": bool(torch.equal(i, pi) and torch.equal(d, pd))})
    scores_ms, _ = timed(torch, lambda: mlp.scores(qry), args.router_reps)
    top1_ms, _ = timed(torch, lambda: mlp.top1(qry), args.router_reps)
    return {"rows": rows, "router_inference_ms": {"score_matrix": scores_ms, "fused_first_probe": top1_ms}}


def cpu_baseline(args, knnhip, oracle, np, ref, qry, d_gpu, i_gpu, k):
    """The reference's knn_c single-threaded on a bounded sample, the oracle on all cores, and the
    bit-exact gate of the sample against the GPU output."""
    ref_h = ref.cpu().numpy()
    qry_h = qry.cpu().numpy()
    use_ref = oracle.reference_lib() is not None

    def one_thread(n):
        q = np.ascontiguousarray(qry_h[:, :n])
        if use_ref:
            return oracle.reference_knn_c(ref_h, q, k)
        return oracle.knn_c_queries(ref_h, q, k, 0, n)

    t = time.perf_counter()
    one_thread(1)
    one = time.perf_counter() - t
    n = int(max(2, min(64, args.cpu_seconds / max(one, 1e-6))))
    t = time.perf_counter()
    gd, gi = one_thread(n)
    cpu_s = time.perf_counter() - t
    ok = np.array_equal(gi, i_gpu[:, :n]) and np.array_equal(gd.view(np.uint32), d_gpu[:, :n].view(np.uint32))
    info = cpu_info()
    # all cores: the oracle's restatement (same results), one query per thread at a time
    m = max(n, 4 * info["omp_threads"])
    qm = np.ascontiguousarray(qry_h[:, :m])
    t = time.perf_counter()
    od, oi = oracle.knn_c(ref_h, qm, k)
    all_s = time.perf_counter() - t
    ok = ok and np.array_equal(oi, i_gpu[:, :m]) and np.array_equal(od.view(np.uint32),
                                                                    d_gpu[:, :m].view(np.uint32))
    R, D = ref_h.shape[1], ref_h.shape[0]
    what = ("the reference's own knn_c (code/test.cpp:129-175) compiled from its sources" if use_ref
            else "the oracle's restatement of knn_c")
    return ok, n, {
        "value": n / cpu_s, "unit": "queries/s", "cores": 1, "kind": "reference" if use_ref else "port",
        "sample": f"first {n} of {qry_h.shape[1]} queries against all {R} refs (dim {D}, k {k}), "
                  f"single-threaded {what}, {cpu_s:.1f} s",
        "cpu": info,
        "all_cores": {"value": m / all_s, "unit": "queries/s", "cores": info["omp_threads"], "kind": "port",
                      "sample": f"first {m} of {qry_h.shape[1]} queries, the oracle's restatement with queries "
                                f"spread over {info['omp_threads']} threads (OpenMP), {all_s:.1f} s"},
    }


def plain_form(args, knnhip, torch):
    """Plain fused FP32 scan at 1,048,576 x 16,384, dim 128, k 32 (uniform [0,1)): queries/s and the
    scan's FP32 MFMA roofline from the library's events."""
    R, Q, D, k = 1 << 20, 16384, 128, 32
    ref = knnhip.generate_uniform(R, D, seed=0x6B4E4E02)
    qry = knnhip.generate_uniform(Q, D, seed=0x6B4E4E12)
    dist = torch.empty((k, Q), dtype=torch.float32, device=qry.device)
    idx = torch.empty((k, Q), dtype=torch.int32, device=qry.device)
    stream = torch.cuda.current_stream(qry.device)
    for _ in range(2):
        knnhip.search_device(ref, qry, k, dist, idx, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    steps = 10
    stage = [[] for _ in range(4)]
    t0 = time.perf_counter()
    for _ in range(steps):
        knnhip.search_device(ref, qry, k, dist, idx, timing=True, stream=stream.cuda_stream)
        for s in range(4):
            stage[s].append(knnhip.last_stage_ms(s))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    plan = knnhip.last_plan()
    scan = sum(stage[1]) / steps
    flops = 2.0 * R * Q * D
    sel = sum(stage[2]) / steps
    kp, slices = plan.get("kp", k), plan.get("slices", 1)
    sel_bytes = 4.0 * Q * kp * D + 8.0 * slices * Q * kp + 8.0 * k * Q
    return {"workload": f"{R} x {Q}, dim {D}, k {k}, uniform [0,1)", "ms_per_search": ms,
            "queries_per_s": Q / ms * 1e3,
            "roofline": {"bound": "mfma", "achieved": flops / (scan * 1e-3) / 1e12, "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": flops / (scan * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                         "traffic": load_traffic(PLAIN_TRAFFIC_FILE, [R, Q, D, k]), "kernel": "scan_f32_kernel",
                         "kernel_ms": scan, "flops_per_launch": flops},
            "select_stage": {"kernel": "select_rerank_kernel", "ms": sel, "algorithmic_bytes": sel_bytes,
                             "achieved_gb_per_s": sel_bytes / (sel * 1e-3) / 1e9,
                             "hbm_peak_frac": sel_bytes / (sel * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "stages_ms": {n: sum(v) / steps for n, v in zip(("norms", "scan", "select_rerank", "fallback"), stage)},
            "exact_scan_queries": knnhip.last_fallback_count(), "plan": plan}, ref, qry, dist, idx, ms, scan


def gap_vs_bound(np, ix, ref, qry, k, margin, sample=64):
    """Per sampled query: the gap between its k-th and (k+margin)-th exact distances, against twice
    the residual rounding bound delta (||q - c|| + rho) of the block holding its nearest neighbour
    (c the block's centroid, rho its radius), for bfloat16 (delta = 2^-8 + 2^-23) and FP8-e4m3
    (delta = 2^-4 + 2^-23) residuals.  A query can certify only where gap > 2 eps."""
    import torch

    import knnhip

    Q = qry.shape[1]
    cols = np.unique(np.linspace(0, Q - 1, sample).astype(np.int64))
    q = qry[:, torch.from_numpy(cols).to(qry.device)].contiguous()
    d, i = knnhip.search_device(ref, q, k + margin)
    d, i = d.cpu().numpy().astype(np.float64), i.cpu().numpy()
    start, count, ref_block = ix.layout()
    rb = ref_block.cpu().numpy()
    qh = q.cpu().numpy().astype(np.float64)
    gaps, eb, e8 = [], [], []
    cache = {}
    for j in range(len(cols)):
        b = int(rb[i[0, j]])
        if b not in cache:
            members = np.nonzero(rb == b)[0]
            pts = ref[:, torch.from_numpy(members).to(ref.device)].cpu().numpy().astype(np.float64)
            c = pts.mean(axis=1)
            cache[b] = (c, float(np.sqrt(((pts - c[:, None]) ** 2).sum(axis=0)).max()))
        c, rho = cache[b]
        eq = float(np.linalg.norm(qh[:, j] - c))
        gaps.append(d[k + margin - 1, j] - d[k - 1, j])
        eb.append(2 * (2.0 ** -8 + 2.0 ** -23) * (eq + rho))
        e8.append(2 * (2.0 ** -4 + 2.0 ** -23) * (eq + rho))
    gaps, eb, e8 = map(np.array, (gaps, eb, e8))
    pct = lambda a: {p: float(np.percentile(a, p)) for p in (10, 50, 90)}
    return {"queries": int(len(cols)), "k": k, "margin": margin, "gap_distance": pct(gaps),
            "two_eps_bf16": pct(eb), "two_eps_fp8_e4m3": pct(e8),
            "share_gap_above_two_eps_bf16": float(np.mean(gaps > eb)),
            "share_gap_above_two_eps_fp8": float(np.mean(gaps > e8))}


def precision_form(args, knnhip, torch, np, ref, qry, dist, idx, fp32_ms, fp32_scan_ms, c5):
    """bfloat16 three-term scan on the plain workload's data (equal to the FP32 result); the dense-
    cluster workload (1,048,576 x 8,192, dim 256, k 64, mixture of 1,024 Gaussians, sigma 0.05)
    through the plain FP32 scan and through the block index with single-pass bfloat16 products on
    block residuals and the k + 64 list re-ranked in FP32; gap-against-bound tables."""
    R, D = ref.shape[1], ref.shape[0]
    Q, k = qry.shape[1], dist.shape[0]
    d16 = torch.empty_like(dist)
    i16 = torch.empty_like(idx)
    stream = torch.cuda.current_stream()
    knnhip.search_device(ref, qry, k, d16, i16, precision="bf16x3", stream=stream.cuda_stream)
    torch.cuda.synchronize()
    scan = []
    t = time.perf_counter()
    for _ in range(10):
        knnhip.search_device(ref, qry, k, d16, i16, precision="bf16x3", timing=True, stream=stream.cuda_stream)
        scan.append(knnhip.last_stage_ms(1))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / 10
    plan = knnhip.last_plan()
    scan_ms = sum(scan) / len(scan)
    flops = 2.0 * R * Q * D
    out = {"uniform": {"workload": f"{R} x {Q}, dim {D}, k {k}, uniform [0,1) (the plain workload's data)",
                       "ms_per_search": ms, "queries_per_s": Q / ms * 1e3, "speedup_vs_fp32": fp32_ms / ms,
                       "scan_ms": scan_ms, "scan_speedup_vs_fp32": fp32_scan_ms / scan_ms,
                       "bf16_tflops_executed": 3.0 * flops / scan_ms / 1e9,
                       "bf16_peak_frac": 3.0 * flops / scan_ms / 1e9 / BF16_PEAK_TFLOPS,
                       "exact_scan_queries": knnhip.last_fallback_count(),
                       "equal_to_fp32": bool(torch.equal(d16, dist) and torch.equal(i16, idx)),
                       "precision": plan.get("precision")}}
    # dense clusters
    R3, Q3, D3, k3 = 1 << 20, 8192, 256, 64
    ref3 = knnhip.generate_mixture(R3, D3, 1024, 0.05, 0x6B4E4E03, 0x6B4E4E13)
    qry3 = knnhip.generate_mixture(Q3, D3, 1024, 0.05, 0x6B4E4E03, 0x6B4E4E23)
    torch.cuda.synchronize()
    ms3, (pd3, pi3) = timed(torch, lambda: knnhip.search_device(ref3, qry3, k3), 3)
    p3 = knnhip.last_plan()
    t = time.perf_counter()
    ix3 = knnhip.BlockIndex(ref3, nblocks=1024, iters=8)
    torch.cuda.synchronize()
    build3 = time.perf_counter() - t
    rows = []
    for prec in ("bf16", "fp32"):
        for P in (1, 2):
            msi, (d, i) = timed(torch, lambda: ix3.search(qry3, k3, probes=P, precision=prec), 3)
            ix3.search(qry3, k3, probes=P, precision=prec, timing=True)
            st = ix3.stage_ms()
            work = ix3.last_work()
            rows.append({"precision": work["precision"], "probes": P, "ms": msi, "queries_per_s": Q3 / msi * 1e3,
                         "speedup_vs_plain_fp32": ms3 / msi, "scan_ms": st["scan"],
                         "certificate_failures": knnhip.last_fallback_count(),
                         "equal_to_plain_fp32": bool(torch.equal(i, pi3) and torch.equal(d, pd3))})
    out["dense_clusters"] = {
        "workload": f"{R3} x {Q3}, dim {D3}, k {k3}, mixture of 1,024 Gaussians (sigma 0.05)",
        "plain_fp32_ms": ms3, "plain_fp32_queries_per_s": Q3 / ms3 * 1e3, "plain_kp": p3.get("kp"),
        "index_blocks": 1024, "index_build_s": build3, "index_searches": rows,
        "gap_vs_bound": gap_vs_bound(np, ix3, ref3, qry3, k3, 64)}
    ix3.close()
    del ref3, qry3, pd3, pi3
    torch.cuda.empty_cache()
    c5ix, c5ref, c5qry = c5
    out["router_workload_gap_vs_bound"] = gap_vs_bound(np, c5ix, c5ref, c5qry, C5["k"], 24)
    return out


def reference_form(knnhip, np):
    """The reference's own benchmark (code/test.cpp:295-298, 344-347; README.md:61-85): 16,384 refs,
    4,096 queries, dim 128, k 16, data 10*U[0,1) as host arrays, each of the three drop-in entry
    points timed per call host to host (H2D, search, D2H), next to the published per-call times
    (NVIDIA Titan X, README.md:76-78).  All three outputs are compared for equality."""
    import torch

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
    with open(os.path.join(ROOT, "tests", "golden", "c4_problems.json")) as f:
        P = json.load(f)["problems"]
    items = [(knnhip.generate_uniform(p["ref_nb"], p["dim"], seed=p["ref_seed"]),
              knnhip.generate_uniform(p["query_nb"], p["dim"], seed=p["query_seed"]), p["k"]) for p in P]
    torch.cuda.synchronize()
    flops = sum(2.0 * r.shape[1] * q.shape[1] * r.shape[0] for r, q, _ in items)
    nq = sum(q.shape[1] for _, q, _ in items)

    def once(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3, out

    seq_ms, seq = once(lambda: [knnhip.search_device(r, q, k) for r, q, k in items])
    for c in knnhip.batched_table():
        knnhip.batched_set_slice(c["dim"], c["k"], 0)
    unt_ms, out = once(lambda: knnhip.search_batched(items))
    equal = all(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]) for a, b in zip(out, seq))
    t = time.perf_counter()
    knnhip.batched_autotune(items)
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - t
    tun_ms, out = once(lambda: knnhip.search_batched(items))
    equal = equal and all(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]) for a, b in zip(out, seq))
    return {"workload": "512 problems (tests/golden/c4_problems.json), uniform [0,1)",
            "flops": flops, "queries": nq, "sequential_ms": seq_ms, "batched_untuned_ms": unt_ms,
            "batched_tuned_ms": tun_ms, "autotune_s": tune_s,
            "queries_per_s": nq / tun_ms * 1e3, "tflops": flops / tun_ms / 1e9,
            "fp32_peak_frac": flops / tun_ms / 1e9 / FP32_PEAK_TFLOPS,
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

    torch.cuda.synchronize()
    if args.copy_calibration_mib > 0:
        # one device copy of known size, the reference for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    head, ref, qry, dist, idx, ix, mlp = headline(args, knnhip, torch)
    R, Q, D, k = C5["R"], C5["Q"], C5["D"], C5["k"]

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    cpu, gate_q = None, 0
    if not args.no_cpu_baseline:
        ok, gate_q, cpu = cpu_baseline(args, knnhip, oracle, np, ref, qry, dist.cpu().numpy(),
                                       idx.cpu().numpy(), k)
        if not ok:
            print(json.dumps({"error": "parity gate failed", "queries_checked": gate_q}), flush=True)
            sys.exit(1)
    if not head["equal_to_plain_fp32"]:
        print(json.dumps({"error": "headline output differs from the plain FP32 search"}), flush=True)
        sys.exit(1)

    plain = precision = None
    if not args.no_plain_form or not args.no_precision_form:
        plain, pref, pqry, pdist, pidx, pms, pscan = plain_form(args, knnhip, torch)
        if not args.no_precision_form:
            precision = precision_form(args, knnhip, torch, np, pref, pqry, pdist, pidx, pms, pscan,
                                       (ix, ref, qry))
        del pref, pqry, pdist, pidx
    mlp.close()
    ix.close()
    del ref, qry
    torch.cuda.empty_cache()
    batched = None if args.no_batched_form else batched_form(knnhip, torch)
    reference = None if args.no_reference_form else reference_form(knnhip, np)

    out = {
        "metric": METRIC,
        "value": head["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if head["precision"] != "fp32" else "f32",
        "data": "synthetic (seeded Gaussian mixture, generated on device)",
        "config": {"workload": "learned-router form: exact kNN through the block index with BF16 products, "
                               "FP32 re-rank and the MLP router's first probe",
                   "ref_nb": R, "query_nb": Q, "dim": D, "k": k,
                   "mixture": "4,096 Gaussians, sigma 0.1, centres U[-1,1]^96", "blocks": C5["nblocks"],
                   "probes": head["probes"], "precision": head["precision"], "scan_units": head["units"],
                   "balls": head["balls"], "rows_per_query": head["rows_per_query"],
                   "router": [D, 256, 256, C5["nblocks"]], "layout": "dim x n row-major", "parallelism": "1 GPU"},
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "parity_gate_queries": gate_q,
        "speedup_vs_plain_fp32": head["speedup_vs_plain_fp32"],
        "plain_fp32_queries_per_s": head["plain_fp32_queries_per_s"],
        "plain_fp32_ms": head["plain_fp32_ms"],
        "index_build_s": head["index_build_s"],
        "router_train_s": head["router_train_s"],
        "setup_in_plain_searches": head["setup_in_plain_searches"],
        "equal_to_plain_fp32": head["equal_to_plain_fp32"],
        "certificate_failures": head["certificate_failures"],
        "stages_ms": head["stages_ms"],
        "forms": head["forms"],
        "router_table": head["router_table"],
        "plain_fp32": plain,
        "precision_form": precision,
        "batched_form": batched,
        "reference_form": reference,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
