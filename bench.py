#!/usr/bin/env python3
"""Falconn++ on one MI355X: query throughput at recall@20 >= 0.9, FP32 plain
against BF16 / FP8 + FP32 refine, plus index build seconds.

Workload (default, the precision configuration the metric is quoted on):
N = 10,000,000 points x d = 128 (Gaussian mixture of 5,000 centres, within
sigma 0.3, lognormal(0, 0.25) norms), 10,000 queries from the same mixture;
n_tables = 64, n_proj = 128, iProbes = 2, alpha = 0.01, bucketLimit = 20,
qProbes = 2,000, k = 20, BF16 / FP8 refine of the top 80, index seed 42.
One step = one query() call over the whole batch with data, queries and
outputs already in HBM.  Recall@k is measured against an exact FP64 brute
force on the GPU.

JSON fields beside the contract's:
  variants      BF16 / FP8 + FP32 refine at the same qProbes: QPS, recall,
                score error against FP64, their own gather roofline
  sweep         per precision, recall@20 and QPS over a qProbes grid and the
                first qProbes reaching recall@20 >= 0.9 (the operating point)
  roofline      FP32 gather/dedup/rerank kernel: algorithmic bytes per launch
                over its HIP-event time on the library's stream, against 8 TB/s
  cpu_baseline  the CPU oracle (C++/OpenMP restatement of the reference) on
                this host's cores: query QPS on a sample of the queries
                against the same index tables, build seconds of one table per
                core (measured), and with --cpu-full-build every table
  extra         the N = 1e6 x d = 200 workload (FP32 and both variants)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "QPS at recall@20≥0.9, BF16/FP8+refine vs FP32 plain, 1 GPU; index build sec"
HBM_PEAK_GBS = 8000.0
CPU_THREADS = 16  # host cores of one GPU's share on the benchmark hosts
T0 = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def profiled_traffic(config, kernel):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, each scaled by a plain
    copy's reading) of a gather kernel as recorded under profiles/ for this
    workload; None when it was not profiled."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_gather_{config}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
            return int(line.rsplit(":", 1)[1])
    return None


def gather_roofline(kind, nq, qprobes, d, k, st, ms, refine=0, scale_bytes=0, config=None):
    """Algorithmic bytes of one gather launch over its time: per query
    8 qProbes + 4 C_tot + C_uniq (d s + s_scale) + 4 R d + 4 d + 4 k, with
    s = 4 (FP32), 2 (BF16), 1 (FP8) and the measured C_tot, C_uniq sums."""
    s = {"f32": 4, "bf16": 2, "fp8": 1}[kind]
    alg = (8 * qprobes * nq + 4 * st["ctot"] + st["cuniq"] * (d * s + scale_bytes)
           + nq * (4 * refine * d + 4 * d + 4 * k))
    achieved = alg / (ms * 1e-3) / 1e9
    kernel = "k_gather_score_f32" if kind == "f32" else f"k_gather_score_lp_{kind}"
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(config, kernel) if config else None,
            "kernel": kernel, "kernel_ms": round(ms, 3), "alg_bytes_per_launch": int(alg)}


class Workload:
    """One configuration built on the GPU: data, queries, index."""

    def __init__(self, name, args, E, synth, torch):
        self.name, self.E, self.torch = name, E, torch
        t = time.perf_counter()
        self.X, self.Q, self.cfg = synth.make_config(name, args.scale)
        if args.qprobes and name == args.config:
            self.cfg["qProbes"] = args.qprobes
        self.n, self.d = self.X.shape
        self.nq = len(self.Q)
        self.k = self.cfg["k"]
        log(f"{name}: data {self.n}x{self.d}, {self.nq} queries generated in {time.perf_counter() - t:.1f}s")
        self.dX = torch.from_numpy(self.X).cuda()
        self.dQ = torch.from_numpy(self.Q).cuda()
        torch.cuda.synchronize()
        self.free0, self.total_mem = torch.cuda.mem_get_info()
        c = self.cfg
        self.idx = E.Index(self.n, self.d)
        self.idx.setIndexParam(c["L"], c["n_proj"], c["bucketLimit"], c["alpha"], c["iProbes"], CPU_THREADS,
                               synth.INDEX_SEED)
        t0 = time.perf_counter()
        self.idx.build_device(self.dX.data_ptr(), self.d)
        torch.cuda.synchronize()
        self.build_s = time.perf_counter() - t0
        self.free_built = torch.cuda.mem_get_info()[0]
        sh, bst = self.idx.shape(), self.idx.stats()
        log(f"{name}: build {self.build_s:.3f}s, {sh['n_ids']} ids, hash ties {bst['hash_ties']}, "
            f"bucket ties {bst['bucket_ties']}")
        self.out = torch.zeros((self.nq, self.k), dtype=torch.int32, device="cuda")
        self.scores = torch.zeros((self.nq, self.k), dtype=torch.float32, device="cuda")
        if args.threads:
            self.idx.set_tuning(E.TUNE_GATHER_THREADS, args.threads)
        if args.persistent >= 0:
            self.idx.set_tuning(E.TUNE_PERSISTENT, args.persistent)
        self.truth = None

    def call(self):
        # on torch's stream: calls queue back to back, results ordered with
        # torch's own work
        self.idx.query_device(self.dQ.data_ptr(), self.d, self.nq, self.k, self.out.data_ptr(),
                              self.scores.data_ptr(), self.torch.cuda.current_stream().cuda_stream)

    def timed(self, steps, warmup, events=False):
        """QPS and ms per step of `steps` calls after `warmup` untimed ones;
        with events the per-stage HIP-event times of every step."""
        self.idx.set_profiling(1 if events else 0)
        for _ in range(max(1, warmup)):
            self.call()
        self.torch.cuda.synchronize()
        stage = []
        t0 = time.perf_counter()
        for _ in range(steps):
            self.call()
            if events:
                stage.append(self.idx.timings())
        self.torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        self.idx.set_profiling(0)
        return self.nq * steps / dt, dt / steps * 1e3, stage

    def gather_ms(self):
        """The gather kernel's HIP-event time of one more call."""
        self.idx.set_profiling(1)
        self.call()
        ms = float(self.idx.timings()["gather"])
        self.idx.set_profiling(0)
        return ms

    def stage_times(self, calls=3):
        """Per-stage HIP-event times of `calls` more calls (after the timed
        loop, which runs without events: reading them synchronises each step)."""
        self.idx.set_profiling(1)
        out = []
        for _ in range(calls):
            self.call()
            out.append(self.idx.timings())
        self.idx.set_profiling(0)
        return out

    def phases(self):
        self.idx.set_profiling(2)
        self.call()
        ph = self.idx.timings()
        self.idx.set_profiling(0)
        return {key: round(float(ph[key]), 2) for key in ("ph_gather_us", "ph_compact_us", "ph_score_us",
                                                          "ph_topk_us")}

    def answer(self):
        return self.out.cpu().numpy()

    def compute_truth(self, seeds):
        t = time.perf_counter()
        self.truth, _ = self.E.bruteforce_f64_device(self.dX.data_ptr(), self.d, self.n, self.d,
                                                     self.dQ.data_ptr(), self.d, self.nq, self.k, seeds)
        log(f"{self.name}: FP64 brute force over {self.nq} queries in {time.perf_counter() - t:.1f}s")

    def recall(self):
        return round(float(self.E.recall(self.truth, self.answer(), self.k)), 4) if self.truth is not None else None

    def mark_fp32_workspace(self):
        """Device memory taken by the FP32 calls (the query workspace and the
        output arrays), before any reduced-precision copy exists."""
        self.free_fp32 = self.torch.cuda.mem_get_info()[0]

    def memory(self):
        free_end = self.torch.cuda.mem_get_info()[0]
        m = {"device_total_gb": round(self.total_mem / 1e9, 1),
             "data_and_queries_gb": round((self.X.nbytes + self.Q.nbytes) / 1e9, 2),
             "index_gb": round((self.free0 - self.free_built) / 1e9, 2)}
        if getattr(self, "free_fp32", None) is not None:
            m["query_workspace_fp32_gb"] = round((self.free_built - self.free_fp32) / 1e9, 2)
        m["query_workspace_and_copies_gb"] = round((self.free_built - free_end) / 1e9, 2)
        m["used_at_end_gb"] = round((self.total_mem - free_end) / 1e9, 2)
        return m

    def release(self):
        self.idx.close()
        del self.dX, self.dQ, self.out, self.scores
        self.torch.cuda.empty_cache()


PRECISIONS = (("fp32", 0), ("bf16_refine", 1), ("fp8_refine", 2))


def run_precisions(w, args, config_key, with_truth):
    """FP32 at the configured qProbes (the headline), then BF16 and FP8 +
    refine at the same qProbes."""
    E, c = w.E, w.cfg
    w.idx.set_qProbes(c["qProbes"])
    qps, ms, _ = w.timed(args.steps, args.warmup)
    stage = w.stage_times()
    w.mark_fp32_workspace()
    st = w.idx.stats()
    log(f"{w.name} fp32: {ms:.2f} ms/step, {qps:.0f} QPS, C_tot/q {st['ctot'] / w.nq:.0f}, "
        f"C_uniq/q {st['cuniq'] / w.nq:.0f}, tie-replayed queries {st['query_ties']}")
    gms = float(sum(s["gather"] for s in stage) / len(stage))
    res = {"qps": qps, "ms": ms, "stats": st,
           "stage_ms": {key: round(float(sum(s[key] for s in stage) / len(stage)), 3)
                        for key in ("hash", "probe", "replay", "gather")},
           "roofline": gather_roofline("f32", w.nq, c["qProbes"], w.d, w.k, st, gms, config=config_key)}
    res["stage_ms"].update(w.phases())
    if with_truth:
        w.compute_truth(w.answer())
        res["recall"] = w.recall()
        log(f"{w.name} fp32: recall@{w.k} = {res['recall']}")
    variants = {}
    if not args.no_variants:
        refine = int(c.get("refine", 4 * w.k))
        for name, mode in PRECISIONS[1:]:
            w.idx.set_precision(mode, refine)
            # the FP8 copy with one scale per row where the configuration
            # asks for it (else one shared scale when the rows allow it)
            w.idx.set_tuning(E.TUNE_FP8_PER_ROW, 1 if (mode == 2 and c.get("fp8_per_row")) else 0)
            vq, vms, _ = w.timed(args.steps, args.warmup)
            vst = w.idx.stats()
            shared = w.idx.fp8_shared_scale() if mode == 2 else 0.0
            sb = 4 if (mode == 2 and shared == 0.0) else 0
            kind = "bf16" if mode == 1 else "fp8"
            v = {"value": round(vq, 1), "ms_per_step": round(vms, 3), "refine": refine,
                 "speedup_vs_fp32": round(vq / qps, 3),
                 "roofline": gather_roofline(kind, w.nq, c["qProbes"], w.d, w.k, vst, w.gather_ms(), refine, sb,
                                             config=config_key),
                 "phases_us": w.phases()}
            if mode == 2:
                v["fp8_scale"] = "shared" if shared > 0 else "per row"
            if w.truth is not None:
                v["recall_at_k"] = w.recall()
                v["score_rel_err"] = approx_score_error(w.X, w.Q, w.idx.center(), w.truth, mode, shared)
            variants[name] = v
            log(f"{w.name} {name}: {v}")
        w.idx.set_precision(0, 0)
    res["variants"] = variants
    return res


def run_sweep(w, args):
    """Recall@k and QPS over a qProbes grid, per precision; the operating
    point of a precision is the first qProbes reaching recall@k >= 0.9."""
    grid = [int(x) for x in args.sweep.split(",") if x]
    refine = int(w.cfg.get("refine", 4 * w.k))
    out = {}
    for name, mode in PRECISIONS:
        w.idx.set_precision(mode, refine if mode else 0)
        pts = []
        op = None
        for qp in grid:
            w.idx.set_qProbes(qp)
            qps, ms, _ = w.timed(max(2, args.steps // 2), 1)
            r = w.recall()
            pts.append({"qprobes": qp, "recall": r, "qps": round(qps, 1)})
            log(f"sweep {name} qProbes={qp}: recall@{w.k} {r}, {qps:.0f} QPS")
            if r >= 0.9:
                op = pts[-1]
                break
        out[name] = {"points": pts, "operating_point": op}
    w.idx.set_precision(0, 0)
    w.idx.set_qProbes(w.cfg["qProbes"])
    base = out["fp32"]["operating_point"]
    for name, _ in PRECISIONS[1:]:
        op = out[name]["operating_point"]
        out[name]["speedup_vs_fp32_at_recall_0.9"] = (round(op["qps"] / base["qps"], 3)
                                                      if op and base else None)
    return out


def _bf16(a):
    import numpy as np
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def _e4m3(a):
    """Round to the nearest e4m3 value (3 mantissa bits, min normal 2^-6), ties to even."""
    import numpy as np
    a = np.asarray(a, np.float64)
    e = np.floor(np.log2(np.maximum(np.abs(a), 2.0 ** -6)))
    step = 2.0 ** (e - 3)
    return np.round(a / step) * step


def approx_score_error(X, Q, center, truth, mode, shared, sample=200):
    """Relative error of the reduced-precision scores of the true top-k rows
    against exact FP64 scores q.(x - c), on a sample of queries (host
    emulation of the engine's rounding; shared = the FP8 scale every row
    uses, 0 when rows keep their own)."""
    import numpy as np
    errs = []
    for j in range(min(sample, len(Q))):
        rows = (X[truth[j]] - center).astype(np.float32)
        q = Q[j].astype(np.float32)
        exact = rows.astype(np.float64) @ q.astype(np.float64)
        if mode == 1:
            approx = _bf16(rows).astype(np.float64) @ _bf16(q).astype(np.float64)
        else:
            sr = np.full(len(rows), shared) if shared > 0 else np.maximum(np.abs(rows).max(1), 1e-30) / 240.0
            sq = max(float(np.abs(q).max()), 1e-30) / 240.0
            approx = (_e4m3(rows / sr[:, None]) @ _e4m3(q / sq)) * sr * sq
        errs.append(np.abs(approx - exact) / np.maximum(np.abs(exact), 1e-12))
    e = np.concatenate(errs)
    return {"mean": round(float(e.mean()), 6), "p99": round(float(np.quantile(e, 0.99)), 6)}


def cpu_baseline(args, w, gpu_ans):
    """The oracle on CPU_THREADS host threads: its query path on the engine's
    own tables for a sample of the queries, and its build of one table per
    thread (every table with --cpu-full-build)."""
    import numpy as np

    from falconnpp_amd import synth
    from oracle import pyoracle as P

    cores = min(CPU_THREADS, os.cpu_count() or 1)
    c = w.cfg
    center = w.idx.center()
    off, ids = w.idx.buckets()
    o = P.OracleIndex(w.n, w.d)
    o.setIndexParam(c["L"], c["n_proj"], c["bucketLimit"], c["alpha"], c["iProbes"], cores, synth.INDEX_SEED)
    o.load_index(w.X, center, off, ids)
    o.set_qProbes(c["qProbes"])
    del off, ids
    m = min(args.cpu_queries, w.nq)
    o.query(w.Q[:8], w.k)  # warm-up
    t = time.perf_counter()
    topk, _, _ = o.query(w.Q[:m], w.k)
    qdt = time.perf_counter() - t
    agree = float(np.mean([len(set(a.tolist()) & set(b.tolist())) / w.k for a, b in zip(topk, gpu_ans[:m])]))
    log(f"cpu query: {m} queries in {qdt:.2f}s on {cores} threads ({m / qdt:.1f} QPS), top-k agreement {agree:.4f}")
    del o
    lt = c["L"] if args.cpu_full_build else min(cores, c["L"])
    log(f"cpu build: {lt} of {c['L']} tables on {cores} threads ...")
    ob = P.OracleIndex(w.n, w.d)
    ob.setIndexParam(lt, c["n_proj"], c["bucketLimit"], c["alpha"], c["iProbes"], cores, synth.INDEX_SEED)
    ob.build(w.X, center)
    bs = ob.build_seconds
    del ob
    # tables are independent and the reference builds them in parallel
    # (src/FalconnPP.cpp:62-64): L tables take ceil(L / cores) rounds of a
    # round of `cores` tables
    rounds = -(-c["L"] // cores)
    est = bs if lt == c["L"] else bs * rounds
    log(f"cpu build: {lt} tables in {bs:.2f}s -> {est:.1f}s for {c['L']} tables")
    res = {"value": round(m / qdt, 2), "unit": "queries/s", "cores": cores, "kind": "port",
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
           "sample": (f"oracle query path on the engine's tables for {m} of {w.nq} queries at qProbes "
                      f"{c['qProbes']}; build of {lt} of {c['L']} tables on {cores} threads"
                      + ("" if lt == c["L"] else f" (one table per thread), x{rounds} rounds for all tables")),
           "build_s_measured": round(bs, 3), "build_tables_measured": lt, "build_s": round(est, 2),
           "build_speedup_gpu": round(est / w.build_s, 1), "topk_agreement": round(agree, 4)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--scale", type=float, default=1.0, help="shrink N and nq (testing only)")
    ap.add_argument("--qprobes", type=int, default=0, help="override qProbes")
    ap.add_argument("--threads", type=int, default=0, help="gather workgroup size (0: engine default)")
    ap.add_argument("--persistent", type=int, default=-1, help="persistent gather launch 0/1 (-1: engine default)")
    ap.add_argument("--sweep", default="2,5,10,20,30,50,75,100,150,200,300,500,1000,2000",
                    help="qProbes grid of the recall sweep")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-recall", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip the BF16/FP8 + refine runs")
    ap.add_argument("--no-extra", action="store_true", help="skip the N = 1e6 x d = 200 workload")
    ap.add_argument("--cpu-queries", type=int, default=2000)
    ap.add_argument("--cpu-full-build", action="store_true", help="time the CPU build of every table")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs on one GPU (sharding over several is described "
              "in the design notes, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_mib": args.calibrate_copy, "copies": 3}))
        return

    from falconnpp_amd import engine as E
    from falconnpp_amd import synth

    # one stream of torch's own (default priority) for every call: the
    # engine's calls on it are asynchronous and ordered with torch's work
    torch.cuda.set_stream(torch.cuda.Stream())
    if "qp_range" in synth.CONFIGS[args.config]:
        serve_batched(args, E, synth, torch)
        return

    profiled = args.config if args.scale == 1.0 and not args.qprobes else None
    w = Workload(args.config, args, E, synth, torch)
    r = run_precisions(w, args, profiled, not args.no_recall)
    sweep = None
    if not args.no_sweep and w.truth is not None:
        sweep = run_sweep(w, args)
    memory = w.memory()
    log(f"memory: {memory}")
    cpu = None
    if not args.no_cpu:
        w.idx.set_qProbes(w.cfg["qProbes"])
        w.call()
        cpu = cpu_baseline(args, w, w.answer())
    c = w.cfg
    workload = (f"{args.config}: N={w.n} x d={w.d} "
                + ("gaussian mixture" if c["kind"] == "mixture" else "lognormal sphere")
                + f", Q={w.nq}, L={c['L']}, n_proj={c['n_proj']}, iProbes={c['iProbes']}, alpha={c['alpha']}, "
                  f"bucketLimit={c['bucketLimit']}, qProbes={c['qProbes']}, k={w.k}, plain FP32 rerank; variants "
                  f"BF16/FP8 + FP32 refine of the top {int(c.get('refine', 4 * w.k))}")
    build_s = w.build_s
    st = r["stats"]
    w.release()
    extra = None
    if not args.no_extra and args.config != "c2" and args.scale == 1.0:
        w2 = Workload("c2", args, E, synth, torch)
        r2 = run_precisions(w2, args, "c2", not args.no_recall)
        extra = {"workload": "c2: N=1000000 x d=200 gaussian mixture, Q=10000, L=100, n_proj=256, iProbes=3, "
                             "alpha=0.005, bucketLimit=20, qProbes=4000, k=20",
                 "value": round(r2["qps"], 1), "unit": "queries/s", "ms_per_step": round(r2["ms"], 3),
                 "recall_at_k": r2.get("recall"), "build_s": round(w2.build_s, 3), "roofline": r2["roofline"],
                 "variants": r2["variants"]}
        w2.release()
    line = {
        "metric": METRIC,
        "value": round(r["qps"], 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "global_batch": w.nq, "parallelism": "single GPU"},
        "build_s": round(build_s, 3),
        "recall_at_k": r.get("recall"),
        "stage_ms": r["stage_ms"],
        "candidates_per_query": {"c_tot": round(st["ctot"] / w.nq, 1), "c_uniq": round(st["cuniq"] / w.nq, 1)},
        "memory": memory,
        "roofline": r["roofline"],
        "variants": r["variants"],
        "sweep": sweep,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


def serve_batched(args, E, synth, torch):
    """Batched serving: one call over all queries, each with its own qProbes
    and k, captured in a hipGraph; the gather's launch shape picked by the
    engine's autotuner on the batch.  Reports QPS and per-query latency."""
    import numpy as np

    t = time.perf_counter()
    X, Q, cfg = synth.make_config(args.config, args.scale)
    n, d = X.shape
    nq = len(Q)
    qps_q, ks = synth.per_query_params(cfg, nq)
    kmax = int(max(cfg["k_choices"]))
    log(f"data {n}x{d}, {nq} queries generated in {time.perf_counter() - t:.1f}s")
    dX = torch.from_numpy(X).cuda()
    dQ = torch.from_numpy(Q).cuda()
    torch.cuda.synchronize()
    idx = E.Index(n, d)
    idx.setIndexParam(cfg["L"], cfg["n_proj"], cfg["bucketLimit"], cfg["alpha"], cfg["iProbes"], CPU_THREADS,
                      synth.INDEX_SEED)
    t0 = time.perf_counter()
    idx.build_device(dX.data_ptr(), d)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    log(f"build {build_s:.3f}s")
    idx.set_tuning(E.TUNE_VERBOSE, 1)

    def tune():
        log("autotune on the batch ...")
        t = idx.autotune_batch(dQ.data_ptr(), d, nq, qps_q, ks, kmax, torch.cuda.current_stream().cuda_stream)
        if args.persistent >= 0:
            idx.set_tuning(E.TUNE_PERSISTENT, args.persistent)
            t["persistent"] = args.persistent
        log(f"autotune: {t}")
        return t

    tuned = tune()
    idx.set_tuning(E.TUNE_GRAPH, 1)
    out = torch.zeros((nq, kmax), dtype=torch.int32, device="cuda")
    sc = torch.zeros((nq, kmax), dtype=torch.float32, device="cuda")

    def call():
        idx.query_batch_device(dQ.data_ptr(), d, nq, qps_q, ks, kmax, out.data_ptr(), sc.data_ptr(),
                               torch.cuda.current_stream().cuda_stream)

    def timed_run():
        """Warm-up, the timed steps, then one more call with per-query
        stamps; returns (QPS, ms per step, stage ms, latency)."""
        idx.set_profiling(0)
        for _ in range(max(1, args.warmup)):
            call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            call()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        # stage times from two more calls with events (reading them
        # synchronises, so the timed loop runs without)
        idx.set_profiling(1)
        stage = []
        for _ in range(2):
            call()
            stage.append(idx.timings())
        stages = {key: round(float(np.mean([s[key] for s in stage])), 3)
                  for key in ("hash", "probe", "replay", "gather")}
        idx.set_profiling(2)
        call()
        service, done = idx.query_times(nq)
        idx.set_profiling(1)
        pre_ms = stages["hash"] + stages["probe"] + stages["replay"]
        lat = {"service_us_p50": round(float(np.percentile(service, 50)), 1),
               "service_us_p99": round(float(np.percentile(service, 99)), 1),
               "completion_ms_p50": round(pre_ms + float(np.percentile(done, 50)) / 1e3, 3),
               "completion_ms_p99": round(pre_ms + float(np.percentile(done, 99)) / 1e3, 3)}
        return nq * args.steps / dt, dt / args.steps * 1e3, stages, lat

    qps, ms_step, stages, lat = timed_run()
    st = idx.stats()
    log(f"batch {ms_step:.1f} ms, {qps:.0f} QPS, C_tot/q {st['ctot'] / nq:.0f}, "
        f"C_uniq/q {st['cuniq'] / nq:.0f}, tie-replayed queries {st['query_ties']}")
    log(f"latency: {lat}")
    res = {}
    truth_by_k = {}

    def recall_by_k(ans):
        return {kv: round(float(E.recall(truth, ans[sel, :kv], kv)), 4) for kv, (sel, truth) in truth_by_k.items()}

    if not args.no_recall:
        # recall@k_q against FP64 brute force, per k on a sample of at most
        # 2000 queries each, seeded with the answer's first k ids
        ans = out.cpu().numpy()
        t = time.perf_counter()
        for kv in cfg["k_choices"]:
            sel = np.where(ks == kv)[0][:2000]
            if len(sel) == 0:
                continue
            dQs = torch.from_numpy(np.ascontiguousarray(Q[sel])).cuda()
            truth, _ = E.bruteforce_f64_device(dX.data_ptr(), d, n, d, dQs.data_ptr(), d, len(sel), int(kv),
                                               np.ascontiguousarray(ans[sel, :kv]))
            truth_by_k[int(kv)] = (sel, truth)
        res["recall_at_k_by_k"] = recall_by_k(ans)
        log(f"recall@k by k = {res['recall_at_k_by_k']} (FP64 brute force {time.perf_counter() - t:.1f}s)")
    variants = {}
    if not args.no_variants:
        for name, mode in (("bf16_refine", E.PREC_BF16), ("fp8_refine", E.PREC_FP8)):
            idx.set_precision(mode, 0)
            vtuned = tune()
            vq, vms, vst, vlat = timed_run()
            v = {"value": round(vq, 1), "ms_per_step": round(vms, 3), "refine": "min(256, 4 k_q)",
                 "speedup_vs_fp32": round(vq / qps, 3), "autotune": vtuned, "stage_ms": vst, "latency": vlat}
            if truth_by_k:
                v["recall_at_k_by_k"] = recall_by_k(out.cpu().numpy())
            variants[name] = v
            log(f"{name}: {v}")
        idx.set_precision(E.PREC_FP32, 0)
    line = {
        "metric": "QPS of one batched call, per-query qProbes U[100,8000] and k in {1,10,100}; "
                  "p50/p99 per-query latency",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"{args.config}: N={n} x d={d} gaussian mixture, Q={nq}, L={cfg['L']}, "
                                f"n_proj={cfg['n_proj']}, iProbes={cfg['iProbes']}, alpha={cfg['alpha']}, "
                                f"bucketLimit={cfg['bucketLimit']}, per-query qProbes in {list(cfg['qp_range'])}, "
                                f"k in {list(cfg['k_choices'])}, gather launch autotuned on the batch, hipGraph"),
                   "global_batch": nq, "parallelism": "single GPU"},
        "build_s": round(build_s, 3),
        **res,
        "latency": lat,
        "autotune": tuned,
        "stage_ms": stages,
        "candidates_per_query": {"c_tot": round(st["ctot"] / nq, 1), "c_uniq": round(st["cuniq"] / nq, 1)},
        "variants": variants,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
