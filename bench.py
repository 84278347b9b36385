#!/usr/bin/env python3
"""Benchmark: RC nets/s (Elmore + slew, 4 corners) and levelize edges/s.

One step = one solve of every net of the batching workload (4M heavy-tailed
RC trees, ~208M nodes) with the size-binned persistent kernel, inputs and
outputs resident in HBM.  Also measured on the same data: the plain
one-net-per-thread kernel, the per-net BFS ordering, and sink-first
levelization of a 10M-pin / 25M-arc DAG.  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RC nets/s (Elmore+slew, 4 corners) & levelize edges/s; batched vs plain"
PEAK_HBM_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nets", type=int, default=4_000_000)
    ap.add_argument("--method", default="batched", choices=["batched", "compensated"])
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-levelize", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--tuning", default=os.path.join(ROOT, "opentimer_amd", "tuning.json"))
    ap.add_argument("--only", default="", help="rct | plain | levelize (profiling runs)")
    return ap.parse_args()


def cpu_baseline(forest_host, topo, budget_s):
    """Oracle (CPU ot::Timer arithmetic, FP32) on a bounded prefix of the same
    nets, over the host cores with a thread pool of 32-net chunks
    (Timer::_build_rc_timing_tasks, timer.cpp:973-975)."""
    import concurrent.futures as cf

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    ns = topo["nodes_start"]
    pid, pres, cap, si = topo["pid"], topo["pres"], topo["cap"], topo["slew_in"]
    cores = min(16, len(os.sched_getaffinity(0)))

    def run(a, b):
        s, e = int(ns[a]), int(ns[b])
        sub_ns = (ns[a: b + 1] - s).astype(np.int32)
        O.rct_all(sub_ns, pid[s:e], pres[s:e], cap[4 * s: 4 * e], si[4 * a: 4 * b], "dfs_f32")

    # calibrate on a small prefix, then size the sample to the budget
    t0 = time.perf_counter()
    run(0, 20000)
    per_net = (time.perf_counter() - t0) / 20000
    sample = int(min(len(ns) - 1, max(20000, budget_s * cores / max(per_net, 1e-9))))
    chunk = 32 * 256
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        list(ex.map(lambda a: run(a, min(sample, a + chunk)), range(0, sample, chunk)))
    dt = time.perf_counter() - t0
    return {"value": sample / dt, "unit": "nets/s", "cores": cores, "kind": "port",
            "sample": f"first {sample} nets of the same workload "
                      f"({int(ns[sample])} nodes), oracle dfs_f32 + slew, {dt:.2f}s"}


def run_levelize(ctx, stream, sh, reps=5):
    import torch

    from opentimer_amd import workloads as W
    from opentimer_amd.levelize import Levelizer
    dcfg = W.config2()
    dag = W.generate_dag(dcfg, nthreads=16)
    csr = W.fanin_csr(dag)
    L = Levelizer(csr, ctx)
    for _ in range(2):
        L.run(stream=sh)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        L.run(stream=sh)
    torch.cuda.synchronize()
    lms = (time.perf_counter() - t0) * 1e3 / reps  # includes the per-batch host checks
    levels = L.num_ends - 1
    m = len(csr.fanin)
    lev = {"workload": dcfg.name, "pins": dcfg.num_pins, "arcs": m, "levels": levels,
           "ms": lms, "edges_per_s": m / (lms / 1e3), "us_per_level": lms * 1e3 / levels,
           "algorithmic_GBps": (20 * dcfg.num_pins + 12 * m + 4) / (lms / 1e3) / 1e9}
    # exactness on the full-size graph: levels vs the oracle
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    _, ends, lv = L.host()
    _, oends, olv = O.levelize(csr.n, csr.start, csr.fanin, csr.out_degree, csr.frontier)
    lev["levels_bitexact_vs_oracle"] = bool(np.array_equal(lv, olv) and np.array_equal(ends, oends))
    log(f"levelize: {lev}")
    return lev


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(nets shard trivially across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    from opentimer_amd import native as N
    from opentimer_amd import workloads as W
    from opentimer_amd.rct import RcSolver

    torch.cuda.init()
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    ctx = N.Context(0)
    if os.path.exists(args.tuning) and not args.autotune:
        import ctypes as C
        ctx.check(N.lib.otg_rct_tuning_load(ctx.ptr, args.tuning.encode(), None))

    if args.only == "levelize":
        print(json.dumps(run_levelize(ctx, stream, sh)))
        return
    if args.only == "copy":
        # a plain device copy of known size: the yardstick that converts the
        # FETCH_SIZE / WRITE_SIZE counters of the RC kernel into HBM bytes
        nbytes = 1 << 30
        a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"only": "copy", "bytes_read": nbytes, "bytes_written": nbytes,
                          "launches": args.steps}))
        return

    cfg = W.config1(num_nets=args.nets)
    t0 = time.perf_counter()
    forest = W.generate_rc(cfg, nthreads=16)
    log(f"generated {cfg.name}: {forest.num_nets} nets {forest.num_nodes} nodes "
        f"in {time.perf_counter() - t0:.1f}s")
    V, Nn = forest.num_nodes, forest.num_nets

    # ---- per-net BFS ordering on the GPU ------------------------------------
    s = RcSolver.from_edges(forest, ctx)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    import ctypes as C
    edges_d = torch.from_numpy(forest.edges).to(dev)
    roots_d = torch.from_numpy(forest.roots).to(dev)
    res_d = torch.from_numpy(forest.edge_res).to(dev)
    ncap_d = torch.from_numpy(forest.node_cap).to(dev)
    torch.cuda.synchronize()
    e0.record(stream)
    s.bfs(edges_d, roots_d, res_d, ncap_d, stream=sh)
    e1.record(stream)
    torch.cuda.synchronize()
    bfs_ms = e0.elapsed_time(e1)
    del edges_d, roots_d, res_d, ncap_d
    log(f"bfs ordering: {bfs_ms:.2f} ms")

    if args.autotune:
        s.alloc_outputs()
        best = N.RctTuning()
        ctx.check(N.lib.otg_rct_autotune(ctx.ptr, C.byref(s.tree()), C.byref(s.result()),
                                         N.BATCHED if args.method == "batched" else
                                         N.BATCHED_COMPENSATED, args.tuning.encode(),
                                         C.byref(best)))
        log(f"autotuned: ring {best.ring_slots} target {best.small_target} "
            f"grid/cu {best.grid_per_cu}")

    t0 = time.perf_counter()
    s.plan(args.method)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    stats = s.plan_stats(args.method)
    log(f"plan {plan_s * 1e3:.1f} ms: {stats}")
    s.alloc_outputs()

    def timed(method, steps, warmup):
        for _ in range(warmup):
            s.solve(method, stream=sh)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(steps):
            s.solve(method, stream=sh)
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / max(steps, 1)

    result = {}
    if args.only in ("", "rct"):
        ms = timed(args.method, args.steps, args.warmup)
        log(f"{args.method}: {ms:.3f} ms/step")
        result["ms"] = ms
    if args.only == "rct":
        print(json.dumps({"only": "rct", "ms_per_step": result["ms"]}))
        return
    plain_ms = None
    if not args.no_plain and args.only in ("", "plain"):
        plain_ms = timed("plain", args.plain_steps, 1)
        log(f"plain: {plain_ms:.3f} ms/step")
    if args.only == "plain":
        print(json.dumps({"only": "plain", "ms_per_step": plain_ms}))
        return

    # ---- sample parity check of the timed outputs ----------------------------
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    parity = None
    topo = None
    if args.only == "":
        s.solve(args.method, stream=sh)
        torch.cuda.synchronize()
        topo = s.host_topology()
        import oracle as O
        ns = topo["nodes_start"]
        rng = np.random.default_rng(0)
        pick = np.sort(rng.choice(Nn, min(Nn, 3000), replace=False))
        # always include the largest nets
        sizes = np.diff(ns)
        pick = np.unique(np.concatenate([pick, np.argsort(sizes)[-20:]]))
        errs = {}
        outs = {k: v for k, v in s.out.items()}
        for k in ("delay", "slew"):
            errs[k] = 0.0
        for a in pick:
            st, ed = int(ns[a]), int(ns[a + 1])
            sub = np.array([0, ed - st], np.int32)
            r = O.rct_all(sub, topo["pid"][st:ed], topo["pres"][st:ed], topo["cap"][4 * st: 4 * ed],
                          topo["slew_in"][4 * a: 4 * a + 4], "dfs_f32")
            for k in errs:
                g = outs[k][4 * st: 4 * ed].cpu().numpy().astype(np.float64)
                ref = r[k].astype(np.float64)
                fl = max(np.max(np.abs(ref)) * 1e-7, 1e-30)
                errs[k] = max(errs[k], float(np.max(np.abs(g - ref) / np.maximum(np.abs(ref), fl))))
        parity = {"nets_checked": int(len(pick)), "max_rel_err_vs_cpu_fp32": errs}
        log(f"parity sample: {parity}")

    # ---- levelization ---------------------------------------------------------
    lev = None
    if not args.no_levelize:
        lev = run_levelize(ctx, stream, sh)

    ms = result["ms"]
    alg_bytes = 120.0 * V + 20.0 * Nn
    achieved = alg_bytes / (ms / 1e3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "rct_batched_cfg1_traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(forest, topo if topo is not None else s.host_topology(), args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
    out = {
        "metric": METRIC,
        "value": Nn / (ms / 1e3),
        "unit": "nets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name} (heavy-tailed RC trees, BFS order, 4 corners)",
                   "nets": Nn, "nodes": V, "method": args.method,
                   "plan": {k: int(v) for k, v in stats.items()}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "algorithmic_bytes": alg_bytes},
        "cpu_baseline": cpu,
        "plain": None if plain_ms is None else {"ms_per_step": plain_ms,
                                                "nets_per_s": Nn / (plain_ms / 1e3),
                                                "batched_speedup": plain_ms / ms},
        "bfs_ordering_ms": bfs_ms,
        "plan_build_ms": plan_s * 1e3,
        "levelize": lev,
        "parity": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
