#!/usr/bin/env python3
"""Benchmark: RC nets/s (Elmore + slew, 4 corners) and levelize edges/s.

One step = one solve of every net of the batching workload (4M heavy-tailed
RC trees, ~208M nodes) with the size-binned persistent kernels (packed groups,
streamed nets), inputs and outputs resident in HBM, timed with HIP events on
the launch stream.  The headline method is the FP32-compensated one (--method);
the FP32 method, the plain one-net-per-thread kernel, the per-net BFS
ordering and sink-first levelization of a 10M-pin / 25M-arc DAG are measured
beside it.  Prints ONE JSON line.
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
    ap.add_argument("--method", default="compensated", choices=["batched", "compensated"],
                    help="headline method (the other batched method is timed beside it)")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-levelize", action="store_true")
    ap.add_argument("--no-fprop", action="store_true")
    ap.add_argument("--no-timer", action="store_true",
                    help="skip the device timing update leg (otg_timer on the cfg2 design)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="solve with the built-in default tuning")
    ap.add_argument("--autotune", action="store_true",
                    help="search each method's tuning on this forest (otg_rct_autotune) instead "
                         "of loading the shipped MI355X tuning for the workload")
    ap.add_argument("--tuning", default="",
                    help="a tuning file (otg_rct_tuning_load) used for both methods")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the host-pointer drop-in leg (otg_rct_bfs_and_compute at cfg1)")
    ap.add_argument("--only", default="",
                    help="rct | plain | levelize | fprop | timer | copy (profiling runs) | precision (cfg3 report) "
                         "| incremental (cfg4 report) | spef (SPEF loader + GPU solve)")
    ap.add_argument("--spef-nets", type=int, default=400_000,
                    help="nets of the cfg1 distribution written as SPEF for --only spef")
    ap.add_argument("--precision-nets", type=int, default=20_000,
                    help="nets of the cfg3 distribution for --only precision")
    ap.add_argument("--precision-ref", choices=("auto", "host", "device"), default="auto",
                    help="FP64 reference of --only precision: the host oracle at every node, or "
                         "the device FP64 kernel over batches (auto: device above 20k nets)")
    ap.add_argument("--precision-batch", type=int, default=16_384,
                    help="nets per batch of a device-reference precision run")
    return ap.parse_args()


def kernel_sources_hash():
    """sha256 over the HIP sources of the P1 solve (what a traffic
    measurement depends on)."""
    import hashlib
    h = hashlib.sha256()
    for f in ("otg_rct.hip", "otg_plan.hip", "otg_internal.hpp"):
        with open(os.path.join(ROOT, "opentimer_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def run_dropin(ctx, forest, method):
    """The host-pointer drop-in (otg_rct_bfs_and_compute, what the bridge's
    rct_bfs_and_compute_cuda calls) on the whole cfg1 forest: one call from
    pageable host arrays, end to end, with its stage breakdown
    (otg_rct_last_breakdown), after one warm-up call."""
    import ctypes as C

    from opentimer_amd import native as N
    V, Nn = forest.num_nodes, forest.num_nets
    out = {k: np.empty(4 * V, np.float32) for k in ("load", "delay", "ldelay", "impulse")}
    n2b = np.empty(V, np.int32)
    pid = np.empty(V, np.int32)
    m = {"batched": N.BATCHED, "compensated": N.BATCHED_COMPENSATED}[method]
    res = {}
    for rep in range(2):
        edges = forest.edges.copy()
        a = N.RctEdgeArray(num_nets=Nn, total_num_nodes=V, total_num_edges=V - Nn,
                           rct_nodes_start=N.addr(forest.nodes_start), rct_edges=N.addr(edges),
                           rct_roots=N.addr(forest.roots), rct_edges_res=N.addr(forest.edge_res),
                           rct_nodes_cap=N.addr(forest.node_cap), rct_node2bfs_order=N.addr(n2b),
                           rct_pid=N.addr(pid), load=N.addr(out["load"]),
                           delay=N.addr(out["delay"]), ldelay=N.addr(out["ldelay"]),
                           impulse=N.addr(out["impulse"]))
        t0 = time.perf_counter()
        ctx.check(N.lib.otg_rct_bfs_and_compute(ctx.ptr, C.byref(a), m))
        wall = (time.perf_counter() - t0) * 1e3
        b = (C.c_double * 8)()
        ctx.check(N.lib.otg_rct_last_breakdown(ctx.ptr, b))
        names = ("alloc", "h2d", "bfs", "plan", "solve", "d2h", "free", "total")
        res = {"ms": wall, "breakdown_ms": {k: float(b[i]) for i, k in enumerate(names)}}
    h2d = 4 * (Nn + 1) + 8 * (V - Nn) + 4 * Nn + 4 * (V - Nn) + 16 * V
    d2h = 8 * (V - Nn) + 8 * V + 4 * 16 * V
    res.update({"method": method, "nets": Nn, "nodes": V, "nets_per_s": Nn / (res["ms"] / 1e3),
                "h2d_bytes": h2d, "d2h_bytes": d2h,
                "h2d_GBps": h2d / (res["breakdown_ms"]["h2d"] / 1e3) / 1e9,
                "d2h_GBps": d2h / (res["breakdown_ms"]["d2h"] / 1e3) / 1e9,
                "what": "otg_rct_bfs_and_compute from pageable host arrays: allocate, copy in, "
                        "BFS-order, plan, solve, copy out, free (second of two calls)"})
    log(f"drop-in: {res}")
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Threads for the CPU baselines: the host cores this process may run on,
    capped by OMP_NUM_THREADS when it is set (the share of the host's cores
    that goes with one GPU on a shared box)."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = min(aff, int(env)) if env.isdigit() and int(env) > 0 else aff
    return n, aff


def cpu_baseline(forest_host, topo, budget_s):
    """Oracle (CPU ot::Timer arithmetic, FP32 DFS + slew) over the host cores,
    a thread pool over chunks of 32*256 nets (Timer::_build_rc_timing_tasks
    runs 32-net chunks, timer.cpp:973-975).  The whole workload is solved
    repeatedly until `budget_s` has passed; the median pass is reported."""
    import concurrent.futures as cf

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    ns = topo["nodes_start"]
    pid, pres, cap, si = topo["pid"], topo["pres"], topo["cap"], topo["slew_in"]
    cores, aff = host_threads()
    nnets = len(ns) - 1

    def run(a, b):
        s, e = int(ns[a]), int(ns[b])
        sub_ns = (ns[a: b + 1] - s).astype(np.int32)
        O.rct_all(sub_ns, pid[s:e], pres[s:e], cap[4 * s: 4 * e], si[4 * a: 4 * b], "dfs_f32")

    chunk = 32 * 256
    times = []
    t_all = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        while True:
            t0 = time.perf_counter()
            list(ex.map(lambda a: run(a, min(nnets, a + chunk)), range(0, nnets, chunk)))
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_all >= budget_s or len(times) >= 50:
                break
    med = float(np.median(times))
    return {"value": nnets / med, "unit": "nets/s", "cores": cores, "kind": "port",
            "cpu": cpu_model(), "affinity_cores": aff,
            "sample": f"the whole workload ({nnets} nets, {int(ns[-1])} nodes), "
                      f"{len(times)} passes in {sum(times):.1f}s, median pass {med:.3f}s; "
                      f"oracle FP32 DFS + slew, {cores} threads"}


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
    t0 = time.perf_counter()
    _, oends, olv = O.levelize(csr.n, csr.start, csr.fanin, csr.out_degree, csr.frontier)
    serial_ms = (time.perf_counter() - t0) * 1e3
    lev["levels_bitexact_vs_oracle"] = bool(np.array_equal(lv, olv) and np.array_equal(ends, oends))
    # the host-pointer drop-in (otg_toposort_compute, what toposort_compute_cuda
    # is replaced by): copies in, levelization, copies out; second of two calls
    from opentimer_amd.levelize import toposort_compute
    for _ in range(2):
        t0 = time.perf_counter()
        _, hends, hlv = toposort_compute(ctx, csr)
        host_ms = (time.perf_counter() - t0) * 1e3
    lev["dropin"] = {"ms": host_ms, "levels_bitexact_vs_oracle":
                     bool(np.array_equal(hlv, olv) and np.array_equal(hends, oends)),
                     "what": "otg_toposort_compute from pageable host arrays (fanin CSR, fanout "
                             "counts, seeded frontier in; frontiers, ends and per-pin levels out)"}
    # CPU baselines: the sink-first Kahn serially (the run above) and over the
    # host cores (median of 3 runs, levels checked against the serial run)
    cores, aff = host_threads()
    par = []
    for _ in range(3):
        t0 = time.perf_counter()
        _, pends, plv = O.levelize_par(csr.n, csr.start, csr.fanin, csr.out_degree, csr.frontier,
                                       cores)
        par.append((time.perf_counter() - t0) * 1e3)
    par_ms = float(np.median(par))
    lev["cpu_baseline"] = {"serial_ms": serial_ms, "parallel_ms": par_ms, "threads": cores,
                           "affinity_cores": aff, "cpu": cpu_model(),
                           "parallel_levels_equal": bool(np.array_equal(plv, olv)
                                                         and np.array_equal(pends, oends)),
                           "gpu_speedup_vs_parallel": par_ms / lms}
    log(f"levelize: {lev}")
    return lev


def run_precision(ctx, nets):
    """cfg3 precision report: every output of the plain kernel (reference
    arithmetic), the FP32 batched method and the compensated method against
    the FP64 oracle over every node of `nets` nets of the cfg3 distribution
    (50% chains of 1,000-20,000 nodes, 50% balanced trees of 64-1,024 nodes,
    R 1e-3..1e4 ohm, C 1e-3..1e2 fF).  Relative errors use a floor of 1e-7 x
    the output's largest magnitude (exact zeros at drivers); impulse is
    compared against 1e-6 x the largest beta (it is a difference that can
    vanish)."""
    import concurrent.futures as cf

    import torch

    from opentimer_amd import workloads as W
    from opentimer_amd.rct import RcSolver
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cfg = W.config3(num_nets=nets)
    t0 = time.perf_counter()
    forest = W.generate_rc(cfg, nthreads=16)
    log(f"generated {cfg.name} x{nets}: {forest.num_nodes} nodes in {time.perf_counter() - t0:.1f}s")
    s = RcSolver.from_edges(forest, ctx)
    del forest
    topo = s.host_topology()
    ns = topo["nodes_start"]
    V, Nn = int(ns[-1]), len(ns) - 1
    outs = ("load", "delay", "ldelay", "beta", "impulse", "slew")
    ref = {k: np.empty(4 * V, np.float64) for k in outs}

    def run(a, b):
        st, ed = int(ns[a]), int(ns[b])
        sub = (ns[a: b + 1] - st).astype(np.int32)
        r = O.rct_all(sub, topo["pid"][st:ed], topo["pres"][st:ed], topo["cap"][4 * st: 4 * ed],
                      topo["slew_in"][4 * a: 4 * b], "f64")
        for k in outs:
            ref[k][4 * st: 4 * ed] = r[k]

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(min(16, len(os.sched_getaffinity(0)))) as ex:
        list(ex.map(lambda a: run(a, min(Nn, a + 64)), range(0, Nn, 64)))
    log(f"fp64 oracle: {time.perf_counter() - t0:.1f}s")
    floors = {k: max(float(np.max(np.abs(ref[k]))) * 1e-7, 1e-300) for k in outs}
    floors["impulse"] = max(float(np.max(np.abs(ref["beta"]))) * 1e-6, 1e-300)
    res = {"workload": f"{cfg.name}: {Nn} nets, {V} nodes (cfg3 distribution)", "nets": Nn,
           "nodes": V, "reference": "FP64 oracle (CPU restatement of net.cpp:101-193), every node"}
    for method in ("plain", "batched", "compensated"):
        s.solve(method)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        s.solve(method)
        b.record()
        torch.cuda.synchronize()
        errs = {}
        for k in outs:
            g = s.out[k].cpu().numpy().astype(np.float64)
            e = np.abs(g - ref[k]) / np.maximum(np.abs(ref[k]), floors[k])
            errs[k] = {"max": float(e.max()), "mean": float(e.mean())}
            del g, e
        res[method] = {"ms": a.elapsed_time(b), "rel_err_vs_fp64": errs}
        log(f"precision {method}: {res[method]}")
    return res


def run_precision_streamed(ctx, nets, batch):
    """cfg3 precision report at full size: the nets of the cfg3 distribution
    are generated, ordered and solved in batches of `batch` nets (each net
    drawn from its global index, so the batches make up the whole
    population), and every output of every node of the plain kernel, the
    FP32 batched method and the compensated method is compared on the device
    with the device FP64 reference (otg_rct_reference_f64).  Relative errors
    use a floor of 1e-7 x the batch's largest reference magnitude (1e-6 x the
    largest beta for impulse); a batch's largest values are at most the
    population's, so these errors bound the ones with population-wide
    floors from above.  The device reference itself is checked against the
    host FP64 oracle on a sample of nets of the first batch (the largest
    included)."""
    import torch

    from opentimer_amd import workloads as W
    from opentimer_amd.rct import RcSolver, OUTPUTS
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    methods = ("plain", "batched", "compensated")
    cfg_all = W.config3(num_nets=nets)
    agg = {m: {k: [0.0, 0.0] for k in OUTPUTS} for m in methods}
    ms = {m: 0.0 for m in methods}
    ref_ms = 0.0
    nvals = 0
    nodes = 0
    check = {}
    t_start = time.perf_counter()
    for a in range(0, nets, batch):
        b = min(nets, a + batch)
        cfg = cfg_all.scaled(b - a)
        cfg.first_net = a
        forest = W.generate_rc(cfg, nthreads=16)
        s = RcSolver.from_edges(forest, ctx)
        del forest
        V = s.num_nodes
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ref = s.reference_f64()
        e1.record()
        torch.cuda.synchronize()
        ref_ms += e0.elapsed_time(e1)
        if a == 0:
            topo = s.host_topology()
            ns = topo["nodes_start"]
            sizes = np.diff(ns)
            pick = np.unique(np.concatenate([np.arange(min(b - a, 24)), np.argsort(sizes)[-8:]]))
            idx = np.concatenate([np.arange(4 * ns[k], 4 * ns[k + 1]) for k in pick])
            nd = np.concatenate([np.arange(ns[k], ns[k + 1]) for k in pick])
            sub = np.concatenate([[0], np.cumsum(sizes[pick])]).astype(np.int32)
            n4 = np.concatenate([np.arange(4 * k, 4 * k + 4) for k in pick])
            r64 = O.rct_all(sub, topo["pid"][nd], topo["pres"][nd], topo["cap"][idx],
                            topo["slew_in"][n4], "f64")
            it = torch.from_numpy(idx).to(s.device)
            for k in OUTPUTS:
                g = ref[k][it].cpu().numpy()
                fl = max(float(np.max(np.abs(r64[k if k != "impulse" else "beta"])))
                          * (1e-6 if k == "impulse" else 1e-7), 1e-300)
                check[k] = float(np.max(np.abs(g - r64[k]) / np.maximum(np.abs(r64[k]), fl)))
            check = {"nets": int(len(pick)), "nodes": int(len(nd)), "max_rel_diff": check}
            log(f"device FP64 reference vs host FP64 oracle on {len(pick)} nets: {check}")
            del topo
        floor = {k: max(float(ref[k].abs().max()) * 1e-7, 1e-300) for k in OUTPUTS}
        floor["impulse"] = max(float(ref["beta"].abs().max()) * 1e-6, 1e-300)
        for m in methods:
            s.solve(m)  # builds the batch's plan
            e0.record()
            s.solve(m)
            e1.record()
            torch.cuda.synchronize()
            ms[m] += e0.elapsed_time(e1)
            for k in OUTPUTS:
                e = (s.out[k].double() - ref[k]).abs_().div_(ref[k].abs().clamp_(min=floor[k]))
                agg[m][k][0] = max(agg[m][k][0], float(e.max()))
                agg[m][k][1] += float(e.sum())
                del e
        nvals += 4 * V
        nodes += V
        del ref
        s.close()
        del s
        torch.cuda.empty_cache()
        log(f"precision batch nets {a}..{b}: {nodes} nodes so far, "
            f"compensated max {max(v[0] for v in agg['compensated'].values()):.3g}, "
            f"{time.perf_counter() - t_start:.0f}s")
    res = {"workload": f"{cfg_all.name}: {nets} nets, {nodes} nodes (cfg3 distribution), "
                       f"batches of {batch} nets",
           "nets": nets, "nodes": nodes,
           "reference": "device FP64 reference (otg_rct_reference_f64), every node",
           "reference_check_vs_host_fp64": check, "reference_ms": ref_ms,
           "wall_s": time.perf_counter() - t_start}
    for m in methods:
        res[m] = {"ms": ms[m], "nodes_per_s": nodes / (ms[m] * 1e-3),
                  "rel_err_vs_fp64": {k: {"max": v[0], "mean": v[1] / nvals}
                                      for k, v in agg[m].items()}}
        log(f"precision {m}: {res[m]}")
    return res


def _closure(fo_start, fo, seeds, n):
    """Forward closure over a CSR (vectorised frontier expansion)."""
    mark = np.zeros(n, bool)
    fr = np.unique(seeds)
    mark[fr] = True
    while len(fr):
        st, ln = fo_start[fr], fo_start[fr + 1] - fo_start[fr]
        tot = int(ln.sum())
        if tot == 0:
            break
        off = np.repeat(st - (np.cumsum(ln) - ln), ln) + np.arange(tot)
        nxt = np.unique(fo[off])
        nxt = nxt[~mark[nxt]]
        mark[nxt] = True
        fr = nxt
    return np.flatnonzero(mark)


def run_incremental(ctx, method, pins=5_000_000, nets=1_000_000):
    """cfg4: a 5M-pin timing graph (cfg2's generator scaled: 50k pins per
    layer, 2.5 arcs per pin) and 1M RC nets (cfg1 size distribution) resident
    on the device.  For batches of 1, 100, 10k and 100k changed nets: R and C
    of the changed nets are perturbed by U(0.8, 1.2) (before the timed
    region), then one hipGraph launch re-solves them and re-levelizes the
    fan-out cone of their driver pins.  Reports p50/p99 per batch size."""
    import torch

    from opentimer_amd import workloads as W
    from opentimer_amd.incremental import Incremental
    from opentimer_amd.levelize import Levelizer
    from opentimer_amd.rct import RcSolver
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    dcfg = W.DagConfig("cfg4-dag-5M", pins, int(2.5 * pins), max(2, pins // 50_000), seed=9)
    dag = W.generate_dag(dcfg, nthreads=16)
    fin = W.fanin_csr(dag)
    fout = W.fanin_csr(W.Dag(dag.n, dag.arc_to, dag.arc_from))
    L = Levelizer(fin, ctx)
    L.run()
    _, ends, lv = L.host()
    nlev = len(ends) - 1
    forest = W.generate_rc(W.config1(num_nets=nets), nthreads=16)
    s = RcSolver.from_edges(forest, ctx)
    ns = forest.nodes_start
    del forest
    rng = np.random.default_rng(11)
    drivers = rng.choice(np.flatnonzero(fin.out_degree > 0), nets).astype(np.int32)
    inc = Incremental(s, fin, fout, drivers, lv, nlev)
    s.plan(method)
    s.alloc_outputs()
    dev = s.device
    side = torch.cuda.Stream(dev)
    # full re-timing for comparison: one full solve + one full levelization
    # (warm; median of five, wall clock: the levelization reads its level
    # count on the host)
    fulls = []
    for rep in range(6):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.solve(method)
        L.run()
        torch.cuda.synchronize()
        if rep:
            fulls.append((time.perf_counter() - t0) * 1e3)
    full_ms = float(np.median(fulls))
    res = {"workload": f"{dcfg.name}: {dag.n} pins, {len(dag.arc_from)} arcs, {nlev} levels; "
                       f"{nets} RC nets (cfg1 sizes), {int(ns[-1])} nodes",
           "method": method, "full_retime_ms": full_ms, "batches": {}}

    def perturb(ids):
        lens = ns[ids + 1] - ns[ids]
        idx = np.repeat(ns[ids] - (np.cumsum(lens) - lens), lens) + np.arange(int(lens.sum()))
        f = torch.from_numpy(rng.uniform(0.8, 1.2, len(idx)).astype(np.float32)).to(dev)
        it = torch.from_numpy(idx.astype(np.int64)).to(dev)
        s.pres[it] *= f
        s.cap.view(-1, 4)[it] *= f[:, None]

    for B in (1, 100, 10_000, 100_000):
        buf = torch.zeros(B, dtype=torch.int32, device=dev)
        g = inc.capture(buf, method, side)
        rounds = 50 if B <= 100 else (30 if B <= 10_000 else 20)
        times, cones, levels = [], [], []
        verified = None
        for r in range(rounds):
            ids = np.sort(rng.choice(nets, B, replace=False)).astype(np.int32)
            buf.copy_(torch.from_numpy(ids))
            perturb(ids)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(side)
            g.launch(side)
            e1.record(side)
            side.synchronize()
            times.append(e0.elapsed_time(e1))
            cone, fr, kends = inc.query(side.cuda_stream)
            cones.append(len(cone))
            levels.append(len(kends) - 1)
            if r == 0 and B <= 100:
                ref = _closure(fout.start, fout.fanin, drivers[ids], dag.n)
                cand = np.zeros(dag.n, np.uint8)
                cand[ref] = 1
                csr = W.fanin_csr(dag, cand=cand)
                _, oends, olv = O.levelize(csr.n, csr.start, csr.fanin, csr.out_degree, csr.frontier)
                lvg = inc.level.cpu().numpy()
                verified = bool(np.array_equal(np.sort(cone), ref) and np.array_equal(kends, oends)
                                and np.array_equal(lvg[ref], olv[ref]))
        g.close()
        t = np.array(times)
        res["batches"][str(B)] = {
            "rounds": rounds, "p50_ms": float(np.percentile(t, 50)), "p99_ms": float(np.percentile(t, 99)),
            "mean_ms": float(t.mean()), "nets_per_s": float(B / (t.mean() / 1e3)),
            "mean_cone_pins": float(np.mean(cones)), "cone_pins_per_s": float(np.mean(cones) / (t.mean() / 1e3)),
            "mean_cone_levels": float(np.mean(levels)), "cone_and_levels_match_oracle": verified}
        log(f"incremental B={B}: {res['batches'][str(B)]}")
    inc.close()
    return res


def run_timer(ctx, method, reps=5):
    """Timer::update_timing on the device (otg_timer) over a whole design: the
    cfg2 graph (10M pins, 25M arcs, 200 levels) whose pins with fanout drive
    RC nets with probability 0.4 (about 4M nets of the cfg1 size
    distribution), the synthetic NLDM library, primary inputs and outputs on
    the graph's boundary.  The first update also orders the RC trees,
    levelizes the graph and builds the plans; the later ones (timed with HIP
    events on the launch stream, median of `reps`) re-solve every net, gather
    the net-arc values and propagate forward and backward.  Consistency: the
    RC delays the timer gathers at 2,000 sampled net-arc sinks equal, bit for
    bit, an independent solve of the same forest queried at the same pins
    (otg_rct_query)."""
    import ctypes as C

    import torch

    from opentimer_amd import native as N
    from opentimer_amd import workloads as W
    from opentimer_amd.rct import RcSolver
    t0 = time.perf_counter()
    lib = W.prop_library()
    D = W.timer_design(W.config2(), lib, nthreads=16)
    f = D.forest
    log(f"timer design: {f.num_nets} nets {f.num_nodes} nodes, {len(D.fanin)} arcs, "
        f"generated in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda", 0)
    tt = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    T = {"nodes_start": tt(f.nodes_start), "edges": tt(f.edges), "roots": tt(f.roots),
         "edge_res": tt(f.edge_res), "node_cap": tt(f.node_cap), "fanin_start": tt(D.fanin_start),
         "fanin": tt(D.fanin), "arc_kind": tt(D.arc_kind), "arc_timing": tt(D.arc_timing),
         "arc_net": tt(D.arc_net), "arc_node": tt(D.arc_node), "pin_net": tt(D.pin_net),
         "pin_load": tt(D.pin_load), "pi_slew": tt(D.pi_slew), "pi_at": tt(D.pi_at),
         "po_rat": tt(D.po_rat)}
    LB = {k: tt(getattr(lib, k)) for k in ("lut_hdr", "lut_val", "timing_sense", "timing_edge",
                                            "timing_lut", "timing_has")}
    a = lambda d, k: d[k].data_ptr()  # noqa: E731
    n, m = len(D.pin_net), len(D.fanin)
    design = N.TimerDesign(f.num_nets, f.num_nodes, a(T, "nodes_start"), a(T, "edges"),
                           a(T, "roots"), a(T, "edge_res"), a(T, "node_cap"), n, m,
                           a(T, "fanin_start"), a(T, "fanin"), a(T, "arc_kind"),
                           a(T, "arc_timing"), a(T, "arc_net"), a(T, "arc_node"),
                           a(T, "pin_net"), a(T, "pin_load"), a(T, "pi_slew"), a(T, "pi_at"),
                           a(T, "po_rat"))
    plib = N.PropLib(len(lib.lut_hdr) // 4, len(lib.timing_sense), len(lib.lut_val),
                     a(LB, "lut_hdr"), a(LB, "lut_val"), a(LB, "timing_sense"),
                     a(LB, "timing_edge"), a(LB, "timing_lut"), a(LB, "timing_has"))
    meth = {"batched": N.BATCHED, "compensated": N.BATCHED_COMPENSATED}[method]
    timer = N.timer_p()
    ctx.check(N.lib.otg_timer_create(ctx.ptr, C.byref(design), C.byref(plib), meth,
                                     C.byref(timer)))
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    try:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.check(N.lib.otg_timer_update(timer, sh))
        torch.cuda.synchronize()
        first_ms = (time.perf_counter() - t1) * 1e3
        ctx.check(N.lib.otg_timer_update(timer, sh))
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            ctx.check(N.lib.otg_timer_update(timer, sh))
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        upd_ms = float(np.median(ms))
        r = N.TimerResults()
        ctx.check(N.lib.otg_timer_results_get(timer, C.byref(r)))
        levels = int(r.num_levels)
        # consistency of the gathered RC delays with an independent solve
        rng = np.random.default_rng(9)
        rc_arcs = np.flatnonzero(D.arc_net >= 0)
        pick = np.sort(rng.choice(rc_arcs, min(2000, len(rc_arcs)), replace=False))
        gathered = torch.empty(4 * m, dtype=torch.float32, device=dev)
        gathered.copy_(torch.as_tensor(_DevView(r.arc_rc_delay, 4 * m), device="cuda"))
        got = gathered.view(m, 4)[torch.from_numpy(pick).to(dev)].cpu().numpy()
    finally:
        N.lib.otg_timer_destroy(timer)
    del T
    s = RcSolver.from_edges(f, ctx)
    s.solve(method)
    q = s.query(torch.from_numpy(D.arc_net[pick]).to(dev), torch.from_numpy(D.arc_node[pick]).to(dev))
    ref = q["delay"].cpu().numpy()
    same = bool(np.array_equal(got.view(np.int32), ref.view(np.int32)))
    s.close()
    return {"workload": f"cfg2 graph ({n} pins, {m} arcs, {levels} levels) with {f.num_nets} RC nets "
                        f"({f.num_nodes} nodes, cfg1 sizes) on the pins that drive nets; synthetic "
                        "7x8 NLDM library", "method": method, "pins": n, "arcs": m,
            "nets": int(f.num_nets), "nodes": int(f.num_nodes), "levels": levels,
            "first_update_ms": first_ms, "update_ms": upd_ms,
            "updates_per_s": 1e3 / upd_ms,
            "what": "otg_timer_update: RC solve of every net, net-arc delay/impulse and cell loads "
                    "gathered from it, forward propagation, backward propagation (the first "
                    "update also BFS-orders, levelizes and builds the plans)",
            "gathered_rc_delay_equals_independent_solve": same}


class _DevView:
    """A raw device array seen by torch through __cuda_array_interface__."""

    def __init__(self, ptr, count):
        self.__cuda_array_interface__ = {"shape": (count,), "typestr": "<f4",
                                         "data": (ptr, False), "version": 3}


def run_spef(ctx, nets, method):
    """SPEF -> flattened forest (host loader, all host threads) -> GPU BFS
    ordering -> solve.  The SPEF is written from a cfg1-distribution forest
    (per-node caps equal on the four corners, as SPEF has them), so the loaded
    forest must equal the generated one net for net; the GPU results are
    compared with the oracle on a sample of nets."""
    import tempfile

    import torch

    from opentimer_amd import workloads as W
    from opentimer_amd.rct import RcSolver
    from opentimer_amd.spef import SpefForest, write

    g = W.generate_rc(W.config1(num_nets=nets, seed=4), nthreads=16)
    c4 = g.node_cap.reshape(-1, 4)
    c4[:] = c4[:, :1]
    tmpdir = tempfile.mkdtemp(prefix="otg_spef_")
    path = os.path.join(tmpdir, "cfg1.spef")
    try:
        t0 = time.time()
        write(path, g)
        wr_s = time.time() - t0
        size = os.path.getsize(path)
        log(f"spef: wrote {size/1e6:.0f} MB ({nets} nets, {g.num_nodes} nodes) in {wr_s:.1f}s")
        threads = min(16, os.cpu_count() or 1)
        runs = []
        f = None
        for _ in range(3):
            if f is not None:
                f.close()
            t0 = time.time()
            f = SpefForest.read(path, threads=threads)
            runs.append(time.time() - t0)
        parse_s = float(np.median(runs))
        same = (np.array_equal(f.nodes_start, g.nodes_start) and np.array_equal(f.edge_res, g.edge_res)
                and np.array_equal(np.sort(f.node_cap[0::4]), np.sort(g.node_cap[0::4])))
        # GPU: upload + BFS ordering, then one solve
        torch.cuda.synchronize()
        t0 = time.time()
        s = RcSolver.from_edges(f, ctx)
        torch.cuda.synchronize()
        bfs_s = time.time() - t0
        s.solve(method)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(5):
            s.solve(method)
        b.record()
        torch.cuda.synchronize()
        solve_ms = a.elapsed_time(b) / 5
        # parity on a sample: the loaded forest's GPU results vs the oracle
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        topo = s.host_topology()
        ns = topo["nodes_start"]
        pick = np.sort(np.random.default_rng(1).choice(f.num_nets, min(2000, f.num_nets),
                                                       replace=False))
        idx = np.concatenate([np.arange(4 * ns[k], 4 * ns[k + 1]) for k in pick])
        nodes = np.concatenate([np.arange(ns[k], ns[k + 1]) for k in pick])
        sub = np.concatenate([[0], np.cumsum(np.diff(ns)[pick])]).astype(np.int32)
        n4 = np.concatenate([np.arange(4 * k, 4 * k + 4) for k in pick])
        r64 = O.rct_all(sub, topo["pid"][nodes], topo["pres"][nodes], topo["cap"][idx],
                        topo["slew_in"][n4], "f64")
        out = s.host()
        err = {}
        for k in ("load", "delay", "slew"):
            r = r64[k].astype(np.float64)
            fl = max(float(np.max(np.abs(r))) * 1e-7, 1e-30)
            err[k] = float(np.max(np.abs(out[k][idx] - r) / np.maximum(np.abs(r), fl)))
        # pin queries (otg_rct_query): 1M random (net, node) pins
        rng = np.random.default_rng(2)
        qn = rng.integers(0, f.num_nets, 1_000_000).astype(np.int32)
        ql = (rng.random(1_000_000) * np.diff(ns)[qn]).astype(np.int32)
        qn_t = torch.from_numpy(qn).to(s.device)
        ql_t = torch.from_numpy(ql).to(s.device)
        s.query(qn_t, ql_t)
        torch.cuda.synchronize()
        a.record()
        for _ in range(10):
            s.query(qn_t, ql_t)
        b.record()
        torch.cuda.synchronize()
        query_ms = a.elapsed_time(b) / 10
        res = {"only": "spef", "nets": int(f.num_nets), "nodes": int(f.num_nodes),
               "file_MB": size / 1e6, "parse_threads": threads, "parse_s": parse_s,
               "parse_MB_per_s": size / 1e6 / parse_s, "parse_nets_per_s": f.num_nets / parse_s,
               "loaded_equals_written": bool(same), "gpu_upload_and_bfs_s": bfs_s,
               "solve_method": method, "solve_ms": solve_ms,
               "rel_err_vs_fp64_sample": err, "undriven_nets": int(f.num_undriven),
               "query_1M_pins_ms": query_ms}
        log(f"spef: {res}")
        return res
    finally:
        try:
            os.remove(path)
            os.rmdir(tmpdir)
        except OSError:
            pass


def run_fprop(ctx, stream, sh, reps=10, cpu=True):
    """Level-wise forward propagation (slew, arc delay, arrival with NLDM
    lookups) over the levelization graph: timed with HIP events on the launch
    stream around `reps` whole propagations (one launch per level), checked
    bit for bit against the oracle on the whole graph; the oracle's run is
    the single-core CPU baseline."""
    import torch

    from opentimer_amd import workloads as W
    from opentimer_amd.levelize import Levelizer
    from opentimer_amd.prop import Propagator
    dcfg = W.config2()
    dag = W.generate_dag(dcfg, nthreads=16)
    csr = W.fanin_csr(dag)
    del dag
    lib = W.prop_library()
    g = W.prop_graph(csr, lib)
    L = Levelizer(csr, ctx)
    L.run(stream=sh)
    levels = L.num_ends - 1
    P = Propagator(g, lib, ctx)

    def timed(fn):
        for _ in range(2):
            fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps
    # pin order, one launch per level (otg_fprop)
    direct_ms = timed(lambda: P.run(L.frontiers, L.ends, levels, stream=sh))
    # the level-ordered plan, built once per levelization (otg_fprop_plan_*)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P.build_plan(L.frontiers, L.ends, levels, stream=sh)
    plan_ms = (time.perf_counter() - t0) * 1e3
    ms = timed(lambda: P.run_plan(stream=sh))
    n, m = csr.n, len(csr.fanin)
    # required times and slacks (otg_bprop) from primary-output values on the
    # sinks, over the forward pass's results
    rng = np.random.default_rng(5)
    po = np.full((n, 4), np.nan, np.float32)
    sinks = csr.out_degree == 0
    po[sinks] = rng.uniform(3000.0, 6000.0, (int(sinks.sum()), 4)).astype(np.float32)
    P.required(po, L.frontiers, L.ends, levels, stream=sh)
    bprop_ms = timed(lambda: P.rerun_required(L.frontiers, L.ends, levels, stream=sh))
    # per pin: fanin_start, frontier entry, load, PI slew and arrival, slew
    # and arrival out (4 + 4 + 5 x 16 B); per arc: source, kind, timing,
    # RC delay and impulse, the source's slew and arrival, 8 arc delays out
    # (3 x 4 + 4 x 16 + 32 B)
    alg = 88 * n + 108 * m
    out = {"workload": f"{dcfg.name} graph, synthetic 7x8 NLDM library ({lib.num_timings} "
                       f"timings, {lib.num_luts} tables)",
           "pins": n, "arcs": m, "levels": levels, "ms": ms, "pins_per_s": n / (ms / 1e3),
           "arcs_per_s": m / (ms / 1e3), "us_per_level": ms * 1e3 / levels,
           "algorithmic_bytes": alg, "achieved_GBps": alg / (ms / 1e3) / 1e9,
           "frac": alg / (ms / 1e3) / 1e9 / PEAK_HBM_GBS, "path": "level-ordered plan",
           "plan_build_ms": plan_ms, "pin_order_ms": direct_ms,
           "bprop": {"ms": bprop_ms, "pins_per_s": n / (bprop_ms / 1e3),
                     "what": "required times and slacks, primary-output values on the sinks"}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        # the outputs of one plan run alone (the pin-order runs wrote them too)
        for t in (P.slew, P.at, P.arc_delay):
            t.fill_(float("nan"))
        P.run_plan(stream=sh)
        rat, slack = P.required(po, L.frontiers, L.ends, levels, stream=sh)
        rat, slack = rat.cpu().numpy(), slack.cpu().numpy()
        slew, at, ad = P.host()
        fr, ends, _ = L.host()
        gk = ("fanin_start", "fanin", "arc_kind", "arc_timing", "arc_rc_delay", "arc_rc_impulse",
              "pin_load", "pi_slew", "pi_at")
        lk = ("lut_hdr", "lut_val", "timing_sense", "timing_edge", "timing_lut", "timing_has")
        cores, aff = host_threads()
        t0 = time.perf_counter()
        os_, oa, od = O.fprop({k: getattr(g, k) for k in gk}, {k: getattr(lib, k) for k in lk},
                              fr, ends, nthreads=cores)
        cms = (time.perf_counter() - t0) * 1e3

        def same(a, b):
            return bool(np.array_equal(np.isnan(a), np.isnan(b))
                        and np.array_equal(a[~np.isnan(a)].view(np.int32),
                                           b[~np.isnan(b)].view(np.int32)))
        t0 = time.perf_counter()
        orat, oslack = O.bprop(g.fanin_start, g.fanin, od, po, oa, fr, ends)
        out["bprop"]["cpu_ms"] = (time.perf_counter() - t0) * 1e3
        out["bitexact_vs_oracle"] = {"slew": same(slew, os_), "at": same(at, oa),
                                     "arc_delay": same(ad, od), "rat": same(rat, orat),
                                     "slack": same(slack, oslack)}
        out["cpu_baseline"] = {"value": n / (cms / 1e3), "unit": "pins/s", "cores": cores,
                               "affinity_cores": aff, "kind": "port", "cpu": cpu_model(),
                               "ms": cms, "sample": "the whole graph, one propagation, the pins "
                                                    "of each level split over the threads"}
    log(f"fprop: {out}")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: nets shard "
              "into independent contiguous ranges, one per GPU, and the figure of one GPU "
              "is what is measured", file=sys.stderr)
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
    if args.tuning:
        ctx.check(N.lib.otg_rct_tuning_load(ctx.ptr, args.tuning.encode(), None))

    if args.only == "levelize":
        print(json.dumps(run_levelize(ctx, stream, sh)))
        return
    if args.only == "fprop":
        print(json.dumps(run_fprop(ctx, stream, sh, reps=args.steps, cpu=not args.no_cpu)))
        return
    if args.only == "precision":
        dev_ref = args.precision_ref == "device" or (args.precision_ref == "auto"
                                                     and args.precision_nets > 20_000)
        if dev_ref:
            print(json.dumps(run_precision_streamed(ctx, args.precision_nets, args.precision_batch)))
        else:
            print(json.dumps(run_precision(ctx, args.precision_nets)))
        return
    if args.only == "incremental":
        print(json.dumps(run_incremental(ctx, args.method)))
        return
    if args.only == "spef":
        print(json.dumps(run_spef(ctx, args.spef_nets, args.method)))
        return
    if args.only == "timer":
        print(json.dumps(run_timer(ctx, args.method)))
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

    import ctypes as C
    # ---- tuning: by default each batched method takes the tuning shipped
    # for this workload on MI355X (opentimer_amd/tunings/, written by
    # otg_rct_autotune on this forest); --autotune searches it anew
    other = "batched" if args.method == "compensated" else "compensated"
    tunings = {}
    tuning_info = {}
    s.alloc_outputs()
    tdir = os.path.join(ROOT, "gpurun_out") if os.path.isdir(os.path.join(ROOT, "gpurun_out")) \
        else "/tmp"
    # (loaded or searched on a context of its own: the solver's context keeps
    # the built-in defaults, and each plan gets its method's tuning explicitly)
    tctx = N.Context(0)
    for m in (args.method, other):
        if args.tuning or args.no_autotune:
            tunings[m] = None
            tuning_info[m] = {"source": args.tuning or "defaults"}
            continue
        shipped = os.path.join(ROOT, "opentimer_amd", "tunings", f"{cfg.name}_{m}.json")
        if not args.autotune and args.nets == 4_000_000 and os.path.exists(shipped):
            t = N.RctTuning()
            tctx.check(N.lib.otg_rct_tuning_load(tctx.ptr, shipped.encode(), C.byref(t)))
            tunings[m] = tuple(getattr(t, k) for k, _ in N.RctTuning._fields_)
            tuning_info[m] = {"source": os.path.relpath(shipped, ROOT),
                              **{k: getattr(t, k) for k, _ in N.RctTuning._fields_}}
            log(f"tuning {m}: {tuning_info[m]}")
            continue
        best = N.RctTuning()
        path = os.path.join(tdir, f"tuning_{m}.json")
        t0 = time.perf_counter()
        tctx.check(N.lib.otg_rct_autotune(tctx.ptr, C.byref(s.tree()), C.byref(s.result()),
                                         N.BATCHED if m == "batched" else N.BATCHED_COMPENSATED,
                                         path.encode(), C.byref(best)))
        tunings[m] = tuple(getattr(best, k) for k, _ in N.RctTuning._fields_)
        rec = json.load(open(path))
        tuning_info[m] = {"source": "autotuned on this forest", "ms_search": (time.perf_counter() - t0) * 1e3,
                          "candidates": rec.get("candidates"),
                          **{k: getattr(best, k) for k, _ in N.RctTuning._fields_}}
        log(f"autotuned {m}: {tuning_info[m]}")
    tctx.close()

    plans = {}
    for m in (args.method, other):
        t0 = time.perf_counter()
        s.plan(m, tunings[m])
        torch.cuda.synchronize()
        plans[m] = {"build_ms": (time.perf_counter() - t0) * 1e3, "stats": s.plan_stats(m, tunings[m])}
        log(f"plan {m}: {plans[m]['build_ms']:.1f} ms {plans[m]['stats']}")
    s.alloc_outputs()

    def timed(method, steps, warmup):
        tun = tunings.get(method)
        for _ in range(warmup):
            s.solve(method, stream=sh, tuning=tun)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(steps):
            s.solve(method, stream=sh, tuning=tun)
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / max(steps, 1)

    alg_bytes = 120.0 * V + 20.0 * Nn
    ms = timed(args.method, args.steps, args.warmup)
    log(f"{args.method}: {ms:.3f} ms/step")
    if args.only == "rct":
        print(json.dumps({"only": "rct", "method": args.method, "ms_per_step": ms}))
        return
    ms_other = timed(other, max(3, args.steps // 2), 1)
    log(f"{other}: {ms_other:.3f} ms/step")
    # ---- the streaming ceiling of a solve's traffic on this device: the same
    # algorithmic bytes over the same arrays with no tree work
    # (otg_rct_stream_ceiling; it overwrites the outputs, which the parity
    # leg below solves again)
    for _ in range(2):
        s.stream_ceiling(stream=sh)
    torch.cuda.synchronize()
    ca, cb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ca.record(stream)
    for _ in range(max(5, args.steps)):
        s.stream_ceiling(stream=sh)
    cb.record(stream)
    torch.cuda.synchronize()
    ceil_ms = ca.elapsed_time(cb) / max(5, args.steps)
    log(f"streaming ceiling: {ceil_ms:.3f} ms ({alg_bytes / ceil_ms / 1e6:.0f} GB/s)")
    plain_ms = None
    if not args.no_plain and args.only in ("", "plain"):
        plain_ms = timed("plain", args.plain_steps, 1)
        log(f"plain: {plain_ms:.3f} ms/step")
    if args.only == "plain":
        print(json.dumps({"only": "plain", "ms_per_step": plain_ms}))
        return

    # ---- sample parity of the timed outputs ------------------------------------
    # FP32 batched vs the CPU timer arithmetic (FP32 DFS, 1e-5 on delay and
    # slew); compensated vs FP64 (1e-6 on every output).
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    topo = s.host_topology()
    ns = topo["nodes_start"]
    rng = np.random.default_rng(0)
    pick = np.sort(rng.choice(Nn, min(Nn, 3000), replace=False))
    sizes = np.diff(ns)
    pick = np.unique(np.concatenate([pick, np.argsort(sizes)[-20:]]))  # the largest nets too
    idx = np.concatenate([np.arange(4 * ns[a], 4 * ns[a + 1]) for a in pick])
    sub_ns = np.concatenate([[0], np.cumsum(sizes[pick])]).astype(np.int32)
    nodes = np.concatenate([np.arange(ns[a], ns[a + 1]) for a in pick])
    nidx4 = np.concatenate([np.arange(4 * a, 4 * a + 4) for a in pick])
    args_o = (sub_ns, topo["pid"][nodes], topo["pres"][nodes], topo["cap"][idx], topo["slew_in"][nidx4])
    ref32 = O.rct_all(*args_o, "dfs_f32")
    ref64 = O.rct_all(*args_o, "f64")
    idx_t = torch.from_numpy(idx).to(dev)
    parity = {"nets_checked": int(len(pick))}
    for m in ("batched", "compensated"):
        s.solve(m, stream=sh, tuning=tunings.get(m))
        torch.cuda.synchronize()
        errs = {}
        for k in ("load", "delay", "ldelay", "beta", "impulse", "slew"):
            g = s.out[k][idx_t].cpu().numpy().astype(np.float64)
            r64 = ref64[k].astype(np.float64)
            scale = float(np.max(np.abs(r64))) if r64.size else 1.0
            fl = max(scale * 1e-7, 1e-30)
            if k == "impulse":
                fl = max(float(np.max(np.abs(ref64["beta"]))) * 1e-6, 1e-30)
            errs[k + "_vs_fp64"] = float(np.max(np.abs(g - r64) / np.maximum(np.abs(r64), fl)))
            if k in ("delay", "slew"):
                r32 = ref32[k].astype(np.float64)
                errs[k + "_vs_cpu_fp32"] = float(np.max(np.abs(g - r32) / np.maximum(np.abs(r32), fl)))
        parity[m] = errs
    log(f"parity sample: {parity}")

    # ---- levelization ---------------------------------------------------------
    lev = None
    if not args.no_levelize:
        lev = run_levelize(ctx, stream, sh)
    fprop = None
    if not args.no_fprop:
        fprop = run_fprop(ctx, stream, sh, cpu=not args.no_cpu)
    timer_leg = None
    if not args.no_timer:
        timer_leg = run_timer(ctx, args.method)
        log(f"timer: {timer_leg}")

    achieved = alg_bytes / (ms / 1e3) / 1e9
    # HBM traffic per step from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # summarised in profiles/, used only if it was measured on the kernel
    # sources built here (their hash is stamped in the file)
    traffic, traffic_note = None, "no traffic measurement in profiles/"
    tfile = os.path.join(ROOT, "profiles", f"rct_{args.method}_cfg1_traffic.json")
    if os.path.exists(tfile):
        try:
            rec = json.load(open(tfile))
            if rec.get("sources_sha256") == kernel_sources_hash():
                traffic = rec.get("hbm_bytes_per_launch")
                traffic_note = f"rocprofv3 PMC, measured at commit {rec.get('commit')}"
            else:
                traffic_note = "the traffic in profiles/ was measured on other kernel sources"
        except Exception:
            traffic = None
    dropin = None
    if not args.no_dropin:
        dropin = run_dropin(ctx, forest, args.method)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(forest, topo, args.cpu_seconds)
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
        "config": {"workload": f"{cfg.name} (4M heavy-tailed RC trees, BFS order, 4 corners)",
                   "nets": Nn, "nodes": V, "method": args.method,
                   "precision": ("fp32 inputs and outputs; subtree and path sums and the final "
                                 "per-node combination in fp64") if args.method == "compensated" else "fp32",
                   "plan": {k: int(v) for k, v in plans[args.method]["stats"].items()},
                   "tuning": tuning_info[args.method]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "ceiling": {"ms": ceil_ms, "GBps": alg_bytes / ceil_ms / 1e6,
                                 "frac_of_peak": alg_bytes / ceil_ms / 1e6 / PEAK_HBM_GBS,
                                 "step_frac_of_ceiling": ceil_ms / ms,
                                 "what": "otg_rct_stream_ceiling: the step's algorithmic bytes "
                                         "(24 B in, 96 B out per node; 20 B per net) moved over "
                                         "the same arrays, coalesced, with no tree work"},
                     "traffic_source": traffic_note, "algorithmic_bytes": alg_bytes},
        "cpu_baseline": cpu,
        other: {"ms_per_step": ms_other, "nets_per_s": Nn / (ms_other / 1e3),
                "tuning": tuning_info[other],
                "achieved_GBps": alg_bytes / (ms_other / 1e3) / 1e9,
                "frac": alg_bytes / (ms_other / 1e3) / 1e9 / PEAK_HBM_GBS},
        "plain": None if plain_ms is None else {"ms_per_step": plain_ms,
                                                "nets_per_s": Nn / (plain_ms / 1e3),
                                                "speedup_of_" + args.method: plain_ms / ms,
                                                "speedup_of_" + other: plain_ms / ms_other},
        "bfs_ordering_ms": bfs_ms,
        "dropin": dropin,
        "plan_build_ms": {m: plans[m]["build_ms"] for m in plans},
        "levelize": lev,
        "fprop": fprop,
        "timer": timer_leg,
        "parity": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
