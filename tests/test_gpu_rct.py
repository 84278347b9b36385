"""GPU parity of P1 (RC-tree Elmore) and the per-net BFS ordering.

Bars:
  * BFS ordering (node2bfs, pid, oriented edges), pres and cap in BFS order:
    bit-exact vs the oracle (integer / copy work).
  * plain kernel: bit-exact vs the oracle's restatement of the reference
    kernel arithmetic (rct.cu:11-59, FP32, one rounding per operation).
  * batched FP32: within 1e-5 relative of the CPU ot::Timer arithmetic
    (net.cpp:101-193) on delay and slew, and of the FP64 oracle on all
    outputs (impulse measured against max(|impulse|, 1e-6*beta) since it is
    a difference that can vanish).
  * batched compensated: within 1e-6 relative of the FP64 oracle.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

import oracle as O
from opentimer_amd import native as N
from opentimer_amd import workloads as W
from opentimer_amd.rct import RcSolver

OUT = ("load", "delay", "ldelay", "beta", "impulse", "slew")


def rel_err(got, ref, floor):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor))) if ref.size else 0.0


def floors(ref):
    """Per-output floor for relative errors: exact zeros (driver delay, beta,
    impulse) are compared absolutely against the output's scale."""
    f = {}
    for k in OUT:
        scale = float(np.max(np.abs(ref[k]))) if ref[k].size else 1.0
        f[k] = max(scale * 1e-7, 1e-30)
    f["impulse"] = max(float(np.max(np.abs(ref["beta"]))) * 1e-6, 1e-30)
    return f


def oracle_inputs(forest):
    e, n2b, pid = O.rct_bfs(forest.nodes_start, forest.edges, forest.roots)
    pres, cap = O.rct_prepare(forest.nodes_start, e, n2b, forest.edge_res, forest.node_cap)
    return e, n2b, pid, pres, cap


@pytest.fixture(scope="module")
def cfg0(ctx):
    f = W.generate_rc(W.config0())
    s = RcSolver.from_edges(f, ctx)
    return f, s


def test_bfs_bitexact(cfg0):
    f, s = cfg0
    e, n2b, pid, pres, cap = oracle_inputs(f)
    h = s.host_topology()
    assert np.array_equal(h["node2bfs"], n2b)
    assert np.array_equal(h["pid"], pid)
    assert np.array_equal(s.edges_directed.cpu().numpy(), e)
    assert np.array_equal(h["pres"].view(np.uint32), pres.view(np.uint32))
    assert np.array_equal(h["cap"].view(np.uint32), cap.view(np.uint32))


@pytest.mark.parametrize("seed", [31, 32])
def test_bfs_bitexact_mixed_sizes(ctx, seed):
    # every BFS path against the oracle's ordering: packed small nets (up to
    # 64 nodes, one-node nets included), large nets on chip and nets larger
    # than the on-chip scratch (chains, balanced trees, random trees)
    cfg = W.RcConfig("bfsmix", 1200, [W.SizeClass(0.45, 1, 12), W.SizeClass(0.25, 13, 64),
                                      W.SizeClass(0.2, 65, 3000, N.SHAPE_CHAIN, 0.1),
                                      W.SizeClass(0.1, 500, 20000, N.SHAPE_RANDOM)], seed=seed)
    f = W.generate_rc(cfg)
    s = RcSolver.from_edges(f, ctx)
    e, n2b, pid, pres, cap = oracle_inputs(f)
    h = s.host_topology()
    assert np.array_equal(h["node2bfs"], n2b)
    assert np.array_equal(h["pid"], pid)
    assert np.array_equal(s.edges_directed.cpu().numpy(), e)
    assert np.array_equal(h["pres"].view(np.uint32), pres.view(np.uint32))
    assert np.array_equal(h["cap"].view(np.uint32), cap.view(np.uint32))


def test_not_a_tree_rejected_in_each_bfs_path(ctx):
    # a broken net among small nets, and one among large nets: the smallest
    # failing net is reported whichever path met it
    cfg = W.RcConfig("bfsbad", 300, [W.SizeClass(0.8, 2, 20), W.SizeClass(0.2, 100, 2000, N.SHAPE_CHAIN, 0.1)],
                     seed=7)
    for pick in ("small", "large"):
        f = W.generate_rc(cfg)
        sizes = np.diff(f.nodes_start)
        k = int(np.flatnonzero(sizes <= 20)[5] if pick == "small" else np.flatnonzero(sizes >= 100)[0])
        e0 = int(f.nodes_start[k] - k)
        bad = f.edges.copy()
        bad[2 * e0], bad[2 * e0 + 1] = bad[2 * e0 + 2], bad[2 * e0 + 3]  # a duplicated edge
        f.edges = bad
        with pytest.raises(N.OtgError) as ei:
            RcSolver.from_edges(f, ctx)
        assert ei.value.status == 2
        assert f"net {k}" in str(ei.value), (pick, k, str(ei.value))


@pytest.mark.parametrize("lo,hi", [(65, 1024), (1025, 2048), (2049, 4096), (4097, 8192),
                                   (8193, 12000)])
def test_not_a_tree_rejected_in_each_size_class(ctx, lo, hi):
    # one broken net (a duplicated edge, or a cycle with a node cut off) in
    # each size class of the large-net kernels, among valid nets of all sizes
    cfg = W.RcConfig("bfsclass", 120, [W.SizeClass(0.5, 2, 60), W.SizeClass(0.5, lo, hi, N.SHAPE_CHAIN, 0.1)],
                     seed=lo)
    for mistake in ("duplicate", "cycle"):
        f = W.generate_rc(cfg)
        sizes = np.diff(f.nodes_start)
        k = int(np.flatnonzero(sizes >= lo)[1])
        e0 = int(f.nodes_start[k] - k)
        bad = f.edges.copy()
        if mistake == "duplicate":
            bad[2 * e0], bad[2 * e0 + 1] = bad[2 * e0 + 2], bad[2 * e0 + 3]
        else:
            # the net's edge 0 re-attached to one of its own endpoint's other neighbours
            ed = bad[2 * e0: 2 * (e0 + sizes[k] - 1)].reshape(-1, 2)
            a = ed[0, 0]
            other = [int(t if s == a else s) for s, t in ed[1:] if a in (s, t)]
            if not other:
                a = ed[0, 1]
                other = [int(t if s == a else s) for s, t in ed[1:] if a in (s, t)]
            nb = [int(t if s == other[0] else s) for s, t in ed[1:] if other[0] in (s, t) and a not in (s, t)]
            ed[0] = (a, nb[0])  # a - other[0] - nb[0] - a: a cycle; ed[0]'s far side cut off
        f.edges = bad
        with pytest.raises(N.OtgError) as ei:
            RcSolver.from_edges(f, ctx)
        assert ei.value.status == 2
        assert f"net {k}" in str(ei.value), (lo, mistake, k, str(ei.value))


def test_plain_bitexact_vs_reference_arithmetic(cfg0):
    f, s = cfg0
    h = s.host_topology()
    s.solve("plain")
    g = s.host()
    ref = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "flat_f32")
    for k in ("load", "delay", "ldelay", "beta", "impulse"):
        assert np.array_equal(g[k].view(np.uint32), ref[k].view(np.uint32)), k
    assert rel_err(g["slew"], ref["slew"], 1e-30) < 1e-6


@pytest.mark.parametrize("method,tol64", [("batched", 1e-5), ("compensated", 1e-6)])
def test_batched_cfg0(cfg0, method, tol64):
    f, s = cfg0
    h = s.host_topology()
    s.solve(method)
    g = s.host()
    r32 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "dfs_f32")
    r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
    fl = floors(r64)
    for k in ("delay", "slew"):
        assert rel_err(g[k], r32[k], fl[k]) < 1e-5, (k, "vs cpu timer fp32")
    for k in OUT:
        assert rel_err(g[k], r64[k], fl[k]) < tol64, (k, "vs fp64")


def _forest(cfg, ctx):
    f = W.generate_rc(cfg)
    return f, RcSolver.from_edges(f, ctx)


@pytest.mark.parametrize("ring", [256, 512, 1024])
@pytest.mark.parametrize("method,tol64", [("batched", 1e-5), ("compensated", 1e-6)])
def test_batched_mixed_sizes(ctx, ring, method, tol64):
    # heavy-tailed mix (cfg1 distribution, scaled down) exercises packed groups,
    # streamed nets and global-window nets
    cfg = W.config1(num_nets=3000, seed=21)
    cfg.classes[2].weight = 0.02
    cfg.classes[2].max_nodes = 3000
    f, s = _forest(cfg, ctx)
    h = s.host_topology()
    s.solve(method, tuning=(ring, 1, 0, 0))
    st = s.plan_stats(method, (ring, 1, 0, 0))
    assert st["packed_nets"] > 0 and st["stream_nets"] + st["global_nets"] > 0
    g = s.host()
    r32 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "dfs_f32")
    r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
    fl = floors(r64)
    for k in ("delay", "slew"):
        assert rel_err(g[k], r32[k], fl[k]) < 1e-5, k
    for k in OUT:
        assert rel_err(g[k], r64[k], fl[k]) < tol64, k


def test_global_window_nets(ctx):
    # wide nets (stars / balanced trees) larger than the window go through the
    # global-memory window
    cfg = W.RcConfig("wide", 40, [W.SizeClass(0.5, 600, 1500, N.SHAPE_BALANCED),
                                  W.SizeClass(0.5, 600, 900, N.SHAPE_RANDOM)], seed=5)
    f, s = _forest(cfg, ctx)
    h = s.host_topology()
    for method, tol in (("batched", 1e-5), ("compensated", 1e-6)):
        s.solve(method, tuning=(256, 1, 0, 0))
        assert s.plan_stats(method, (256, 1, 0, 0))["global_nets"] > 0
        g = s.host()
        r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
        fl = floors(r64)
        for k in OUT:
            assert rel_err(g[k], r64[k], fl[k]) < tol, (method, k)


def test_deep_chain_precision(ctx):
    cfg = W.config3(num_nets=6, seed=4)
    cfg.classes[0].max_nodes = 6000
    f, s = _forest(cfg, ctx)
    h = s.host_topology()
    r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
    fl = floors(r64)
    s.solve("compensated")
    g = s.host()
    for k in OUT:
        assert rel_err(g[k], r64[k], fl[k]) < 1e-6, k


def test_edge_cases(ctx):
    # single-node nets, two-node nets, empty net, net of one chain
    ns = np.array([0, 1, 1, 3, 4, 9], np.int32)
    pid = np.array([0, 0, 1, 0, 0, 1, 1, 2, 3], np.int32)
    rng = np.random.default_rng(0)
    pres = rng.uniform(0.01, 1, 9).astype(np.float32)
    pres[[0, 1, 3, 4]] = 0
    cap = rng.uniform(0.1, 2, 36).astype(np.float32)
    si = rng.uniform(-20, 20, 20).astype(np.float32)
    s = RcSolver.from_bfs_arrays(ns, pid, pres, cap, si, ctx)
    r64 = O.rct_all(ns, pid, pres, cap, si, "f64")
    fl = floors(r64)
    for method in ("plain", "batched", "compensated"):
        s.solve(method)
        g = s.host()
        for k in OUT:
            assert rel_err(g[k], r64[k], fl[k]) < 1e-5, (method, k)


def test_repeat_and_determinism(cfg0):
    f, s = cfg0
    s.solve("batched")
    a = {k: v.copy() for k, v in s.host().items()}
    for _ in range(3):
        s.solve("batched")
    b = s.host()
    for k in OUT:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


def test_subset_solve(ctx):
    cfg = W.config1(num_nets=2000, seed=31)
    f, s = _forest(cfg, ctx)
    s.solve("batched")
    full = {k: v.copy() for k, v in s.host().items()}
    # perturb caps of some nets on the device, re-solve only those
    rng = np.random.default_rng(1)
    ids = np.sort(rng.choice(cfg.num_nets, 100, replace=False)).astype(np.int32)
    ns = s.nodes_start.cpu().numpy()
    cap = s.cap.clone()
    for a in ids:
        cap[ns[a] * 4: ns[a + 1] * 4] *= 1.1
    s.cap = cap
    s.solve_subset(torch.from_numpy(ids).to(s.device), "batched")
    sub = s.host()
    h = s.host_topology()
    r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
    fl = floors(r64)
    mask = np.zeros(s.num_nodes * 4, bool)
    for a in ids:
        mask[ns[a] * 4: ns[a + 1] * 4] = True
    for k in OUT:
        assert rel_err(sub[k][mask], r64[k][mask], fl[k]) < 1e-5, k
        assert np.array_equal(sub[k][~mask], full[k][~mask]), k


def test_bfs_and_compute_host_api(ctx):
    f = W.generate_rc(W.config0(num_nets=300, seed=9))
    V, Nn = f.num_nodes, f.num_nets
    a = N.RctEdgeArray()
    bufs = dict(edges=f.edges.copy(), n2b=np.zeros(V, np.int32), pid=np.zeros(V, np.int32))
    outs = {k: np.zeros(4 * V, np.float32) for k in OUT}
    a.num_nets, a.total_num_nodes, a.total_num_edges = Nn, V, V - Nn
    a.rct_nodes_start = N.addr(f.nodes_start)
    a.rct_edges = N.addr(bufs["edges"])
    a.rct_roots = N.addr(f.roots)
    a.rct_edges_res = N.addr(f.edge_res)
    a.rct_nodes_cap = N.addr(f.node_cap)
    a.rct_node2bfs_order = N.addr(bufs["n2b"])
    a.rct_pid = N.addr(bufs["pid"])
    for k in OUT:
        setattr(a, k, N.addr(outs[k]))
    a.slew_in = N.addr(f.slew_in)
    import ctypes as C
    ctx.check(N.lib.otg_rct_bfs_and_compute(ctx.ptr, C.byref(a), N.BATCHED))
    e, n2b, pid, pres, cap = oracle_inputs(f)
    assert np.array_equal(bufs["n2b"], n2b) and np.array_equal(bufs["pid"], pid)
    assert np.array_equal(bufs["edges"], e)
    r64 = O.rct_all(f.nodes_start, pid, pres, cap, f.slew_in, "f64")
    fl = floors(r64)
    for k in OUT:
        assert rel_err(outs[k], r64[k], fl[k]) < 1e-5, k


def test_host_api_large_arrays(ctx):
    """Host arrays of 60-130 MB (7.9M nodes) go through the library's staged
    pinned copies, several chunks per copy thread: both host entry points
    must return exactly what the device-pointer path returns for the same
    forest (same kernels, so bit-identical)."""
    import ctypes as C
    f = W.generate_rc(W.config1(num_nets=150000, seed=5))
    V, Nn = f.num_nodes, f.num_nets
    assert 16 * V > (64 << 20)
    s = RcSolver.from_edges(f, ctx)
    ref = {k: v.cpu().numpy() for k, v in s.solve("compensated").items()}
    a = N.RctEdgeArray()
    bufs = dict(edges=f.edges.copy(), n2b=np.full(V, -1, np.int32), pid=np.full(V, -1, np.int32))
    outs = {k: np.full(4 * V, np.nan, np.float32) for k in OUT}
    a.num_nets, a.total_num_nodes, a.total_num_edges = Nn, V, V - Nn
    a.rct_nodes_start = N.addr(f.nodes_start)
    a.rct_edges = N.addr(bufs["edges"])
    a.rct_roots = N.addr(f.roots)
    a.rct_edges_res = N.addr(f.edge_res)
    a.rct_nodes_cap = N.addr(f.node_cap)
    a.rct_node2bfs_order = N.addr(bufs["n2b"])
    a.rct_pid = N.addr(bufs["pid"])
    for k in OUT:
        setattr(a, k, N.addr(outs[k]))
    a.slew_in = N.addr(f.slew_in)
    ctx.check(N.lib.otg_rct_bfs_and_compute(ctx.ptr, C.byref(a), N.BATCHED_COMPENSATED))
    assert np.array_equal(bufs["n2b"], s.node2bfs.cpu().numpy())
    assert np.array_equal(bufs["pid"], s.pid.cpu().numpy())
    assert np.array_equal(bufs["edges"], s.edges_directed.cpu().numpy())
    for k in OUT:
        assert np.array_equal(outs[k].view(np.uint32), ref[k].reshape(-1).view(np.uint32)), k
    pres = s.pres.cpu().numpy()
    cap = s.cap.cpu().numpy()
    outs2 = {k: np.full(4 * V, np.nan, np.float32) for k in OUT}
    h = N.RctHost(num_nets=Nn, total_num_nodes=V, rct_nodes_start=N.addr(f.nodes_start),
                  rct_pid=N.addr(bufs["pid"]), pres=N.addr(pres), cap=N.addr(cap),
                  load=N.addr(outs2["load"]), delay=N.addr(outs2["delay"]),
                  ldelay=N.addr(outs2["ldelay"]), impulse=N.addr(outs2["impulse"]),
                  beta=N.addr(outs2["beta"]), slew=N.addr(outs2["slew"]),
                  slew_in=N.addr(f.slew_in))
    ctx.check(N.lib.otg_rct_compute(ctx.ptr, C.byref(h), N.BATCHED_COMPENSATED))
    for k in OUT:
        assert np.array_equal(outs2[k].view(np.uint32), ref[k].reshape(-1).view(np.uint32)), k


@pytest.mark.parametrize("method", ["plain", "batched", "compensated"])
def test_rct_compute_host_api(ctx, method):
    """otg_rct_compute: host-pointer drop-in for rct_compute_cuda(RctCUDA)."""
    import ctypes as C
    f = W.generate_rc(W.config1(num_nets=400, seed=13))
    e, n2b, pid, pres, cap = oracle_inputs(f)
    V, Nn = f.num_nodes, f.num_nets
    outs = {k: np.zeros(4 * V, np.float32) for k in OUT}
    h = N.RctHost(num_nets=Nn, total_num_nodes=V, rct_nodes_start=N.addr(f.nodes_start),
                  rct_pid=N.addr(pid), pres=N.addr(pres), cap=N.addr(cap),
                  load=N.addr(outs["load"]), delay=N.addr(outs["delay"]),
                  ldelay=N.addr(outs["ldelay"]), impulse=N.addr(outs["impulse"]),
                  beta=N.addr(outs["beta"]), slew=N.addr(outs["slew"]), slew_in=N.addr(f.slew_in))
    m = {"plain": N.PLAIN, "batched": N.BATCHED, "compensated": N.BATCHED_COMPENSATED}[method]
    ctx.check(N.lib.otg_rct_compute(ctx.ptr, C.byref(h), m))
    r64 = O.rct_all(f.nodes_start, pid, pres, cap, f.slew_in, "f64")
    fl = floors(r64)
    for k in OUT:
        assert rel_err(outs[k], r64[k], fl[k]) < 1e-5, (method, k)


def test_bad_topology_rejected(ctx):
    ns = np.array([0, 3], np.int32)
    pid = np.array([0, 1, 3], np.int32)  # node 2 points before the net start
    s = RcSolver.from_bfs_arrays(ns, pid, np.ones(3, np.float32), np.ones(12, np.float32),
                                 np.ones(4, np.float32), ctx)
    with pytest.raises(N.OtgError) as ei:
        s.solve("batched")
    assert ei.value.status == 3


def test_not_a_tree_rejected(ctx):
    f = W.generate_rc(W.config0(num_nets=10, seed=3))
    bad = f.edges.copy()
    bad[0], bad[1] = bad[2], bad[3]  # duplicate an edge -> cycle + disconnected node
    f.edges = bad
    with pytest.raises(N.OtgError) as ei:
        RcSolver.from_edges(f, ctx)
    assert ei.value.status == 2


@pytest.mark.parametrize("method,tol64", [("batched", 1e-5), ("compensated", 1e-6)])
def test_wide_nets_take_large_rings(ctx, method, tol64):
    # balanced trees of up to 3000 nodes have BFS levels up to ~1500 wide,
    # so their parent offsets need the 1024- and 2048-slot rings (default
    # tuning: no net falls back to the global-window path)
    cfg = W.RcConfig("wide", 400, [W.SizeClass(0.5, 2, 64), W.SizeClass(0.5, 600, 3000, N.SHAPE_BALANCED)],
                     seed=13)
    f, s = _forest(cfg, ctx)
    h = s.host_topology()
    st = s.plan_stats(method)
    assert st["global_nets"] == 0 and st["stream_nets"] > 0
    ns, pid = h["nodes_start"], h["pid"]
    maxpid = max(int(pid[ns[k]:ns[k + 1]].max()) for k in range(f.num_nets))
    assert maxpid > 960  # beyond the 1024-slot ring: the 2048 ring is used
    s.solve(method)
    g = s.host()
    r32 = O.rct_all(ns, pid, h["pres"], h["cap"], h["slew_in"], "dfs_f32")
    r64 = O.rct_all(ns, pid, h["pres"], h["cap"], h["slew_in"], "f64")
    fl = floors(r64)
    for k in ("delay", "slew"):
        assert rel_err(g[k], r32[k], fl[k]) < 1e-5, k
    for k in OUT:
        assert rel_err(g[k], r64[k], fl[k]) < tol64, k


@pytest.mark.parametrize("method,tol64", [("batched", 1e-5), ("compensated", 1e-6)])
def test_chains(ctx, method, tol64):
    # pure chains (every 64-position window is a path: the compensated
    # kernels take it with wave-wide scans) next to branching chains
    cfg = W.RcConfig("chains", 300, [W.SizeClass(0.6, 100, 5000, N.SHAPE_CHAIN, 0.0),
                                     W.SizeClass(0.4, 100, 3000, N.SHAPE_CHAIN, 0.05)],
                     seed=17, r_range=(1e-3, 1e4), c_range=(1e-3, 1e2))
    f, s = _forest(cfg, ctx)
    h = s.host_topology()
    s.solve(method)
    g = s.host()
    r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
    fl = floors(r64)
    for k in OUT:
        assert rel_err(g[k], r64[k], fl[k]) < (1e-6 if method == "compensated" else 5e-3), k


def test_subset_solve_compensated_all_bins(ctx):
    # small nets (packed groups), branching chains (ordinary streamed bins),
    # pure chains (chain bins) and wide balanced trees (large rings)
    cfg = W.RcConfig("mix", 1500, [W.SizeClass(0.6, 2, 40), W.SizeClass(0.2, 100, 3000, N.SHAPE_CHAIN, 0.1),
                                   W.SizeClass(0.15, 200, 3000, N.SHAPE_CHAIN, 0.0),
                                   W.SizeClass(0.05, 600, 2500, N.SHAPE_BALANCED)], seed=41)
    f, s = _forest(cfg, ctx)
    s.solve("compensated")
    full = {k: v.copy() for k, v in s.host().items()}
    rng = np.random.default_rng(3)
    ids = np.sort(rng.choice(cfg.num_nets, 300, replace=False)).astype(np.int32)
    ns = s.nodes_start.cpu().numpy()
    cap = s.cap.clone()
    for a in ids:
        cap[ns[a] * 4: ns[a + 1] * 4] *= 1.25
    s.cap = cap
    for _ in range(2):  # a second call starts from the state the first one left
        s.solve_subset(torch.from_numpy(ids).to(s.device), "compensated")
    sub = s.host()
    h = s.host_topology()
    r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
    fl = floors(r64)
    mask = np.zeros(s.num_nodes * 4, bool)
    for a in ids:
        mask[ns[a] * 4: ns[a + 1] * 4] = True
    for k in OUT:
        assert rel_err(sub[k][mask], r64[k][mask], fl[k]) < 1e-6, k
        assert np.array_equal(sub[k][~mask], full[k][~mask]), k


def test_reference_f64_matches_fp64_oracle(ctx):
    # the device FP64 reference (used for precision reports of workloads too
    # large for the host oracle) against the host FP64 oracle: small trees,
    # branching chains and balanced trees over the cfg3 value ranges; sums
    # agree to FP64 summation-order noise, impulse (2 beta - delay^2 cancels)
    # to that noise relative to its floor
    cfg = W.RcConfig("ref64", 600, [W.SizeClass(0.5, 2, 64), W.SizeClass(0.3, 100, 3000, N.SHAPE_CHAIN, 0.1),
                                    W.SizeClass(0.2, 64, 1024, N.SHAPE_BALANCED)], seed=23,
                     r_range=(1e-3, 1e4), c_range=(1e-3, 1e2))
    f, s = _forest(cfg, ctx)
    h = s.host_topology()
    ref = {k: v.cpu().numpy() for k, v in s.reference_f64().items()}
    r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
    fl = floors(r64)
    for k in OUT:
        tol = 1e-9 if k in ("impulse", "slew") else 1e-12
        assert rel_err(ref[k], r64[k], fl[k]) < tol, k


# launch arrangements of one solve: streams for the streaming launches (1, 3,
# 4), packed groups first or from a queue beside them (group_alongside 1 or
# 4 workgroups per CU), the largest ring on the caller's stream
ARRANGEMENTS = [
    (0, 0, 0, 0, 1),
    (0, 0, 0, 0, 3),
    (0, 0, 0, 0, 4),
    (0, 0, 0, 0, 3, 0, 0, 0, 0, 1),
    (0, 0, 0, 0, 3, 0, 0, 0, 0, 4),
    (0, 0, 0, 0, 4, 0, 0, 1, 0, 2),
]


@pytest.mark.parametrize("method", ["batched", "compensated"])
def test_launch_arrangements_do_not_change_results(ctx, method):
    # every arrangement gives identical bits (every net is solved by one
    # workgroup with the same code, whatever runs beside it), for full and
    # subset solves
    cfg = W.RcConfig("mixstreams", 2500, [W.SizeClass(0.7, 2, 40), W.SizeClass(0.2, 100, 3000, N.SHAPE_CHAIN, 0.1),
                                          W.SizeClass(0.05, 200, 3000, N.SHAPE_CHAIN, 0.0),
                                          W.SizeClass(0.05, 600, 2500, N.SHAPE_BALANCED)], seed=43)
    f, s = _forest(cfg, ctx)
    st = s.plan_stats(method, ARRANGEMENTS[1])
    assert st["stream_nets"] > 0 and st["packed_groups"] > 0
    ids = torch.arange(0, cfg.num_nets, 7, dtype=torch.int32, device=s.device)
    s.alloc_outputs()
    full, sub = [], []
    for tun in ARRANGEMENTS:
        for k in OUT:
            s.out[k].fill_(float("nan"))
        s.solve(method, tuning=tun)
        full.append({k: v.clone() for k, v in s.out.items()})
        for k in OUT:
            s.out[k].zero_()
        s.solve_subset(ids, method, tuning=tun)
        sub.append({k: v.clone() for k, v in s.out.items()})
    for i, tun in enumerate(ARRANGEMENTS[1:], 1):
        for k in OUT:
            assert torch.equal(full[0][k], full[i][k]), (tun, k)
            assert torch.equal(sub[0][k], sub[i][k]), (tun, k)


@pytest.mark.parametrize("method", ["batched", "compensated"])
def test_groups_beside_no_streamed_nets(ctx, method):
    # packed groups from a queue on their own stream when the forest has no
    # net larger than a group (no streaming launch to run beside): same bits
    cfg = W.RcConfig("smallonly", 3000, [W.SizeClass(1.0, 2, 60)], seed=44)
    f, s = _forest(cfg, ctx)
    assert s.plan_stats(method)["stream_nets"] == 0
    s.alloc_outputs()
    ids = torch.arange(0, cfg.num_nets, 5, dtype=torch.int32, device=s.device)
    res = []
    for tun in ((0, 0, 0, 0, 3), (0, 0, 0, 0, 3, 0, 0, 0, 0, 2)):
        for k in OUT:
            s.out[k].fill_(float("nan"))
        s.solve(method, tuning=tun)
        s.solve_subset(ids, method, tuning=tun)
        res.append({k: v.clone() for k, v in s.out.items()})
    for k in OUT:
        assert torch.equal(res[0][k], res[1][k]), k


@pytest.mark.parametrize("tun", [
    (0, 0, 0, 0, 3, 64),             # nets of 65..256 nodes streamed instead of packed
    (0, 0, 0, 0, 3, 0, 0b11101),     # no 320-slot ring
    (0, 0, 0, 0, 3, 0, 0, 1),        # largest ring on the caller's stream
    (0, 0, 0, 0, 3, 0, 0, 0, 4),     # smallest-ring launch held to 4 workgroups per CU
    (0, 320, 128, 0, 2, 100, 0b11110, 1, 3),
])
@pytest.mark.parametrize("method,tol64", [("batched", 1e-5), ("compensated", 1e-6)])
def test_tuning_knobs_keep_results(ctx, tun, method, tol64):
    cfg = W.RcConfig("knobs", 2500, [W.SizeClass(0.55, 2, 40), W.SizeClass(0.25, 41, 256),
                                     W.SizeClass(0.15, 257, 3000, N.SHAPE_CHAIN, 0.1),
                                     W.SizeClass(0.05, 600, 2500, N.SHAPE_BALANCED)], seed=47)
    f, s = _forest(cfg, ctx)
    h = s.host_topology()
    s.solve(method, tuning=tun)
    st = s.plan_stats(method, tun)
    if len(tun) > 5 and tun[5]:
        assert st["packed_nets"] == int(np.sum((np.diff(h["nodes_start"]) >= 1)
                                               & (np.diff(h["nodes_start"]) <= tun[5])))
    g = s.host()
    r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
    fl = floors(r64)
    for k in OUT:
        assert rel_err(g[k], r64[k], fl[k]) < tol64, (tun, k)


def test_batched_solves_reject_unaligned_corner_arrays(ctx):
    # the batched kernels move a node's four corners as one 16-byte vector:
    # a result array (or cap) 4 bytes off a 16-byte boundary is refused
    # before any launch, by the full and the subset solve; the plain method
    # takes it
    import ctypes as C
    cfg = W.RcConfig("align", 300, [W.SizeClass(0.7, 2, 40), W.SizeClass(0.3, 300, 700)], seed=5)
    f, s = _forest(cfg, ctx)
    good = s.result()
    V = s.num_nodes
    spare = torch.empty(4 * V + 4, dtype=torch.float32, device=s.device)
    ids = torch.arange(s.num_nets, dtype=torch.int32, device=s.device)
    for method in ("batched", "compensated"):
        m = {"batched": N.BATCHED, "compensated": N.BATCHED_COMPENSATED}[method]
        plan = s.plan(method)
        for field in ("load", "delay", "slew"):
            bad = N.RctResult(*[getattr(good, k) for k, _ in N.RctResult._fields_])
            setattr(bad, field, spare.data_ptr() + 4)
            st = N.lib.otg_rct_solve(ctx.ptr, plan, C.byref(s.tree()), C.byref(bad), m, None)
            assert st == 1, (method, field, st)  # OTG_ERR_INVALID_ARGUMENT
            st = N.lib.otg_rct_solve_subset(ctx.ptr, plan, C.byref(s.tree()), C.byref(bad), m,
                                            ids.data_ptr(), s.num_nets, None)
            assert st == 1, (method, field, st)
        tree = s.tree()
        cap2 = torch.empty(4 * V + 4, dtype=torch.float32, device=s.device)
        cap2[1:4 * V + 1].copy_(s.cap)
        tree.cap = cap2.data_ptr() + 4
        assert N.lib.otg_rct_solve(ctx.ptr, plan, C.byref(tree), C.byref(good), m, None) == 1
    # the plain kernel reads scalars: an offset result array is fine
    bad = N.RctResult(*[getattr(good, k) for k, _ in N.RctResult._fields_])
    bad.delay = spare.data_ptr() + 4
    ctx.check(N.lib.otg_rct_solve(ctx.ptr, None, C.byref(s.tree()), C.byref(bad), N.PLAIN, None))
    torch.cuda.synchronize()
    s.solve("plain")
    assert torch.equal(spare[1:4 * V + 1], s.out["delay"][:4 * V])


def test_stream_ceiling_instrument(ctx):
    """otg_rct_stream_ceiling writes its documented function of the node
    inputs into every output (load = cap, delay = cap * (pres + pid), ...),
    so every load it is timed on is real; nets' data are read, not written."""
    f = W.generate_rc(W.config0(num_nets=700, seed=9))
    s = RcSolver.from_edges(f, ctx)
    s.alloc_outputs()
    for k in OUT:
        s.out[k].fill_(float("nan"))
    s.stream_ceiling()
    torch.cuda.synchronize()
    t = s.host_topology()
    g = s.host()
    fv = (t["pres"] + t["pid"].astype(np.float32)).astype(np.float32)
    c = t["cap"].reshape(-1, 4)
    fc = fv[:, None]
    want = {
        "load": c,
        "delay": c * fc,
        "ldelay": c + fc,
        "beta": c - fc,
        "impulse": np.stack([fv, c[:, 0], c[:, 1], c[:, 2]], 1),
        "slew": np.stack([c[:, 3], fv, c[:, 0], c[:, 1]], 1),
    }
    for k in OUT:
        assert np.array_equal(g[k].reshape(-1, 4), want[k].astype(np.float32)), k
