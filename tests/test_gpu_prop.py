"""GPU parity of level-wise forward propagation (otg_fprop) against the
oracle's restatement of the reference (timer.cpp:733-783, arc.cpp:82-147,
lut.cpp:161-241).

Bar: bit-exact.  Slew, arrival and every arc delay equal the oracle's FP32
values exactly, NaN (unset) in the same places: the kernel performs the same
operations in the same order with contraction off.  The level order comes
from the GPU levelizer; the oracle is run on the same frontiers (the order
of pins inside a level does not change any value).
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

import oracle as O
from opentimer_amd import native as N
from opentimer_amd import workloads as W
from opentimer_amd.levelize import Levelizer
from opentimer_amd.prop import Propagator

GKEYS = ("fanin_start", "fanin", "arc_kind", "arc_timing", "arc_rc_delay", "arc_rc_impulse",
         "pin_load", "pi_slew", "pi_at")
LKEYS = ("lut_hdr", "lut_val", "timing_sense", "timing_edge", "timing_lut", "timing_has")


def _oracle(g, lib, fr, ends):
    return O.fprop({k: getattr(g, k) for k in GKEYS}, {k: getattr(lib, k) for k in LKEYS},
                   fr, ends)


def _same(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape
    assert np.array_equal(np.isnan(a), np.isnan(b)), int((np.isnan(a) != np.isnan(b)).sum())
    np.testing.assert_array_equal(a[~np.isnan(a)].view(np.int32), b[~np.isnan(b)].view(np.int32))


def _run(ctx, dcfg, lib_seed=11, graph_seed=12, num_timings=64, **kw):
    print(f"fprop case {dcfg.name} pins={dcfg.num_pins} arcs={dcfg.num_arcs} "
          f"levels={dcfg.num_layers} lib_seed={lib_seed} graph_seed={graph_seed} "
          f"timings={num_timings} {kw}", flush=True)
    dag = W.generate_dag(dcfg, nthreads=8)
    csr = W.fanin_csr(dag)
    lib = W.prop_library(num_timings=num_timings, seed=lib_seed)
    g = W.prop_graph(csr, lib, seed=graph_seed, **kw)
    L = Levelizer(csr, ctx)
    L.run()
    P = Propagator(g, lib, ctx)
    P.run(L.frontiers, L.ends, L.num_ends - 1)
    slew, at, ad = P.host()
    fr, ends, _ = L.host()
    oslew, oat, oad = _oracle(g, lib, fr, ends)
    _same(slew, oslew)
    _same(at, oat)
    _same(ad, oad)
    # the level-ordered plan: the same values
    pos, slots = P.build_plan(L.frontiers, L.ends, L.num_ends - 1)
    assert (pos, slots) == (csr.n, len(csr.fanin))
    for t in (P.slew, P.at, P.arc_delay):
        t.fill_(float("nan"))
    P.run_plan()
    ps, pa, pd = P.host()
    _same(ps, oslew)
    _same(pa, oat)
    _same(pd, oad)
    return slew, at, ad, P, L


@pytest.mark.parametrize("cfg", [W.config0_dag(),
                                 W.DagConfig("dag-300k", 300_000, 750_000, 60, seed=21)],
                         ids=lambda c: c.name)
def test_fprop_bitexact(ctx, cfg):
    slew, at, _, _, _ = _run(ctx, cfg)
    assert np.isfinite(at).mean() > 0.9


def _lds_words(lib):
    return 4 * lib.num_luts + 12 * lib.num_timings + len(lib.lut_val)


@pytest.mark.parametrize("num_timings", [3, 160])
def test_fprop_library_in_lds_and_in_device_memory(ctx, num_timings):
    # 36864 words is the LDS copy's limit: the first library is read from
    # LDS, the second from device memory
    lib = W.prop_library(num_timings=num_timings, seed=11)
    assert (_lds_words(lib) <= 36864) == (num_timings == 3)
    _run(ctx, W.config0_dag(seed=13), num_timings=num_timings)


def test_fprop_all_arc_kinds_and_dense_misses(ctx):
    # many net arcs without parasitics, loop breakers and sinks outside the tree
    cfg = W.DagConfig("dag-50k-mixed", 50_000, 140_000, 30, seed=22)
    _run(ctx, cfg, lib_seed=5, graph_seed=6, net_fraction=0.3, empty_fraction=0.2,
         skip_fraction=0.1, missing_sink_fraction=0.2)


def test_fprop_rerun_and_no_arc_delay(ctx):
    # a second run over the same outputs gives the same values; without the
    # arc-delay output the pin values do not change
    cfg = W.config0_dag(seed=9)
    slew, at, ad, P, L = _run(ctx, cfg)
    P.arc_delay.fill_(1234.0)
    P.run(L.frontiers, L.ends, L.num_ends - 1, arc_delay=False)
    s2, a2, ad2 = P.host()
    _same(s2, slew)
    _same(a2, at)
    assert (ad2 == 1234.0).all()


def test_fprop_plan_with_sources_outside_it(ctx):
    # a plan over all levels but the last (the sources): their values are
    # read from the caller's arrays, which hold a previous run's results
    cfg = W.DagConfig("dag-40k", 40_000, 100_000, 20, seed=23)
    slew, at, ad, P, L = _run(ctx, cfg)
    lv = L.num_ends - 1
    fr, ends, _ = L.host()
    pos, slots = P.build_plan(L.frontiers, L.ends, lv - 1)
    assert pos == ends[lv - 1]
    P.at[torch.from_numpy(fr[: ends[lv - 1]]).long().cuda()] = 0.0  # overwritten by the run
    P.run_plan()
    s2, a2, d2 = P.host()
    _same(s2, slew)
    _same(a2, at)
    _same(d2, ad)
    # rerunning a plan gives the same values; a graph of another size is refused
    P.run_plan()
    _same(P.host()[1], at)
    bad = N.PropGraph(P.graph.num_pins + 1, P.graph.num_arcs, *([None] * 9))
    out = N.PropResult(P.slew.data_ptr(), P.at.data_ptr(), None)
    assert N.lib.otg_fprop_run(ctx.ptr, P._plan, C.byref(bad), C.byref(P.lib), C.byref(out),
                               None) == 1
    P.close_plan()


def test_fprop_plan_unaligned_arrays(ctx):
    # per-pin and per-arc arrays 4 bytes off a 16-byte boundary take the
    # plan's scalar gather and scatter
    cfg = W.config0_dag(seed=14)
    slew, at, ad, P, L = _run(ctx, cfg)

    def shifted(t):
        buf = torch.empty(t.numel() + 1, dtype=t.dtype, device=t.device)
        buf[1:] = t.reshape(-1)
        return buf, buf.data_ptr() + 4
    keep = []
    g = N.PropGraph.from_buffer_copy(P.graph)
    for name in ("arc_rc_delay", "arc_rc_impulse", "pin_load", "pi_slew", "pi_at"):
        buf, ptr = shifted(P._t[name])
        keep.append(buf)
        setattr(g, name, ptr)
    so, sp = shifted(torch.full_like(P.slew, float("nan")))
    ao, ap = shifted(torch.full_like(P.at, float("nan")))
    P.arc_delay.fill_(float("nan"))
    out = N.PropResult(sp, ap, P.arc_delay.data_ptr())
    assert sp % 16 == 4 and P._t["pin_load"].data_ptr() % 16 == 0
    ctx.check(N.lib.otg_fprop_run(ctx.ptr, P._plan, C.byref(g), C.byref(P.lib), C.byref(out),
                                  None))
    torch.cuda.synchronize()
    _same(so[1:].reshape(-1, 4)[: P.n].cpu().numpy(), slew)
    _same(ao[1:].reshape(-1, 4)[: P.n].cpu().numpy(), at)
    _same(P.arc_delay[: P.m].cpu().numpy(), ad)
    P.close_plan()


@pytest.mark.parametrize("design", ("c17", "s27_spef", "s400"))
def test_tau15_whole_chain_on_gpu(ctx, design):
    # SPEF -> GPU BFS and compensated RC -> GPU levelization -> GPU
    # propagation over a plan, against the reference's printed arrival
    # times (3 decimals; tolerance of tests/test_prop_tau15.py) and, on the
    # same RC inputs, bit-exact vs the oracle
    import tau15_case as T
    from opentimer_amd.rct import RcSolver
    print(f"tau15 design {design}", flush=True)

    def gpu_rc(forest):
        s = RcSolver.from_edges(forest, ctx)
        s.solve("compensated")
        h = s.host()
        return s.host_topology()["node2bfs"], h["load"], h["delay"], h["impulse"]
    g, lib, pins, gold = T.build(gpu_rc, design)
    n = len(pins)
    graph = W.PropGraphData(*(g[k] for k in GKEYS))
    library = W.PropLibrary(*(lib[k] for k in LKEYS))
    csr = W.FaninCsr(n, g["fanin_start"], g["fanin"],
                     np.bincount(g["fanin"], minlength=n).astype(np.int32), None)
    csr.frontier = np.flatnonzero(csr.out_degree == 0).astype(np.int32)
    L = Levelizer(csr, ctx)
    L.run()
    P = Propagator(graph, library, ctx)
    P.build_plan(L.frontiers, L.ends, L.num_ends - 1)
    P.run_plan()
    slew, at, ad = P.host()
    m = ~np.isnan(gold)
    err = np.abs(at.astype(np.float64)[m] - gold[m])
    assert (err <= 0.0005 + 1e-5 + 1e-6 * np.abs(gold[m])).all(), float(err.max())
    fr, ends, _ = L.host()
    os_, oa, od = O.fprop(g, lib, fr, ends)
    _same(slew, os_)
    _same(at, oa)
    _same(ad, od)
    # required times and slacks (c17: every required time comes from a
    # primary output; the sequential designs' come from timing checks too)
    po_rat, gr, gs = T.required(design, pins)
    rat, slack = P.required(po_rat, L.frontiers, L.ends, L.num_ends - 1)
    orat, oslack = O.bprop(g["fanin_start"], g["fanin"], ad, po_rat, at, fr, ends)
    _same(rat.cpu().numpy(), orat)
    _same(slack.cpu().numpy(), oslack)
    if design == "c17":
        for got, want in ((orat, gr), (oslack, gs)):
            err = np.abs(got.astype(np.float64) - want)
            assert (err <= 0.0005 + 1e-5 + 1e-6 * np.abs(want)).all()


@pytest.mark.parametrize("cfg", [W.config0_dag(seed=15),
                                 W.DagConfig("dag-200k", 200_000, 500_000, 40, seed=24)],
                         ids=lambda c: c.name)
def test_bprop_bitexact(ctx, cfg):
    # required times and slacks (otg_bprop) from primary-output values on the
    # sinks and 5% of the other pins, over the forward pass's results
    slew, at, ad, P, L = _run(ctx, cfg)
    fr, ends, _ = L.host()
    g = {k: P._t[k].cpu().numpy() for k in ("fanin_start", "fanin")}
    rng = np.random.default_rng(3)
    n = len(g["fanin_start"]) - 1
    out_deg = np.bincount(g["fanin"], minlength=n)
    po = np.full((n, 4), np.nan, np.float32)
    pick = (out_deg == 0) | (rng.random(n) < 0.05)
    po[pick] = rng.uniform(500, 3000, (int(pick.sum()), 4))
    rat, slack = P.required(po, L.frontiers, L.ends, L.num_ends - 1)
    orat, oslack = O.bprop(g["fanin_start"], g["fanin"], ad, po, at, fr, ends)
    _same(rat.cpu().numpy(), orat)
    _same(slack.cpu().numpy(), oslack)
    assert np.isfinite(orat).mean() > 0.5


def test_fprop_plan_high_fanin_positions(ctx):
    # positions with more than 16 fanin arcs are wavefront chunks of their
    # own, taken 16 arc slots at a time: fanins of 100, 17..40 and 11 (the
    # generated DAGs stay below 16)
    from test_prop_oracle import _graph, _levels
    rng = np.random.default_rng(41)
    n_src = 100
    arcs, kind, timing = [], [], []

    def add(u, v):
        arcs.append((u, v))
        kind.append(int(rng.choice(4, p=[0.4, 0.1, 0.45, 0.05])))
        timing.append(int(rng.integers(0, 64)))
    for u in range(n_src):
        add(u, n_src)                               # pin 100: fanin 100
    for k in range(10):                             # pins 101..110: fanin 17..40
        for u in rng.choice(n_src, int(rng.integers(17, 41)), replace=False):
            add(int(u), n_src + 1 + k)
    for k in range(11):                             # pin 111: fanin 11
        add(n_src + k, n_src + 11)
    n = n_src + 12
    m = len(arcs)
    pis = np.full((n, 4), np.nan, np.float32)
    pia = np.full((n, 4), np.nan, np.float32)
    pis[:n_src] = rng.uniform(5, 20, (n_src, 4))
    pia[:n_src] = rng.uniform(0, 10, (n_src, 4))
    rcd = rng.uniform(0, 5, (m, 4)).astype(np.float32)
    rci = rng.uniform(0, 40, (m, 4)).astype(np.float32)
    g, _ = _graph(n, arcs, kind, timing, rcd, rci, rng.uniform(1, 200, (n, 4)), pis, pia)
    lib = W.prop_library(seed=11)
    fr, ends = _levels(g)
    P = Propagator(W.PropGraphData(**g), lib, ctx)
    frt = torch.from_numpy(fr).cuda()
    ent = torch.from_numpy(ends).cuda()
    P.build_plan(frt, ent, len(ends) - 1)
    P.run_plan()
    slew, at, ad = P.host()
    os_, oa, od = O.fprop(g, {k: getattr(lib, k) for k in LKEYS}, fr, ends)
    _same(slew, os_)
    _same(at, oa)
    _same(ad, od)
    assert np.isfinite(oa[n_src:]).all()
    P.close_plan()


@pytest.mark.parametrize("fanout", [5, 40])
def test_bprop_signed_zero_ties_follow_fanout_order(ctx, fanout):
    # a hub pin drives `fanout` sinks through net arcs without parasitics
    # (delay 0) whose required times are +0 and -0 mixed: the sequential
    # relaxation keeps the first of the equal candidates in fanout order, and
    # so must the thread path (5 arcs) and the wavefront path (40 arcs)
    from test_prop_oracle import _graph, _levels, _scalar_lib
    n = fanout + 1
    arcs = [(0, k + 1) for k in range(fanout)]
    pis = np.full((n, 4), np.nan, np.float32)
    pis[0] = 5.0
    pia = np.full((n, 4), np.nan, np.float32)
    pia[0] = 0.0
    g, _ = _graph(n, arcs, [1] * fanout, pis=pis, pia=pia)
    lib, _ = _scalar_lib(1, np.random.default_rng(0))
    fr, ends = _levels(g)
    P = Propagator(W.PropGraphData(**g), W.PropLibrary(**lib), ctx)
    frt = torch.from_numpy(fr).cuda()
    ent = torch.from_numpy(ends).cuda()
    P.run(frt, ent, len(ends) - 1)
    _, at, ad = P.host()
    rng = np.random.default_rng(fanout)
    for trial in range(4):
        po = np.full((n, 4), np.nan, np.float32)
        po[1:] = np.where(rng.random((fanout, 4)) < 0.5, np.float32(-0.0), np.float32(0.0))
        rat, slack = P.required(po, frt, ent, len(ends) - 1)
        orat, oslack = O.bprop(g["fanin_start"], g["fanin"], ad, po, at, fr, ends)
        _same(rat.cpu().numpy(), orat)
        _same(slack.cpu().numpy(), oslack)
        assert np.signbit(orat[0]).tolist() == np.signbit(po[1, [0, 1, 2, 3]]).tolist()


def test_fprop_and_bprop_replayed_in_a_hip_graph(ctx):
    # a plan run and a backward pass allocate nothing and clear nothing with
    # memset: captured once into a hipGraph, every replay (with eager work and
    # synchronisation between replays) gives the eager results bit for bit
    from opentimer_amd.incremental import Graph
    cfg = W.config0_dag(seed=16)
    slew, at, ad, P, L = _run(ctx, cfg)
    lv = L.num_ends - 1
    n = P.n
    out_deg = np.bincount(P._t["fanin"].cpu().numpy()[: P.m], minlength=n)
    po = np.full((n, 4), np.nan, np.float32)
    po[out_deg == 0] = 900.0
    rat0, slack0 = (t.cpu().numpy() for t in P.required(po, L.frontiers, L.ends, lv))
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    h = stream.cuda_stream
    ctx.check(N.lib.otg_graph_begin(ctx.ptr, h))
    try:
        P.run_plan(stream=h)
        P.rerun_required(L.frontiers, L.ends, lv, stream=h)
    finally:
        gh = C.c_void_p()
        st = N.lib.otg_graph_end(ctx.ptr, h, C.byref(gh))
    ctx.check(st)
    G = Graph(ctx, gh)
    try:
        for rep in range(3):
            for t in (P.slew, P.at, P.arc_delay, P._bp[2], P._bp[3]):
                t.fill_(float("nan"))
            torch.cuda.synchronize()
            G.launch(stream)
            stream.synchronize()
            s2, a2, d2 = P.host()
            _same(s2, slew)
            _same(a2, at)
            _same(d2, ad)
            _same(P._bp[2][:n].cpu().numpy(), rat0)
            _same(P._bp[3][:n].cpu().numpy(), slack0)
    finally:
        G.close()
        P.close_plan()


def test_fprop_empty_and_invalid(ctx):
    lib = W.prop_library(num_timings=2)
    g = W.PropGraphData(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32),
                        np.zeros(0, np.int32), np.zeros((0, 4), np.float32),
                        np.zeros((0, 4), np.float32), np.zeros((0, 4), np.float32),
                        np.zeros((0, 4), np.float32), np.zeros((0, 4), np.float32))
    P = Propagator(g, lib, ctx)
    z = torch.zeros(2, dtype=torch.int32, device="cuda")
    P.run(z, z, 0)
    torch.cuda.synchronize()
    out = N.PropResult(P.slew.data_ptr(), P.at.data_ptr(), None)
    st = N.lib.otg_fprop(ctx.ptr, C.byref(P.graph), C.byref(P.lib), z.data_ptr(), z.data_ptr(),
                         -1, C.byref(out), None)
    assert st == 1  # OTG_ERR_INVALID_ARGUMENT
    st = N.lib.otg_fprop(ctx.ptr, None, C.byref(P.lib), z.data_ptr(), z.data_ptr(), 1,
                         C.byref(out), None)
    assert st == 1


def test_fprop_plan_rejects_malformed_levels(ctx):
    # level ends that do not start at 0, decrease or pass the pin count, a
    # frontier pin out of range and a fanin source out of range are refused
    # with OTG_ERR_INVALID_ARGUMENT; the well-formed levels build a plan
    dag = W.generate_dag(W.DagConfig("dag-small", 2000, 5000, 12, seed=31), nthreads=4)
    csr = W.fanin_csr(dag)
    L = Levelizer(csr, ctx)
    L.run()
    fr, ends, _ = L.host()
    nl = len(ends) - 1
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)  # noqa: E731
    fs, fi = t(csr.start), t(csr.fanin)

    def create(frontiers, level_ends, fanin=fi):
        plan = C.c_void_p()
        dfr, dends = t(frontiers), t(level_ends)  # held until the call returns
        st = N.lib.otg_fprop_plan_create(ctx.ptr, csr.n, len(csr.fanin), fs.data_ptr(),
                                         fanin.data_ptr(), dfr.data_ptr(), dends.data_ptr(), nl,
                                         None, C.byref(plan))
        if st == 0:
            N.lib.otg_fprop_plan_destroy(plan)
        return st

    assert create(fr, ends) == 0
    e = ends.copy()
    e[0] = 1
    assert create(fr, e) == 1
    e = ends.copy()
    e[nl // 2] = e[nl // 2 + 1] + 1  # a level end above the next one
    assert create(fr, e) == 1
    e = ends.copy()
    e[-1] = csr.n + 1
    assert create(fr, e) == 1
    f2 = fr.copy()
    f2[7] = csr.n + 5
    assert create(f2, ends) == 1
    f2[7] = -2
    assert create(f2, ends) == 1
    bad_fanin = csr.fanin.copy()
    bad_fanin[len(bad_fanin) // 3] = csr.n
    dbad = t(bad_fanin)
    assert create(fr, ends, dbad) == 1
    assert create(fr, ends) == 0  # the context stays usable


@pytest.mark.parametrize("use_plan", [False, True])
def test_fprop_partial_cone_over_the_full_fanin_csr(ctx, use_plan):
    # Incremental update: one pin's load changes; its fanout cone is
    # levelized from the candidate CSR (arcs between candidates only), and
    # propagated over the FULL fanin CSR, starting from the previous results
    # (a cone pin still relaxes over its fanin arcs from outside the cone,
    # timer.cpp:733-783).  Equal, bit for bit, to a full propagation of the
    # changed design.
    import dataclasses
    dag = W.generate_dag(W.DagConfig("dag-cone", 20_000, 50_000, 30, seed=41), nthreads=4)
    csr = W.fanin_csr(dag)
    lib = W.prop_library(num_timings=64, seed=11)
    g = W.prop_graph(csr, lib, seed=12)
    L = Levelizer(csr, ctx)
    L.run()
    P1 = Propagator(g, lib, ctx)
    P1.run(L.frontiers, L.ends, L.num_ends - 1)
    before = P1.host()
    # the changed input: the load of one pin (the cell arcs driving it see
    # it), whose fanout cone holds a few hundred pins (bounded searches)
    n = csr.n
    fo = [[] for _ in range(n)]
    for a, b in zip(dag.arc_from, dag.arc_to):
        fo[a].append(b)
    rng = np.random.default_rng(5)
    cell_driven = np.zeros(n, bool)
    cell_driven[np.repeat(np.arange(n), np.diff(csr.start))[g.arc_kind == 2]] = True
    best = None
    for p0 in rng.permutation(np.flatnonzero(cell_driven))[:400]:
        seen = {int(p0)}
        stack = [int(p0)]
        while stack and len(seen) <= 5000:
            for v in fo[stack.pop()]:
                if v not in seen:
                    seen.add(v)
                    stack.append(v)
        if 100 <= len(seen) <= 5000 and not stack:
            best = (int(p0), seen)
            break
    assert best is not None
    p0, cone = best
    load = g.pin_load.copy()
    load[p0] *= np.float32(1.5)
    g2 = dataclasses.replace(g, pin_load=load)
    # expected: a full propagation of the changed design
    P2 = Propagator(g2, lib, ctx)
    P2.run(L.frontiers, L.ends, L.num_ends - 1)
    want = P2.host()
    # incremental: candidate levelization, full CSR, previous results
    mask = np.zeros(n, np.uint8)
    mask[list(cone)] = 1
    cc = W.fanin_csr(dag, cand=mask)
    Lc = Levelizer(cc, ctx)
    Lc.run()
    fr_c, ends_c, _ = Lc.host()
    assert set(fr_c[: ends_c[-1]].tolist()) == cone
    # some cone pin has a fanin arc from outside the cone (the case the
    # candidate CSR would lose)
    fs = csr.start
    assert any(int(u) not in cone for v in cone for u in csr.fanin[fs[v]:fs[v + 1]])
    P3 = Propagator(g2, lib, ctx)
    P3.slew.copy_(P1.slew)
    P3.at.copy_(P1.at)
    P3.arc_delay.copy_(P1.arc_delay)
    if use_plan:
        P3.build_plan(Lc.frontiers, Lc.ends, Lc.num_ends - 1)
        P3.run_plan()
        P3.close_plan()
    else:
        P3.run(Lc.frontiers, Lc.ends, Lc.num_ends - 1)
    got = P3.host()
    for a, b in zip(got, want):
        _same(a, b)
    assert not np.array_equal(np.nan_to_num(want[1]), np.nan_to_num(before[1]))
