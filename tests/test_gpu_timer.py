"""The device timing update in one call (otg_timer, include/otg_timer.h) on
the TAU 2015 designs c17, s27_spef and s400.

The design goes in as the reference's timer holds it before an update: RC
trees in original node numbering with the sink pins' capacitances, net arcs
that name their net and sink node, per-pin driven nets.  One
otg_timer_update runs BFS ordering, the compensated RC solve, the gather of
net values at the arcs' sinks, levelization, forward and backward
propagation.  Checked against:
  * the reference's printed arrival times (3 decimals, the tolerance of
    tests/test_prop_tau15.py), and on c17 its required times and slacks;
  * the oracle's propagation over the timer's own levels with the net values
    that an RC solve of the same forest gives through the test's host
    gather (tests/tau15_case.build): bit for bit;
  * a second update with unchanged inputs (bit-identical), and an update
    after the caps change against a new timer built from the changed caps
    (bit-identical: the kept BFS order and the fresh ordering agree).
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

import oracle as O
from opentimer_amd import native as N
from opentimer_amd.rct import RcSolver


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    assert np.array_equal(np.isnan(a), np.isnan(b)), int((np.isnan(a) != np.isnan(b)).sum())
    np.testing.assert_array_equal(a[~np.isnan(a)].view(np.int32), b[~np.isnan(b)].view(np.int32))


class DeviceDesign:
    """Device copies of an assembled design and the C structs over them."""

    def __init__(self, D, cap_scale=1.0):
        dev = torch.device("cuda:0")
        f, arcs = D["forest"], D["arcs"]
        n = len(D["pins"])
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.t = {
            "nodes_start": t(f.nodes_start), "edges": t(f.edges), "roots": t(f.roots),
            "edge_res": t(f.edge_res),
            "node_cap": t((f.node_cap * np.float32(cap_scale)).astype(np.float32)),
            "fanin_start": t(D["fanin_start"]),
            "fanin": t(np.array([a[0] for a in arcs], np.int32)),
            "arc_kind": t(np.array([a[2] for a in arcs], np.int32)),
            "arc_timing": t(np.array([a[3] for a in arcs], np.int32)),
            "arc_net": t(np.array([a[4] for a in arcs], np.int32)),
            "arc_node": t(np.array([a[5] for a in arcs], np.int32)),
            "pin_net": t(D["pin_net"]), "pin_load": t(D["load_norc"]),
            "pi_slew": t(D["pi_slew"]), "pi_at": t(D["pi_at"]),
            "po_rat": t(D["po_rat"]),
        }
        for k, v in D["lib"].items():
            self.t["lib_" + k] = t(v)
        a = lambda k: self.t[k].data_ptr()  # noqa: E731
        self.design = N.TimerDesign(
            f.num_nets, int(f.nodes_start[-1]), a("nodes_start"), a("edges"), a("roots"),
            a("edge_res"), a("node_cap"), n, len(arcs), a("fanin_start"), a("fanin"),
            a("arc_kind"), a("arc_timing"), a("arc_net"), a("arc_node"), a("pin_net"),
            a("pin_load"), a("pi_slew"), a("pi_at"), a("po_rat"))
        L = D["lib"]
        self.lib = N.PropLib(len(L["lut_hdr"]) // 4, len(L["timing_sense"]), len(L["lut_val"]),
                             a("lib_lut_hdr"), a("lib_lut_val"), a("lib_timing_sense"),
                             a("lib_timing_edge"), a("lib_timing_lut"), a("lib_timing_has"))


class _Dev:
    """A raw device array seen by torch through __cuda_array_interface__."""

    def __init__(self, ptr, count, typestr):
        self.__cuda_array_interface__ = {"shape": (count,), "typestr": typestr,
                                         "data": (ptr, False), "version": 3}


def _host(ptr, count, typestr="<f4"):
    if count <= 0:
        return np.empty(0, np.float32 if typestr == "<f4" else np.int32)
    return torch.as_tensor(_Dev(ptr, count, typestr), device="cuda").cpu().numpy().copy()


def _results(timer, n, m):
    r = N.TimerResults()
    N.check(N.lib.otg_timer_results_get(timer, C.byref(r)))
    torch.cuda.synchronize()
    out = {}
    for k, count in (("slew", 4 * n), ("at", 4 * n), ("rat", 4 * n), ("slack", 4 * n),
                     ("arc_delay", 8 * m), ("arc_rc_delay", 4 * m), ("arc_rc_impulse", 4 * m),
                     ("pin_load", 4 * n)):
        out[k] = _host(getattr(r, k), count)
    out["frontiers"] = _host(r.frontiers, n, "<i4")
    out["ends"] = _host(r.frontier_ends, r.num_levels + 1, "<i4")
    return out


def _assemble(design):
    import tau15_case as T
    D = T.assemble(design)
    po_rat, gr, gs = T.required(design, D["pins"])
    D["po_rat"], D["gold_rat"], D["gold_slack"] = po_rat, gr, gs
    return D


@pytest.mark.parametrize("design", ("c17", "s27_spef", "s400"))
def test_timer_update_matches_the_reference_and_the_oracle(ctx, design):
    import tau15_case as T
    print(f"timer design {design}", flush=True)
    D = _assemble(design)
    n, m = len(D["pins"]), len(D["arcs"])
    dd = DeviceDesign(D)
    timer = N.timer_p()
    ctx.check(N.lib.otg_timer_create(ctx.ptr, C.byref(dd.design), C.byref(dd.lib),
                                     N.BATCHED_COMPENSATED, C.byref(timer)))
    try:
        ctx.check(N.lib.otg_timer_update(timer, None))
        torch.cuda.synchronize()
        R = _results(timer, n, m)
        at = R["at"].reshape(-1, 4)
        gold = D["gold"]
        mk = ~np.isnan(gold)
        err = np.abs(at.astype(np.float64)[mk] - gold[mk])
        assert (err <= 0.0005 + 1e-5 + 1e-6 * np.abs(gold[mk])).all(), float(err.max())
        # the net values equal the host gather from an RC solve of the same
        # forest (RcSolver: the same BFS and compensated kernels)

        def gpu_rc(forest):
            s = RcSolver.from_edges(forest, ctx)
            s.solve("compensated")
            h = s.host()
            return s.host_topology()["node2bfs"], h["load"], h["delay"], h["impulse"]
        g, lib, _, _ = T.build(gpu_rc, design)
        rc = np.array([a[2] == 0 for a in D["arcs"]])  # OTG_ARC_NET_RC
        _same(R["arc_rc_delay"].reshape(-1, 4)[rc], g["arc_rc_delay"][rc])
        _same(R["arc_rc_impulse"].reshape(-1, 4)[rc], g["arc_rc_impulse"][rc])
        drives = D["pin_net"] >= 0
        drives[D["inputs"]] = False
        _same(R["pin_load"].reshape(-1, 4)[drives], g["pin_load"][drives])
        # propagation: bit for bit against the oracle on the timer's levels
        g["pin_load"] = R["pin_load"].reshape(-1, 4)
        fr, ends = R["frontiers"], R["ends"]
        assert ends[-1] == n
        os_, oa, od = O.fprop(g, lib, fr, ends)
        _same(R["slew"].reshape(-1, 4), os_)
        _same(at, oa)
        _same(R["arc_delay"].reshape(-1, 8), od.reshape(-1, 8))
        orat, oslack = O.bprop(g["fanin_start"], g["fanin"], od, D["po_rat"], oa, fr, ends)
        _same(R["rat"].reshape(-1, 4), orat)
        _same(R["slack"].reshape(-1, 4), oslack)
        if design == "c17":
            for got, want in ((orat, D["gold_rat"]), (oslack, D["gold_slack"])):
                e2 = np.abs(got.astype(np.float64) - want)
                assert (e2 <= 0.0005 + 1e-5 + 1e-6 * np.abs(want)).all()
        # a second update on the kept topology: the same values
        ctx.check(N.lib.otg_timer_update(timer, None))
        torch.cuda.synchronize()
        R2 = _results(timer, n, m)
        for k in ("slew", "at", "rat", "slack", "arc_delay", "arc_rc_delay", "arc_rc_impulse"):
            _same(R2[k], R[k])
        # reports: the values per pin, NaN out of range
        pins = np.array([0, n - 1, n, -1], np.int32)
        dp = torch.from_numpy(pins).cuda()
        out = torch.empty(16, dtype=torch.float32, device="cuda:0")
        for what, key in ((N.REPORT_SLEW, "slew"), (N.REPORT_AT, "at"), (N.REPORT_RAT, "rat"),
                          (N.REPORT_SLACK, "slack")):
            ctx.check(N.lib.otg_timer_report(timer, what, 4, dp.data_ptr(), out.data_ptr(), None))
            o = out.cpu().numpy().reshape(4, 4)
            v = R[key].reshape(-1, 4)
            _same(o[:2], v[[0, n - 1]])
            assert np.isnan(o[2:]).all()
    finally:
        N.lib.otg_timer_destroy(timer)
        D["forest"].close()


def test_timer_update_after_the_caps_change(ctx):
    # s27_spef: update, scale every node capacitance in place, update again;
    # a new timer built from the scaled caps gives the same values
    D = _assemble("s27_spef")
    n, m = len(D["pins"]), len(D["arcs"])
    dd = DeviceDesign(D)
    fresh = DeviceDesign(D, cap_scale=1.25)
    t1, t2 = N.timer_p(), N.timer_p()
    ctx.check(N.lib.otg_timer_create(ctx.ptr, C.byref(dd.design), C.byref(dd.lib),
                                     N.BATCHED_COMPENSATED, C.byref(t1)))
    ctx.check(N.lib.otg_timer_create(ctx.ptr, C.byref(fresh.design), C.byref(fresh.lib),
                                     N.BATCHED_COMPENSATED, C.byref(t2)))
    try:
        ctx.check(N.lib.otg_timer_update(t1, None))
        torch.cuda.synchronize()
        before = _results(t1, n, m)
        dd.t["node_cap"].mul_(1.25)
        ctx.check(N.lib.otg_timer_update(t1, None))
        ctx.check(N.lib.otg_timer_update(t2, None))
        torch.cuda.synchronize()
        a, b = _results(t1, n, m), _results(t2, n, m)
        for k in ("slew", "at", "rat", "slack", "arc_delay", "arc_rc_delay", "arc_rc_impulse"):
            _same(a[k], b[k])
        assert not np.array_equal(np.nan_to_num(a["at"]), np.nan_to_num(before["at"]))
    finally:
        N.lib.otg_timer_destroy(t1)
        N.lib.otg_timer_destroy(t2)
        D["forest"].close()


class _SynDesign:
    """Device copies of a workloads.TimerDesignData and the C structs over
    them (the arrays may be replaced before .build())."""

    def __init__(self, D, lib):
        self.D, self.lib = D, lib
        self.h = {"nodes_start": D.forest.nodes_start, "edges": D.forest.edges,
                  "roots": D.forest.roots, "edge_res": D.forest.edge_res,
                  "node_cap": D.forest.node_cap, "fanin_start": D.fanin_start, "fanin": D.fanin,
                  "arc_kind": D.arc_kind, "arc_timing": D.arc_timing, "arc_net": D.arc_net,
                  "arc_node": D.arc_node, "pin_net": D.pin_net, "pin_load": D.pin_load,
                  "pi_slew": D.pi_slew, "pi_at": D.pi_at, "po_rat": D.po_rat}
        self.num_nets = D.forest.num_nets

    def build(self):
        dev = torch.device("cuda:0")
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.T = {k: t(v) for k, v in self.h.items()}
        lk = ("lut_hdr", "lut_val", "timing_sense", "timing_edge", "timing_lut", "timing_has")
        self.LB = {k: t(getattr(self.lib, k)) for k in lk}
        a = lambda d, k: d[k].data_ptr()  # noqa: E731
        T = self.T
        self.n, self.m = len(self.h["pin_net"]), len(self.h["fanin"])
        self.design = N.TimerDesign(
            self.num_nets, int(self.h["nodes_start"][-1]) if self.num_nets else 0,
            a(T, "nodes_start"), a(T, "edges"), a(T, "roots"), a(T, "edge_res"),
            a(T, "node_cap"), self.n, self.m, a(T, "fanin_start"), a(T, "fanin"),
            a(T, "arc_kind"), a(T, "arc_timing"), a(T, "arc_net"), a(T, "arc_node"),
            a(T, "pin_net"), a(T, "pin_load"), a(T, "pi_slew"), a(T, "pi_at"), a(T, "po_rat"))
        self.plib = N.PropLib(len(self.lib.lut_hdr) // 4, len(self.lib.timing_sense),
                              len(self.lib.lut_val), *[a(self.LB, k) for k in lk])
        return self

    def oracle_check(self, R):
        """The timer's slews, arrivals, arc delays, required times and
        slacks against the oracle's propagation over the timer's own levels
        and gathered net values, bit for bit."""
        h, lib = self.h, self.lib
        lk = ("lut_hdr", "lut_val", "timing_sense", "timing_edge", "timing_lut", "timing_has")
        g = {"fanin_start": h["fanin_start"], "fanin": h["fanin"], "arc_kind": h["arc_kind"],
             "arc_timing": h["arc_timing"], "arc_rc_delay": R["arc_rc_delay"].reshape(-1, 4),
             "arc_rc_impulse": R["arc_rc_impulse"].reshape(-1, 4),
             "pin_load": R["pin_load"].reshape(-1, 4), "pi_slew": h["pi_slew"], "pi_at": h["pi_at"]}
        fr, ends = R["frontiers"], R["ends"]
        assert ends[-1] == self.n
        os_, oa, od = O.fprop(g, {k: getattr(lib, k) for k in lk}, fr, ends)
        _same(R["slew"].reshape(-1, 4), os_)
        _same(R["at"].reshape(-1, 4), oa)
        _same(R["arc_delay"].reshape(-1, 8), od.reshape(-1, 8))
        orat, oslack = O.bprop(h["fanin_start"], h["fanin"], od, h["po_rat"], oa, fr, ends)
        _same(R["rat"].reshape(-1, 4), orat)
        _same(R["slack"].reshape(-1, 4), oslack)


def _last_error(ctx):
    return (N.lib.otg_last_error(ctx.ptr) or b"").decode(errors="replace")


def _run_timer(ctx, S, updates=1):
    timer = N.timer_p()
    ctx.check(N.lib.otg_timer_create(ctx.ptr, C.byref(S.design), C.byref(S.plib),
                                     N.BATCHED_COMPENSATED, C.byref(timer)))
    try:
        for _ in range(updates):
            ctx.check(N.lib.otg_timer_update(timer, None))
        torch.cuda.synchronize()
        return _results(timer, S.n, S.m)
    finally:
        N.lib.otg_timer_destroy(timer)


def test_timer_net_arcs_out_of_range_and_designs_without_nets(ctx):
    """Net arcs naming a net or node outside the forest get NaN net values
    (otg_timer.h: out-of-range ids give NaN), which then propagate as the
    oracle propagates them; a design whose pins drive no RC net (no forest at
    all) takes every cell load from pin_load; an empty design updates."""
    from opentimer_amd import workloads as W
    lib = W.prop_library()
    D = W.timer_design(W.DagConfig("timer-oor", 6_000, 15_000, 12, seed=15), lib, seed=16)
    S = _SynDesign(D, lib)
    rc = np.flatnonzero(D.arc_net >= 0)
    rng = np.random.default_rng(17)
    pick = rng.choice(rc, 40, replace=False)
    arc_net, arc_node = D.arc_net.copy(), D.arc_node.copy()
    sizes = np.diff(D.forest.nodes_start)
    arc_net[pick[:10]] = D.forest.num_nets + 3            # net past the forest
    arc_net[pick[10:20]] = -7                             # negative net on an RC arc
    arc_node[pick[20:30]] = sizes[arc_net[pick[20:30]]]   # node one past its net
    arc_node[pick[30:40]] = -1                            # negative node
    S.h["arc_net"], S.h["arc_node"] = arc_net, arc_node
    R = _run_timer(ctx, S.build())
    d = R["arc_rc_delay"].reshape(-1, 4)
    assert np.isnan(d[pick]).all()
    good = np.setdiff1d(rc, pick)
    assert not np.isnan(d[good]).any()
    S.oracle_check(R)
    # no RC nets: every arc a cell arc, no pin drives a net
    S2 = _SynDesign(D, lib)
    S2.num_nets = 0
    S2.h["arc_kind"] = np.where(D.arc_kind == 0, 2, D.arc_kind).astype(np.int32)
    S2.h["arc_net"] = np.full_like(D.arc_net, -1)
    S2.h["arc_node"] = np.full_like(D.arc_node, -1)
    S2.h["pin_net"] = np.full_like(D.pin_net, -1)
    S2.h["nodes_start"] = np.zeros(1, np.int32)
    R2 = _run_timer(ctx, S2.build(), updates=2)
    _same(R2["pin_load"].reshape(-1, 4), D.pin_load)
    S2.oracle_check(R2)
    # an empty design
    S3 = _SynDesign(D, lib)
    S3.num_nets = 0
    for k in ("fanin", "arc_kind", "arc_timing", "arc_net", "arc_node"):
        S3.h[k] = np.zeros(1, np.int32)
    S3.h["nodes_start"] = np.zeros(1, np.int32)
    S3.build()
    S3.design.num_pins = S3.design.num_arcs = 0
    timer = N.timer_p()
    ctx.check(N.lib.otg_timer_create(ctx.ptr, C.byref(S3.design), C.byref(S3.plib),
                                     N.BATCHED_COMPENSATED, C.byref(timer)))
    ctx.check(N.lib.otg_timer_update(timer, None))
    N.lib.otg_timer_destroy(timer)


def test_timer_rejects_malformed_designs(ctx):
    """Argument and design errors come back as status codes with a message,
    before any kernel reads out of range: NULL arguments, negative sizes, a
    fanin CSR that decreases or does not end at num_arcs, a fanin arc naming
    a pin out of range (OTG_ERR_INVALID_ARGUMENT), an RC net that is not a
    tree (OTG_ERR_NOT_A_TREE)."""
    INVALID, NOT_A_TREE = 1, 2
    from opentimer_amd import workloads as W
    lib = W.prop_library()
    D = W.timer_design(W.DagConfig("timer-bad", 4_000, 10_000, 10, seed=25), lib, seed=26)

    def status(S, update=True):
        timer = N.timer_p()
        st = N.lib.otg_timer_create(ctx.ptr, C.byref(S.design), C.byref(S.plib),
                                    N.BATCHED_COMPENSATED, C.byref(timer))
        if st != 0 or not update:
            return st, _last_error(ctx) if st else ""
        try:
            st = N.lib.otg_timer_update(timer, None)
            torch.cuda.synchronize()
            return st, _last_error(ctx) if st else ""
        finally:
            N.lib.otg_timer_destroy(timer)

    S = _SynDesign(D, lib).build()
    timer = N.timer_p()
    assert N.lib.otg_timer_create(ctx.ptr, None, C.byref(S.plib), N.BATCHED_COMPENSATED,
                                  C.byref(timer)) == INVALID
    S.design.num_pins = -1
    assert status(S)[0] == INVALID
    # fanin CSR decreasing, and not ending at num_arcs
    for fix in ("decrease", "end"):
        S = _SynDesign(D, lib)
        fs = D.fanin_start.copy()
        if fix == "decrease":
            fs[100] = fs[102] + 1
        else:
            fs[-1] = fs[-1] - 1
        S.h["fanin_start"] = fs
        st, msg = status(S.build())
        assert st == INVALID and "fanin_start" in msg, (fix, st, msg)
    # a fanin arc from a pin out of range
    S = _SynDesign(D, lib)
    fi = D.fanin.copy()
    fi[5] = len(D.pin_net) + 9
    S.h["fanin"] = fi
    st, msg = status(S.build())
    assert st == INVALID and "out of range" in msg, msg
    # an RC net that is not a tree (a duplicated edge)
    S = _SynDesign(D, lib)
    f = D.forest
    k = int(np.flatnonzero(np.diff(f.nodes_start) >= 4)[0])
    e0 = int(f.nodes_start[k] - k)
    bad = f.edges.copy()
    bad[2 * e0], bad[2 * e0 + 1] = bad[2 * e0 + 2], bad[2 * e0 + 3]
    S.h["edges"] = bad
    st, msg = status(S.build())
    assert st == NOT_A_TREE and f"net {k}" in msg, (st, msg)


def test_timer_synthetic_design(ctx):
    """A generated design (workloads.timer_design: a layered graph whose
    driving pins own RC trees of the cfg1 size distribution): the RC delays
    and impulses the timer gathers at every net-arc sink equal an independent
    solve of the same forest queried at those pins (otg_rct_query), and the
    timer's slews, arrivals, arc delays, required times and slacks equal the
    oracle's propagation over the timer's own levels and gathered values, bit
    for bit."""
    from opentimer_amd import workloads as W
    lib = W.prop_library()
    D = W.timer_design(W.DagConfig("timer-syn", 30_000, 75_000, 30, seed=5), lib, seed=6)
    f = D.forest
    n, m = len(D.pin_net), len(D.fanin)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    T = {k: t(v) for k, v in (("nodes_start", f.nodes_start), ("edges", f.edges), ("roots", f.roots),
                              ("edge_res", f.edge_res), ("node_cap", f.node_cap),
                              ("fanin_start", D.fanin_start), ("fanin", D.fanin),
                              ("arc_kind", D.arc_kind), ("arc_timing", D.arc_timing),
                              ("arc_net", D.arc_net), ("arc_node", D.arc_node),
                              ("pin_net", D.pin_net), ("pin_load", D.pin_load),
                              ("pi_slew", D.pi_slew), ("pi_at", D.pi_at), ("po_rat", D.po_rat))}
    lk = ("lut_hdr", "lut_val", "timing_sense", "timing_edge", "timing_lut", "timing_has")
    LB = {k: t(getattr(lib, k)) for k in lk}
    a = lambda d, k: d[k].data_ptr()  # noqa: E731
    design = N.TimerDesign(f.num_nets, f.num_nodes, a(T, "nodes_start"), a(T, "edges"),
                           a(T, "roots"), a(T, "edge_res"), a(T, "node_cap"), n, m,
                           a(T, "fanin_start"), a(T, "fanin"), a(T, "arc_kind"),
                           a(T, "arc_timing"), a(T, "arc_net"), a(T, "arc_node"),
                           a(T, "pin_net"), a(T, "pin_load"), a(T, "pi_slew"), a(T, "pi_at"),
                           a(T, "po_rat"))
    plib = N.PropLib(len(lib.lut_hdr) // 4, len(lib.timing_sense), len(lib.lut_val),
                     *[a(LB, k) for k in ("lut_hdr", "lut_val", "timing_sense", "timing_edge",
                                          "timing_lut", "timing_has")])
    timer = N.timer_p()
    ctx.check(N.lib.otg_timer_create(ctx.ptr, C.byref(design), C.byref(plib),
                                     N.BATCHED_COMPENSATED, C.byref(timer)))
    try:
        ctx.check(N.lib.otg_timer_update(timer, None))
        torch.cuda.synchronize()
        R = _results(timer, n, m)
        # a second update re-orders R and C through the kept BFS order: the
        # same bits (the first update takes them from the ordering itself)
        ctx.check(N.lib.otg_timer_update(timer, None))
        torch.cuda.synchronize()
        R2 = _results(timer, n, m)
        for k in ("slew", "at", "rat", "slack", "arc_delay", "arc_rc_delay", "arc_rc_impulse",
                  "pin_load"):
            _same(R2[k], R[k])
    finally:
        N.lib.otg_timer_destroy(timer)
    rc = D.arc_net >= 0
    assert rc.any() and (~rc).any()
    s = RcSolver.from_edges(f, ctx)
    s.solve("compensated")
    q = s.query(t(D.arc_net[rc]), t(D.arc_node[rc]))
    _same(R["arc_rc_delay"].reshape(-1, 4)[rc], q["delay"].cpu().numpy())
    # a driving pin carries its net's load (every pin of a net reports it)
    drv = np.flatnonzero(D.pin_net >= 0)
    ql = s.query(t(D.pin_net[drv]), t(np.zeros(len(drv), np.int32)))["load"]
    _same(R["pin_load"].reshape(-1, 4)[drv], ql.cpu().numpy())
    s.close()
    g = {"fanin_start": D.fanin_start, "fanin": D.fanin, "arc_kind": D.arc_kind,
         "arc_timing": D.arc_timing, "arc_rc_delay": R["arc_rc_delay"].reshape(-1, 4),
         "arc_rc_impulse": R["arc_rc_impulse"].reshape(-1, 4),
         "pin_load": R["pin_load"].reshape(-1, 4), "pi_slew": D.pi_slew, "pi_at": D.pi_at}
    fr, ends = R["frontiers"], R["ends"]
    assert ends[-1] == n
    os_, oa, od = O.fprop(g, {k: getattr(lib, k) for k in lk}, fr, ends)
    _same(R["slew"].reshape(-1, 4), os_)
    _same(R["at"].reshape(-1, 4), oa)
    _same(R["arc_delay"].reshape(-1, 8), od.reshape(-1, 8))
    orat, oslack = O.bprop(D.fanin_start, D.fanin, od, D.po_rat, oa, fr, ends)
    _same(R["rat"].reshape(-1, 4), orat)
    _same(R["slack"].reshape(-1, 4), oslack)
