"""CPU tests of the host side: the C-ABI library loads and exports every
function the headers declare, the status/error surface, and the seeded
workload generators and fanin-CSR builder (no GPU compute calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"\b(otg_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    from opentimer_amd import native as N
    declared = _declared()
    assert len(declared) >= 20
    assert declared == set(N.EXPORTED), sorted(declared ^ set(N.EXPORTED))
    for name in declared:
        assert hasattr(N.lib, name), name


def test_status_strings_and_version():
    from opentimer_amd import native as N
    seen = set()
    for s in range(9):
        txt = N.lib.otg_status_string(s)
        assert txt, s
        seen.add(txt)
    assert len(seen) == 9  # one distinct message per status
    assert N.lib.otg_version() > 0


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from opentimer_amd import native as N
    p = N.ctx_p()
    st = N.lib.otg_create(0, C.byref(p))
    assert st != N.OK
    assert N.STATUS[st] in ("OTG_ERR_NO_DEVICE", "OTG_ERR_DEVICE")


def _forest_parent(f):
    """Parent of each node (net-local ids) from the undirected edges + roots."""
    ns = f.nodes_start
    par = np.full(f.num_nodes, -2, np.int64)
    for net in range(f.num_nets):
        a, b = int(ns[net]), int(ns[net + 1])
        n = b - a
        e = f.edges[2 * (a - net): 2 * (b - net - 1)].reshape(-1, 2)
        adj = [[] for _ in range(n)]
        for s, t in e:
            adj[s].append(t)
            adj[t].append(s)
        r = int(f.roots[net])
        par[a + r] = -1
        stack = [r]
        seen = {r}
        while stack:
            u = stack.pop()
            for v in adj[u]:
                if v not in seen:
                    seen.add(v)
                    par[a + v] = u
                    stack.append(v)
        assert len(seen) == n, f"net {net} is not connected"
    return par


def test_rc_generator_is_seeded_and_well_formed():
    from opentimer_amd import workloads as W
    cfg = W.config0(num_nets=300)
    f1 = W.generate_rc(cfg, nthreads=1)
    f2 = W.generate_rc(cfg, nthreads=4)
    for k in ("nodes_start", "edges", "roots", "edge_res", "node_cap", "slew_in"):
        assert np.array_equal(getattr(f1, k), getattr(f2, k)), k
    f3 = W.generate_rc(cfg.scaled(300, seed=99), nthreads=1)
    assert not np.array_equal(f1.edge_res, f3.edge_res)
    sizes = np.diff(f1.nodes_start)
    assert sizes.min() >= 2 and sizes.max() <= 16
    assert np.all((f1.roots >= 0) & (f1.roots < sizes))
    # every net is a tree spanning its nodes
    par = _forest_parent(f1)
    assert np.all(par >= -1)
    # R in kohm from 0.1-100 ohm, ground C 0.1-10 fF, driver slew 5-50 ps
    assert f1.edge_res.min() >= 0.1e-3 * 0.999 and f1.edge_res.max() <= 100e-3 * 1.001
    assert f1.slew_in.min() >= 5.0 and f1.slew_in.max() <= 50.0
    cap = f1.node_cap.reshape(-1, 4)
    # corners are fixed multiples of one base on sinks: MIN/MAX x RISE/FALL
    assert np.all(cap > 0)


def test_rc_generator_batches_match_the_whole_population():
    # nets first_net.. of a population are drawn from their global indices, so
    # batches concatenate to the population generated at once
    from opentimer_amd import workloads as W
    cfg = W.config1(num_nets=900, seed=5)
    whole = W.generate_rc(cfg, nthreads=2)
    parts = []
    for a, b in ((0, 250), (250, 600), (600, 900)):
        c = cfg.scaled(b - a)
        c.first_net = a
        parts.append(W.generate_rc(c, nthreads=2))
    sizes = np.concatenate([np.diff(f.nodes_start) for f in parts])
    assert np.array_equal(sizes, np.diff(whole.nodes_start))
    for k in ("edges", "roots", "edge_res", "node_cap", "slew_in"):
        assert np.array_equal(np.concatenate([getattr(f, k) for f in parts]), getattr(whole, k)), k


def test_rc_generator_heavy_tail_classes():
    from opentimer_amd import workloads as W
    f = W.generate_rc(W.config1(num_nets=20000), nthreads=2)
    sizes = np.diff(f.nodes_start)
    assert sizes.min() >= 2 and sizes.max() <= 8192
    frac_small = np.mean(sizes <= 8)
    frac_big = np.mean(sizes >= 65)
    assert 0.77 < frac_small < 0.83
    assert 0.005 < frac_big < 0.016


def test_dag_generator_and_fanin_csr():
    from opentimer_amd import workloads as W
    cfg = W.config0_dag()
    d1 = W.generate_dag(cfg, nthreads=1)
    d2 = W.generate_dag(cfg, nthreads=3)
    assert np.array_equal(d1.arc_from, d2.arc_from) and np.array_equal(d1.arc_to, d2.arc_to)
    assert len(d1.arc_from) == cfg.num_arcs
    csr = W.fanin_csr(d1)
    n = cfg.num_pins
    # CSR of fanins: for pin v, the sources u of arcs u->v
    ref = [[] for _ in range(n)]
    for u, v in zip(d1.arc_from, d1.arc_to):
        ref[v].append(u)
    assert csr.start[0] == 0 and csr.start[-1] == len(csr.fanin)
    for v in range(0, n, 97):
        got = sorted(csr.fanin[csr.start[v]: csr.start[v + 1]].tolist())
        assert got == sorted(ref[v]), v
    outdeg = np.bincount(d1.arc_from, minlength=n)
    assert np.array_equal(csr.out_degree, outdeg)
    assert np.array_equal(np.sort(csr.frontier), np.flatnonzero(outdeg == 0))


def test_fanin_csr_candidates_and_loop_breakers():
    """Arcs from non-candidates and loop-breaker arcs are dropped, and out
    degrees count only the kept arcs (timer.cpp:1011-1047)."""
    from opentimer_amd import workloads as W
    d = W.generate_dag(W.config0_dag(seed=11), nthreads=1)
    n, m = d.n, len(d.arc_from)
    rng = np.random.default_rng(5)
    cand = (rng.random(n) < 0.8).astype(np.uint8)
    skip = (rng.random(m) < 0.1).astype(np.uint8)
    csr = W.fanin_csr(d, cand, skip)
    keep = (cand[d.arc_from] == 1) & (cand[d.arc_to] == 1) & (skip == 0)
    assert len(csr.fanin) == int(keep.sum())
    outdeg = np.bincount(d.arc_from[keep], minlength=n)
    assert np.array_equal(csr.out_degree[cand == 1], outdeg[cand == 1])
    front = np.flatnonzero((cand == 1) & (outdeg == 0))
    assert np.array_equal(np.sort(csr.frontier), front)


def test_fanout_csr_matches_a_stable_sort_of_the_fanin_csr():
    from opentimer_amd import native as N
    from opentimer_amd import workloads as W
    import oracle as O
    csr = W.fanin_csr(W.generate_dag(W.config0_dag(seed=31)))
    n, m = csr.n, len(csr.fanin)
    fo_start = np.zeros(n + 1, np.int32)
    fo_arc = np.zeros(m, np.int32)
    arc_to = np.zeros(m, np.int32)
    N.check(N.lib.otg_build_fanout_csr(n, N.addr(csr.start), N.addr(csr.fanin), N.addr(fo_start),
                                       N.addr(fo_arc), N.addr(arc_to)))
    s, a, t = O.fanout_csr(csr.start, csr.fanin)
    assert np.array_equal(fo_start, s) and np.array_equal(fo_arc, a) and np.array_equal(arc_to, t)
    # every arc once, under its source, in arc order
    assert np.array_equal(np.sort(fo_arc), np.arange(m))
    for u in (0, n // 2, n - 1):
        arcs = fo_arc[fo_start[u]:fo_start[u + 1]]
        assert np.all(np.diff(arcs) > 0) and np.all(csr.fanin[arcs] == u)
    # a source outside the pins is refused
    bad = csr.fanin.copy()
    bad[0] = n
    assert N.lib.otg_build_fanout_csr(n, N.addr(csr.start), N.addr(bad), N.addr(fo_start),
                                      N.addr(fo_arc), N.addr(arc_to)) == 1
