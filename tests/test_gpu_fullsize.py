"""Parity at the full sizes of the benchmark configurations.

* cfg1 (4,000,000 nets, heavy-tailed sizes up to 8192 nodes): every node of
  the whole forest is checked on the device by properties the mathematics
  fixes -- the driver's load equals the net's total capacitance, delay never
  decreases from a node's parent to the node, impulse >= 0, |slew| >= |slew
  in| -- and a sample of 3,000 random nets plus the 20 largest is compared
  with the oracle: plain bit-exact vs the reference kernel arithmetic,
  batched within 1e-5 of the CPU timer arithmetic (delay, slew) and of FP64
  (every output), compensated within 1e-6 of FP64.
* cfg2 (10M pins, 25M arcs, 200 levels): levels, level ends and the pins of
  every level bit-exact vs the oracle's sink-first Kahn levelization.
* cfg3 (2M-net precision population): two batches of 16,384 nets drawn by
  global index, one of them holding the population's largest chain; every
  output of every node of the compensated method within 1e-6 of the device
  FP64 reference, which is itself pinned against the host FP64 oracle.
* cfg4 (5M-pin graph + 1M resident nets): one update each of 1, 100, 10,000
  and 100,000 changed nets; the cone equals the CPU forward closure, its
  levels and level ends equal the oracle's levelization of the candidate
  subgraph, and changed nets' outputs match FP64.
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
ULP = 2.0 ** -23


def _floors(ref):
    f = {}
    for k in OUT:
        f[k] = max(float(np.max(np.abs(ref[k]))) * 1e-7, 1e-30)
    f["impulse"] = max(float(np.max(np.abs(ref["beta"]))) * 1e-6, 1e-30)
    return f


def _rel(got, ref, floor):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor)))


def _sample(topo, nets):
    ns = topo["nodes_start"]
    sizes = np.diff(ns)
    idx = np.concatenate([np.arange(4 * ns[a], 4 * ns[a + 1]) for a in nets])
    nodes = np.concatenate([np.arange(ns[a], ns[a + 1]) for a in nets])
    n4 = np.concatenate([np.arange(4 * a, 4 * a + 4) for a in nets])
    sub = np.concatenate([[0], np.cumsum(sizes[nets])]).astype(np.int32)
    return idx, (sub, topo["pid"][nodes], topo["pres"][nodes], topo["cap"][idx], topo["slew_in"][n4])


# ---------------------------------------------------------------------------
# cfg1
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cfg1(ctx):
    f = W.generate_rc(W.config1(), nthreads=16)
    assert f.num_nets == 4_000_000
    s = RcSolver.from_edges(f, ctx)
    del f
    topo = s.host_topology()
    ns = topo["nodes_start"]
    sizes = np.diff(ns)
    rng = np.random.default_rng(0)
    pick = np.unique(np.concatenate([rng.choice(len(sizes), 3000, replace=False),
                                     np.argsort(sizes, kind="stable")[-20:]]))
    assert sizes[pick].max() == sizes.max()
    idx, args = _sample(topo, pick)
    yield s, idx, args
    s.close()


def _properties(s, tol_load):
    """Whole-forest checks on the device."""
    V, Nn = s.num_nodes, s.num_nets
    ns = s.nodes_start.long()
    cap = s.cap[: 4 * V].view(V, 4).double()
    cs = torch.cumsum(cap, 0)
    last = cs[ns[1:] - 1]
    first = torch.where((ns[:-1] > 0)[:, None], cs[(ns[:-1] - 1).clamp_(min=0)], 0.0)
    total = last - first
    del cs, cap
    load = s.out["load"][: 4 * V].view(V, 4)
    root_load = load[ns[:-1]].double()
    err = float(((root_load - total).abs() / total.abs().clamp(min=1e-30)).max())
    assert err < tol_load, ("root load vs total cap", err)
    # node-wise checks in slices of 2^24 nodes (gathers over whole-forest
    # index tensors of ~10^9 elements gave wrong results in torch)
    d = s.out["delay"][: 4 * V].view(V, 4)
    sl_all = s.out["slew"][: 4 * V].view(V, 4)
    imp_all = s.out["impulse"][: 4 * V].view(V, 4)
    si_all = s.slew_in.view(Nn, 4)
    nsc = ns.contiguous()
    step = 1 << 24
    bad = neg = low = 0
    for a in range(0, V, step):
        b = min(V, a + step)
        pos = torch.arange(a, b, device=s.device)
        pid = s.pid[a:b].long()
        nonroot = pid > 0
        dc = d[a:b][nonroot]
        dp = d.index_select(0, (pos - pid)[nonroot])
        bad += int((dc < dp - dp.abs() * ULP).sum())
        neg += int((imp_all[a:b] < 0).sum())
        owner = torch.searchsorted(nsc, pos, right=True) - 1
        si = si_all.index_select(0, owner).abs()
        low += int((sl_all[a:b].abs() < si - si * 2 * ULP).sum())
    assert bad == 0, ("delay decreases from a parent to its child", bad)
    assert neg == 0, ("negative impulse", neg)
    assert low == 0, ("|slew| below |slew in|", low)


@pytest.mark.parametrize("method", ["plain", "batched", "compensated"])
def test_cfg1_full_forest(cfg1, method):
    s, idx, args = cfg1
    s.solve(method)
    torch.cuda.synchronize()
    _properties(s, {"plain": 1e-5, "batched": 1e-5, "compensated": 1e-6}[method])
    it = torch.from_numpy(idx).to(s.device)
    g = {k: s.out[k][it].cpu().numpy() for k in OUT}
    if method == "plain":
        ref = O.rct_all(*args, "flat_f32")
        for k in ("load", "delay", "ldelay", "beta", "impulse"):
            assert np.array_equal(g[k].view(np.uint32), ref[k].view(np.uint32)), k
        return
    r64 = O.rct_all(*args, "f64")
    fl = _floors(r64)
    tol = 1e-5 if method == "batched" else 1e-6
    for k in OUT:
        assert _rel(g[k], r64[k], fl[k]) < tol, (method, k)
    if method == "batched":
        r32 = O.rct_all(*args, "dfs_f32")
        for k in ("delay", "slew"):
            assert _rel(g[k], r32[k], fl[k]) < 1e-5, (method, k, "vs cpu fp32")


# ---------------------------------------------------------------------------
# cfg2
# ---------------------------------------------------------------------------
def test_cfg2_levelize_full(ctx):
    from opentimer_amd.levelize import Levelizer
    dag = W.generate_dag(W.config2(), nthreads=16)
    csr = W.fanin_csr(dag)
    del dag
    assert csr.n == 10_000_000 and len(csr.fanin) == 25_000_000
    L = Levelizer(csr, ctx)
    L.run()
    fr, ends, lv = L.host()
    _, oends, olv = O.levelize(csr.n, csr.start, csr.fanin, csr.out_degree, csr.frontier)
    assert len(ends) - 1 == 200
    assert np.array_equal(ends, oends)
    assert np.array_equal(lv, olv)
    # the pins of level k are frontiers[ends[k]:ends[k+1]], each pin once
    lvl_of_pos = np.repeat(np.arange(len(ends) - 1, dtype=np.int32), np.diff(ends))
    assert np.array_equal(olv[fr], lvl_of_pos)
    assert np.array_equal(np.sort(fr), np.arange(csr.n, dtype=np.int32))
    # the host-pointer drop-in on the same graph (its 40-100 MB arrays go
    # through the library's staged pinned copies)
    from opentimer_amd.levelize import toposort_compute
    hfr, hends, hlv = toposort_compute(ctx, csr)
    assert np.array_equal(hends, oends)
    assert np.array_equal(hlv, olv)
    assert np.array_equal(olv[hfr], lvl_of_pos)
    assert np.array_equal(np.sort(hfr), np.arange(csr.n, dtype=np.int32))


# ---------------------------------------------------------------------------
# cfg3
# ---------------------------------------------------------------------------
def test_cfg3_precision_batches(ctx):
    import ctypes as C
    population = W.config3()
    # every net's size is drawn from its global index: the population's sizes
    # in slices whose node counts fit int32
    sizes = np.empty(population.num_nets, np.int64)
    step = 65536
    for a in range(0, population.num_nets, step):
        part = population.scaled(min(step, population.num_nets - a))
        part.first_net = a
        p = W._params(part)
        ns_part = np.zeros(part.num_nets + 1, np.int32)
        tot = C.c_int64()
        N.check(N.lib.otg_gen_rc_sizes(C.byref(p), N.addr(ns_part), C.byref(tot)))
        sizes[a: a + part.num_nets] = np.diff(ns_part)
    assert int(sizes.sum()) > 10_000_000_000
    biggest = int(np.argmax(sizes))
    batch = 16384
    firsts = sorted({0, biggest - biggest % batch})
    for first in firsts:
        cfg = population.scaled(min(batch, population.num_nets - first))
        cfg.first_net = first
        f = W.generate_rc(cfg, nthreads=16)
        assert np.array_equal(np.diff(f.nodes_start), sizes[first: first + cfg.num_nets])
        s = RcSolver.from_edges(f, ctx)
        del f
        ref = s.reference_f64()
        s.solve("compensated")
        torch.cuda.synchronize()
        floor = {k: max(float(ref[k].abs().max()) * 1e-7, 1e-300) for k in OUT}
        floor["impulse"] = max(float(ref["beta"].abs().max()) * 1e-6, 1e-300)
        for k in OUT:
            e = float(((s.out[k][: ref[k].numel()].double() - ref[k]).abs()
                       / ref[k].abs().clamp(min=floor[k])).max())
            assert e < 1e-6, (first, k, e)
        if first == firsts[-1]:
            # the device reference vs the host FP64 oracle on the largest chain
            topo = s.host_topology()
            local = biggest - first
            idx, args = _sample(topo, np.array([local]))
            r64 = O.rct_all(*args, "f64")
            it = torch.from_numpy(idx).to(s.device)
            fl = _floors(r64)
            for k in OUT:
                tol = 1e-9 if k in ("impulse", "slew") else 1e-12
                assert _rel(ref[k][it].cpu().numpy(), r64[k], fl[k]) < tol, k
        del ref
        s.close()
        del s
        torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# cfg4
# ---------------------------------------------------------------------------
def _closure(fo_start, fo, seeds, n):
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


def test_cfg4_incremental_full(ctx):
    from opentimer_amd.incremental import Incremental
    from opentimer_amd.levelize import Levelizer
    pins, nets = 5_000_000, 1_000_000
    dag = W.generate_dag(W.DagConfig("cfg4-dag-5M", pins, int(2.5 * pins), pins // 50_000, seed=9),
                         nthreads=16)
    fin = W.fanin_csr(dag)
    fout = W.fanin_csr(W.Dag(dag.n, dag.arc_to, dag.arc_from))
    L = Levelizer(fin, ctx)
    L.run()
    _, ends, lv = L.host()
    forest = W.generate_rc(W.config1(num_nets=nets), nthreads=16)
    s = RcSolver.from_edges(forest, ctx)
    ns = forest.nodes_start
    del forest
    rng = np.random.default_rng(11)
    drivers = rng.choice(np.flatnonzero(fin.out_degree > 0), nets).astype(np.int32)
    inc = Incremental(s, fin, fout, drivers, lv, len(ends) - 1)
    s.solve("compensated")
    dev = s.device
    for B in (1, 100, 10_000, 100_000):
        ids = np.sort(rng.choice(nets, B, replace=False)).astype(np.int32)
        lens = ns[ids + 1] - ns[ids]
        idx = np.repeat(ns[ids] - (np.cumsum(lens) - lens), lens) + np.arange(int(lens.sum()))
        fct = torch.from_numpy(rng.uniform(0.8, 1.2, len(idx)).astype(np.float32)).to(dev)
        itn = torch.from_numpy(idx.astype(np.int64)).to(dev)
        s.pres[itn] *= fct
        s.cap.view(-1, 4)[itn] *= fct[:, None]
        inc.step(torch.from_numpy(ids).to(dev), "compensated")
        torch.cuda.synchronize()
        cone, fr, kends = inc.query()
        ref = _closure(fout.start, fout.fanin, drivers[ids], dag.n)
        assert np.array_equal(np.sort(cone), ref), B
        cand = np.zeros(dag.n, np.uint8)
        cand[ref] = 1
        csr = W.fanin_csr(dag, cand=cand)
        _, oends, olv = O.levelize(csr.n, csr.start, csr.fanin, csr.out_degree, csr.frontier)
        assert np.array_equal(kends, oends), B
        lvg = inc.level.cpu().numpy()
        assert np.array_equal(lvg[ref], olv[ref]), B
        lvl_of_pos = np.repeat(np.arange(len(kends) - 1, dtype=np.int32), np.diff(kends))
        assert np.array_equal(olv[fr[: kends[-1]]], lvl_of_pos), B
        # changed nets' outputs vs FP64 (a sample, the largest changed net included)
        pick = np.unique(np.concatenate([ids[:16], ids[np.argsort(lens)[-4:]]]))
        st = torch.from_numpy(np.concatenate([np.arange(ns[a], ns[a + 1]) for a in pick])).to(dev)
        sizes = ns[pick + 1] - ns[pick]
        sub = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        i4 = (st[:, None] * 4 + torch.arange(4, device=dev)).reshape(-1)
        n4 = torch.from_numpy(np.concatenate([np.arange(4 * a, 4 * a + 4) for a in pick])).to(dev)
        args = (sub, s.pid[st].cpu().numpy(), s.pres[st].cpu().numpy(), s.cap[i4].cpu().numpy(),
                s.slew_in[n4].cpu().numpy())
        r64 = O.rct_all(*args, "f64")
        fl = _floors(r64)
        for k in OUT:
            assert _rel(s.out[k][i4].cpu().numpy(), r64[k], fl[k]) < 1e-6, (B, k)
    inc.close()
    s.close()
