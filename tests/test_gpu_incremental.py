"""GPU parity of incremental re-timing: subset solve + fan-out-cone
re-levelization, eager and replayed from a hipGraph.

Bars: the cone is exactly the forward closure of the changed nets' driver
pins (CPU BFS); the levels of cone pins and the level ends equal the
oracle's sink-first levelization of the candidate subgraph (the cone), as
Timer::_build_prop_tasks_cuda builds it (timer.cpp:1011-1071); outputs of the
changed nets match the FP64 oracle on their current inputs; graph replays
equal eager updates.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

import oracle as O
from opentimer_amd import workloads as W
from opentimer_amd.incremental import Incremental
from opentimer_amd.levelize import Levelizer
from opentimer_amd.rct import RcSolver


def forward_closure(fo_start, fo, seeds, n):
    mark = np.zeros(n, bool)
    fr = np.unique(seeds)
    mark[fr] = True
    while len(fr):
        nxt = np.concatenate([fo[fo_start[u]: fo_start[u + 1]] for u in fr]) if len(fr) else fr
        nxt = np.unique(nxt)
        nxt = nxt[~mark[nxt]]
        mark[nxt] = True
        fr = nxt
    return np.flatnonzero(mark)


@pytest.fixture(scope="module")
def session(ctx):
    dag = W.generate_dag(W.DagConfig("incr-test", 30_000, 75_000, 30, seed=21))
    fin = W.fanin_csr(dag)
    fout = W.fanin_csr(W.Dag(dag.n, dag.arc_to, dag.arc_from))
    L = Levelizer(fin, ctx)
    L.run()
    _, ends, lv = L.host()
    forest = W.generate_rc(W.config1(num_nets=3000, seed=22))
    s = RcSolver.from_edges(forest, ctx)
    # every output defined before the subset updates: the output arrays are
    # uninitialised device memory (possibly NaN left by earlier tests), and
    # the comparisons below are bitwise over whole arrays
    s.solve("compensated")
    torch.cuda.synchronize()
    rng = np.random.default_rng(3)
    drivers = rng.choice(np.flatnonzero(fin.out_degree > 0), forest.num_nets).astype(np.int32)
    inc = Incremental(s, fin, fout, drivers, lv, len(ends) - 1)
    return dag, fin, fout, s, drivers, inc, len(ends) - 1


def check_update(session, ids):
    dag, fin, fout, s, drivers, inc, nlev = session
    cone, fr, ends = inc.query()
    ref_cone = forward_closure(fout.start, fout.fanin, drivers[ids], dag.n)
    assert np.array_equal(np.sort(cone), ref_cone)
    cand = np.zeros(dag.n, np.uint8)
    cand[ref_cone] = 1
    csr = W.fanin_csr(dag, cand=cand)
    _, oends, olv = O.levelize(csr.n, csr.start, csr.fanin, csr.out_degree, csr.frontier)
    assert np.array_equal(ends, oends)
    lv = inc.level.cpu().numpy()
    assert np.array_equal(lv[ref_cone], olv[ref_cone])
    # the pins of each level, as sets
    for k in range(len(ends) - 1):
        got = np.sort(fr[ends[k]: ends[k + 1]])
        assert np.array_equal(got, np.flatnonzero((olv == k) & (cand == 1))), k


def perturb(s, ids, rng):
    ns = s.nodes_start.cpu().numpy()
    idx = np.concatenate([np.arange(ns[a], ns[a + 1]) for a in ids])
    f = torch.from_numpy(rng.uniform(0.8, 1.2, len(idx)).astype(np.float32)).to(s.device)
    it = torch.from_numpy(idx).to(s.device)
    s.pres[it] *= f
    cv = s.cap.view(-1, 4)
    cv[it] *= f[:, None]


def check_outputs(s, ids):
    h = s.host_topology()
    ns = h["nodes_start"]
    g = s.host()
    for a in ids[:20]:
        st, ed = int(ns[a]), int(ns[a + 1])
        sub = np.array([0, ed - st], np.int32)
        r = O.rct_all(sub, h["pid"][st:ed], h["pres"][st:ed], h["cap"][4 * st: 4 * ed],
                      h["slew_in"][4 * a: 4 * a + 4], "f64")
        for k in ("delay", "slew", "load"):
            ref = r[k]
            fl = max(float(np.abs(ref).max()) * 1e-7, 1e-30)
            err = np.max(np.abs(g[k][4 * st: 4 * ed] - ref) / np.maximum(np.abs(ref), fl))
            assert err < 1e-6, (a, k, err)


@pytest.mark.parametrize("batch", [1, 37, 500])
def test_incremental_eager(session, batch):
    dag, fin, fout, s, drivers, inc, nlev = session
    rng = np.random.default_rng(batch)
    ids = np.sort(rng.choice(s.num_nets, batch, replace=False)).astype(np.int32)
    perturb(s, ids, rng)
    inc.step(torch.from_numpy(ids).to(s.device), "compensated")
    torch.cuda.synchronize()
    check_update(session, ids)
    check_outputs(s, ids)


def test_incremental_graph_replay(session):
    dag, fin, fout, s, drivers, inc, nlev = session
    side = torch.cuda.Stream()
    buf = torch.empty(64, dtype=torch.int32, device=s.device)
    g = inc.capture(buf, "compensated", side)
    rng = np.random.default_rng(7)
    for rep in range(3):
        ids = np.sort(rng.choice(s.num_nets, 64, replace=False)).astype(np.int32)
        buf.copy_(torch.from_numpy(ids))
        perturb(s, ids, rng)
        torch.cuda.synchronize()
        g.launch(side)
        side.synchronize()
        check_update(session, ids)
        check_outputs(s, ids)
        lv_graph = inc.level.cpu().numpy().copy()
        out_graph = s.host()["delay"].copy()
        # the same update run eagerly gives the same levels and outputs
        inc.step(buf, "compensated")
        torch.cuda.synchronize()
        assert np.array_equal(inc.level.cpu().numpy(), lv_graph)
        assert np.array_equal(s.host()["delay"], out_graph)
    g.close()
