"""Pins the CPU oracle (oracle/otg_oracle.c) before anything is checked against it.

Each oracle function is checked against something it does not restate:
closed forms, an O(n^2) brute force of the Elmore moments written from their
definition (shared-path resistance), invariants that a dropped term, a wrong
sign or a transposed operand would break, the reference's own counting-sort
test vector (ot/cuda/test.cu), and a longest-path definition of the levels.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


# ----------------------------------------------------------------------------
# helpers: small forests in BFS-ordered flat form, built directly
# ----------------------------------------------------------------------------
def bfs_forest_from_parents(parents_list, rng, r_range=(0.01, 1.0), c_range=(0.1, 10.0)):
    """parents_list: per net, parent index in generation order (parent < child).

    Returns flat BFS-ordered arrays via the oracle's own BFS (tested separately)."""
    ns = [0]
    edges, roots, res, caps = [], [], [], []
    for par in parents_list:
        n = len(par)
        for k in range(1, n):
            edges += [par[k], k]
            res.append(rng.uniform(*r_range))
        roots.append(0)
        for _ in range(n):
            caps += list(rng.uniform(*c_range, size=4))
        ns.append(ns[-1] + n)
    ns = np.array(ns, np.int32)
    e, n2b, pid = O.rct_bfs(ns, np.array(edges, np.int32), np.array(roots, np.int32))
    pres, cap = O.rct_prepare(ns, e, n2b, np.array(res, np.float32), np.array(caps, np.float32))
    return ns, pid, pres, cap


def brute_force(ns, pid, pres, cap):
    """Elmore moments from their definition, per net, in float64.

    delay_i = sum_k C_k Rc(i,k), ldelay_i = sum_{d in sub(i)} C_d delay_d,
    beta_i = sum_k C_k delay_k Rc(i,k), with Rc the resistance shared by the
    driver->i and driver->k paths."""
    V = int(ns[-1])
    out = {k: np.zeros(4 * V) for k in ["load", "delay", "ldelay", "beta", "impulse"]}
    for a in range(len(ns) - 1):
        st, ed = int(ns[a]), int(ns[a + 1])
        n = ed - st
        par = [-1] + [i - int(pid[st + i]) for i in range(1, n)]
        paths, anc = [], []
        for i in range(n):
            p, s = i, set()
            while p > 0:
                s.add(p)
                p = par[p]
            paths.append(s)          # nodes whose incoming edge is on driver->i
            anc.append(s | {0})      # ancestors of i, i included
        R = [float(pres[st + i]) for i in range(n)]
        for c in range(4):
            C = [float(cap[(st + i) * 4 + c]) for i in range(n)]
            Rc = [[sum(R[e] for e in paths[i] & paths[k]) for k in range(n)] for i in range(n)]
            d = [sum(C[k] * Rc[i][k] for k in range(n)) for i in range(n)]
            sub = [[k for k in range(n) if i in anc[k]] for i in range(n)]
            for i in range(n):
                out["load"][(st + i) * 4 + c] = sum(C[k] for k in sub[i])
                out["delay"][(st + i) * 4 + c] = d[i]
                out["ldelay"][(st + i) * 4 + c] = sum(C[k] * d[k] for k in sub[i])
                b = sum(C[k] * d[k] * Rc[i][k] for k in range(n))
                out["beta"][(st + i) * 4 + c] = b
                out["impulse"][(st + i) * 4 + c] = 2 * b - d[i] ** 2
    return out


def rel(a, b, floor=1e-30):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)) if a.size else 0.0


ALL = ("flat_f32", "dfs_f32", "f64")


# ----------------------------------------------------------------------------
# closed forms
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ALL)
def test_two_node_closed_form(kind):
    R, C0, C1 = 0.5, 2.0, 3.0
    ns = np.array([0, 2], np.int32)
    pid = np.array([0, 1], np.int32)
    pres = np.array([0.0, R], np.float32)
    cap = np.array([C0] * 4 + [C1] * 4, np.float32)
    si = np.array([4.0, -4.0, 4.0, 4.0], np.float32)
    r = O.rct_all(ns, pid, pres, cap, si, kind)
    assert np.allclose(r["load"][:4], C0 + C1) and np.allclose(r["load"][4:], C1)
    assert np.allclose(r["delay"][:4], 0) and np.allclose(r["delay"][4:], R * C1)
    assert np.allclose(r["ldelay"][4:], R * C1 * C1)
    assert np.allclose(r["ldelay"][:4], C1 * R * C1)
    assert np.allclose(r["beta"][4:], R * R * C1 * C1)
    assert np.allclose(r["impulse"][4:], R * R * C1 * C1)
    assert np.allclose(r["impulse"][:4], 0)
    s = np.sqrt(16 + R * R * C1 * C1)
    assert np.allclose(r["slew"][4:], [s, -s, s, s])  # sign follows the driver slew
    assert np.allclose(r["slew"][:4], [4, -4, 4, 4])  # impulse 0 -> slew = si


@pytest.mark.parametrize("kind", ALL)
def test_chain_closed_form(kind):
    rng = np.random.default_rng(5)
    n = 40
    R = rng.uniform(0.01, 1.0, n)
    R[0] = 0
    Cv = rng.uniform(0.1, 5.0, n)
    ns = np.array([0, n], np.int32)
    pid = np.array([0] + [1] * (n - 1), np.int32)
    cap = np.repeat(Cv, 4).astype(np.float32)
    r = O.rct_all(ns, pid, R.astype(np.float32), cap, np.zeros(4, np.float32), kind)
    Rf, Cf = R.astype(np.float32).astype(np.float64), Cv.astype(np.float32).astype(np.float64)
    suffix = np.cumsum(Cf[::-1])[::-1]
    delay = np.cumsum(Rf * suffix)  # delay_k = sum_{i<=k} R_i sum_{j>=i} C_j
    tol = 1e-5 if kind != "f64" else 1e-12
    assert rel(r["load"][0::4], suffix) < tol
    assert rel(r["delay"][0::4][1:], delay[1:]) < tol


@pytest.mark.parametrize("kind", ALL)
def test_star_branches_independent(kind):
    # driver with k leaves: each leaf sees only its own branch
    k = 7
    rng = np.random.default_rng(9)
    R = np.concatenate([[0], rng.uniform(0.1, 1, k)]).astype(np.float32)
    Cv = rng.uniform(0.5, 2, k + 1).astype(np.float32)
    ns = np.array([0, k + 1], np.int32)
    pid = np.array([0] + list(range(1, k + 1)), np.int32)  # parent of i is 0
    r = O.rct_all(ns, pid, R, np.repeat(Cv, 4), np.zeros(4, np.float32), kind)
    d = r["delay"][4::4].astype(np.float64)
    assert rel(d, R[1:].astype(np.float64) * Cv[1:]) < 1e-6
    assert rel(r["impulse"][4::4], (R[1:].astype(np.float64) * Cv[1:]) ** 2) < 1e-5


# ----------------------------------------------------------------------------
# brute force on random small trees
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_brute_force_random_trees(seed):
    rng = np.random.default_rng(seed)
    parents = []
    for _ in range(20):
        n = int(rng.integers(1, 12))
        parents.append([-1] + [int(rng.integers(0, k)) for k in range(1, n)])
    ns, pid, pres, cap = bfs_forest_from_parents(parents, rng)
    bf = brute_force(ns, pid, pres, cap)
    r64 = O.rct_dfs_f64(ns, pid, pres, cap)
    for k in bf:
        floor = 1e-9 * np.max(np.abs(bf[k])) if k == "impulse" else 1e-30
        assert rel(r64[k], bf[k], floor) < 1e-10, k
    for kind in ("flat_f32", "dfs_f32"):
        r = O.rct_all(ns, pid, pres, cap, np.zeros(4 * (len(ns) - 1), np.float32), kind)
        for k in ("load", "delay", "ldelay", "beta"):
            assert rel(r[k], bf[k]) < 2e-6, (kind, k)
        scale = np.max(np.abs(bf["beta"]))
        assert np.max(np.abs(r["impulse"] - bf["impulse"])) < 1e-5 * scale


def test_brute_force_detects_planted_mistake():
    # a transposed operand (beta from load instead of ldelay) must be caught
    rng = np.random.default_rng(3)
    parents = [[-1, 0, 1, 1, 0, 4]]
    ns, pid, pres, cap = bfs_forest_from_parents(parents, rng)
    bf = brute_force(ns, pid, pres, cap)
    r = O.rct_dfs_f64(ns, pid, pres, cap)
    wrong = np.zeros_like(r["beta"])
    for i in range(1, 6):
        for c in range(4):
            wrong[i * 4 + c] = wrong[(i - pid[i]) * 4 + c] + pres[i] * r["load"][i * 4 + c]
    assert rel(wrong, bf["beta"]) > 1e-3


# ----------------------------------------------------------------------------
# invariants
# ----------------------------------------------------------------------------
def _random_forest(seed, nets=30, nmax=40):
    rng = np.random.default_rng(seed)
    parents = []
    for _ in range(nets):
        n = int(rng.integers(1, nmax))
        parents.append([-1] + [int(rng.integers(0, k)) for k in range(1, n)])
    return bfs_forest_from_parents(parents, rng)


def test_invariants_f64():
    ns, pid, pres, cap = _random_forest(11)
    r = O.rct_dfs_f64(ns, pid, pres, cap)
    for a in range(len(ns) - 1):
        st, ed = ns[a], ns[a + 1]
        for c in range(4):
            assert np.isclose(r["load"][st * 4 + c], cap[st * 4 + c: ed * 4: 4].astype(np.float64).sum())
            assert r["delay"][st * 4 + c] == 0
            for i in range(st + 1, ed):
                p = i - pid[i]
                assert r["delay"][i * 4 + c] >= r["delay"][p * 4 + c]
    assert np.all(r["impulse"] >= -1e-12 * np.max(r["beta"]))
    # scaling C by s: load, delay x s; beta, impulse x s^2
    s = 3.0
    r2 = O.rct_dfs_f64(ns, pid, pres, (cap * s).astype(np.float32))
    for k, e in [("load", 1), ("delay", 1), ("ldelay", 2), ("beta", 2)]:
        assert rel(r2[k], r[k] * s ** e) < 1e-6
    # scaling R by s: delay x s, beta x s^2, load unchanged
    r3 = O.rct_dfs_f64(ns, pid, (pres * s).astype(np.float32), cap)
    assert rel(r3["load"], r["load"]) < 1e-12
    assert rel(r3["delay"], r["delay"] * s) < 1e-6
    assert rel(r3["beta"], r["beta"] * s * s) < 1e-6
    # corners are independent
    cap2 = cap.copy()
    cap2[1::4] *= 2.0
    r4 = O.rct_dfs_f64(ns, pid, pres, cap2)
    for c in (0, 2, 3):
        assert rel(r4["delay"][c::4], r["delay"][c::4]) < 1e-15
    assert rel(r4["delay"][1::4], 2 * r["delay"][1::4]) < 1e-6


def test_flat_and_dfs_agree():
    ns, pid, pres, cap = _random_forest(12, nets=50, nmax=200)
    a = O.rct_flat_f32(ns, pid, pres, cap)
    b = O.rct_dfs_f32(ns, pid, pres, cap)
    c = O.rct_dfs_f64(ns, pid, pres, cap)
    for k in ("load", "delay", "ldelay", "beta"):
        assert rel(a[k], c[k]) < 1e-5 and rel(b[k], c[k]) < 1e-5


# ----------------------------------------------------------------------------
# BFS ordering
# ----------------------------------------------------------------------------
def test_counting_sort_reference_vector():
    g = json.load(open(os.path.join(HERE, "golden", "counting_sort_n32_seed123.json")))
    x = np.array(g["inputs"], np.int32)
    orders, outputs, _ = O.counting_argsort(x, g["amin"], g["amax"])
    assert np.all(x == outputs[orders])  # test.cu:73-78
    assert list(outputs) == g["sorted"]
    assert sorted(orders) == list(range(g["n"]))
    # equal values: the later input takes the earlier position (--counts[v])
    for v in set(g["inputs"]):
        idx = np.nonzero(x == v)[0]
        assert list(orders[idx]) == sorted(orders[idx], reverse=True)


def bfs_rule(n, edges, root):
    adj = [[] for _ in range(n)]
    for s, t in edges:
        adj[s].append(t)
        adj[t].append(s)
    dist = [-1] * n
    dist[root] = 0
    q = [root]
    for u in q:
        for v in adj[u]:
            if dist[v] < 0:
                dist[v] = dist[u] + 1
                q.append(v)
    order = sorted(range(n), key=lambda u: (dist[u], -u))
    pos = [0] * n
    for p, u in enumerate(order):
        pos[u] = p
    return dist, pos


@pytest.mark.parametrize("seed", [0, 1])
def test_bfs_order_rule(seed):
    rng = np.random.default_rng(seed)
    ns = [0]
    edges, roots, parts = [], [], []
    for _ in range(40):
        n = int(rng.integers(1, 30))
        perm = rng.permutation(n)
        e = []
        for k in range(1, n):
            p = int(rng.integers(0, k))
            s, t = int(perm[p]), int(perm[k])
            if rng.random() < 0.5:
                s, t = t, s
            e.append((s, t))
        root = int(perm[0])
        parts.append((n, e, root))
        edges += [x for st in e for x in st]
        roots.append(root)
        ns.append(ns[-1] + n)
    ns = np.array(ns, np.int32)
    ed, n2b, pid = O.rct_bfs(ns, np.array(edges, np.int32), np.array(roots, np.int32))
    eoff = 0
    for a, (n, e, root) in enumerate(parts):
        dist, pos = bfs_rule(n, e, root)
        assert list(n2b[ns[a]: ns[a + 1]]) == pos
        for j, (s, t) in enumerate(e):
            ps, pt = ed[2 * (eoff + j)], ed[2 * (eoff + j) + 1]
            assert {ps, pt} == {s, t} and dist[ps] + 1 == dist[pt]  # oriented parent->child
            assert pid[ns[a] + pos[pt]] == pos[pt] - pos[ps]
        eoff += n - 1


def test_bfs_rejects_non_tree():
    ns = np.array([0, 4], np.int32)
    # edges 0-1, 1-0, 2-3: duplicate edge, node set not connected to root 0
    with pytest.raises(ValueError):
        O.rct_bfs(ns, np.array([0, 1, 1, 0, 2, 3], np.int32), np.array([0], np.int32))


# ----------------------------------------------------------------------------
# levelization
# ----------------------------------------------------------------------------
def random_dag(n, m, seed):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    keep = a != b
    a, b = a[keep], b[keep]
    fr = np.minimum(a, b)
    to = np.maximum(a, b)
    return perm[fr].astype(np.int32), perm[to].astype(np.int32)


def csr(n, src, dst):
    order = np.argsort(dst, kind="stable")
    start = np.zeros(n + 1, np.int32)
    np.add.at(start, dst + 1, 1)
    start = np.cumsum(start).astype(np.int32)
    return start, src[order].astype(np.int32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_levelize_matches_longest_path(seed):
    n = 300
    fr, to = random_dag(n, 900, seed)
    fanin_start, fanin = csr(n, fr, to)
    fanout_start, fanout = csr(n, to, fr)
    out = np.diff(fanout_start).astype(np.int32)
    f0 = np.nonzero(out == 0)[0].astype(np.int32)
    frontiers, ends, lv = O.levelize(n, fanin_start, fanin, out, f0)
    ref = O.levels_longest_path(n, fanout_start, fanout)
    assert np.array_equal(lv, ref)
    assert ends[0] == 0 and ends[-1] == n
    for l in range(len(ends) - 1):
        assert np.all(lv[frontiers[ends[l]: ends[l + 1]]] == l)
    # every arc u->v has level(u) >= level(v) + 1
    assert np.all(lv[fr] >= lv[to] + 1)


def test_levelize_chain_and_empty():
    n = 5
    fr = np.array([0, 1, 2, 3], np.int32)
    to = np.array([1, 2, 3, 4], np.int32)
    s, f = csr(n, fr, to)
    out = np.array([1, 1, 1, 1, 0], np.int32)
    frontiers, ends, lv = O.levelize(n, s, f, out, np.array([4], np.int32))
    assert list(lv) == [4, 3, 2, 1, 0] and list(ends) == [0, 1, 2, 3, 4, 5]
    frontiers, ends, lv = O.levelize(0, np.zeros(1, np.int32), np.zeros(0, np.int32),
                                     np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert list(ends) == [0, 0]


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_parallel_levelize_matches_serial(threads):
    # the threaded CPU baseline gives the serial oracle's levels and frontier
    # ends (pin order inside a frontier may differ), on a DAG whose frontiers
    # exceed one release block
    from opentimer_amd import workloads as W
    dag = W.generate_dag(W.DagConfig("par", 60_000, 150_000, 30, seed=3), nthreads=2)
    c = W.fanin_csr(dag)
    fr, ends, lv = O.levelize(c.n, c.start, c.fanin, c.out_degree, c.frontier)
    pfr, pends, plv = O.levelize_par(c.n, c.start, c.fanin, c.out_degree, c.frontier, threads)
    assert np.array_equal(plv, lv)
    assert np.array_equal(pends, ends)
    for l in range(len(ends) - 1):
        assert np.array_equal(np.sort(pfr[ends[l]: ends[l + 1]]), np.sort(fr[ends[l]: ends[l + 1]]))
