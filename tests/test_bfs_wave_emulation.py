"""CPU emulation of the large-net BFS wavefront (bfs_net16 in
opentimer_amd/csrc/otg_bfs.hip): the kernel's phases and loop bounds
executed for 64 lanes at a time on 16-bit LDS arrays, instruction by
instruction (a store from several lanes to one entry keeps the last lane's
value; atomics run in lane order).

What it shows before the kernel runs on a GPU: every loop ends within its
stated bound on malformed nets (duplicated edges, cycles, unreachable nodes,
out-of-range endpoints, self loops, a bad root), each of them is rejected,
and on valid trees the emulated positions equal the oracle's BFS order.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle as O  # noqa: E402

W = 64
UNSEEN = 0xFFFF


class Steps:
    """Counts loop iterations against the kernel's bounds."""

    def __init__(self, n):
        self.n = n
        self.total = 0

    def tick(self, k=1):
        self.total += k
        # every phase is O(n) chunks, the BFS O(n) levels x O(n) neighbour
        # steps at worst: a generous polynomial bound that a hang would exceed
        assert self.total < 64 * (self.n + 64) ** 2


def bfs_net16(n, edges, root):
    """Returns (node2bfs, pid-by-position) or None when the net is rejected."""
    st = Steps(n)
    ne = n - 1
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    off = np.zeros(n + 2, np.int64)
    adj = np.zeros(max(2 * ne, 1), np.int64)
    dist = np.full(n + 1, UNSEEN, np.int64)
    q = np.zeros(n + 1, np.int64)
    bad = not (0 <= root < n)
    # 1. degrees
    for k0 in range(0, ne, W):
        st.tick()
        for k in range(k0, min(ne, k0 + W)):
            s, t = e[k]
            if s < 0 or s >= n or t < 0 or t >= n or s == t:
                bad = True
            else:
                off[s + 1] += 1
                off[t + 1] += 1
    if bad:
        return None
    # 2. offsets
    off[: n + 1] = np.cumsum(off[: n + 1])
    assert off[n] <= 0xFFFF
    # 3. adjacency
    cur = np.zeros(n, np.int64)
    for k0 in range(0, ne, W):
        st.tick()
        for k in range(k0, min(ne, k0 + W)):
            s, t = e[k]
            adj[off[s] + cur[s]] = t
            cur[s] += 1
            adj[off[t] + cur[t]] = s
            cur[t] += 1
    # 4. BFS
    dist[root] = 0
    q[0] = root
    fb, fe, levels, over = 0, 1, 0, False
    while fb < fe and levels < n:
        st.tick()
        tail = fe
        for b in range(fb, fe, W):
            lanes = np.arange(b, b + W)
            u = np.where(lanes < fe, q[np.minimum(lanes, n)], -1)
            jb = np.where(u >= 0, off[np.maximum(u, 0)], 0)
            je = np.where(u >= 0, off[np.maximum(u, 0) + 1], 0)
            for it in range(2 * ne):
                st.tick()
                j = jb + it
                has = j < je
                if not has.any():
                    break
                v = np.where(has, adj[np.where(has, j, 0)], 0)
                take = has & (dist[v] == UNSEEN)  # all lanes read, then write
                idx = np.flatnonzero(take)
                for r, lane in enumerate(idx):
                    dist[v[lane]] = levels + 1
                    if tail + r < n:
                        q[tail + r] = v[lane]
                tail += len(idx)
        if tail > n:
            over = True
            break
        fb, fe, levels = fe, tail, levels + 1
    bad = over or fe != n or bool((dist[:n] == UNSEEN).any())
    if bad:
        return None
    # 5. level starts and positions
    cnt = np.bincount(dist[:n], minlength=levels)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    pos = np.zeros(n, np.int64)
    for top in range(n - 1, -1, -W):
        st.tick()
        u = top - np.arange(W)
        valid = u >= 0
        d = np.where(valid, dist[np.maximum(u, 0)], -1)
        remaining = valid.copy()
        for it in range(W):
            if not remaining.any():
                break
            leader = int(np.flatnonzero(remaining)[0])
            key = d[leader]
            m = valid & (d == key)
            base = start[key]
            start[key] += int(m.sum())
            ranks = np.cumsum(m) - 1
            pos[u[m]] = base + ranks[m]
            remaining &= ~m
        assert not remaining.any()
    # 6. pid by position
    pid = np.zeros(n, np.int64)
    for s, t in e:
        ps, pt = pos[s], pos[t]
        if ps > pt:
            ps, pt = pt, ps
        pid[pt] = pt - ps
    return pos, pid


def random_tree(rng, n, chain):
    par = np.zeros(n, np.int64)
    for i in range(1, n):
        par[i] = i - 1 if (chain and rng.random() >= 0.1) else rng.integers(0, i)
    perm = rng.permutation(n)
    edges = []
    for i in range(1, n):
        a, b = perm[par[i]], perm[i]
        edges.append((a, b) if rng.random() < 0.5 else (b, a))
    order = rng.permutation(n - 1)
    return np.array(edges, np.int64)[order].reshape(-1), int(perm[0])


@pytest.mark.parametrize("n,chain", [(65, True), (300, False), (1024, True), (2500, True)])
def test_emulated_order_matches_oracle(n, chain):
    rng = np.random.default_rng(n)
    edges, root = random_tree(rng, n, chain)
    got = bfs_net16(n, edges, root)
    assert got is not None
    pos, pid = got
    ns = np.array([0, n], np.int32)
    _, n2b, opid = O.rct_bfs(ns, edges.astype(np.int32), np.array([root], np.int32))
    assert np.array_equal(pos, n2b) and np.array_equal(pid, opid)


def _malformed(rng, n, kind):
    edges, root = random_tree(rng, n, True)
    e = edges.reshape(-1, 2).copy()
    if kind == "duplicate":      # the rejection test's mistake: edge 0 := edge 1
        e[0] = e[1]
    elif kind == "cycle":
        # cut edge 5 (a, b); join a to another node of a's side: a cycle there,
        # b's side unreachable
        a, b = e[5]
        rest = np.delete(e, 5, axis=0)
        side, frontier = {int(a)}, [int(a)]
        while frontier:
            x = frontier.pop()
            for s_, t_ in rest:
                for p_, r_ in ((s_, t_), (t_, s_)):
                    if p_ == x and int(r_) not in side:
                        side.add(int(r_))
                        frontier.append(int(r_))
        c = next(x for x in sorted(side) if x != a)
        e[5] = (a, c)
    elif kind == "self_loop":
        e[2] = (e[2][0], e[2][0])
    elif kind == "out_of_range":
        e[4] = (e[4][0], n + 3)
    elif kind == "bad_root":
        root = n
    elif kind == "star_of_duplicates":  # one node reached from many frontier nodes at once
        hub = e[0][0]
        for k in range(1, min(len(e), 60)):
            e[k] = (e[k][0], hub) if e[k][0] != hub else (e[k][1], hub)
    return e.reshape(-1), root


@pytest.mark.parametrize("kind", ["duplicate", "cycle", "self_loop", "out_of_range", "bad_root",
                                  "star_of_duplicates"])
@pytest.mark.parametrize("n", [100, 2000])
def test_emulated_loops_end_on_malformed_nets(kind, n):
    rng = np.random.default_rng(7)
    edges, root = _malformed(rng, n, kind)
    assert bfs_net16(n, edges, root) is None
