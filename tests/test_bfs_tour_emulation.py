"""CPU emulation of the workgroup Euler-tour BFS ordering (tour_net in
opentimer_amd/csrc/otg_bfs.hip) for nets of 1025..8192 nodes.

The emulation follows the kernel's phases with its loop bounds: degrees and
endpoint checks, CSR offsets, adjacency slots and twins (filled in a random
edge order, as the LDS atomics may interleave them), tour successors, the
ruling-set list ranking (walks held to A = 2(n-1) steps, bit_width(R) + 1
rounds of pointer jumping), descents by tour position, the +-1 scan, and the
positions from the sorted (distance, CAP-1-id) keys.

What it shows before the kernel runs on a GPU: on malformed nets (duplicated
edges, cycles with an unreachable part, self loops, out-of-range endpoints,
a bad root, a root without edges) every loop stays within its bound and the
net is rejected; on valid trees the positions and parent offsets equal the
oracle's BFS order whatever the slot order.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle as O  # noqa: E402

from test_bfs_wave_emulation import _malformed, random_tree  # noqa: E402

END = 0xFFFF
UNSEEN = 0xFFFF


def tour_net(n, edges, root, T, CAP, rng):
    """Returns (node2bfs, pid-by-position) or None when the net is rejected."""
    assert 1 < n <= CAP
    ne, A = n - 1, 2 * (n - 1)
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    if not (0 <= root < n):
        return None
    # 1. degrees and endpoint checks
    off = np.zeros(n + 1, np.int64)
    s, t = e[:, 0], e[:, 1]
    if np.any((s < 0) | (s >= n) | (t < 0) | (t >= n) | (s == t)):
        return None
    np.add.at(off, s + 1, 1)
    np.add.at(off, t + 1, 1)
    # 2. offsets (inclusive scan of the shifted degrees)
    off = np.cumsum(off)
    assert off[n] == A
    # 3. slots and twins, edges in an arbitrary interleaving
    cur = np.zeros(n, np.int64)
    adj = np.zeros(A, np.int64)
    tw = np.zeros(A, np.int64)
    for k in rng.permutation(ne):
        a, b = e[k]
        js = off[a] + cur[a]
        cur[a] += 1
        jt = off[b] + cur[b]
        cur[b] += 1
        adj[js], adj[jt] = b, a
        tw[js], tw[jt] = jt, js
    # 4. successors; a root without edges is rejected
    h = off[root]
    if off[root + 1] == h:
        return None
    nxt = np.where(tw + 1 < off[adj + 1], tw + 1, off[adj])
    nx = np.where(nxt == h, END, nxt)
    # 5. ruling set
    sh = 0
    while (T << sh) < A:
        sh += 1
    Sm = (1 << sh) - 1
    Rb = (A + Sm) >> sh
    hr = (h & Sm) != 0
    R = Rb + int(hr)
    assert R <= T + 1
    H = Rb if hr else h >> sh

    def is_ruler(k):
        return (k & Sm) == 0 or k == h

    def rid(k):
        return k >> sh if (k & Sm) == 0 else Rb

    def rstart(r):
        return r << sh if r < Rb else h

    steps = 0
    rsuf = np.zeros(R, np.int64)
    rnext = np.zeros(R, np.int64)
    for r in range(R):
        k, cnt, it = nx[rstart(r)], 1, 0
        while it < A and k != END and not is_ruler(k):
            cnt += 1
            k = nx[k]
            it += 1
        steps += it
        rsuf[r] = cnt
        rnext[r] = -1 if k == END else (rid(k) if is_ruler(k) else -2)
    assert steps <= R * A
    rounds = int(R).bit_length() + 1
    for _ in range(rounds):  # synchronous: every ruler reads, then every ruler writes
        a = rnext.copy()
        ok = a >= 0
        ns = rsuf + np.where(ok, rsuf[np.where(ok, a, 0)], 0)
        nn = np.where(ok, rnext[np.where(ok, a, 0)], a)
        rsuf, rnext = ns, nn
    if rsuf[H] != A or rnext[H] != -1:
        return None
    rk = np.full(A, -1, np.int64)
    for r in range(R):
        j, p = rstart(r), A - rsuf[r]
        rk[j] = p
        k, it = nx[j], 0
        while it < A and k != END and not is_ruler(k):
            p += 1
            rk[k] = p
            k = nx[k]
            it += 1
    assert np.all(rk >= 0) and np.array_equal(np.sort(rk), np.arange(A))
    # 6. descents and distances
    down = rk < rk[tw]
    w = np.zeros(A, np.int64)
    w[rk] = np.where(down, 1, -1)
    depth = np.cumsum(w)
    dist = np.full(n, UNSEEN, np.int64)
    dist[adj[down]] = depth[rk[down]]
    dist[root] = 0
    if np.any(dist == UNSEEN):
        return None
    # 7. positions from the sorted keys (distance, CAP-1-id)
    log = CAP.bit_length() - 1
    keys = (dist << log) | (CAP - 1 - np.arange(n))
    order = np.argsort(keys, kind="stable")
    pos = np.empty(n, np.int64)
    pos[CAP - 1 - (keys[order] & (CAP - 1))] = np.arange(n)
    # 8. parent offsets by position
    pid = np.zeros(n, np.int64)
    ps, pt = pos[s], pos[t]
    lo, hi = np.minimum(ps, pt), np.maximum(ps, pt)
    pid[hi] = hi - lo
    return pos, pid


def star_tree(rng, n):
    perm = rng.permutation(n)
    e = np.array([(perm[0], perm[i]) for i in range(1, n)], np.int64)
    return e[rng.permutation(n - 1)].reshape(-1), int(perm[0])


@pytest.mark.parametrize("n,shape,cap,T", [
    (1025, "chain", 2048, 256), (2048, "random", 2048, 256), (3000, "chain", 4096, 512),
    (4096, "star", 4096, 512), (5000, "random", 8192, 1024), (8192, "chain", 8192, 1024),
])
def test_emulated_tour_order_matches_oracle(n, shape, cap, T):
    rng = np.random.default_rng(n)
    if shape == "star":
        edges, root = star_tree(rng, n)
    else:
        edges, root = random_tree(rng, n, shape == "chain")
    ns = np.array([0, n], np.int32)
    _, n2b, opid = O.rct_bfs(ns, edges.astype(np.int32), np.array([root], np.int32))
    for trial in range(2):  # two slot orders
        got = tour_net(n, edges, root, T, cap, np.random.default_rng(trial))
        assert got is not None
        pos, pid = got
        assert np.array_equal(pos, n2b) and np.array_equal(pid, opid)


@pytest.mark.parametrize("kind", ["duplicate", "cycle", "self_loop", "out_of_range", "bad_root",
                                  "star_of_duplicates", "isolated_root"])
@pytest.mark.parametrize("n,cap,T", [(1500, 2048, 256), (6000, 8192, 1024)])
def test_emulated_tour_rejects_malformed_nets(kind, n, cap, T):
    rng = np.random.default_rng(11)
    if kind == "isolated_root":
        # the root keeps no edge: its edge goes between two other nodes (a
        # cycle there)
        edges, root = random_tree(rng, n, True)
        e = edges.reshape(-1, 2).copy()
        for k in range(len(e)):
            if root in e[k]:
                others = [x for x in range(n) if x != root and x not in e[k]]
                e[k] = (others[0], others[1]) if e[k][0] == root else (others[1], others[0])
        edges = e.reshape(-1)
    else:
        edges, root = _malformed(rng, n, kind)
    assert tour_net(n, edges, root, T, cap, np.random.default_rng(0)) is None
