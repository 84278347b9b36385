"""numpy front-end of the CPU oracle (oracle/otg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and only as the checker.  The product
(opentimer_amd, libotg.so) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return SO


def _load():
    if not os.path.exists(SO):
        build()
    return C.CDLL(SO)


_lib = _load()
vp = C.c_void_p
for name, args, res in [
    ("ora_rct_bfs", [C.c_int, vp, vp, vp, vp, vp], C.c_int),
    ("ora_rct_prepare", [C.c_int, vp, vp, vp, vp, vp, vp, vp], None),
    ("ora_rct_flat_f32", [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp], None),
    ("ora_rct_dfs_f32", [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp], None),
    ("ora_rct_dfs_f64", [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp], None),
    ("ora_slew_f32", [C.c_int, vp, vp, vp, vp], None),
    ("ora_slew_f64", [C.c_int, vp, vp, vp, vp], None),
    ("ora_levelize", [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp], C.c_int),
    ("ora_levelize_par", [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp, C.c_int], C.c_int),
    ("ora_levels_longest_path", [C.c_int, vp, vp, vp], None),
    ("ora_counting_argsort", [C.c_int, vp, C.c_int, C.c_int, vp, vp, vp], None),
    ("ora_lut_eval", [vp, vp, C.c_int, C.c_float, C.c_float], C.c_float),
    ("ora_fprop", [C.c_int] + [vp] * 17 + [C.c_int, vp, vp, vp], None),
    ("ora_fprop_par", [C.c_int] + [vp] * 17 + [C.c_int, vp, vp, vp, C.c_int], None),
    ("ora_bprop", [vp] * 8 + [C.c_int, vp, vp], None),
]:
    f = getattr(_lib, name)
    f.argtypes = args
    f.restype = res


def _p(a):
    return None if a is None else a.ctypes.data


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def rct_bfs(nodes_start, edges, roots):
    """BFS order of every net (rct.cu:120-245 + sort.cuh:63-142).

    Returns (edges_directed, node2bfs, pid)."""
    ns = _i32(nodes_start)
    e = _i32(edges).copy()
    r = _i32(roots)
    V = int(ns[-1])
    n2b = np.zeros(V, np.int32)
    pid = np.zeros(V, np.int32)
    rc = _lib.ora_rct_bfs(len(ns) - 1, _p(ns), _p(e), _p(r), _p(n2b), _p(pid))
    if rc != 0:
        raise ValueError("not a forest of trees")
    return e, n2b, pid


def rct_prepare(nodes_start, edges_directed, node2bfs, edge_res, node_cap):
    """pres / cap in BFS order (rct.cu:314-359)."""
    ns = _i32(nodes_start)
    V = int(ns[-1])
    pres = np.zeros(V, np.float32)
    cap = np.zeros(4 * V, np.float32)
    _lib.ora_rct_prepare(len(ns) - 1, _p(ns), _p(_i32(edges_directed)), _p(_i32(node2bfs)),
                         _p(_f32(edge_res)), _p(_f32(node_cap)), _p(pres), _p(cap))
    return pres, cap


def _moments(fn, dtype, nodes_start, pid, pres, cap):
    ns = _i32(nodes_start)
    V = int(ns[-1])
    outs = [np.zeros(4 * V, dtype) for _ in range(5)]
    fn(len(ns) - 1, _p(ns), _p(_i32(pid)), _p(_f32(pres)), _p(_f32(cap)), *[_p(o) for o in outs])
    return dict(zip(["load", "delay", "ldelay", "beta", "impulse"], outs))


def rct_flat_f32(nodes_start, pid, pres, cap):
    """Reference GPU kernel arithmetic (rct.cu:11-59), FP32."""
    return _moments(_lib.ora_rct_flat_f32, np.float32, nodes_start, pid, pres, cap)


def rct_dfs_f32(nodes_start, pid, pres, cap):
    """CPU ot::Timer arithmetic (net.cpp:101-193), FP32 with double pow."""
    return _moments(_lib.ora_rct_dfs_f32, np.float32, nodes_start, pid, pres, cap)


def rct_dfs_f64(nodes_start, pid, pres, cap):
    """FP64 oracle of the same sweeps."""
    return _moments(_lib.ora_rct_dfs_f64, np.float64, nodes_start, pid, pres, cap)


def slew(nodes_start, slew_in, impulse):
    ns = _i32(nodes_start)
    if impulse.dtype == np.float64:
        out = np.zeros_like(impulse)
        _lib.ora_slew_f64(len(ns) - 1, _p(ns), _p(_f32(slew_in)), _p(impulse), _p(out))
    else:
        out = np.zeros(impulse.shape, np.float32)
        _lib.ora_slew_f32(len(ns) - 1, _p(ns), _p(_f32(slew_in)), _p(_f32(impulse)), _p(out))
    return out


def rct_all(nodes_start, pid, pres, cap, slew_in, kind="dfs_f32"):
    fn = {"dfs_f32": rct_dfs_f32, "flat_f32": rct_flat_f32, "f64": rct_dfs_f64}[kind]
    r = fn(nodes_start, pid, pres, cap)
    r["slew"] = slew(nodes_start, slew_in, r["impulse"])
    return r


def levelize(n, fanin_start, fanin, out_degree, frontier0):
    """Sink-first Kahn (toposort.cu:4-72).  Returns (frontiers, ends, level)."""
    fs = _i32(fanin_start)
    fi = _i32(fanin) if len(fanin) else np.zeros(1, np.int32)
    out = _i32(out_degree).copy()
    fr = np.zeros(max(n, 1), np.int32)
    f0 = _i32(frontier0)
    fr[: len(f0)] = f0
    ends = np.zeros(n + 2, np.int32)
    lv = np.zeros(max(n, 1), np.int32)
    ne = _lib.ora_levelize(n, _p(fs), _p(fi), _p(out), len(f0), _p(fr), _p(ends), _p(lv))
    return fr[:n], ends[:ne], lv[:n]


def levelize_par(n, fanin_start, fanin, out_degree, frontier0, nthreads):
    """The same sink-first Kahn over nthreads threads (the levelizer's CPU
    baseline).  Levels and frontier ends equal levelize's; the pin order
    inside a frontier depends on the interleaving.  Returns (frontiers,
    ends, level)."""
    fs = _i32(fanin_start)
    fi = _i32(fanin) if len(fanin) else np.zeros(1, np.int32)
    out = _i32(out_degree).copy()
    fr = np.zeros(max(n, 1), np.int32)
    f0 = _i32(frontier0)
    fr[: len(f0)] = f0
    ends = np.zeros(n + 2, np.int32)
    lv = np.zeros(max(n, 1), np.int32)
    ne = _lib.ora_levelize_par(n, _p(fs), _p(fi), _p(out), len(f0), _p(fr), _p(ends), _p(lv),
                               int(nthreads))
    return fr[:n], ends[:ne], lv[:n]


def counting_argsort(inputs, amin, amax):
    """block_couting_sort (sort.cuh:63-142).  Returns (orders, outputs, counts)."""
    x = _i32(inputs)
    n = len(x)
    counts = np.zeros(amax - amin + 1, np.int32)
    orders = np.zeros(max(n, 1), np.int32)
    outputs = np.zeros(max(n, 1), np.int32)
    _lib.ora_counting_argsort(n, _p(x), amin, amax, _p(counts), _p(orders), _p(outputs))
    return orders[:n], outputs[:n], counts


def levels_longest_path(n, fanout_start, fanout):
    fs = _i32(fanout_start)
    fo = _i32(fanout) if len(fanout) else np.zeros(1, np.int32)
    lv = np.zeros(max(n, 1), np.int32)
    _lib.ora_levels_longest_path(n, _p(fs), _p(fo), _p(lv))
    return lv[:n]


def lut_eval(hdr, pool, l, val1, val2):
    """Lut::operator() (ot/liberty/lut.cpp:161-241) on table l of a flat pool
    laid out as in include/otg_prop.h."""
    h = _i32(hdr)
    p = _f32(pool)
    return float(_lib.ora_lut_eval(_p(h), _p(p), int(l), float(val1), float(val2)))


def fprop(graph, lib, frontiers, ends, nthreads=1):
    """Level-wise forward propagation (timer.cpp:733-783, arc.cpp:82-147).
    nthreads > 1 splits every level's pins over host threads (the same
    results bit for bit: pins of a level write disjoint values).

    graph: dict with fanin_start, fanin, arc_kind, arc_timing, arc_rc_delay
    [m,4], arc_rc_impulse [m,4], pin_load [n,4], pi_slew [n,4], pi_at [n,4];
    lib: dict with lut_hdr, lut_val, timing_sense, timing_edge, timing_lut,
    timing_has.  Returns (slew [n,4], at [n,4], arc_delay [m,8]); pins outside
    the frontiers keep NaN."""
    fs = _i32(graph["fanin_start"])
    n = len(fs) - 1
    m = int(fs[-1])
    one_i = lambda a: _i32(a) if len(a) else np.zeros(1, np.int32)
    one_f = lambda a: _f32(a).reshape(-1) if np.size(a) else np.zeros(1, np.float32)
    fi, kind, tim = one_i(graph["fanin"]), one_i(graph["arc_kind"]), one_i(graph["arc_timing"])
    rcd, rci = one_f(graph["arc_rc_delay"]), one_f(graph["arc_rc_impulse"])
    load, pis, pia = one_f(graph["pin_load"]), one_f(graph["pi_slew"]), one_f(graph["pi_at"])
    sense, edge = one_i(lib["timing_sense"]), one_i(lib["timing_edge"])
    tlut, thas = one_i(lib["timing_lut"]), one_i(lib["timing_has"])
    hdr, pool = one_i(lib["lut_hdr"]), one_f(lib["lut_val"])
    fr, en = one_i(frontiers), _i32(ends)
    slew = np.full(max(4 * n, 1), np.nan, np.float32)
    at = np.full(max(4 * n, 1), np.nan, np.float32)
    ad = np.full(max(8 * m, 1), np.nan, np.float32)
    args = (n, _p(fs), _p(fi), _p(kind), _p(tim), _p(rcd), _p(rci), _p(load), _p(pis), _p(pia),
            _p(sense), _p(edge), _p(tlut), _p(thas), _p(hdr), _p(pool), _p(fr), _p(en),
            len(en) - 1, _p(slew), _p(at), _p(ad))
    if nthreads > 1:
        _lib.ora_fprop_par(*args, int(nthreads))
    else:
        _lib.ora_fprop(*args)
    return slew[: 4 * n].reshape(n, 4), at[: 4 * n].reshape(n, 4), ad[: 8 * m].reshape(m, 8)


def fanout_csr(fanin_start, fanin):
    """Arcs grouped by their source pin (stable in arc order): (fanout_start
    [n+1], fanout_arc [m], arc_to [m])."""
    fs = _i32(fanin_start)
    n, m = len(fs) - 1, int(fs[-1])
    arc_to = np.repeat(np.arange(n, dtype=np.int32), np.diff(fs)).astype(np.int32)
    src = _i32(fanin)[:m]
    order = np.argsort(src, kind="stable").astype(np.int32)
    start = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(src, minlength=n), out=start[1:])
    return start, order, arc_to


def bprop(fanin_start, fanin, arc_delay, po_rat, at, frontiers, ends):
    """Required times and slacks, levels first to last (timer.cpp:812-841,
    arc.cpp:150-175, pin.cpp:251-256, 362-380).  Returns (rat [n,4], slack [n,4])."""
    fo_start, fo_arc, arc_to = fanout_csr(fanin_start, fanin)
    n = len(fo_start) - 1
    one_i = lambda a: _i32(a) if len(a) else np.zeros(1, np.int32)
    ad = _f32(arc_delay).reshape(-1) if np.size(arc_delay) else np.zeros(1, np.float32)
    pr = _f32(po_rat).reshape(-1) if n else np.zeros(1, np.float32)
    a = _f32(at).reshape(-1) if n else np.zeros(1, np.float32)
    rat = np.full(max(4 * n, 1), np.nan, np.float32)
    slack = np.full(max(4 * n, 1), np.nan, np.float32)
    en = _i32(ends)
    _lib.ora_bprop(_p(fo_start), _p(one_i(fo_arc)), _p(one_i(arc_to)), _p(ad), _p(pr), _p(a),
                   _p(one_i(frontiers)), _p(en), len(en) - 1, _p(rat), _p(slack))
    return rat[: 4 * n].reshape(n, 4), slack[: 4 * n].reshape(n, 4)
