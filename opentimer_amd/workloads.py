"""Seeded synthetic workloads (host numpy arrays) built with libotg's generator.

RC forests come out in the reference's flat-input form (RctEdgeArrayCUDA,
ot/cuda/rct.cuh:20-28): undirected net-local edges with random orientation
and random node numbering, a driver id per net, per-edge R and per-node
4-corner C in original order, plus a driver slew per net and corner.

Units: R is drawn in ohm and stored in kohm (x1e-3) so that R*C is in ps
(kohm * fF = ps, the TAU/SPEF convention of the reference's benchmarks),
matching the driver slew in ps.

Corner c = el*2 + rf; sink (leaf) pin caps are scaled by
MIN 0.9 / MAX 1.1 and RISE 1.0 / FALL 1.05.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import native as N

CORNER_SCALE = (0.9 * 1.0, 0.9 * 1.05, 1.1 * 1.0, 1.1 * 1.05)


@dataclass
class SizeClass:
    weight: float
    min_nodes: int
    max_nodes: int
    shape: int = N.SHAPE_RANDOM
    branch_prob: float = 0.0


@dataclass
class RcConfig:
    name: str
    num_nets: int
    classes: list = field(default_factory=list)
    seed: int = 1
    r_range: tuple = (0.1, 100.0)        # ohm
    c_range: tuple = (0.1, 10.0)         # fF, ground cap per node
    pin_range: tuple = (1.0, 5.0)        # fF, sink pin cap base
    slew_range: tuple = (5.0, 50.0)      # ps
    r_unit: float = 1e-3                 # ohm -> kohm
    relabel: bool = True
    random_orient: bool = True
    first_net: int = 0                   # global index of net 0 (batches of a larger population)

    def scaled(self, num_nets: int, seed: int | None = None) -> "RcConfig":
        c = RcConfig(**{**self.__dict__})
        c.num_nets = num_nets
        if seed is not None:
            c.seed = seed
        return c


def config0(num_nets=4096, seed=1):
    """Plain method, CPU-runnable: 4,096 random trees of 2-16 nodes."""
    return RcConfig("cfg0-plain-4096", num_nets, [SizeClass(1.0, 2, 16)], seed=seed)


def config1(num_nets=4_000_000, seed=2):
    """Batching: heavy-tailed sizes; the 1% tail are chain-like trees."""
    return RcConfig("cfg1-batched-4M", num_nets, [
        SizeClass(0.80, 2, 8),
        SizeClass(0.19, 9, 64),
        SizeClass(0.01, 65, 8192, N.SHAPE_CHAIN, 0.1),
    ], seed=seed)


def config3(num_nets=2_000_000, seed=3):
    """Precision: deep chains and balanced trees, wide R/C ranges."""
    return RcConfig("cfg3-precision", num_nets, [
        SizeClass(0.5, 1000, 20000, N.SHAPE_CHAIN, 0.0),
        SizeClass(0.5, 64, 1024, N.SHAPE_BALANCED),
    ], seed=seed, r_range=(1e-3, 1e4), c_range=(1e-3, 1e2), pin_range=(0.0, 0.0))


def _params(cfg: RcConfig) -> N.GenRcParams:
    p = N.GenRcParams()
    p.seed = cfg.seed
    p.num_nets = cfg.num_nets
    p.num_classes = len(cfg.classes)
    for i, c in enumerate(cfg.classes):
        p.classes[i] = N.GenSizeClass(c.weight, c.min_nodes, c.max_nodes, c.shape, 0, c.branch_prob)
    p.r_min, p.r_max = cfg.r_range
    p.c_min, p.c_max = cfg.c_range
    p.pin_min, p.pin_max = cfg.pin_range
    for i in range(4):
        p.corner_scale[i] = CORNER_SCALE[i]
    p.slew_min, p.slew_max = cfg.slew_range
    p.r_unit = cfg.r_unit
    p.relabel = int(cfg.relabel)
    p.random_orient = int(cfg.random_orient)
    p.first_net = cfg.first_net
    return p


@dataclass
class RcForest:
    """Host arrays in RctEdgeArray form (original node numbering)."""
    nodes_start: np.ndarray   # int32 [N+1]
    edges: np.ndarray         # int32 [2E]
    roots: np.ndarray         # int32 [N]
    edge_res: np.ndarray      # float32 [E]
    node_cap: np.ndarray      # float32 [4V]
    slew_in: np.ndarray       # float32 [4N]
    parent_gen: np.ndarray | None = None

    @property
    def num_nets(self):
        return len(self.nodes_start) - 1

    @property
    def num_nodes(self):
        return int(self.nodes_start[-1])


def generate_rc(cfg: RcConfig, nthreads: int = 0, keep_parent: bool = False) -> RcForest:
    p = _params(cfg)
    ns = np.zeros(cfg.num_nets + 1, np.int32)
    tot = C.c_int64()
    N.check(N.lib.otg_gen_rc_sizes(C.byref(p), N.addr(ns), C.byref(tot)))
    V = tot.value
    E = V - cfg.num_nets
    edges = np.empty(2 * max(E, 0), np.int32)
    roots = np.empty(cfg.num_nets, np.int32)
    res = np.empty(max(E, 0), np.float32)
    cap = np.empty(4 * V, np.float32)
    si = np.empty(4 * cfg.num_nets, np.float32)
    par = np.empty(V, np.int32) if keep_parent else None
    N.check(N.lib.otg_gen_rc_fill(C.byref(p), N.addr(ns), N.addr(edges), N.addr(roots),
                                  N.addr(res), N.addr(cap), N.addr(si), N.addr(par), nthreads))
    return RcForest(ns, edges, roots, res, cap, si, par)


@dataclass
class DagConfig:
    name: str
    num_pins: int
    num_arcs: int
    num_layers: int
    fanout_alpha: float = 2.2
    fanout_cap: int = 512
    skip_fraction: float = 0.05   # of all arcs
    skip_min: int = 2
    skip_max: int = 10
    seed: int = 7
    shuffle_ids: bool = True


def config0_dag(seed=7):
    """20k pins in 40 layers (10 gate stages x 4 pin layers), fanout U{1..4}."""
    return DagConfig("cfg0-dag-20k", 20_000, 48_750, 40, fanout_alpha=0.0, fanout_cap=4,
                     skip_fraction=0.0, seed=seed)


def config2(seed=8):
    """Levelization: 10M pins, 25M arcs, 200 levels, power-law fanout."""
    return DagConfig("cfg2-levelize-10M", 10_000_000, 25_000_000, 200, seed=seed)


@dataclass
class Dag:
    n: int
    arc_from: np.ndarray
    arc_to: np.ndarray


def generate_dag(cfg: DagConfig, nthreads: int = 0) -> Dag:
    p = N.GenDagParams()
    p.seed = cfg.seed
    p.num_pins = cfg.num_pins
    p.num_layers = cfg.num_layers
    p.num_arcs = cfg.num_arcs
    p.fanout_alpha = cfg.fanout_alpha
    p.fanout_cap = cfg.fanout_cap
    p.skip_min, p.skip_max = cfg.skip_min, cfg.skip_max
    src = cfg.num_pins * (cfg.num_layers - 1) // cfg.num_layers
    extra = max(1, cfg.num_arcs - src)
    # the first arc of every pin goes to the next layer; the others skip with
    # this probability so that skip_fraction of all arcs skip
    p.skip_prob = min(1.0, cfg.skip_fraction * cfg.num_arcs / extra)
    p.shuffle_ids = int(cfg.shuffle_ids)
    fr = np.empty(cfg.num_arcs, np.int32)
    to = np.empty(cfg.num_arcs, np.int32)
    N.check(N.lib.otg_gen_dag(C.byref(p), N.addr(fr), N.addr(to), nthreads))
    return Dag(cfg.num_pins, fr, to)


@dataclass
class FaninCsr:
    n: int
    start: np.ndarray      # [n+1]
    fanin: np.ndarray      # [m]
    out_degree: np.ndarray  # [n]
    frontier: np.ndarray   # [first_size]


def fanin_csr(dag: Dag, cand=None, skip_arc=None) -> FaninCsr:
    n, m = dag.n, len(dag.arc_from)
    start = np.empty(n + 1, np.int32)
    fanin = np.empty(max(m, 1), np.int32)
    out = np.empty(max(n, 1), np.int32)
    fr = np.empty(max(n, 1), np.int32)
    fs = C.c_int32()
    kept = C.c_int64()
    c = None if cand is None else np.ascontiguousarray(cand, np.uint8)
    s = None if skip_arc is None else np.ascontiguousarray(skip_arc, np.uint8)
    N.check(N.lib.otg_build_fanin_csr(n, m, N.addr(dag.arc_from), N.addr(dag.arc_to), N.addr(c),
                                      N.addr(s), N.addr(start), N.addr(fanin), N.addr(out),
                                      N.addr(fr), C.byref(fs), C.byref(kept)))
    return FaninCsr(n, start, fanin[: kept.value], out[:n], fr[: fs.value])


# ---------------------------------------------------------------------------
# Forward propagation: a synthetic NLDM library and per-arc / per-pin inputs
# over a levelization DAG.  Layout as in include/otg_prop.h.
# ---------------------------------------------------------------------------

@dataclass
class PropLibrary:
    lut_hdr: np.ndarray       # [4L] int32 {n1, n2, var1, off}
    lut_val: np.ndarray       # float32 pool
    timing_sense: np.ndarray  # [T] int32
    timing_edge: np.ndarray   # [T] int32
    timing_lut: np.ndarray    # [8T] int32
    timing_has: np.ndarray    # [2T] int32

    @property
    def num_luts(self):
        return len(self.lut_hdr) // 4

    @property
    def num_timings(self):
        return len(self.timing_sense)


class _LutPool:
    def __init__(self):
        self.hdr, self.val = [], []

    def add(self, i1, i2, table, var1):
        off = sum(len(v) for v in self.val)
        self.hdr += [len(i1), len(i2), var1, off]
        self.val.append(np.concatenate([np.asarray(i1, np.float32), np.asarray(i2, np.float32),
                                        np.asarray(table, np.float32).reshape(-1)]))
        return len(self.hdr) // 4 - 1

    def arrays(self):
        return (np.asarray(self.hdr, np.int32),
                np.concatenate(self.val) if self.val else np.zeros(1, np.float32))


def prop_library(num_timings: int = 64, seed: int = 11) -> PropLibrary:
    """A seeded library shaped like the TAU 2015 cells: 7x8 slew-by-load
    tables (input transition 5..350 ps, output load 1..200 fF) with delays of
    5..40 ps and output transitions of 3..30 ps, monotone in both operands.
    Mixed in: load-first templates, 1xN / Nx1 tables, scalar tables, absent
    tables, all three senses, rising/falling-edge triggers and timings with
    no view for one split."""
    rng = np.random.default_rng(seed)
    pool = _LutPool()
    slews = np.array([5, 30, 50, 80, 140, 200, 350], np.float32)
    loads = np.array([1, 5, 10, 15, 20, 50, 100, 200], np.float32)

    def table(kind):
        base = rng.uniform(3, 8) if kind == "delay" else rng.uniform(2, 4)
        ks = rng.uniform(0.005, 0.02) if kind == "delay" else rng.uniform(0.002, 0.01)
        kl = rng.uniform(0.05, 0.15) if kind == "delay" else rng.uniform(0.04, 0.12)
        t = base + ks * slews[:, None] + kl * loads[None, :]
        return np.round(t + rng.uniform(0, 0.01, t.shape), 3)

    def make(kind):
        shape = rng.choice(5, p=[0.70, 0.12, 0.06, 0.06, 0.06])
        t = table(kind)
        if shape == 0:
            return pool.add(slews, loads, t, 0)
        if shape == 1:  # load-first template
            return pool.add(loads, slews, t.T, 1)
        if shape == 2:  # 1 x N: one slew, indexed by load
            return pool.add(slews[:1], loads, t[:1], 0)
        if shape == 3:  # N x 1: indexed by slew, one load
            return pool.add(slews, loads[:1], t[:, :1], 0)
        return pool.add([0.0], [0.0], t[:1, :1], -1)  # scalar, no template

    T = num_timings
    sense = rng.choice(3, size=T, p=[0.35, 0.45, 0.20]).astype(np.int32)
    edge = np.zeros(T, np.int32)
    trig = rng.random(T) < 0.08
    edge[trig] = rng.choice([1, 2], size=int(trig.sum())).astype(np.int32)
    sense[trig] = 2
    tlut = np.full(8 * T, -1, np.int32)
    has = np.ones(2 * T, np.int32)
    for t in range(T):
        for el in range(2):
            for k in range(4):
                if rng.random() < 0.02:
                    continue  # an absent table
                tlut[8 * t + 4 * el + k] = make("delay" if k < 2 else "slew")
        if rng.random() < 0.04:
            has[2 * t + int(rng.integers(2))] = 0
    hdr, val = pool.arrays()
    return PropLibrary(hdr, val, sense, edge, tlut, has)


@dataclass
class PropGraphData:
    fanin_start: np.ndarray     # [n+1] int32
    fanin: np.ndarray           # [m] int32
    arc_kind: np.ndarray        # [m] int32
    arc_timing: np.ndarray      # [m] int32
    arc_rc_delay: np.ndarray    # [m,4] float32
    arc_rc_impulse: np.ndarray  # [m,4] float32
    pin_load: np.ndarray        # [n,4] float32
    pi_slew: np.ndarray         # [n,4] float32
    pi_at: np.ndarray           # [n,4] float32

    @property
    def n(self):
        return len(self.fanin_start) - 1

    @property
    def m(self):
        return len(self.fanin)


def prop_graph(csr: "FaninCsr", lib: PropLibrary, seed: int = 12, net_fraction: float = 0.5,
               empty_fraction: float = 0.03, skip_fraction: float = 0.01,
               missing_sink_fraction: float = 0.005) -> PropGraphData:
    """Per-arc and per-pin inputs over a fanin CSR: arcs are RC net arcs
    (delay 0..5 ps, impulse 0..40 ps^2), net arcs without parasitics, cell
    arcs of a random timing, or loop breakers; a few RC net arcs lead to a
    sink outside the RC tree (NaN delay and impulse).  Every pin without
    fanin is a primary input (slew 5..20 ps, arrival 0..10 ps) except 2% left
    unconstrained; pin loads are 1..200 fF per corner."""
    rng = np.random.default_rng(seed)
    n, m = csr.n, len(csr.fanin)
    u = rng.random(m)
    kind = np.full(m, 2, np.int32)
    kind[u < net_fraction] = 0
    kind[(u >= net_fraction) & (u < net_fraction + empty_fraction)] = 1
    kind[u >= 1.0 - skip_fraction] = 3
    timing = rng.integers(0, max(lib.num_timings, 1), size=m, dtype=np.int32)
    rcd = rng.uniform(0.0, 5.0, size=(m, 4)).astype(np.float32)
    rci = rng.uniform(0.0, 40.0, size=(m, 4)).astype(np.float32)
    miss = (kind == 0) & (rng.random(m) < missing_sink_fraction)
    rcd[miss] = np.nan
    rci[miss] = np.nan
    load = rng.uniform(1.0, 200.0, size=(n, 4)).astype(np.float32)
    deg = np.diff(csr.start)
    pi = (deg == 0) & (rng.random(n) >= 0.02)
    pis = np.full((n, 4), np.nan, np.float32)
    pia = np.full((n, 4), np.nan, np.float32)
    k = int(pi.sum())
    pis[pi] = rng.uniform(5.0, 20.0, size=(k, 4)).astype(np.float32)
    pia[pi] = rng.uniform(0.0, 10.0, size=(k, 4)).astype(np.float32)
    return PropGraphData(np.ascontiguousarray(csr.start, np.int32),
                         np.ascontiguousarray(csr.fanin, np.int32), kind, timing, rcd, rci,
                         load, pis, pia)


# ---------------------------------------------------------------------------
# A whole design for the device timer (include/otg_timer.h): a timing graph
# whose net arcs lead into RC trees.
# ---------------------------------------------------------------------------

@dataclass
class TimerDesignData:
    forest: RcForest          # RC trees, original numbering, root = node 0's driver
    fanin_start: np.ndarray   # [n+1]
    fanin: np.ndarray         # [m] source pin of arc j
    arc_kind: np.ndarray      # [m] OTG_ARC_NET_RC (0) or OTG_ARC_CELL (2)
    arc_timing: np.ndarray    # [m]
    arc_net: np.ndarray       # [m] RC net of a net arc, else -1
    arc_node: np.ndarray      # [m] sink node of a net arc in its net, else -1
    pin_net: np.ndarray       # [n] RC net the pin drives, else -1
    pin_load: np.ndarray      # [n, 4]
    pi_slew: np.ndarray       # [n, 4]
    pi_at: np.ndarray         # [n, 4]
    po_rat: np.ndarray        # [n, 4]


def timer_design(dag_cfg: DagConfig, lib: PropLibrary, seed: int = 31, net_fraction: float = 0.4,
                 nthreads: int = 0) -> TimerDesignData:
    """A design over `dag_cfg`'s graph: a pin with fanout drives an RC net with
    probability `net_fraction` (all its fanout arcs are then net arcs of that
    net; every other arc is a cell arc of a random timing).  The nets come
    from the cfg1 size distribution, the largest nets given to the pins with
    the largest fanout; the k-th fanout arc of a driver leads to node 1 + (k
    mod (size - 1)) of its tree (distinct sinks while the tree has more nodes
    than the driver has fanout arcs).  Pins without fanin are primary inputs
    (slew 5..20 ps, arrival 0..10 ps), pins without fanout primary outputs
    (required time 3000..6000 ps); cell-arc loads of pins driving no RC net
    are 1..200 fF."""
    rng = np.random.default_rng(seed)
    csr = fanin_csr(generate_dag(dag_cfg, nthreads=nthreads))
    n, m = csr.n, len(csr.fanin)
    od = csr.out_degree
    drivers = np.flatnonzero((od > 0) & (rng.random(n) < net_fraction)).astype(np.int32)
    forest = generate_rc(config1(num_nets=len(drivers), seed=seed), nthreads=nthreads)
    sizes = np.diff(forest.nodes_start)
    pin_net = np.full(n, -1, np.int32)
    pin_net[drivers[np.argsort(-od[drivers], kind="stable")]] = \
        np.argsort(-sizes, kind="stable").astype(np.int32)
    src = csr.fanin
    arc_net = pin_net[src]
    order = np.argsort(src, kind="stable")
    s_sorted = src[order]
    rank = np.empty(m, np.int64)
    rank[order] = np.arange(m) - np.searchsorted(s_sorted, s_sorted, side="left")
    size_of = sizes[np.maximum(arc_net, 0)].astype(np.int64)
    rc = arc_net >= 0
    arc_node = np.where(rc, 1 + rank % np.maximum(size_of - 1, 1), -1).astype(np.int32)
    arc_kind = np.where(rc, 0, 2).astype(np.int32)
    arc_timing = rng.integers(0, max(lib.num_timings, 1), size=m, dtype=np.int32)
    pin_load = rng.uniform(1.0, 200.0, size=(n, 4)).astype(np.float32)
    deg = np.diff(csr.start)
    pis = np.full((n, 4), np.nan, np.float32)
    pia = np.full((n, 4), np.nan, np.float32)
    pi = deg == 0
    pis[pi] = rng.uniform(5.0, 20.0, size=(int(pi.sum()), 4)).astype(np.float32)
    pia[pi] = rng.uniform(0.0, 10.0, size=(int(pi.sum()), 4)).astype(np.float32)
    po = np.full((n, 4), np.nan, np.float32)
    sink = od == 0
    po[sink] = rng.uniform(3000.0, 6000.0, size=(int(sink.sum()), 4)).astype(np.float32)
    return TimerDesignData(forest, np.ascontiguousarray(csr.start, np.int32),
                           np.ascontiguousarray(src, np.int32), arc_kind, arc_timing, arc_net,
                           arc_node, pin_net, pin_load, pis, pia, po)
