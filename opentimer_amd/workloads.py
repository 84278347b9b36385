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
