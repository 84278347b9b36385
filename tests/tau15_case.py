"""The TAU 2015 c17 design as a forward-propagation case (test helper).

Builds, from tests/golden/prop/c17.json and tests/golden/spef/c17.spef, the
graph and library arrays of include/otg_prop.h the way the reference's
Timer sees the design:

* RC trees from the SPEF (otg_spef_read), with the input capacitance of
  every sink pin added to its node (RctNode::cap, net.cpp:24-27; a cell
  input's rise/fall capacitance, a primary output's load; the driver's pin
  is the root and adds nothing), solved per corner;
* net arcs driver -> sink with the RC delay and impulse at the sink's node
  (Net::_delay / _slew, net.cpp:646-699), cell arcs A1/A2 -> ZN of
  NAND2_X1 with the early and late library's tables, and each ZN's load =
  the root load of the net it drives (Net::_load, net.cpp:626-642);
* primary-input arrival and slew from the .timing file (tau15.cpp:99-118).
"""
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")

SENSE = {"positive_unate": 0, "negative_unate": 1, "non_unate": 2, None: 2}
EDGE = {"rising_edge": 1, "falling_edge": 2}
TABLES = ("cell_rise", "cell_fall", "rise_transition", "fall_transition")


def load():
    return json.load(open(os.path.join(GOLD, "prop", "c17.json")))


def spef_forest():
    from opentimer_amd.spef import SpefForest
    return SpefForest.read(os.path.join(GOLD, "spef", "c17.spef"))


def build(rc_solve):
    """rc_solve(forest) -> (node2bfs, load[4V], delay[4V], impulse[4V]) in the
    BFS order of every net.  Returns (graph dict, lib dict, pin names,
    golden arrival [n,4])."""
    J = load()
    cell = J["cell"]
    pins = list(J["inputs"]) + list(J["outputs"])
    for inst in J["instances"]:
        assert inst["cell"] == cell
        pins += [f"{inst['name']}:{p}" for p in ("A1", "A2", "ZN")]
    pid = {p: i for i, p in enumerate(pins)}
    n = len(pins)

    # net of every pin (ports are named after their net)
    net_pins = {}
    for p in J["inputs"] + J["outputs"]:
        net_pins.setdefault(p, []).append(p)
    for inst in J["instances"]:
        for p, net in inst["pins"].items():
            net_pins.setdefault(net, []).append(f"{inst['name']}:{p}")

    def pin_cap(name):
        if name in J["inputs"]:
            return [0.0] * 4
        if name in J["outputs"]:
            return [J["load"].get(name, 0.0)] * 4
        lp = name.split(":")[1]
        e, l = J["pins"]["early"][lp], J["pins"]["late"][lp]
        return [e["cap_rise"], e["cap_fall"], l["cap_rise"], l["cap_fall"]]

    forest = spef_forest()
    drivers = {}
    for k in range(forest.num_nets):
        net = forest.net_name(k)
        root = int(forest.roots[k])
        for p in net_pins[net]:
            loc = forest.find_node(k, p)
            assert loc >= 0, (net, p)
            if loc == root:
                drivers[net] = (k, p)
            else:
                forest.add_pin_cap(k, loc, np.asarray(pin_cap(p), np.float32))
    n2b, ld, dl, im = rc_solve(forest)
    ns = forest.nodes_start

    def at_node(k, loc, arr):
        return arr[4 * (ns[k] + n2b[ns[k] + loc]): 4 * (ns[k] + n2b[ns[k] + loc]) + 4]

    # library: timing t = index of the arc in the cell's list, per split
    arcs_e, arcs_l = J["arcs"]["early"], J["arcs"]["late"]
    assert [(a["from"], a["to"]) for a in arcs_e] == [(a["from"], a["to"]) for a in arcs_l]
    hdr, pool = [], []
    T = len(arcs_e)
    tlut = np.full(8 * T, -1, np.int32)
    sense = np.zeros(T, np.int32)
    edge = np.zeros(T, np.int32)
    for t in range(T):
        sense[t] = SENSE[arcs_e[t]["timing_sense"]]
        edge[t] = EDGE.get(arcs_e[t]["timing_type"], 0)
        for el, a in enumerate((arcs_e[t], arcs_l[t])):
            for k, tn in enumerate(TABLES):
                tab = a["tables"].get(tn)
                if tab is None:
                    continue
                var1 = {"input_net_transition": 0, "total_output_net_capacitance": 1}.get(
                    tab["variable_1"], -1)
                i1, i2 = tab["index_1"] or [0.0], tab["index_2"] or [0.0]
                assert len(tab["values"]) == len(i1) * len(i2)
                tlut[8 * t + 4 * el + k] = len(hdr) // 4
                hdr += [len(i1), len(i2), var1, sum(len(x) for x in pool)]
                pool.append(np.asarray(i1 + i2 + tab["values"], np.float32))
    lib = {"lut_hdr": np.asarray(hdr, np.int32), "lut_val": np.concatenate(pool),
           "timing_sense": sense, "timing_edge": edge, "timing_lut": tlut,
           "timing_has": np.ones(2 * T, np.int32)}

    # arcs, grouped by their sink pin into a fanin CSR
    arcs = []  # (from, to, kind, timing, rc delay[4], rc impulse[4])
    for net, (k, drv) in drivers.items():
        for p in net_pins[net]:
            if p == drv:
                continue
            loc = forest.find_node(k, p)
            arcs.append((pid[drv], pid[p], 0, 0, at_node(k, loc, dl), at_node(k, loc, im)))
    nan4 = np.full(4, np.nan, np.float32)
    for inst in J["instances"]:
        for t, a in enumerate(arcs_e):
            arcs.append((pid[f"{inst['name']}:{a['from']}"], pid[f"{inst['name']}:{a['to']}"], 2,
                         t, nan4, nan4))
    arcs.sort(key=lambda x: x[1])
    fs = np.zeros(n + 1, np.int32)
    for a in arcs:
        fs[a[1] + 1] += 1
    fs = np.cumsum(fs).astype(np.int32)
    load4 = np.zeros((n, 4), np.float32)
    for net, (k, drv) in drivers.items():
        if drv not in J["inputs"]:
            load4[pid[drv]] = at_node(k, int(forest.roots[k]), ld)
    pis = np.full((n, 4), np.nan, np.float32)
    pia = np.full((n, 4), np.nan, np.float32)
    for p in J["inputs"]:
        pis[pid[p]] = J["slew"][p]
        pia[pid[p]] = J["at"][p]
    graph = {
        "fanin_start": fs,
        "fanin": np.array([a[0] for a in arcs], np.int32),
        "arc_kind": np.array([a[2] for a in arcs], np.int32),
        "arc_timing": np.array([a[3] for a in arcs], np.int32),
        "arc_rc_delay": np.array([a[4] for a in arcs], np.float32).reshape(-1, 4),
        "arc_rc_impulse": np.array([a[5] for a in arcs], np.float32).reshape(-1, 4),
        "pin_load": load4, "pi_slew": pis, "pi_at": pia,
    }
    gold = np.array([J["golden_at"][p] for p in pins], np.float64)
    forest.close()
    return graph, lib, pins, gold


def oracle_rc(forest):
    """The CPU timer's RC arithmetic (oracle, Rct::update_rc_timing in FP32)."""
    import oracle as O
    e, n2b, pid = O.rct_bfs(forest.nodes_start, forest.edges, forest.roots)
    pres, cap = O.rct_prepare(forest.nodes_start, e, n2b, forest.edge_res, forest.node_cap)
    r = O.rct_dfs_f32(forest.nodes_start, pid, pres, cap)
    return n2b, r["load"], r["delay"], r["impulse"]


def levels(graph):
    import oracle as O
    fs, fi = graph["fanin_start"], graph["fanin"]
    n = len(fs) - 1
    out = np.bincount(fi, minlength=n).astype(np.int32)
    fr0 = np.flatnonzero(out == 0).astype(np.int32)
    fr, ends, _ = O.levelize(n, fs, fi, out, fr0)
    return fr, ends
