"""SPEF loader (include/otg_spef.h; opentimer_amd/spef.py), host side.

The loader replaces spef::Spef::read and the flattening of
Net::_make_flat_rct (ot/timer/net.cpp:347-461).  Its rules are pinned here by
an independent line-by-line reading of the fixture files in Python, by
hand-written SPEF text for each rule, and by a round trip through the writer.
"""
import os
import re

import numpy as np
import pytest

from opentimer_amd import workloads as W
from opentimer_amd.spef import SpefError, SpefForest, write

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spef")
FIXTURES = ["c17.spef", "simple.spef", "s27_spef.spef"]


def _independent_read(path):
    """Per net: (name, total cap, driver, [res (n1, n2)], [cap (n1, value)]),
    names with *NAME_MAP indices expanded."""
    nets, cur, sec, nmap, in_map = [], None, None, {}, False

    def ex(tok):
        m = re.match(r"\*(\d+)(.*)$", tok)
        return nmap[int(m.group(1))] + m.group(2) if m and int(m.group(1)) in nmap else tok

    for raw in open(path):
        t = raw.split("//")[0].split()
        if not t:
            continue
        if t[0] == "*NAME_MAP":
            in_map = True
            continue
        if in_map and re.match(r"\*\d+$", t[0]):
            nmap[int(t[0][1:])] = t[1]
            continue
        in_map = False
        t = [t[0]] + [ex(x) for x in t[1:]]
        if t[0] == "*D_NET":
            cur = {"name": t[1], "total": float(t[2]), "driver": None, "res": [], "cap": []}
            nets.append(cur)
            sec = None
        elif cur is None:
            continue
        elif t[0] in ("*CONN", "*CAP", "*RES"):
            sec = t[0]
        elif t[0] == "*END":
            cur, sec = None, None
        elif sec == "*CONN" and t[0] in ("*I", "*P"):
            if cur["driver"] is None and ((t[0] == "*I" and t[2] == "O") or (t[0] == "*P" and t[2] == "I")):
                cur["driver"] = t[1]
        elif sec == "*RES":
            cur["res"].append((t[1], t[2]))
        elif sec == "*CAP":
            cur["cap"].append((t[1], float(t[-1])))
    return nets


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_matches_independent_reading(name):
    path = os.path.join(GOLD, name)
    f = SpefForest.read(path)
    ref = _independent_read(path)
    assert f.num_nets == len(ref)
    assert f.num_undriven == 0 and f.num_dropped_caps == 0
    ns = f.nodes_start
    for k, net in enumerate(ref):
        assert f.net_name(k) == net["name"]
        n = ns[k + 1] - ns[k]
        assert n == len(net["res"]) + 1  # net.cpp:364
        # local ids in order of first appearance in *RES, node1 then node2
        order = []
        for a, b in net["res"]:
            for x in (a, b):
                if x not in order:
                    order.append(x)
        for j, x in enumerate(order):
            assert f.node_name(k, j) == x
        e0 = ns[k] - k
        for i, (a, b) in enumerate(net["res"]):
            assert f.edges[2 * (e0 + i)] == order.index(a)
            assert f.edges[2 * (e0 + i) + 1] == order.index(b)
        assert f.roots[k] == order.index(net["driver"])
        # caps on all four corners; the printed *D_NET total agrees with the
        # node sum to the file's 4-decimal precision
        cap = f.node_cap[4 * ns[k]: 4 * ns[k + 1]].reshape(-1, 4)
        assert np.all(cap == cap[:, :1])
        for x, v in net["cap"]:
            assert cap[order.index(x), 0] == np.float32(v)
        assert abs(float(cap[:, 0].astype(np.float64).sum()) - net["total"]) <= 5e-5 * (n + 1)
        assert f.total_cap[k] == np.float32(net["total"])
    assert f.cap_unit == pytest.approx(1e-15) and f.res_unit == pytest.approx(1e3)


HEAD = """*SPEF "IEEE 1481-1998"
*DESIGN "t"
*DIVIDER /
*DELIMITER :
*BUS_DELIMITER [ ]
*T_UNIT 1 PS
*C_UNIT 1 FF
*R_UNIT 1 KOHM
"""


def test_local_ids_root_and_cap_rules():
    txt = HEAD + """
*D_NET x 9.0
*CONN
*I u2:A I
*I u1:Z O
*CAP
1 x:2 1.0
2 u1:Z 2.0
3 x:1 3.0
4 x:1 u9:Q 0.5
5 u2:A 4.0
*RES
1 x:2 x:1 10.0
2 x:1 u1:Z 20.0
3 x:2 u2:A 30.0
*END
"""
    f = SpefForest.parse(txt)
    # first appearance in *RES: x:2, x:1, u1:Z, u2:A
    assert [f.node_name(0, j) for j in range(4)] == ["x:2", "x:1", "u1:Z", "u2:A"]
    assert list(f.edges) == [0, 1, 1, 2, 0, 3]
    assert list(f.edge_res) == [10.0, 20.0, 30.0]
    assert f.roots[0] == 2  # the cell output pin u1:Z
    # the coupling entry of x:1 overwrites its ground entry (net.cpp:425-431)
    assert list(f.node_cap[0::4]) == [1.0, 0.5, 2.0, 4.0]
    assert f.find_node(0, "u2:A") == 3 and f.find_node(0, "nope") == -1
    assert f.find_net("x") == 0 and f.find_net("y") == -1
    # library pin caps: added to every node but the root (net.cpp:433-450)
    f.add_pin_cap(0, 3, [0.1, 0.2, 0.3, 0.4])
    f.add_pin_cap(0, 2, [9.0, 9.0, 9.0, 9.0])
    assert np.allclose(f.node_cap[12:16], [4.1, 4.2, 4.3, 4.4])
    assert list(f.node_cap[8:12]) == [2.0] * 4


def test_port_driver_name_map_and_comments():
    txt = """*SPEF "IEEE 1481-1998"
*DESIGN "t"
*DELIMITER :
*C_UNIT 1 FF
*R_UNIT 1 KOHM
*NAME_MAP
*1 inp
*2 u7
*D_NET *1 3.0 // header comment
*CONN
*P *1 I
*I *2:A I *L 0.5
*CAP
1 *1 1.0

2 *1:1 2.0
*RES
1 *1 *1:1 1.5
2 *1:1 *2:A 2.5
*END

// between nets
*D_NET lone 1.5
*CONN
*I u8:Z O
*CAP
1 u8:Z 1.5
*END
"""
    f = SpefForest.parse(txt)
    assert f.num_nets == 2
    assert f.net_name(0) == "inp" and f.find_net("inp") == 0
    assert [f.node_name(0, j) for j in range(3)] == ["inp", "inp:1", "u7:A"]
    assert f.roots[0] == 0  # the input port drives the net
    # a net without *RES entries is one node, its driver
    assert f.nodes_start[2] - f.nodes_start[1] == 1
    assert f.roots[1] == 0 and f.node_cap[4 * 3] == 1.5


def test_units_rebased_like_the_timer():
    txt = HEAD.replace("*C_UNIT 1 FF", "*C_UNIT 1 PF").replace("*R_UNIT 1 KOHM", "*R_UNIT 1 OHM") + """
*D_NET a 0.002
*CONN
*I d:Z O
*CAP
1 d:Z 0.001
2 a:1 0.001
*RES
1 d:Z a:1 100
*END
"""
    f = SpefForest.parse(txt, cap_unit=1e-15, res_unit=1e3)
    assert np.allclose(f.node_cap[0::4], [1.0, 1.0])
    assert np.allclose(f.edge_res, [0.1])
    assert np.isclose(f.total_cap[0], 2.0)
    # within 1% the values are kept (unit.cpp:328-366)
    g = SpefForest.parse(txt, cap_unit=1.005e-12, res_unit=0)
    assert np.allclose(g.node_cap[0::4], [0.001, 0.001]) and g.res_unit == 1.0


@pytest.mark.parametrize("bad,line", [
    ("*D_NET a 1.0\n*CONN\n*I d:Z O\n*RES\n1 d:Z a:1\n*END\n", 5),
    ("*D_NET a 1.0\n*CAP\n1 d:Z x\n*END\n", 3),
    ("*D_NET a 1.0\n*CAP\n1 d:Z 1.0\n", 3),
])
def test_errors_name_the_line(bad, line):
    with pytest.raises(SpefError, match=rf"line {line + HEAD.count(chr(10))}\b"):
        SpefForest.parse(HEAD + bad)


def test_round_trip_through_the_writer(tmp_path):
    cfg = W.RcConfig("rt", 300, [W.SizeClass(0.7, 2, 16), W.SizeClass(0.3, 17, 300, 1, 0.2)],
                     seed=5)
    g = W.generate_rc(cfg)
    cap4 = g.node_cap.reshape(-1, 4).copy()
    cap4[:] = cap4[:, :1]  # SPEF caps are per node, not per corner
    g.node_cap = cap4.reshape(-1)
    p = tmp_path / "rt.spef"
    write(str(p), g)
    for threads in (1, 3, 0):
        f = SpefForest.read(str(p), threads=threads)
        assert f.num_nets == g.num_nets
        assert np.array_equal(f.nodes_start, g.nodes_start)
        # node names carry the original ids: map the loader's ids back
        for k in (0, 1, 7, g.num_nets - 1):
            n = g.nodes_start[k + 1] - g.nodes_start[k]
            orig = [int(re.split("[:]", f.node_name(k, j))[1]) if f.node_name(k, j).startswith("n")
                    else int(g.roots[k]) for j in range(n)]
            e0 = g.nodes_start[k] - k
            ge = g.edges[2 * e0: 2 * (e0 + n - 1)]
            fe = f.edges[2 * e0: 2 * (e0 + n - 1)]
            assert [orig[x] for x in fe] == list(ge)
            assert orig[f.roots[k]] == g.roots[k]
            assert np.array_equal(f.node_cap[4 * g.nodes_start[k]: 4 * g.nodes_start[k + 1]].reshape(-1, 4)[np.argsort(orig)],
                                  g.node_cap[4 * g.nodes_start[k]: 4 * g.nodes_start[k + 1]].reshape(-1, 4))
        assert np.array_equal(f.edge_res, g.edge_res)
