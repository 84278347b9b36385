"""The OpenTimer bridge (opentimer_amd/bridge/ot_cuda_bridge.cpp) called as
OpenTimer's timer calls it.

The harness (tests/bridge/harness.cpp, C++) builds a RctEdgeArrayCUDA and
passes it by value to rct_bfs_and_compute_cuda (ot/timer/net.cpp:254-274),
builds a RctCUDA for rct_compute_cuda (ot/cuda/rct.cuh:44), and calls
toposort_compute_cuda with a std::vector<int> that already holds 0 and
first_size (ot/timer/timer.cpp:1058-1071).  Bars: BFS order, pid and the
oriented edges bit-exact vs the oracle; outputs within 1e-6 of FP64 for the
default (compensated) method and bit-exact vs the reference kernel
arithmetic for OTG_RCT_METHOD=plain; frontier ends equal to the oracle's and
the pins of every level equal as sets.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from opentimer_amd import native as N  # loads libotg (and torch's HIP runtime) first
from opentimer_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BRIDGE = os.path.join(ROOT, "opentimer_amd", "libot_cuda_bridge.so")
HARNESS = os.path.join(ROOT, "tests", "bridge", "libbridge_harness.so")
OUT = ("load", "delay", "ldelay", "impulse")

# The reference's declarations (ot/cuda/rct.cuh:44-45, toposort.cuh:3-6) as
# g++ mangles them: the names OpenTimer's objects link against.
MANGLED = ("_Z24rct_bfs_and_compute_cuda16RctEdgeArrayCUDA", "_Z16rct_compute_cuda7RctCUDA",
           "_Z21toposort_compute_cudaiiiPiS_S_S_RSt6vectorIiSaIiEE")


def test_bridge_exports_the_reference_symbols():
    lib = C.CDLL(BRIDGE)
    for name in MANGLED:
        assert hasattr(lib, name), name


def _harness():
    h = C.CDLL(HARNESS)
    vp = C.c_void_p
    h.harness_rct_bfs_and_compute.argtypes = [C.c_int] * 3 + [vp] * 11
    h.harness_rct_compute.argtypes = [C.c_int] * 2 + [vp] * 8
    h.harness_toposort.argtypes = [C.c_int] * 3 + [vp] * 5
    h.harness_toposort.restype = C.c_int
    return h


def _floor(ref):
    f = {k: max(float(np.max(np.abs(ref[k]))) * 1e-7, 1e-30) for k in OUT}
    f["impulse"] = max(float(np.max(np.abs(ref["beta"]))) * 1e-6, 1e-30)
    return f


def _rel(got, ref, fl):
    return float(np.max(np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), fl)))


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["compensated", "plain"])
def test_rct_bfs_and_compute_by_value(method, monkeypatch):
    monkeypatch.setenv("OTG_RCT_METHOD", method)
    h = _harness()
    f = W.generate_rc(W.config0(seed=31))
    V, Nn = f.num_nodes, f.num_nets
    edges = f.edges.copy()
    n2b = np.full(V, -1, np.int32)
    pid = np.full(V, -1, np.int32)
    out = {k: np.zeros(4 * V, np.float32) for k in OUT}
    h.harness_rct_bfs_and_compute(Nn, V, V - Nn, N.addr(f.nodes_start), N.addr(edges),
                                  N.addr(f.roots), N.addr(f.edge_res), N.addr(f.node_cap),
                                  N.addr(n2b), N.addr(pid), *[N.addr(out[k]) for k in OUT])
    e, on2b, opid = O.rct_bfs(f.nodes_start, f.edges, f.roots)
    assert np.array_equal(n2b, on2b) and np.array_equal(pid, opid) and np.array_equal(edges, e)
    pres, cap = O.rct_prepare(f.nodes_start, e, on2b, f.edge_res, f.node_cap)
    zero_si = np.zeros(4 * Nn, np.float32)
    if method == "plain":
        ref = O.rct_all(f.nodes_start, opid, pres, cap, zero_si, "flat_f32")
        for k in OUT:
            assert np.array_equal(out[k].view(np.uint32), ref[k].view(np.uint32)), k
    else:
        ref = O.rct_all(f.nodes_start, opid, pres, cap, zero_si, "f64")
        fl = _floor(ref)
        for k in OUT:
            assert _rel(out[k], ref[k], fl[k]) < 1e-6, k


@pytest.mark.gpu
def test_rct_compute_by_value(monkeypatch):
    monkeypatch.setenv("OTG_RCT_METHOD", "compensated")
    h = _harness()
    f = W.generate_rc(W.config1(num_nets=500, seed=33))
    V, Nn = f.num_nodes, f.num_nets
    e, n2b, pid = O.rct_bfs(f.nodes_start, f.edges, f.roots)
    pres, cap = O.rct_prepare(f.nodes_start, e, n2b, f.edge_res, f.node_cap)
    out = {k: np.zeros(4 * V, np.float32) for k in OUT}
    h.harness_rct_compute(Nn, V, N.addr(f.nodes_start), N.addr(pid), N.addr(pres), N.addr(cap),
                          *[N.addr(out[k]) for k in OUT])
    ref = O.rct_all(f.nodes_start, pid, pres, cap, np.zeros(4 * Nn, np.float32), "f64")
    fl = _floor(ref)
    for k in OUT:
        assert _rel(out[k], ref[k], fl[k]) < 1e-6, k


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [41, 42])
def test_toposort_with_seeded_ends(seed):
    h = _harness()
    csr = W.fanin_csr(W.generate_dag(W.config0_dag(seed=seed)))
    n = csr.n
    fr = np.full(n, -1, np.int32)
    fr[: len(csr.frontier)] = csr.frontier
    ends = np.zeros(n + 2, np.int32)
    outd = csr.out_degree.copy()
    ne = h.harness_toposort(n, len(csr.fanin), len(csr.frontier), N.addr(csr.start),
                            N.addr(csr.fanin), N.addr(outd), N.addr(fr), N.addr(ends))
    ofr, oends, olv = O.levelize(n, csr.start, csr.fanin, csr.out_degree, csr.frontier)
    assert np.array_equal(ends[:ne], oends)
    assert np.array_equal(outd, csr.out_degree)  # the host fanout counts are not modified
    for k in range(ne - 1):
        assert np.array_equal(np.sort(fr[ends[k]: ends[k + 1]]), np.sort(ofr[oends[k]: oends[k + 1]])), k
