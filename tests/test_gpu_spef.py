"""SPEF → GPU: the loader's forests through the GPU BFS ordering and every
RC method, against the oracle on the same flattened input.

Fixtures: three TAU 2015 SPEF files (tests/golden/spef/) and a generated
forest written and re-read through the SPEF writer.  Bars as in
test_gpu_rct.py; in addition the root load of every net equals the sum of
its node capacitances (pin caps 0), which the files' own *D_NET totals pin
to 4 decimals (tests/test_spef.py).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O
from opentimer_amd import workloads as W
from opentimer_amd.rct import RcSolver
from opentimer_amd.spef import SpefForest, write

from test_gpu_rct import OUT, floors, oracle_inputs, rel_err

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spef")


def _check(f, ctx):
    s = RcSolver.from_edges(f, ctx)
    e, n2b, pid, pres, cap = oracle_inputs(f)
    h = s.host_topology()
    assert np.array_equal(h["node2bfs"], n2b)
    assert np.array_equal(h["pid"], pid)
    r32 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "dfs_f32")
    r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
    ref_flat = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "flat_f32")
    fl = floors(r64)
    ns = h["nodes_start"]
    capsum = np.add.reduceat(f.node_cap.reshape(-1, 4).astype(np.float64), f.nodes_start[:-1], axis=0)
    for method, tol64 in (("plain", None), ("batched", 1e-5), ("compensated", 1e-6)):
        s.solve(method)
        g = s.host()
        if method == "plain":
            for k in ("load", "delay", "ldelay", "beta", "impulse"):
                assert np.array_equal(g[k].view(np.uint32), ref_flat[k].view(np.uint32)), k
            continue
        for k in ("delay", "slew"):
            assert rel_err(g[k], r32[k], fl[k]) < 1e-5, (method, k)
        for k in OUT:
            assert rel_err(g[k], r64[k], fl[k]) < tol64, (method, k)
        root_load = g["load"].reshape(-1, 4)[ns[:-1]]
        assert rel_err(root_load, capsum, 1e-30) < (1e-6 if method == "compensated" else 1e-5)


@pytest.mark.parametrize("name", ["c17.spef", "simple.spef", "s27_spef.spef"])
def test_tau15_spef_fixtures(ctx, name):
    _check(SpefForest.read(os.path.join(GOLD, name)), ctx)


def test_written_forest(ctx, tmp_path):
    cfg = W.config1(num_nets=2000, seed=9)
    cfg.classes[2].weight = 0.03
    cfg.classes[2].max_nodes = 2000
    g = W.generate_rc(cfg)
    c4 = g.node_cap.reshape(-1, 4)
    c4[:] = c4[:, :1]
    p = str(tmp_path / "w.spef")
    write(p, g)
    _check(SpefForest.read(p), ctx)
