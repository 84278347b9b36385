"""The autotuner (otg_rct_autotune): a coordinate-descent search over the
batched launch's knobs on a cfg1-distribution forest.  The tuning it writes
loads back (otg_rct_tuning_load) into a fresh context and gives outputs
bit-identical to the autotuned context's, within 1e-6 of FP64."""
import ctypes as C
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O
from opentimer_amd import native as N
from opentimer_amd import workloads as W
from opentimer_amd.rct import RcSolver, OUTPUTS


def test_autotune_writes_a_tuning_that_loads_back(ctx, tmp_path):
    f = W.generate_rc(W.config1(num_nets=20_000, seed=61))
    s = RcSolver.from_edges(f, ctx)
    s.alloc_outputs()
    path = str(tmp_path / "tuning.json")
    best = N.RctTuning()
    ctx.check(N.lib.otg_rct_autotune(ctx.ptr, C.byref(s.tree()), C.byref(s.result()),
                                     N.BATCHED_COMPENSATED, path.encode(), C.byref(best)))
    rec = json.load(open(path))
    assert rec["candidates"] >= 10 and rec["ms"] > 0
    for k in ("ring_slots", "min_ring", "small_target", "small_max", "ring_mask", "streams"):
        assert rec[k] == getattr(best, k), k
    assert rec["min_ring"] in (256, 320, 512) and 2 <= rec["small_max"] <= 256
    # the autotuned context's solve (its tuning is the context default now)
    s.solve("compensated")
    tuned = {k: v.copy() for k, v in s.host().items()}
    # a fresh context that loads the file
    ctx2 = N.Context(0)
    try:
        loaded = N.RctTuning()
        ctx2.check(N.lib.otg_rct_tuning_load(ctx2.ptr, path.encode(), C.byref(loaded)))
        for k, _ in N.RctTuning._fields_:
            assert getattr(loaded, k) == getattr(best, k), k
        s2 = RcSolver.from_edges(f, ctx2)
        s2.solve("compensated")
        again = s2.host()
        for k in OUTPUTS:
            assert np.array_equal(tuned[k].view(np.uint32), again[k].view(np.uint32)), k
        h = s2.host_topology()
        r64 = O.rct_all(h["nodes_start"], h["pid"], h["pres"], h["cap"], h["slew_in"], "f64")
        for k in OUTPUTS:
            ref = r64[k]
            fl = max(float(np.max(np.abs(r64["beta" if k == "impulse" else k])))
                     * (1e-6 if k == "impulse" else 1e-7), 1e-30)
            err = float(np.max(np.abs(again[k] - ref) / np.maximum(np.abs(ref), fl)))
            assert err < 1e-6, (k, err)
        s2.close()
    finally:
        ctx2.close()
