"""Forward propagation on the TAU 2015 c17 design against the reference's
own reported arrival times (benchmark/c17: c17.ops / c17.output, the 100
report_at values before the first design change; fixture
tests/golden/prop/c17.json, made by tests/golden/make_tau15_c17.py).

The whole chain runs on the host: the SPEF loader (otg_spef_read, host
C++), pin capacitances, the oracle's CPU-timer RC arithmetic, the oracle's
levelization and its restatement of the forward propagation.  The reference
prints 3 decimals, so every value must lie within 0.0005 ps of the printed
one (plus 2e-6 of FP32 rounding)."""
import numpy as np

import oracle as O
import tau15_case as T

TOL = 0.0005 + 2e-6


def test_c17_arrival_times_match_reference_output():
    g, lib, pins, gold = T.build(T.oracle_rc)
    fr, ends = T.levels(g)
    slew, at, _ = O.fprop(g, lib, fr, ends)
    assert at.shape == gold.shape == (25, 4)
    err = np.abs(at.astype(np.float64) - gold)
    worst = int(np.argmax(err.max(axis=1)))
    assert err.max() <= TOL, (pins[worst], at[worst], gold[worst])
    # every pin is reached, with a slew, in every corner
    assert np.isfinite(at).all() and np.isfinite(slew).all()


def test_c17_fixture_shape():
    J = T.load()
    assert len(J["instances"]) == 6 and len(J["golden_at"]) == 25
    for split in ("early", "late"):
        arcs = J["arcs"][split]
        assert [(a["from"], a["to"], a["timing_sense"]) for a in arcs] == [
            ("A1", "ZN", "negative_unate"), ("A2", "ZN", "negative_unate")]
        for a in arcs:
            for tab in a["tables"].values():
                assert len(tab["values"]) == len(tab["index_1"]) * len(tab["index_2"]) == 56
                assert np.all(np.diff(tab["index_1"]) > 0) and np.all(np.diff(tab["index_2"]) > 0)
