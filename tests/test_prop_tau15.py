"""Forward propagation on TAU 2015 designs against the reference's own
reported arrival times (benchmark/<design>: the report_at values of the
leading reports of <design>.ops, before the first design change, with the
matching <design>.output lines; fixtures tests/golden/prop/<design>.json made
by tests/golden/make_tau15_prop.py, parasitics tests/golden/spef/).

c17: 6 NAND2 gates, 100 values; s27_spef: 28 values of a sequential design
(flip-flops clocked through edge-triggered CK -> Q arcs); s400: 221 cells of
25 types, 2504 values.

The whole chain runs on the host: the SPEF loader (otg_spef_read, host
C++), pin capacitances, the oracle's CPU-timer RC arithmetic, the oracle's
levelization and its restatement of the forward propagation.  The reference
prints 3 decimals: every value lies within 0.0005 ps of the printed one,
plus 1e-6 of its magnitude for FP32 sums along paths of ~30 stages (the
largest excess measured is 3.6e-7 of the value)."""
import numpy as np
import pytest

import oracle as O
import tau15_case as T


def tolerance(gold):
    return 0.0005 + 1e-6 * np.abs(gold)


@pytest.mark.parametrize("design", T.DESIGNS)
def test_arrival_times_match_reference_output(design):
    g, lib, pins, gold = T.build(T.oracle_rc, design)
    fr, ends = T.levels(g)
    slew, at, _ = O.fprop(g, lib, fr, ends)
    m = ~np.isnan(gold)
    assert m.sum() == 4 * len(T.load(design)["golden_at"])
    err = np.where(m, np.abs(at.astype(np.float64) - gold), 0.0)
    bad = err > np.where(m, tolerance(gold), 0.0)
    worst = int(np.argmax(err.max(axis=1)))
    assert not bad.any(), (int(bad.sum()), pins[worst], at[worst], gold[worst])
    # every reported pin is reached, with a slew, in every corner
    assert np.isfinite(at[m]).all() and np.isfinite(slew[m]).all()


@pytest.mark.parametrize("design", T.DESIGNS)
def test_fixture_shape(design):
    J = T.load(design)
    for split in ("early", "late"):
        for cell in J["cells"][split].values():
            for a in cell["arcs"]:
                assert a["from"] in cell["pins"] and a["to"] in cell["pins"]
                for tab in a["tables"].values():
                    n1, n2 = max(len(tab["index_1"]), 1), max(len(tab["index_2"]), 1)
                    assert len(tab["values"]) == n1 * n2
                    assert np.all(np.diff(tab["index_1"]) > 0)
                    assert np.all(np.diff(tab["index_2"]) > 0)
    for inst in J["instances"]:
        assert inst["cell"] in J["cells"]["early"]


def test_c17_required_times_and_slacks_match_reference_output():
    # backward propagation from the primary outputs' required times (c17 has
    # no timing checks, so every required time comes from a primary output)
    g, lib, pins, gold = T.build(T.oracle_rc, "c17")
    fr, ends = T.levels(g)
    slew, at, ad = O.fprop(g, lib, fr, ends)
    po_rat, gr, gs = T.required("c17", pins)
    rat, slack = O.bprop(g["fanin_start"], g["fanin"], ad, po_rat, at, fr, ends)
    for got, want in ((rat, gr), (slack, gs)):
        assert (~np.isnan(want)).sum() == 100
        err = np.abs(got.astype(np.float64) - want)
        assert (err <= tolerance(want)).all(), float(np.nanmax(err))
