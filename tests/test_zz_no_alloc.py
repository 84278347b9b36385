"""otg_levelize_noalloc (no device allocation) against otg_levelize through
the C ABI: the same frontiers (as sets per level), level ends and levels on
two calls in a row, the second on a smaller graph (the fixed array keeps the
first call's contents, so whatever must start at zero has to be cleared by
the call itself); sizes above the array's maximum are rejected."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from opentimer_amd import native as N
from opentimer_amd import workloads as W


def _run(ctx, fn, csr):
    d = torch.device("cuda", ctx.device)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(d)
    n, m = csr.n, len(csr.fanin)
    start = t(csr.start)
    fanin = t(csr.fanin if m else np.zeros(1, np.int32))
    outd = t(csr.out_degree)
    f0 = t(csr.frontier if len(csr.frontier) else np.zeros(1, np.int32))
    fr = torch.full((max(n, 1),), -7, dtype=torch.int32, device=d)
    ends = torch.full((n + 2,), -7, dtype=torch.int32, device=d)
    lv = torch.full((max(n, 1),), -7, dtype=torch.int32, device=d)
    ne = C.c_int32()
    torch.cuda.synchronize(d)
    ctx.check(fn(ctx.ptr, n, m, start.data_ptr(), fanin.data_ptr(), outd.data_ptr(),
                 f0.data_ptr(), len(csr.frontier), fr.data_ptr(), ends.data_ptr(),
                 lv.data_ptr(), C.byref(ne), torch.cuda.current_stream(d).cuda_stream))
    torch.cuda.synchronize(d)
    # the input out-degree array is not modified
    assert np.array_equal(outd.cpu().numpy(), csr.out_degree)
    return fr[:n].cpu().numpy(), ends[: ne.value].cpu().numpy(), lv[:n].cpu().numpy()


def _same(a, b):
    fa, ea, la = a
    fb, eb, lb = b
    assert np.array_equal(ea, eb)
    assert np.array_equal(la, lb)
    for k in range(len(ea) - 1):
        assert np.array_equal(np.sort(fa[ea[k]:ea[k + 1]]), np.sort(fb[eb[k]:eb[k + 1]])), k


def test_noalloc_matches_levelize_on_two_calls(ctx):
    big = W.fanin_csr(W.generate_dag(W.DagConfig("noalloc-300k", 300_000, 750_000, 80, seed=21)))
    small = W.fanin_csr(W.generate_dag(W.config0_dag(seed=22)))
    assert small.n < big.n
    for csr in (big, small):
        want = _run(ctx, N.lib.otg_levelize, csr)
        got = _run(ctx, N.lib.otg_levelize_noalloc, csr)
        _same(got, want)
        assert int(got[2].max()) == len(got[1]) - 2


def test_noalloc_rejects_oversized_graph(ctx):
    n = N.LEVELIZE_NOALLOC_MAX_PINS + 1
    ne = C.c_int32()
    # rejected before any device array is touched
    st = N.lib.otg_levelize_noalloc(ctx.ptr, n, 0, 1, None, 1, None, 0, 1, 1, None, C.byref(ne),
                                    None)
    assert N.STATUS[st] == "OTG_ERR_INVALID_ARGUMENT"
    assert b"exceed" in N.lib.otg_last_error(ctx.ptr)


def test_noalloc_empty_graph(ctx):
    d = torch.device("cuda", ctx.device)
    ends = torch.full((2,), -7, dtype=torch.int32, device=d)
    ne = C.c_int32()
    ctx.check(N.lib.otg_levelize_noalloc(ctx.ptr, 0, 0, None, None, None, None, 0, None,
                                         ends.data_ptr(), None, C.byref(ne), None))
    assert ne.value == 2 and list(ends.cpu().numpy()) == [0, 0]
