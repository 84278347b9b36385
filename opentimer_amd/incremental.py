"""Device-resident incremental session (host mirror of otg_incr_* and
otg_graph_*): RC nets solved by an RcSolver and a timing graph whose pins
are re-levelized in the fan-out cone of changed nets.

One update = `otg_rct_solve_subset` on the changed nets (the net-timing
update of FlatRctStorage::_update_timing_cuda, ot/timer/net.cpp:239-276,
restricted to them) followed by `otg_incr_relevelize` on their driver pins
(the candidate cone of Timer::_build_prop_cands, ot/timer/timer.cpp:845-923,
levelized as in Timer::_build_prop_tasks_cuda, timer.cpp:988-1071).  Both
are fixed kernel sequences, so an update can be captured once per batch size
in a hipGraph and replayed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import native as N


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(device)


class Incremental:
    def __init__(self, solver, fanin_csr, fanout_csr, net_driver, level, max_levels: int):
        """fanin_csr / fanout_csr: workloads.FaninCsr of the graph and of its
        reverse; net_driver[num_nets]: driver pin of each RC net; level: the
        pins' current levels (from a full levelization); max_levels: the
        graph's number of levels (bound on any cone's depth)."""
        self.solver = solver
        self.ctx = solver.ctx
        d = solver.device
        self.device = d
        self.n = fanin_csr.n
        self.fi_start = _t(fanin_csr.start, d)
        self.fi = _t(fanin_csr.fanin if len(fanin_csr.fanin) else np.zeros(1), d)
        self.fo_start = _t(fanout_csr.start, d)
        self.fo = _t(fanout_csr.fanin if len(fanout_csr.fanin) else np.zeros(1), d)
        self.out_degree = _t(fanin_csr.out_degree, d)
        self.net_driver = _t(net_driver, d)
        self.level = _t(level, d)
        self.max_levels = int(max_levels)
        g = N.IncrGraph(self.n, int(len(net_driver)), self.fo_start.data_ptr(), self.fo.data_ptr(),
                        self.fi_start.data_ptr(), self.fi.data_ptr(), self.out_degree.data_ptr(),
                        self.net_driver.data_ptr(), self.level.data_ptr())
        self._g = g
        self.ptr = C.c_void_p()
        self.ctx.check(N.lib.otg_incr_create(self.ctx.ptr, C.byref(g), self.max_levels,
                                             C.byref(self.ptr)))

    def step(self, net_ids: torch.Tensor, method="compensated", stream=None):
        """Asynchronous update for the nets in net_ids (device int32)."""
        stream = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        self.solver.solve_subset(net_ids, method, stream=stream)
        self.ctx.check(N.lib.otg_incr_relevelize(self.ptr, net_ids.data_ptr(),
                                                 int(net_ids.numel()), stream))

    def capture(self, net_ids: torch.Tensor, method, stream: torch.cuda.Stream):
        """hipGraph of one update on `stream` (a non-default stream); the
        graph reads net_ids at replay time."""
        self.solver.plan(method)
        self.solver.alloc_outputs()
        torch.cuda.synchronize(self.device)
        h = stream.cuda_stream
        self.ctx.check(N.lib.otg_graph_begin(self.ctx.ptr, h))
        try:
            self.step(net_ids, method, stream=h)
        finally:
            g = C.c_void_p()
            st = N.lib.otg_graph_end(self.ctx.ptr, h, C.byref(g))
        self.ctx.check(st)
        return Graph(self.ctx, g)

    def query(self, stream=None):
        """(cone pins, cone pins by level, level ends) of the last update, host."""
        stream = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        size, nl = C.c_int32(), C.c_int32()
        cone, fr, ends = C.c_void_p(), C.c_void_p(), C.c_void_p()
        self.ctx.check(N.lib.otg_incr_query(self.ptr, stream, C.byref(size), C.byref(nl),
                                            C.byref(cone), C.byref(fr), C.byref(ends)))
        return (_download(cone.value, size.value), _download(fr.value, size.value),
                _download(ends.value, nl.value + 1))

    def close(self):
        if self.ptr:
            N.lib.otg_incr_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    def launch(self, stream: torch.cuda.Stream):
        self.ctx.check(N.lib.otg_graph_launch(self.ctx.ptr, self.h, stream.cuda_stream))

    def close(self):
        if self.h:
            N.lib.otg_graph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DevArray:
    """A raw device int32 array seen by torch through __cuda_array_interface__."""

    def __init__(self, ptr, count):
        self.__cuda_array_interface__ = {"shape": (count,), "typestr": "<i4",
                                         "data": (ptr, False), "version": 3}


def _download(ptr, count):
    if count <= 0 or not ptr:
        return np.empty(0, np.int32)
    return torch.as_tensor(_DevArray(ptr, count), device="cuda").cpu().numpy().copy()
