"""Device-resident sink-first levelization (host mirror of otg_levelize).

Mirrors the inputs of toposort_compute_cuda (ot/cuda/toposort.cuh:3-6): a
fanin CSR, the fanout count of every pin and the initial frontier of pins
without fanout; returns the pins grouped by level, the level ends and the
level of every pin.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import native as N


class Levelizer:
    def __init__(self, csr, ctx: N.Context | None = None, device: int = 0):
        self.ctx = ctx or N.Context(device)
        d = torch.device("cuda", self.ctx.device)
        self.device = d
        self.n = csr.n
        self.m = len(csr.fanin)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(d)
        self.start = t(csr.start)
        self.fanin = t(csr.fanin if self.m else np.zeros(1, np.int32))
        self.out_degree = t(csr.out_degree if self.n else np.zeros(1, np.int32))
        self.frontier0 = t(csr.frontier if len(csr.frontier) else np.zeros(1, np.int32))
        self.first_size = len(csr.frontier)
        self.frontiers = torch.empty(max(self.n, 1), dtype=torch.int32, device=d)
        self.ends = torch.empty(self.n + 2, dtype=torch.int32, device=d)
        self.level = torch.empty(max(self.n, 1), dtype=torch.int32, device=d)
        self.num_ends = 0

    def run(self, stream=None):
        stream = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        ne = C.c_int32()
        self.ctx.check(N.lib.otg_levelize(
            self.ctx.ptr, self.n, self.m, self.start.data_ptr(), self.fanin.data_ptr(),
            self.out_degree.data_ptr(), self.frontier0.data_ptr(), self.first_size,
            self.frontiers.data_ptr(), self.ends.data_ptr(), self.level.data_ptr(), C.byref(ne),
            stream))
        self.num_ends = ne.value
        return self.num_ends

    def host(self):
        torch.cuda.synchronize(self.device)
        return (self.frontiers[: self.n].cpu().numpy(), self.ends[: self.num_ends].cpu().numpy(),
                self.level[: self.n].cpu().numpy())


def toposort_compute(ctx: N.Context, csr):
    """Host-pointer drop-in (otg_toposort_compute), like toposort_compute_cuda."""
    torch.cuda.synchronize()
    n = csr.n
    fr = np.zeros(max(n, 1), np.int32)
    fr[: len(csr.frontier)] = csr.frontier
    ends = np.zeros(n + 2, np.int32)
    lv = np.zeros(max(n, 1), np.int32)
    ne = C.c_int32()
    fanin = csr.fanin if len(csr.fanin) else np.zeros(1, np.int32)
    ctx.check(N.lib.otg_toposort_compute(ctx.ptr, n, len(csr.fanin), len(csr.frontier),
                                         N.addr(csr.start), N.addr(fanin),
                                         N.addr(np.ascontiguousarray(csr.out_degree)),
                                         N.addr(fr), N.addr(ends), C.byref(ne), N.addr(lv)))
    return fr[:n], ends[: ne.value], lv[:n]
