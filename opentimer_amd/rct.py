"""Device-resident RC forests and the P1 solve (host mirror of the C ABI).

`RcSolver` keeps a forest on one GPU in the reference's BFS-ordered flat
layout (FlatRctStorage, ot/timer/net.hpp:149-166): nodes_start, pid, pres,
cap[4V], plus the driver slew per net.  `from_edges` runs the per-net BFS
ordering on the GPU (otg_rct_bfs, the replacement of compute_net_bfs +
prepare_res_cap, ot/cuda/rct.cu:120-359).  `solve` runs one of the three
methods and leaves load/delay/ldelay/beta/impulse/slew on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import native as N

OUTPUTS = ("load", "delay", "ldelay", "beta", "impulse", "slew")
METHODS = {"plain": N.PLAIN, "batched": N.BATCHED, "compensated": N.BATCHED_COMPENSATED}


def _dev(a, device, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)


def torch_stream(device):
    """hipStream_t of torch's current stream: kernels are ordered with torch's
    own copies and timed by torch events on that stream."""
    return torch.cuda.current_stream(device).cuda_stream


class RcSolver:
    def __init__(self, ctx: N.Context | None = None, device: int = 0):
        self.ctx = ctx or N.Context(device)
        self.device = torch.device("cuda", self.ctx.device)
        self.plans = {}
        self.out = {}

    # ---- construction -------------------------------------------------------
    @classmethod
    def from_edges(cls, forest, ctx=None, device=0):
        """Upload a forest in RctEdgeArray form and order it on the GPU."""
        self = cls(ctx, device)
        d = self.device
        self.num_nets = forest.num_nets
        self.num_nodes = forest.num_nodes
        self.nodes_start = _dev(forest.nodes_start, d, torch.int32)
        edges = _dev(forest.edges, d, torch.int32)
        roots = _dev(forest.roots, d, torch.int32)
        res = _dev(forest.edge_res, d, torch.float32)
        ncap = _dev(forest.node_cap, d, torch.float32)
        self.slew_in = _dev(forest.slew_in, d, torch.float32)
        V = self.num_nodes
        self.node2bfs = torch.empty(max(V, 1), dtype=torch.int32, device=d)
        self.pid = torch.empty(max(V, 1), dtype=torch.int32, device=d)
        self.pres = torch.empty(max(V, 1), dtype=torch.float32, device=d)
        self.cap = torch.empty(max(4 * V, 1), dtype=torch.float32, device=d)
        self.edges_directed = edges  # oriented in place
        self.bfs(edges, roots, res, ncap)
        return self

    @classmethod
    def from_bfs_arrays(cls, nodes_start, pid, pres, cap, slew_in, ctx=None, device=0):
        """Upload a forest that is already in BFS (topological) order."""
        self = cls(ctx, device)
        d = self.device
        self.num_nets = len(nodes_start) - 1
        self.num_nodes = int(nodes_start[-1])
        self.nodes_start = _dev(nodes_start, d, torch.int32)
        self.pid = _dev(pid, d, torch.int32)
        self.pres = _dev(pres, d, torch.float32)
        self.cap = _dev(cap, d, torch.float32)
        self.slew_in = _dev(slew_in, d, torch.float32)
        return self

    def bfs(self, edges, roots, res, ncap, stream=None):
        stream = torch_stream(self.device) if stream is None else stream
        e = N.RctEdges(self.num_nets, self.num_nodes, self.nodes_start.data_ptr(), edges.data_ptr(),
                       roots.data_ptr(), res.data_ptr(), ncap.data_ptr())
        self.ctx.check(N.lib.otg_rct_bfs(self.ctx.ptr, C.byref(e), self.node2bfs.data_ptr(),
                                         self.pid.data_ptr(), edges.data_ptr(),
                                         self.pres.data_ptr(), self.cap.data_ptr(), stream))

    # ---- P1 -----------------------------------------------------------------
    def tree(self) -> N.RctTree:
        return N.RctTree(self.num_nets, self.num_nodes, self.nodes_start.data_ptr(),
                         self.pid.data_ptr(), self.pres.data_ptr(), self.cap.data_ptr(),
                         self.slew_in.data_ptr() if self.slew_in is not None else None)

    def alloc_outputs(self, names=OUTPUTS):
        V = self.num_nodes
        for k in names:
            if k not in self.out:
                self.out[k] = torch.empty(max(4 * V, 1), dtype=torch.float32, device=self.device)
        return self.out

    def result(self, names=OUTPUTS) -> N.RctResult:
        self.alloc_outputs(names)
        return N.RctResult(*[self.out[k].data_ptr() if k in names else None for k in OUTPUTS])

    def plan(self, method="batched", tuning=None):
        m = METHODS[method]
        key = (m, None if tuning is None else tuple(tuning))
        if key not in self.plans:
            torch.cuda.synchronize(self.device)  # the plan reads nodes_start/pid on its own stream
            p = N.plan_p()
            t = None if tuning is None else C.byref(N.RctTuning(*tuning))
            self.ctx.check(N.lib.otg_rct_plan_create(self.ctx.ptr, C.byref(self.tree()), t, m,
                                                     C.byref(p)))
            self.plans[key] = p
        return self.plans[key]

    def plan_stats(self, method="batched", tuning=None):
        s = (C.c_int64 * 8)()
        self.ctx.check(N.lib.otg_rct_plan_stats(self.plan(method, tuning), s))
        keys = ["packed_groups", "packed_nets", "stream_nets", "global_nets", "max_net",
                "items", "scratch_floats", "ring_slots"]
        return dict(zip(keys, list(s)))

    def solve(self, method="batched", stream=None, names=OUTPUTS, tuning=None):
        m = METHODS[method]
        plan = None if m == N.PLAIN else self.plan(method, tuning)
        stream = torch_stream(self.device) if stream is None else stream
        res = self.result(names)
        self.ctx.check(N.lib.otg_rct_solve(self.ctx.ptr, plan, C.byref(self.tree()), C.byref(res),
                                           m, stream))
        return self.out

    def stream_ceiling(self, stream=None):
        """otg_rct_stream_ceiling over this forest and its output arrays: the
        algorithmic bytes of one solve moved with no tree work (overwrites the
        outputs)."""
        stream = torch_stream(self.device) if stream is None else stream
        self.ctx.check(N.lib.otg_rct_stream_ceiling(self.ctx.ptr, C.byref(self.tree()),
                                                    C.byref(self.result()), stream))
        return self.out

    def reference_f64(self, stream=None) -> dict:
        """FP64 reference values of every output (otg_rct_reference_f64): the
        reference's sweeps (ot/timer/net.cpp:101-193) in double on the device,
        for precision reports on workloads too large for a host check."""
        V = self.num_nodes
        ref = {k: torch.empty(max(4 * V, 1), dtype=torch.float64, device=self.device)
               for k in OUTPUTS}
        stream = torch_stream(self.device) if stream is None else stream
        self.ctx.check(N.lib.otg_rct_reference_f64(
            self.ctx.ptr, C.byref(self.tree()), *[ref[k].data_ptr() for k in OUTPUTS], stream))
        return ref

    def solve_subset(self, net_ids: torch.Tensor, method="batched", stream=None, names=OUTPUTS,
                     tuning=None):
        m = METHODS[method]
        plan = self.plan(method, tuning)
        stream = torch_stream(self.device) if stream is None else stream
        res = self.result(names)
        self.ctx.check(N.lib.otg_rct_solve_subset(self.ctx.ptr, plan,
                                                  C.byref(self.tree()), C.byref(res), m,
                                                  net_ids.data_ptr(), int(net_ids.numel()),
                                                  stream))

    def sync(self):
        self.ctx.sync()
        torch.cuda.synchronize(self.device)

    def host(self, names=OUTPUTS):
        self.sync()
        V = self.num_nodes
        return {k: self.out[k][: 4 * V].cpu().numpy() for k in names if k in self.out}

    def host_topology(self):
        self.sync()
        V = self.num_nodes
        return dict(nodes_start=self.nodes_start.cpu().numpy(), pid=self.pid[:V].cpu().numpy(),
                    pres=self.pres[:V].cpu().numpy(), cap=self.cap[: 4 * V].cpu().numpy(),
                    slew_in=self.slew_in.cpu().numpy(),
                    node2bfs=self.node2bfs[:V].cpu().numpy() if hasattr(self, "node2bfs") else None)

    def query(self, net_ids: torch.Tensor, local_ids: torch.Tensor, stream=None):
        """delay, slew and net load [count, 4] at pins (net, original local id)
        from the last solve (otg_rct_query; FlatRct::delay/slew and
        Net::_load, ot/timer/net.cpp:280-287,626-642)."""
        stream = torch_stream(self.device) if stream is None else stream
        n = int(net_ids.numel())
        out = {k: torch.empty(max(4 * n, 1), dtype=torch.float32, device=self.device)
               for k in ("delay", "slew", "load")}
        res = self.result()
        self.ctx.check(N.lib.otg_rct_query(self.ctx.ptr, self.num_nets, self.nodes_start.data_ptr(),
                                           self.node2bfs.data_ptr(), C.byref(res), n,
                                           net_ids.data_ptr(), local_ids.data_ptr(),
                                           out["delay"].data_ptr(), out["slew"].data_ptr(),
                                           out["load"].data_ptr(), stream))
        return {k: v[: 4 * n].view(n, 4) for k, v in out.items()}

    def close(self):
        for p in self.plans.values():
            N.lib.otg_rct_plan_destroy(p)
        self.plans = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
