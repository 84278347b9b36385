"""Device-resident forward propagation (host mirror of otg_fprop).

Mirrors the forward tasks of the reference's level-wise flow
(Timer::_build_prop_tasks_cuda, ot/timer/timer.cpp:1079-1086): for every
level from the sources to the sinks, Timer::_fprop_slew / _fprop_delay /
_fprop_at (timer.cpp:733-783) on each pin of the level.  The graph, the
library and the results stay in device memory; the level order comes from
the levelizer's device arrays (opentimer_amd.levelize.Levelizer).

`run` propagates in pin order (otg_fprop); `build_plan` + `run_plan`
renumber the levels' pins and arcs once into level order (otg_fprop_plan_*)
and then propagate over contiguous data, with the same results.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import native as N


class Propagator:
    def __init__(self, graph, lib, ctx: N.Context | None = None, device: int = 0):
        self.ctx = ctx or N.Context(device)
        d = torch.device("cuda", self.ctx.device)
        self.device = d
        self.n, self.m = graph.n, graph.m
        ti = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32).reshape(-1)
                                        if np.size(a) else np.zeros(1, np.int32)).to(d)
        tf = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32).reshape(-1)
                                        if np.size(a) else np.zeros(1, np.float32)).to(d)
        self._t = {
            "fanin_start": ti(graph.fanin_start), "fanin": ti(graph.fanin),
            "arc_kind": ti(graph.arc_kind), "arc_timing": ti(graph.arc_timing),
            "arc_rc_delay": tf(graph.arc_rc_delay), "arc_rc_impulse": tf(graph.arc_rc_impulse),
            "pin_load": tf(graph.pin_load), "pi_slew": tf(graph.pi_slew), "pi_at": tf(graph.pi_at),
            "lut_hdr": ti(lib.lut_hdr), "lut_val": tf(lib.lut_val),
            "timing_sense": ti(lib.timing_sense), "timing_edge": ti(lib.timing_edge),
            "timing_lut": ti(lib.timing_lut), "timing_has": ti(lib.timing_has),
        }
        t = self._t
        self.graph = N.PropGraph(self.n, self.m, *(t[k].data_ptr() for k in (
            "fanin_start", "fanin", "arc_kind", "arc_timing", "arc_rc_delay", "arc_rc_impulse",
            "pin_load", "pi_slew", "pi_at")))
        self.lib = N.PropLib(lib.num_luts, lib.num_timings, int(np.size(lib.lut_val)),
                             *(t[k].data_ptr() for k in (
                "lut_hdr", "lut_val", "timing_sense", "timing_edge", "timing_lut",
                "timing_has")))
        nan = float("nan")
        self.slew = torch.full((max(self.n, 1), 4), nan, dtype=torch.float32, device=d)
        self.at = torch.full((max(self.n, 1), 4), nan, dtype=torch.float32, device=d)
        self.arc_delay = torch.full((max(self.m, 1), 8), nan, dtype=torch.float32, device=d)

    def run(self, frontiers, ends, num_levels: int, stream=None, arc_delay: bool = True):
        """frontiers / ends: device int32 tensors from the levelizer (ends
        holds num_levels + 1 entries)."""
        stream = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        out = N.PropResult(self.slew.data_ptr(), self.at.data_ptr(),
                           self.arc_delay.data_ptr() if arc_delay else None)
        self.ctx.check(N.lib.otg_fprop(self.ctx.ptr, C.byref(self.graph), C.byref(self.lib),
                                       frontiers.data_ptr(), ends.data_ptr(), int(num_levels),
                                       C.byref(out), stream))

    def build_plan(self, frontiers, ends, num_levels: int, stream=None):
        stream = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        self.close_plan()
        h = C.c_void_p()
        self.ctx.check(N.lib.otg_fprop_plan_create(
            self.ctx.ptr, self.n, self.m, self._t["fanin_start"].data_ptr(),
            self._t["fanin"].data_ptr(), frontiers.data_ptr(), ends.data_ptr(), int(num_levels),
            stream, C.byref(h)))
        self._plan = h
        pos, slots = C.c_int32(), C.c_int32()
        self.ctx.check(N.lib.otg_fprop_plan_sizes(h, C.byref(pos), C.byref(slots)))
        return pos.value, slots.value

    def run_plan(self, stream=None, arc_delay: bool = True):
        stream = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        out = N.PropResult(self.slew.data_ptr(), self.at.data_ptr(),
                           self.arc_delay.data_ptr() if arc_delay else None)
        self.ctx.check(N.lib.otg_fprop_run(self.ctx.ptr, self._plan, C.byref(self.graph),
                                           C.byref(self.lib), C.byref(out), stream))

    def close_plan(self):
        if getattr(self, "_plan", None):
            torch.cuda.synchronize(self.device)
            N.lib.otg_fprop_plan_destroy(self._plan)
        self._plan = None

    def __del__(self):
        try:
            self.close_plan()
        except Exception:
            pass

    def host(self):
        torch.cuda.synchronize(self.device)
        return (self.slew[: self.n].cpu().numpy(), self.at[: self.n].cpu().numpy(),
                self.arc_delay[: self.m].cpu().numpy())

    def required(self, po_rat, frontiers, ends, num_levels: int, stream=None):
        """Required times and slacks over the same levels (otg_bprop), from
        the primary-output required times po_rat [n,4] (NaN elsewhere) and
        this propagator's arrival times and arc delays.  Returns device
        tensors (rat, slack), [n,4] each."""
        stream = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        if getattr(self, "_fanout", None) is None:
            fs = self._t["fanin_start"].cpu().numpy()
            fi = self._t["fanin"].cpu().numpy()
            fo_start = np.zeros(self.n + 1, np.int32)
            fo_arc = np.zeros(max(self.m, 1), np.int32)
            arc_to = np.zeros(max(self.m, 1), np.int32)
            N.check(N.lib.otg_build_fanout_csr(self.n, N.addr(fs), N.addr(fi), N.addr(fo_start),
                                               N.addr(fo_arc), N.addr(arc_to)))
            fo_to = arc_to[fo_arc[: self.m]] if self.m else np.zeros(1, np.int32)
            self._fanout = tuple(torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(self.device)
                                 for a in (fo_start, fo_arc, arc_to, fo_to))
        pr = torch.from_numpy(np.ascontiguousarray(po_rat, np.float32).reshape(-1, 4)).to(self.device)
        rat = torch.full((max(self.n, 1), 4), float("nan"), dtype=torch.float32, device=self.device)
        slack = torch.full_like(rat, float("nan"))
        fo_start, fo_arc, arc_to, fo_to = self._fanout
        g = N.BpropGraph(self.n, self.m, fo_start.data_ptr(), fo_arc.data_ptr(),
                         arc_to.data_ptr(), self.arc_delay.data_ptr(), pr.data_ptr(),
                         self.at.data_ptr(), fo_to.data_ptr())
        self._bp = (g, pr, rat, slack)
        self.rerun_required(frontiers, ends, num_levels, stream)
        return rat[: self.n], slack[: self.n]

    def rerun_required(self, frontiers, ends, num_levels: int, stream=None):
        """otg_bprop again over the buffers of the last `required` call (for
        timing and after a new forward pass)."""
        stream = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        g, _, rat, slack = self._bp
        self.ctx.check(N.lib.otg_bprop(self.ctx.ptr, C.byref(g), frontiers.data_ptr(),
                                       ends.data_ptr(), int(num_levels), rat.data_ptr(),
                                       slack.data_ptr(), stream))
