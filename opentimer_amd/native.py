"""ctypes binding of libotg.so (include/otg.h, include/otg_gen.h,
include/otg_spef.h).

The library is the product: every compute call goes to its HIP kernels.  If
the library is missing this module raises at import time -- there is no CPU
fallback.  Device buffers are passed as raw pointers (PyTorch tensors are used
for allocation and streams, nothing more).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libotg.so")

try:
    # libotg.so needs libamdhip64.so.7.  When PyTorch is present its bundled
    # HIP runtime must be the one loaded (one runtime per process): importing
    # torch first makes the dynamic linker bind libotg to it.
    import torch  # noqa: F401
except ImportError:  # the C ABI works without torch
    torch = None

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -m opentimer_amd.build` "
        "(or __graft_entry__.build()) first")

lib = C.CDLL(LIB_PATH)

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f32p = C.POINTER(C.c_float)
vp = C.c_void_p

OK = 0
STATUS = {
    0: "OTG_OK", 1: "OTG_ERR_INVALID_ARGUMENT", 2: "OTG_ERR_NOT_A_TREE",
    3: "OTG_ERR_BAD_TOPOLOGY", 4: "OTG_ERR_DEVICE", 5: "OTG_ERR_OUT_OF_MEMORY",
    6: "OTG_ERR_NO_DEVICE", 7: "OTG_ERR_CYCLE", 8: "OTG_ERR_UNSUPPORTED",
}

PLAIN, BATCHED, BATCHED_COMPENSATED = 0, 1, 2
SHAPE_RANDOM, SHAPE_CHAIN, SHAPE_BALANCED = 0, 1, 2


class OtgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class RctTree(C.Structure):
    _fields_ = [("num_nets", C.c_int32), ("total_num_nodes", C.c_int32),
                ("nodes_start", vp), ("pid", vp), ("pres", vp), ("cap", vp),
                ("slew_in", vp)]


class RctResult(C.Structure):
    _fields_ = [("load", vp), ("delay", vp), ("ldelay", vp), ("beta", vp),
                ("impulse", vp), ("slew", vp)]


class RctTuning(C.Structure):
    _fields_ = [("ring_slots", C.c_int32), ("min_ring", C.c_int32),
                ("small_target", C.c_int32), ("grid_per_cu", C.c_int32),
                ("streams", C.c_int32), ("small_max", C.c_int32), ("ring_mask", C.c_int32),
                ("first_on_caller", C.c_int32), ("small_grid_per_cu", C.c_int32),
                ("group_alongside", C.c_int32)]


class RctEdges(C.Structure):
    _fields_ = [("num_nets", C.c_int32), ("total_num_nodes", C.c_int32),
                ("nodes_start", vp), ("edges", vp), ("roots", vp), ("edge_res", vp),
                ("node_cap", vp)]


class RctEdgeArray(C.Structure):
    _fields_ = [("num_nets", C.c_int32), ("total_num_nodes", C.c_int32),
                ("total_num_edges", C.c_int32),
                ("rct_nodes_start", vp), ("rct_edges", vp), ("rct_roots", vp),
                ("rct_edges_res", vp), ("rct_nodes_cap", vp),
                ("rct_node2bfs_order", vp), ("rct_pid", vp),
                ("load", vp), ("delay", vp), ("ldelay", vp), ("impulse", vp),
                ("beta", vp), ("slew", vp), ("slew_in", vp)]


class RctHost(C.Structure):
    _fields_ = [("num_nets", C.c_int32), ("total_num_nodes", C.c_int32),
                ("rct_nodes_start", vp), ("rct_pid", vp), ("pres", vp), ("cap", vp),
                ("load", vp), ("delay", vp), ("ldelay", vp), ("impulse", vp),
                ("beta", vp), ("slew", vp), ("slew_in", vp)]


class IncrGraph(C.Structure):
    _fields_ = [("num_pins", C.c_int32), ("num_nets", C.c_int32),
                ("fanout_start", vp), ("fanout", vp), ("fanin_start", vp), ("fanin", vp),
                ("out_degree", vp), ("net_driver", vp), ("level", vp)]


class PropLib(C.Structure):
    _fields_ = [("num_luts", C.c_int32), ("num_timings", C.c_int32),
                ("num_lut_vals", C.c_int32), ("lut_hdr", vp),
                ("lut_val", vp), ("timing_sense", vp), ("timing_edge", vp),
                ("timing_lut", vp), ("timing_has", vp)]


class PropGraph(C.Structure):
    _fields_ = [("num_pins", C.c_int32), ("num_arcs", C.c_int32), ("fanin_start", vp),
                ("fanin", vp), ("arc_kind", vp), ("arc_timing", vp), ("arc_rc_delay", vp),
                ("arc_rc_impulse", vp), ("pin_load", vp), ("pi_slew", vp), ("pi_at", vp)]


class PropResult(C.Structure):
    _fields_ = [("slew", vp), ("at", vp), ("arc_delay", vp)]


class BpropGraph(C.Structure):
    _fields_ = [("num_pins", C.c_int32), ("num_arcs", C.c_int32), ("fanout_start", vp),
                ("fanout_arc", vp), ("arc_to", vp), ("arc_delay", vp), ("po_rat", vp),
                ("at", vp), ("fanout_to", vp)]


class TimerDesign(C.Structure):
    _fields_ = [("num_nets", C.c_int32), ("total_num_nodes", C.c_int32),
                ("nodes_start", vp), ("edges", vp), ("roots", vp), ("edge_res", vp),
                ("node_cap", vp), ("num_pins", C.c_int32), ("num_arcs", C.c_int32),
                ("fanin_start", vp), ("fanin", vp), ("arc_kind", vp), ("arc_timing", vp),
                ("arc_net", vp), ("arc_node", vp), ("pin_net", vp), ("pin_load", vp),
                ("pi_slew", vp), ("pi_at", vp), ("po_rat", vp)]


class TimerResults(C.Structure):
    _fields_ = [("slew", vp), ("at", vp), ("rat", vp), ("slack", vp), ("arc_delay", vp),
                ("arc_rc_delay", vp), ("arc_rc_impulse", vp), ("pin_load", vp),
                ("frontiers", vp), ("frontier_ends", vp), ("num_levels", C.c_int32)]


REPORT_SLEW, REPORT_AT, REPORT_RAT, REPORT_SLACK = 0, 1, 2, 3


class GenSizeClass(C.Structure):
    _fields_ = [("weight", C.c_double), ("min_nodes", C.c_int32), ("max_nodes", C.c_int32),
                ("shape", C.c_int32), ("pad", C.c_int32), ("branch_prob", C.c_double)]


class GenRcParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("num_nets", C.c_int32), ("num_classes", C.c_int32),
                ("classes", GenSizeClass * 4),
                ("r_min", C.c_double), ("r_max", C.c_double),
                ("c_min", C.c_double), ("c_max", C.c_double),
                ("pin_min", C.c_double), ("pin_max", C.c_double),
                ("corner_scale", C.c_double * 4),
                ("slew_min", C.c_double), ("slew_max", C.c_double),
                ("r_unit", C.c_double), ("relabel", C.c_int32), ("random_orient", C.c_int32),
                ("first_net", C.c_int64)]


class GenDagParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("num_pins", C.c_int32), ("num_layers", C.c_int32),
                ("num_arcs", C.c_int64), ("fanout_alpha", C.c_double),
                ("fanout_cap", C.c_int32), ("skip_min", C.c_int32), ("skip_max", C.c_int32),
                ("pad", C.c_int32), ("skip_prob", C.c_double), ("shuffle_ids", C.c_int32)]


class SpefOptions(C.Structure):
    _fields_ = [("cap_unit", C.c_double), ("res_unit", C.c_double), ("threads", C.c_int32),
                ("reserved", C.c_int32)]


class SpefForest(C.Structure):
    _fields_ = [("num_nets", C.c_int32), ("total_num_nodes", C.c_int32),
                ("total_num_edges", C.c_int32), ("num_undriven", C.c_int32),
                ("num_dropped_caps", C.c_int32), ("reserved", C.c_int32),
                ("nodes_start", vp), ("edges", vp), ("roots", vp), ("edges_res", vp),
                ("nodes_cap", vp), ("total_cap", vp), ("cap_unit", C.c_double),
                ("res_unit", C.c_double)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


ctx_p = C.c_void_p
plan_p = C.c_void_p
_sig("otg_create", C.c_int, C.c_int, C.POINTER(ctx_p))
_sig("otg_destroy", C.c_int, ctx_p)
_sig("otg_synchronize", C.c_int, ctx_p, vp)
_sig("otg_status_string", C.c_char_p, C.c_int)
_sig("otg_last_error", C.c_char_p, ctx_p)
_sig("otg_version", C.c_int)
_sig("otg_rct_plan_create", C.c_int, ctx_p, C.POINTER(RctTree), C.POINTER(RctTuning), C.c_int,
     C.POINTER(plan_p))
_sig("otg_rct_plan_destroy", C.c_int, plan_p)
_sig("otg_rct_plan_stats", C.c_int, plan_p, i64p)
_sig("otg_rct_solve", C.c_int, ctx_p, plan_p, C.POINTER(RctTree), C.POINTER(RctResult), C.c_int, vp)
_sig("otg_rct_reference_f64", C.c_int, ctx_p, C.POINTER(RctTree), vp, vp, vp, vp, vp, vp, vp)
_sig("otg_rct_stream_ceiling", C.c_int, ctx_p, C.POINTER(RctTree), C.POINTER(RctResult), vp)
_sig("otg_rct_solve_subset", C.c_int, ctx_p, plan_p, C.POINTER(RctTree), C.POINTER(RctResult),
     C.c_int, vp, C.c_int32, vp)
_sig("otg_rct_bfs", C.c_int, ctx_p, C.POINTER(RctEdges), vp, vp, vp, vp, vp, vp)
_sig("otg_rct_bfs_and_compute", C.c_int, ctx_p, C.POINTER(RctEdgeArray), C.c_int)
_sig("otg_rct_compute", C.c_int, ctx_p, C.POINTER(RctHost), C.c_int)
_sig("otg_rct_last_breakdown", C.c_int, ctx_p, C.POINTER(C.c_double))
_sig("otg_rct_query", C.c_int, ctx_p, C.c_int32, vp, vp, C.POINTER(RctResult), C.c_int32, vp, vp,
     vp, vp, vp, vp)
_sig("otg_rct_autotune", C.c_int, ctx_p, C.POINTER(RctTree), C.POINTER(RctResult), C.c_int,
     C.c_char_p, C.POINTER(RctTuning))
_sig("otg_rct_tuning_load", C.c_int, ctx_p, C.c_char_p, C.POINTER(RctTuning))
_sig("otg_toposort_compute", C.c_int, ctx_p, C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp, vp,
     i32p, vp)
_sig("otg_levelize", C.c_int, ctx_p, C.c_int32, C.c_int32, vp, vp, vp, vp, C.c_int32, vp, vp, vp,
     i32p, vp)
_sig("otg_levelize_noalloc", C.c_int, ctx_p, C.c_int32, C.c_int32, vp, vp, vp, vp, C.c_int32, vp, vp,
     vp, i32p, vp)
LEVELIZE_NOALLOC_MAX_PINS = 1 << 21
_sig("otg_incr_create", C.c_int, ctx_p, C.POINTER(IncrGraph), C.c_int32, C.POINTER(C.c_void_p))
_sig("otg_incr_destroy", C.c_int, C.c_void_p)
_sig("otg_incr_relevelize", C.c_int, C.c_void_p, vp, C.c_int32, vp)
_sig("otg_incr_query", C.c_int, C.c_void_p, vp, i32p, i32p, C.POINTER(C.c_void_p),
     C.POINTER(C.c_void_p), C.POINTER(C.c_void_p))
_sig("otg_fprop", C.c_int, ctx_p, C.POINTER(PropGraph), C.POINTER(PropLib), vp, vp, C.c_int32,
     C.POINTER(PropResult), vp)
_sig("otg_fprop_plan_create", C.c_int, ctx_p, C.c_int32, C.c_int32, vp, vp, vp, vp, C.c_int32, vp,
     C.POINTER(C.c_void_p))
_sig("otg_fprop_plan_destroy", C.c_int, C.c_void_p)
_sig("otg_fprop_plan_sizes", C.c_int, C.c_void_p, i32p, i32p)
_sig("otg_fprop_run", C.c_int, ctx_p, C.c_void_p, C.POINTER(PropGraph), C.POINTER(PropLib),
     C.POINTER(PropResult), vp)
_sig("otg_bprop", C.c_int, ctx_p, C.POINTER(BpropGraph), vp, vp, C.c_int32, vp, vp, vp)
_sig("otg_build_fanout_csr", C.c_int, C.c_int32, vp, vp, vp, vp, vp)
timer_p = C.c_void_p
_sig("otg_timer_create", C.c_int, ctx_p, C.POINTER(TimerDesign), C.POINTER(PropLib), C.c_int,
     C.POINTER(timer_p))
_sig("otg_timer_destroy", C.c_int, timer_p)
_sig("otg_timer_update", C.c_int, timer_p, vp)
_sig("otg_timer_results_get", C.c_int, timer_p, C.POINTER(TimerResults))
_sig("otg_timer_report", C.c_int, timer_p, C.c_int, C.c_int32, vp, vp, vp)
_sig("otg_graph_begin", C.c_int, ctx_p, vp)
_sig("otg_graph_end", C.c_int, ctx_p, vp, C.POINTER(C.c_void_p))
_sig("otg_graph_launch", C.c_int, ctx_p, C.c_void_p, vp)
_sig("otg_graph_destroy", C.c_int, C.c_void_p)
_sig("otg_gen_rc_sizes", C.c_int, C.POINTER(GenRcParams), vp, i64p)
_sig("otg_gen_rc_fill", C.c_int, C.POINTER(GenRcParams), vp, vp, vp, vp, vp, vp, vp, C.c_int)
_sig("otg_gen_dag", C.c_int, C.POINTER(GenDagParams), vp, vp, C.c_int)
_sig("otg_build_fanin_csr", C.c_int, C.c_int32, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp, i32p,
     i64p)
spef_p = C.c_void_p
_sig("otg_spef_read", C.c_int, C.c_char_p, C.POINTER(SpefOptions), C.POINTER(spef_p))
_sig("otg_spef_read_buffer", C.c_int, C.c_char_p, C.c_size_t, C.POINTER(SpefOptions),
     C.POINTER(spef_p))
_sig("otg_spef_last_error", C.c_char_p)
_sig("otg_spef_forest_get", C.c_int, spef_p, C.POINTER(SpefForest))
_sig("otg_spef_net_name", C.c_int32, spef_p, C.c_int32, C.c_char_p, C.c_size_t)
_sig("otg_spef_node_name", C.c_int32, spef_p, C.c_int32, C.c_int32, C.c_char_p, C.c_size_t)
_sig("otg_spef_find_net", C.c_int32, spef_p, C.c_char_p)
_sig("otg_spef_find_node", C.c_int32, spef_p, C.c_int32, C.c_char_p)
_sig("otg_spef_add_pin_cap", C.c_int, spef_p, C.c_int32, C.c_int32, f32p)
_sig("otg_spef_write", C.c_int, C.c_char_p, C.c_int32, vp, vp, vp, vp, vp)
_sig("otg_spef_destroy", C.c_int, spef_p)

# Every symbol the C headers declare; tests check the library exports each.
EXPORTED = [
    "otg_create", "otg_destroy", "otg_synchronize", "otg_status_string", "otg_last_error",
    "otg_version", "otg_rct_plan_create", "otg_rct_plan_destroy", "otg_rct_plan_stats",
    "otg_rct_solve", "otg_rct_reference_f64", "otg_rct_stream_ceiling", "otg_rct_solve_subset", "otg_rct_bfs", "otg_rct_bfs_and_compute",
    "otg_rct_compute", "otg_rct_query", "otg_rct_last_breakdown",
    "otg_rct_autotune", "otg_rct_tuning_load", "otg_toposort_compute", "otg_levelize", "otg_levelize_noalloc",
    "otg_incr_create", "otg_incr_destroy", "otg_incr_relevelize", "otg_incr_query",
    "otg_fprop", "otg_fprop_plan_create", "otg_fprop_plan_destroy", "otg_fprop_plan_sizes",
    "otg_fprop_run", "otg_bprop", "otg_build_fanout_csr", "otg_timer_create", "otg_timer_destroy",
    "otg_timer_update", "otg_timer_results_get", "otg_timer_report", "otg_graph_begin", "otg_graph_end", "otg_graph_launch", "otg_graph_destroy",
    "otg_gen_rc_sizes", "otg_gen_rc_fill", "otg_gen_dag", "otg_build_fanin_csr",
    "otg_spef_read", "otg_spef_read_buffer", "otg_spef_last_error", "otg_spef_forest_get",
    "otg_spef_net_name", "otg_spef_node_name", "otg_spef_find_net", "otg_spef_find_node",
    "otg_spef_add_pin_cap", "otg_spef_write", "otg_spef_destroy",
]


def check(status: int, ctx=None):
    if status != OK:
        msg = lib.otg_last_error(ctx)
        raise OtgError(status, (msg or b"").decode(errors="replace"))


def addr(x):
    """Raw address of a numpy array or torch tensor (None -> NULL)."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data


class Context:
    """Owns an otg_context (device, stream)."""

    def __init__(self, device: int = 0):
        self.ptr = ctx_p()
        check(lib.otg_create(device, C.byref(self.ptr)))
        self.device = device

    def close(self):
        if self.ptr:
            lib.otg_destroy(self.ptr)
            self.ptr = ctx_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, status):
        check(status, self.ptr)

    def sync(self, stream=None):
        self.check(lib.otg_synchronize(self.ptr, stream))
