"""SPEF parasitics to an RC forest (host mirror of include/otg_spef.h).

`SpefForest.read(path)` parses and flattens a SPEF file with the library's
multi-threaded C++ loader (the replacement of spef::Spef::read and
Net::_make_flat_rct, ot/parser-spef/parser-spef.hpp and
ot/timer/net.cpp:347-461).  The forest has the fields of
`workloads.RcForest`, so `RcSolver.from_edges(forest)` orders it on the GPU
(otg_rct_bfs) and `solve` computes every net's moments and slews.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import native as N


class SpefError(RuntimeError):
    pass


def _view(ptr, count, ctype, dtype):
    if count <= 0 or not ptr:
        return np.empty(0, dtype)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).view(dtype)


class SpefForest:
    """A flattened SPEF.  Array fields are views of the loader's memory
    (kept alive by this object); node_cap may be changed through
    add_pin_cap."""

    def __init__(self, handle):
        self._h = handle
        f = N.SpefForest()
        N.check(N.lib.otg_spef_forest_get(self._h, C.byref(f)))
        self.info = f
        n, v, e = f.num_nets, f.total_num_nodes, f.total_num_edges
        self.nodes_start = _view(f.nodes_start, n + 1, C.c_int32, np.int32)
        self.edges = _view(f.edges, 2 * e, C.c_int32, np.int32)
        self.roots = _view(f.roots, n, C.c_int32, np.int32)
        self.edge_res = _view(f.edges_res, e, C.c_float, np.float32)
        self.node_cap = _view(f.nodes_cap, 4 * v, C.c_float, np.float32)
        self.total_cap = _view(f.total_cap, n, C.c_float, np.float32)
        self.slew_in = np.zeros(4 * n, np.float32)
        self.cap_unit = f.cap_unit
        self.res_unit = f.res_unit
        self.num_undriven = f.num_undriven
        self.num_dropped_caps = f.num_dropped_caps

    @classmethod
    def read(cls, path, cap_unit: float = 0.0, res_unit: float = 0.0, threads: int = 0):
        opt = N.SpefOptions(cap_unit, res_unit, threads, 0)
        h = N.spef_p()
        st = N.lib.otg_spef_read(os.fsencode(path), C.byref(opt), C.byref(h))
        if st != N.OK:
            raise SpefError(N.lib.otg_spef_last_error().decode(errors="replace"))
        return cls(h)

    @classmethod
    def parse(cls, text: str | bytes, cap_unit: float = 0.0, res_unit: float = 0.0,
              threads: int = 0):
        b = text.encode() if isinstance(text, str) else text
        opt = N.SpefOptions(cap_unit, res_unit, threads, 0)
        h = N.spef_p()
        st = N.lib.otg_spef_read_buffer(b, len(b), C.byref(opt), C.byref(h))
        if st != N.OK:
            raise SpefError(N.lib.otg_spef_last_error().decode(errors="replace"))
        return cls(h)

    @property
    def num_nets(self):
        return len(self.nodes_start) - 1

    @property
    def num_nodes(self):
        return int(self.nodes_start[-1])

    def _name(self, fn, *args):
        n = fn(self._h, *args, None, 0)
        if n < 0:
            raise IndexError(args)
        buf = C.create_string_buffer(n + 1)
        fn(self._h, *args, buf, n + 1)
        return buf.value.decode()

    def net_name(self, net: int) -> str:
        return self._name(N.lib.otg_spef_net_name, net)

    def node_name(self, net: int, local: int) -> str:
        return self._name(N.lib.otg_spef_node_name, net, local)

    def find_net(self, name: str) -> int:
        return N.lib.otg_spef_find_net(self._h, name.encode())

    def find_node(self, net: int, name: str) -> int:
        return N.lib.otg_spef_find_node(self._h, net, name.encode())

    def add_pin_cap(self, net: int, local: int, cap):
        c = (C.c_float * 4)(*[float(x) for x in cap])
        N.check(N.lib.otg_spef_add_pin_cap(self._h, net, local, c))

    def close(self):
        if self._h:
            N.lib.otg_spef_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write(path, forest):
    """Writes a forest (workloads.RcForest fields) as SPEF text."""
    N.check(N.lib.otg_spef_write(os.fsencode(path), forest.num_nets, N.addr(forest.nodes_start),
                                 N.addr(forest.edges), N.addr(forest.roots),
                                 N.addr(forest.edge_res), N.addr(forest.node_cap)))
