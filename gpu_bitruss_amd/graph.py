"""Python mirror of the reference's ``class Graph`` (src/graph/graph.h:30-52).

Arrays are numpy (uint32 CSR), results are uint64. Loading, conversion and
generation are done by libbitruss_hip.so (include/bt_graph.h); this module
only moves arrays across the C-ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

WORKLOADS = {
    "small_uniform": 1,    # 1,000 x 1,500 uniform, m = 10,000, seed 0
    "powerlaw_2m": 2,      # Chung-Lu 200k x 100k, gamma 2.3, m ~ 2M, cap 5,000, seed 1
    "skewed_60m": 3,       # Chung-Lu 4M (2.5) x 1M (2.0), m ~ 60M, cap 200k, seed 2
    "mix_graph": 4,        # graph `index` of the 4,096-graph mix (seed 1000+index)
    "large_request": 5,    # one large serving request (Chung-Lu 2.2, m in [1M, 5M])
}


def _as_u32p(a: np.ndarray):
    return a.ctypes.data_as(_lib.u32p)


def _copy(ptr, count: int) -> np.ndarray:
    if count == 0:
        return np.zeros(0, dtype=np.uint32)
    return np.ctypeslib.as_array(ptr, shape=(count,)).copy()


class Graph:
    """Bipartite CSR graph with the reference's public members.

    ``Graph(filename, is_to_bin)`` follows graph.cpp:261-277: with
    ``is_to_bin`` true the file is a text edge list to convert, otherwise it
    must be a ``.bin`` file (the name must contain ".bin").
    """

    def __init__(self, filename: str | None = None, is_to_bin: bool = False):
        self.u_num = self.l_num = self.u_max = self.m = self.n = 0
        self.support_max = 0
        self.neighbors = self.offsets = self.degrees = np.zeros(0, np.uint32)
        self.edges = self.edge_ids = np.zeros(0, np.uint32)
        self.edge_support = np.zeros(0, np.uint64)
        self.bitruss = np.zeros(0, np.uint64)
        self.butterflies = None
        if filename is None:
            return
        if is_to_bin:
            self._take(lambda hg: _lib.lib().bt_graph_read_text(filename.encode(), hg), "bt_graph_read_text")
        else:
            if ".bin" not in filename:
                raise ValueError("graph file should be binary file")
            self.load_graph_bin(filename)

    # -- construction -----------------------------------------------------
    def _take(self, fill, name: str) -> None:
        hg = _lib.BtHostGraph()
        _lib.check(name, fill(C.byref(hg)))
        try:
            self.u_num, self.l_num, self.u_max = hg.u_num, hg.l_num, hg.u_max
            self.m, self.n, self.support_max = hg.m, hg.n, hg.support_max
            m, n = hg.m, hg.n
            self.neighbors = _copy(hg.neighbors, 2 * m)
            self.offsets = _copy(hg.offsets, n + 1) if n or hg.offsets else np.zeros(1, np.uint32)
            self.degrees = _copy(hg.degrees, n)
            self.edges = _copy(hg.edges, 2 * m)
            self.edge_ids = _copy(hg.edge_ids, 2 * m)
        finally:
            _lib.lib().bt_graph_release(C.byref(hg))
        if self.offsets.size == 0:
            self.offsets = np.zeros(1, np.uint32)
        self.edge_support = np.zeros(self.m, np.uint64)
        self.bitruss = np.zeros(self.m, np.uint64)

    @classmethod
    def from_edges(cls, upper, lower) -> "Graph":
        """CSR from raw (upper, lower) pairs with process_graph semantics."""
        u = np.ascontiguousarray(upper, dtype=np.uint32)
        v = np.ascontiguousarray(lower, dtype=np.uint32)
        if u.shape != v.shape:
            raise ValueError("upper and lower must have the same length")
        g = cls()
        g._take(lambda hg: _lib.lib().bt_graph_from_edges(_as_u32p(u), _as_u32p(v), u.size, hg),
                "bt_graph_from_edges")
        return g

    def priority_wedges(self) -> int:
        """Wedges butterfly counting enumerates (bt_graph_wedges)."""
        w = C.c_uint64()
        _lib.check("bt_graph_wedges", _lib.lib().bt_graph_wedges(C.byref(self.host_struct()), C.byref(w)))
        return w.value

    @classmethod
    def union(cls, graphs) -> tuple["Graph", np.ndarray]:
        """Disjoint union in the batch layout (bt_graph_union): the upper
        vertices of every graph, then the lower ones, edge ids offset per
        graph. Returns (union, edge offsets[k + 1])."""
        graphs = list(graphs)
        hs = (_lib.BtHostGraph * max(len(graphs), 1))()
        for i, g in enumerate(graphs):
            hs[i] = g.host_struct()
        e_off = np.zeros(len(graphs) + 1, np.uint64)
        u = cls()
        u._take(lambda hg: _lib.lib().bt_graph_union(hs, len(graphs), hg, e_off.ctypes.data_as(_lib.u64p)),
                "bt_graph_union")
        return u, e_off

    @classmethod
    def generate(cls, workload: str, index: int = 0) -> "Graph":
        """One of the named seeded workloads (see WORKLOADS)."""
        el = _lib.BtEdgeList()
        _lib.check("bt_gen_workload", _lib.lib().bt_gen_workload(WORKLOADS[workload], index, C.byref(el)))
        try:
            u = _copy(el.upper, el.count)
            v = _copy(el.lower, el.count)
        finally:
            _lib.lib().bt_edge_list_release(C.byref(el))
        return cls.from_edges(u, v)

    @classmethod
    def chung_lu(cls, nu, nv, gamma_u, gamma_v, m_target, cap_u, cap_v, seed) -> "Graph":
        el = _lib.BtEdgeList()
        _lib.check("bt_gen_chung_lu", _lib.lib().bt_gen_chung_lu(
            nu, nv, gamma_u, gamma_v, m_target, cap_u, cap_v, seed, C.byref(el)))
        try:
            u = _copy(el.upper, el.count)
            v = _copy(el.lower, el.count)
        finally:
            _lib.lib().bt_edge_list_release(C.byref(el))
        return cls.from_edges(u, v)

    @classmethod
    def uniform(cls, nu, nv, m, seed) -> "Graph":
        el = _lib.BtEdgeList()
        _lib.check("bt_gen_uniform", _lib.lib().bt_gen_uniform(nu, nv, m, seed, C.byref(el)))
        try:
            u = _copy(el.upper, el.count)
            v = _copy(el.lower, el.count)
        finally:
            _lib.lib().bt_edge_list_release(C.byref(el))
        return cls.from_edges(u, v)

    # -- the C view ---------------------------------------------------------
    def host_struct(self) -> _lib.BtHostGraph:
        """A bt_host_graph viewing this object's arrays (keep self alive)."""
        for name in ("neighbors", "offsets", "degrees", "edges", "edge_ids"):
            a = getattr(self, name)
            if a.dtype != np.uint32 or not a.flags["C_CONTIGUOUS"]:
                setattr(self, name, np.ascontiguousarray(a, dtype=np.uint32))
        hg = _lib.BtHostGraph()
        hg.u_num, hg.l_num, hg.u_max = self.u_num, self.l_num, self.u_max
        hg.m, hg.n, hg.support_max = self.m, self.n, min(int(self.support_max), 0xFFFFFFFF)
        hg.neighbors = _as_u32p(self.neighbors)
        hg.offsets = _as_u32p(self.offsets)
        hg.degrees = _as_u32p(self.degrees)
        hg.edges = _as_u32p(self.edges)
        hg.edge_ids = _as_u32p(self.edge_ids)
        return hg

    # -- I/O (graph.cpp:293-375) ------------------------------------------------
    def graph_to_bin(self, filename: str) -> None:
        hg = self.host_struct()
        _lib.check("bt_graph_write_bin", _lib.lib().bt_graph_write_bin(filename.encode(), C.byref(hg)))

    def load_graph_bin(self, filename: str) -> None:
        self._take(lambda hg: _lib.lib().bt_graph_read_bin(filename.encode(), hg), "bt_graph_read_bin")

    def check_graph(self) -> None:
        hg = self.host_struct()
        _lib.check("bt_graph_check", _lib.lib().bt_graph_check(C.byref(hg)))

    def __repr__(self) -> str:
        return f"Graph(upper={self.u_num}, lower={self.l_num}, vertices={self.n}, edges={self.m})"
