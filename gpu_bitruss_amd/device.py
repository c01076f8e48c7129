"""Device-side API: contexts, resident graphs and the reference's four
algorithm entry points (src/bfc/bfc.cuh:17-23, src/bitruss/bitruss.cuh:17-24)
on top of the C-ABI in include/bitruss_hip.h.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from . import _lib
from .graph import Graph

# Live handles, released in dependency order when the interpreter exits:
# at shutdown module globals (this module's `_lib` among them) may be cleared
# before the objects' __del__ runs, which would leave device memory, streams
# and server threads to the HIP runtime's own teardown.
_live: "weakref.WeakSet" = weakref.WeakSet()


def _release_all() -> None:
    for kind, method in (("Server", "stop"), ("Batch", "free"), ("DeviceGraph", "free"), ("Device", "close")):
        for obj in list(_live):
            if type(obj).__name__ == kind:
                try:
                    getattr(obj, method)()
                except Exception:
                    pass


atexit.register(_release_all)


def tune_class(edges: int, graphs: int = 1) -> int:
    """Graph-size class of the counting settings: floor(log4(edges per graph))."""
    return _lib.lib().bt_tune_class(edges, graphs)


def tune_default(size_class: int) -> tuple:
    """The counting settings tuned on MI355X for a size class."""
    c = _lib.BtCountConfig()
    _lib.check("bt_tune_default", _lib.lib().bt_tune_default(size_class, C.byref(c)))
    return c.as_tuple()


def tune_candidates() -> list:
    n = C.c_int(0)
    _lib.check("bt_tune_candidates", _lib.lib().bt_tune_candidates(None, 0, C.byref(n)))
    cands = (_lib.BtCountConfig * n.value)()
    _lib.check("bt_tune_candidates", _lib.lib().bt_tune_candidates(cands, n.value, C.byref(n)))
    return [c.as_tuple() for c in cands]


def device_count() -> int:
    c = C.c_int(0)
    _lib.check("bt_device_count", _lib.lib().bt_device_count(C.byref(c)))
    return c.value


class Device:
    """A bt_ctx: one HIP stream plus scratch on one GPU."""

    def __init__(self, device: int = 0):
        self._ctx = C.c_void_p()
        _lib.check("bt_ctx_create", _lib.lib().bt_ctx_create(device, C.byref(self._ctx)))
        self.device = device
        _live.add(self)

    @property
    def handle(self) -> C.c_void_p:
        if not self._ctx:
            raise RuntimeError("device context is closed")
        return self._ctx

    def set_msp_path(self, path: str) -> None:
        """"auto" (bloom-index peel when it fits), "merge" or "index"."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(
            self.handle, _lib.OPT_MSP_PATH, _lib.MSP_PATHS[path]))

    def set_batch_wedges(self, wedges: int) -> None:
        """Bound of priority wedges per batch union (0: as one index holds)."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_BATCH_WEDGES, wedges))

    def set_batch_levels(self, mode: str) -> None:
        """"union" (default) or "per_graph": how msp walks a batch union's levels."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(
            self.handle, _lib.OPT_BATCH_LEVELS, _lib.BATCH_LEVELS[mode]))

    def set_batch_lanes(self, lanes: int) -> None:
        """Lanes (contexts on equal CU shares) that batches uploaded afterwards run on."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_BATCH_LANES, lanes))

    def set_peel_blocks(self, cus: int) -> None:
        """CUs the persistent msp peel occupies (0: all)."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_PEEL_BLOCKS, cus))

    def set_peel_slice_ms(self, ms: int) -> None:
        """Time slice of one persistent msp peel launch (0: one launch)."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_PEEL_SLICE_MS, ms))

    def set_index_mem_pct(self, pct: int) -> None:
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_INDEX_MEM_PCT, pct))

    def set_peel_max_level(self, level: int) -> None:
        """msp stops after this level (0: never); msp() then reports the
        prefix state with ``partial`` set on the device graph."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_PEEL_MAX_LEVEL, level))

    def set_peel_c2cap(self, cap: int) -> None:
        """Both-frontier middles summed on chip (0..128; 0 forces the sorted lists)."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_PEEL_C2CAP, cap))

    def set_peel_window_edges(self, edges: int) -> None:
        """Live edges from which the level collect uses a support window."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_PEEL_WINDOW_EDGES, edges))

    def set_peel_big_middles(self, middles: int) -> None:
        """Blooms with more middles are walked by a whole workgroup."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_PEEL_BIG_MIDDLES, middles))

    def set_hidx_graph_rounds(self, rounds: int) -> None:
        """h-index rounds per captured HIP graph (0: one by one, with a
        read-back after each round)."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_HIDX_GRAPH_ROUNDS, rounds))

    def set_hidx_sorted(self, on: bool) -> None:
        """Whether h-index rounds count over sorted chunks of the larger
        blooms where that is cheaper (the default) or every edge streams its
        values. Results and round counts are the same."""
        _lib.check("bt_ctx_set_option", _lib.lib().bt_ctx_set_option(self.handle, _lib.OPT_HIDX_SORTED, int(bool(on))))

    def tune_get(self, size_class: int) -> tuple:
        """Counting settings of a graph-size class: (wave_wedges,
        block_wedges, block_threads, hub_threads)."""
        c = _lib.BtCountConfig()
        _lib.check("bt_tune_get", _lib.lib().bt_tune_get(self.handle, size_class, C.byref(c)))
        return c.as_tuple()

    def tune_set(self, size_class: int, cfg) -> None:
        c = _lib.BtCountConfig(*cfg)
        _lib.check("bt_tune_set", _lib.lib().bt_tune_set(self.handle, size_class, C.byref(c)))

    def autotune(self, dgraph: "DeviceGraph", reps: int = 3):
        """Times every counting candidate on a resident graph, keeps the
        fastest for its size class; returns (best settings, best ms,
        {settings: ms})."""
        n = C.c_int(0)
        _lib.check("bt_tune_candidates", _lib.lib().bt_tune_candidates(None, 0, C.byref(n)))
        cands = (_lib.BtCountConfig * n.value)()
        _lib.check("bt_tune_candidates", _lib.lib().bt_tune_candidates(cands, n.value, C.byref(n)))
        ms = (C.c_double * n.value)()
        best, best_ms = _lib.BtCountConfig(), C.c_double()
        _lib.check("bt_autotune", _lib.lib().bt_autotune(self.handle, dgraph.handle, reps, C.byref(best),
                                                          C.byref(best_ms), ms))
        return best.as_tuple(), best_ms.value, {c.as_tuple(): ms[i] for i, c in enumerate(cands)}

    def trim(self) -> None:
        """Give back the device memory the context caches between calls."""
        _lib.check("bt_ctx_trim", _lib.lib().bt_ctx_trim(self.handle))

    def upload(self, g: Graph) -> "DeviceGraph":
        return DeviceGraph(self, g)

    def upload_batch(self, graphs) -> "Batch":
        return Batch(self, list(graphs))

    def last_timings(self) -> dict:
        t = _lib.BtTimings()
        _lib.check("bt_ctx_last_timings", _lib.lib().bt_ctx_last_timings(self.handle, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def close(self) -> None:
        if self._ctx:
            _lib.lib().bt_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    """A graph resident in device memory (bt_dgraph)."""

    def __init__(self, dev: Device, g: Graph):
        self.dev = dev
        self.m = g.m
        self._h = C.c_void_p()
        hg = g.host_struct()
        _lib.check("bt_graph_upload", _lib.lib().bt_graph_upload(dev.handle, C.byref(hg), C.byref(self._h)))
        _live.add(self)

    @property
    def handle(self) -> C.c_void_p:
        if not self._h:
            raise RuntimeError("device graph is freed")
        return self._h

    def _out(self, copy: bool):
        if not copy:
            return None, None
        a = np.zeros(max(self.m, 1), dtype=np.uint64)
        return a, a.ctypes.data_as(_lib.u64p)

    def bfc(self) -> int:
        t = C.c_uint64()
        _lib.check("bt_bfc", _lib.lib().bt_bfc(self.dev.handle, self.handle, C.byref(t)))
        return t.value

    def ebfc(self, copy: bool = True):
        a, p = self._out(copy)
        t, s = C.c_uint64(), C.c_uint64()
        _lib.check("bt_ebfc", _lib.lib().bt_ebfc(self.dev.handle, self.handle, p, C.byref(t), C.byref(s)))
        return (a[: self.m] if a is not None else None), t.value, s.value

    def hidx(self, copy: bool = True):
        a, p = self._out(copy)
        r = C.c_uint32()
        _lib.check("bt_hidx", _lib.lib().bt_hidx(self.dev.handle, self.handle, p, C.byref(r)))
        return (a[: self.m] if a is not None else None), r.value

    def msp(self, copy: bool = True):
        """bt_msp. With a level bound on the device (set_peel_max_level) a
        decomposition cut short returns its prefix state and sets
        ``self.partial``."""
        a, p = self._out(copy)
        lv = C.c_uint32()
        st = _lib.lib().bt_msp(self.dev.handle, self.handle, p, C.byref(lv))
        self.partial = st == _lib.BT_PARTIAL
        if not self.partial:
            _lib.check("bt_msp", st)
        return (a[: self.m] if a is not None else None), lv.value

    def mspp(self, copy: bool = True):
        """bt_mspp: the reference's baseline msp (bitruss_mspp)."""
        a, p = self._out(copy)
        lv = C.c_uint32()
        _lib.check("bt_mspp", _lib.lib().bt_mspp(self.dev.handle, self.handle, p, C.byref(lv)))
        return (a[: self.m] if a is not None else None), lv.value

    def peel(self, copy: bool = True):
        """bt_peel: true bitruss numbers by peeling (bitruss_peeling_cpu)."""
        a, p = self._out(copy)
        lv = C.c_uint32()
        _lib.check("bt_peel", _lib.lib().bt_peel(self.dev.handle, self.handle, p, C.byref(lv)))
        return (a[: self.m] if a is not None else None), lv.value

    def free(self) -> None:
        if self._h:
            _lib.lib().bt_graph_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many graphs resident on the device as one disjoint union (bt_batch);
    each algorithm runs once over all of them and returns per-graph results
    equal to the plain ones."""

    def __init__(self, dev: Device, graphs):
        self.dev = dev
        self.ms = [g.m for g in graphs]
        self._h = C.c_void_p()
        hs = (_lib.BtHostGraph * max(len(graphs), 1))()
        for i, g in enumerate(graphs):
            hs[i] = g.host_struct()
        self._keep = graphs  # host arrays stay alive during the upload
        _lib.check("bt_batch_upload", _lib.lib().bt_batch_upload(dev.handle, hs, len(graphs), C.byref(self._h)))
        self._keep = None
        _live.add(self)

    @property
    def handle(self) -> C.c_void_p:
        if not self._h:
            raise RuntimeError("batch is freed")
        return self._h

    def __len__(self) -> int:
        return len(self.ms)

    def chunks(self) -> int:
        """Number of unions the batch is held as."""
        n = C.c_size_t()
        _lib.check("bt_batch_chunks", _lib.lib().bt_batch_chunks(self.handle, C.byref(n)))
        return n.value

    def run(self, algo: str, copy: bool = True):
        """(per-graph uint64 arrays or None, per-graph scalar dicts)."""
        k = len(self.ms)
        res = (_lib.BtBatchResult * max(k, 1))()
        outs = None
        arrays = None
        if copy and algo != "bfc":
            arrays = [np.zeros(max(m, 1), np.uint64) for m in self.ms]
            outs = (_lib.u64p * max(k, 1))(*[a.ctypes.data_as(_lib.u64p) for a in arrays])
        _lib.check("bt_batch_run", _lib.lib().bt_batch_run(self.dev.handle, self.handle, _lib.ALGOS[algo],
                                                            outs, res))
        scal = [{"total": r.total, "support_max": r.support_max, "levels": r.levels, "rounds": r.rounds}
                for r in res[:k]]
        if arrays is not None:
            arrays = [a[:m] for a, m in zip(arrays, self.ms)]
        return arrays, scal

    def free(self) -> None:
        if self._h:
            _lib.lib().bt_batch_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (bt_serve_*): light, heavy and
    large requests are batched on three lanes with disjoint CU shares
    (`cus` = light, heavy, large; 0 for the defaults), under a device-memory
    admission budget."""

    def __init__(self, device: int = 0, budget_bytes: int = 0, large_edges: int = 0, max_batch: int = 0,
                 light_wedges: int = 0, batch_wedges: int = 0, cus: tuple = (0, 0, 0), light_workers: int = 0):
        cfg = _lib.BtServeConfig(device, budget_bytes, large_edges, max_batch, light_wedges, batch_wedges, *cus,
                                 light_workers)
        self._h = C.c_void_p()
        _lib.check("bt_serve_start", _lib.lib().bt_serve_start(C.byref(cfg), C.byref(self._h)))
        self._m = {}
        _live.add(self)

    def submit(self, g: Graph, algo: str = "msp") -> int:
        t = C.c_uint64()
        _lib.check("bt_serve_submit", _lib.lib().bt_serve_submit(self._h, C.byref(g.host_struct()),
                                                                  _lib.ALGOS[algo], C.byref(t)))
        self._m[t.value] = (g.m, algo)
        return t.value

    def done(self, ticket: int) -> bool:
        d = C.c_int()
        _lib.check("bt_serve_poll", _lib.lib().bt_serve_poll(self._h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int):
        """(per-edge uint64 array or None, scalars dict, latency in ms)."""
        m, algo = self._m.pop(ticket)
        out = np.zeros(max(m, 1), np.uint64) if algo != "bfc" else None
        res = _lib.BtBatchResult()
        lat = C.c_double()
        _lib.check("bt_serve_wait", _lib.lib().bt_serve_wait(
            self._h, ticket, out.ctypes.data_as(_lib.u64p) if out is not None else None, C.byref(res),
            C.byref(lat)))
        scal = {"total": res.total, "support_max": res.support_max, "levels": res.levels, "rounds": res.rounds}
        return (out[:m] if out is not None else None), scal, lat.value

    def stats(self) -> dict:
        st = _lib.BtServeStats()
        _lib.check("bt_serve_stats_get", _lib.lib().bt_serve_stats_get(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def stop(self) -> None:
        if self._h:
            _lib.lib().bt_serve_stop(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


_default: Device | None = None


def default_device() -> Device:
    global _default
    if _default is None:
        _default = Device(0)
    return _default


# ---- the reference's free functions on a Graph ---------------------------

def bfc_vpp(g: Graph, dev: Device | None = None) -> int:
    """Total butterflies (bfc_vpp.cu:321-385); stored in g.butterflies."""
    dg = (dev or default_device()).upload(g)
    try:
        g.butterflies = dg.bfc()
    finally:
        dg.free()
    return g.butterflies


def bfc_evpp(g: Graph, dev: Device | None = None) -> None:
    """Per-edge supports into g.edge_support and g.support_max
    (bfc_vpp.cu:233-319)."""
    dg = (dev or default_device()).upload(g)
    try:
        g.edge_support, g.butterflies, g.support_max = dg.ebfc()
    finally:
        dg.free()


def g_bitruss_hindex(g: Graph, dev: Device | None = None) -> int:
    """h-index bitruss numbers into g.bitruss (bitruss_hidx.cu:179-256);
    returns the number of rounds."""
    dg = (dev or default_device()).upload(g)
    try:
        g.bitruss, rounds = dg.hidx()
    finally:
        dg.free()
    return rounds


def bitruss_msp(g: Graph, dev: Device | None = None) -> int:
    """GBiD msp results into g.bitruss (bitruss_msp.cu:440-583); returns the
    number of levels the reference logs."""
    dg = (dev or default_device()).upload(g)
    try:
        g.bitruss, levels = dg.msp()
    finally:
        dg.free()
    return levels


def bitruss_mspp(g: Graph, dev: Device | None = None) -> int:
    """The baseline msp (bitruss_msp.cu:600-722) into g.bitruss; returns the
    number of levels the reference logs."""
    dg = (dev or default_device()).upload(g)
    try:
        g.bitruss, levels = dg.mspp()
    finally:
        dg.free()
    return levels


def bitruss_peeling(g: Graph, dev: Device | None = None) -> int:
    """True bitruss numbers by peeling (the results of bitruss_peeling_cpu,
    cpu_bitruss.cpp:34-122) into g.bitruss; returns the largest."""
    dg = (dev or default_device()).upload(g)
    try:
        g.bitruss, kmax = dg.peel()
    finally:
        dg.free()
    return kmax
