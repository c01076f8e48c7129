"""ctypes binding of libbitruss_hip.so (the C-ABI declared in include/*.h).

The library is built in-tree by ``__graft_entry__.build()`` into
``gpu_bitruss_amd/lib/``. There is no fallback: if the library is missing
every entry point raises, so a run can never silently use another path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libbitruss_hip.so")
COH_PATH = os.path.join(LIB_DIR, "coh")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class BtHostGraph(C.Structure):
    """struct bt_host_graph (include/bt_graph.h)."""

    _fields_ = [
        ("u_num", C.c_uint32),
        ("l_num", C.c_uint32),
        ("u_max", C.c_uint32),
        ("m", C.c_uint32),
        ("n", C.c_uint32),
        ("support_max", C.c_uint32),
        ("neighbors", u32p),
        ("offsets", u32p),
        ("degrees", u32p),
        ("edges", u32p),
        ("edge_ids", u32p),
    ]


class BtEdgeList(C.Structure):
    """struct bt_edge_list (include/bt_graph.h)."""

    _fields_ = [("count", C.c_uint64), ("upper", u32p), ("lower", u32p)]


class BtTimings(C.Structure):
    """struct bt_timings (include/bitruss_hip.h)."""

    _fields_ = [
        ("total_ms", C.c_double),
        ("prep_ms", C.c_double),
        ("count_ms", C.c_double),
        ("peel_ms", C.c_double),
        ("peel_kernel_ms", C.c_double),
        ("peel_kernel_launches", C.c_uint64),
        ("count_kernel_ms", C.c_double),
        ("count_kernel_launches", C.c_uint64),
        ("wedges_count", C.c_uint64),
        ("wedges_peel", C.c_uint64),
        ("sub_rounds", C.c_uint64),
        ("levels_scanned", C.c_uint64),
        ("sorted_blooms", C.c_uint64),
    ]


class BtBatchResult(C.Structure):
    """struct bt_batch_result (include/bitruss_hip.h)."""

    _fields_ = [
        ("total", C.c_uint64),
        ("support_max", C.c_uint64),
        ("levels", C.c_uint32),
        ("rounds", C.c_uint32),
    ]


class BtServeConfig(C.Structure):
    """struct bt_serve_config (include/bitruss_hip.h)."""

    _fields_ = [
        ("device", C.c_int),
        ("budget_bytes", C.c_uint64),
        ("large_edges", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("light_wedges", C.c_uint64),
        ("batch_wedges", C.c_uint64),
        ("light_cus", C.c_uint32),
        ("heavy_cus", C.c_uint32),
        ("large_cus", C.c_uint32),
        ("light_workers", C.c_uint32),
    ]


class BtServeStats(C.Structure):
    """struct bt_serve_stats (include/bitruss_hip.h)."""

    _fields_ = [(k, C.c_uint64) for k in (
        "submitted", "completed", "failed", "batches", "batched_requests", "large_runs", "peak_bytes",
        "in_use_bytes")]


class BtCountConfig(C.Structure):
    """struct bt_count_config (include/bitruss_hip.h)."""

    _fields_ = [("wave_wedges", C.c_uint32), ("block_wedges", C.c_uint32), ("block_threads", C.c_uint32),
                ("hub_threads", C.c_uint32)]

    def as_tuple(self) -> tuple:
        return (self.wave_wedges, self.block_wedges, self.block_threads, self.hub_threads)


TUNE_CLASSES = 16


class BitrussError(RuntimeError):
    """A non-zero status returned by libbitruss_hip.so."""

    def __init__(self, func: str, status: int, text: str):
        super().__init__(f"{func} failed: status {status} ({text})")
        self.status = status


OPT_MSP_PATH = 1
OPT_INDEX_MEM_PCT = 2
OPT_BATCH_WEDGES = 3
OPT_BATCH_LEVELS = 4
BATCH_LEVELS = {"per_graph": 0, "union": 1}
OPT_PEEL_BLOCKS = 5
OPT_PEEL_SLICE_MS = 6
OPT_BATCH_LANES = 7
OPT_PEEL_MAX_LEVEL = 8
OPT_PEEL_C2CAP = 9
OPT_PEEL_WINDOW_EDGES = 10
OPT_PEEL_BIG_MIDDLES = 11
OPT_HIDX_GRAPH_ROUNDS = 12
OPT_HIDX_SORTED = 13
# statuses (enum bt_status, include/bt_graph.h) the wrappers act on
BT_PARTIAL = 10
BT_ERR_BUSY = 11
BT_ERR_SCRATCH = 12
MSP_PATHS = {"auto": 0, "merge": 1, "index": 2}
ALGOS = {"bfc": 1, "ebfc": 2, "hidx": 3, "msp": 4}

_lib = None


def lib() -> C.CDLL:
    """Load libbitruss_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "bt_status_str": (C.c_char_p, [C.c_int]),
        "bt_graph_from_edges": (C.c_int, [u32p, u32p, C.c_uint64, C.POINTER(BtHostGraph)]),
        "bt_graph_read_text": (C.c_int, [C.c_char_p, C.POINTER(BtHostGraph)]),
        "bt_graph_read_bin": (C.c_int, [C.c_char_p, C.POINTER(BtHostGraph)]),
        "bt_graph_write_bin": (C.c_int, [C.c_char_p, C.POINTER(BtHostGraph)]),
        "bt_graph_check": (C.c_int, [C.POINTER(BtHostGraph)]),
        "bt_graph_release": (None, [C.POINTER(BtHostGraph)]),
        "bt_edge_list_release": (None, [C.POINTER(BtEdgeList)]),
        "bt_gen_uniform": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(BtEdgeList)]),
        "bt_gen_chung_lu": (
            C.c_int,
            [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
             C.c_uint64, C.POINTER(BtEdgeList)],
        ),
        "bt_gen_workload": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(BtEdgeList)]),
        "bt_graph_wedges": (C.c_int, [C.POINTER(BtHostGraph), C.POINTER(C.c_uint64)]),
        "bt_graph_union": (C.c_int, [C.POINTER(BtHostGraph), C.c_uint64, C.POINTER(BtHostGraph), u64p]),
        "bt_tune_class": (C.c_int, [C.c_uint64, C.c_uint32]),
        "bt_tune_default": (C.c_int, [C.c_int, C.POINTER(BtCountConfig)]),
        "bt_tune_candidates": (C.c_int, [C.POINTER(BtCountConfig), C.c_int, C.POINTER(C.c_int)]),
    }
    gpu_sig = {
        "bt_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "bt_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "bt_ctx_destroy": (C.c_int, [vp]),
        "bt_ctx_last_timings": (C.c_int, [vp, C.POINTER(BtTimings)]),
        "bt_ctx_set_option": (C.c_int, [vp, C.c_int, C.c_int64]),
        "bt_ctx_trim": (C.c_int, [vp]),
        "bt_tune_get": (C.c_int, [vp, C.c_int, C.POINTER(BtCountConfig)]),
        "bt_tune_set": (C.c_int, [vp, C.c_int, C.POINTER(BtCountConfig)]),
        "bt_autotune": (C.c_int, [vp, vp, C.c_int, C.POINTER(BtCountConfig), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
        "bt_graph_upload": (C.c_int, [vp, C.POINTER(BtHostGraph), C.POINTER(vp)]),
        "bt_graph_free": (C.c_int, [vp]),
        "bt_bfc": (C.c_int, [vp, vp, u64p]),
        "bt_ebfc": (C.c_int, [vp, vp, u64p, u64p, u64p]),
        "bt_ebfc_static": (C.c_int, [vp, vp, u64p, u64p, u64p]),
        "bt_hidx": (C.c_int, [vp, vp, u64p, C.POINTER(C.c_uint32)]),
        "bt_msp": (C.c_int, [vp, vp, u64p, C.POINTER(C.c_uint32)]),
        "bt_mspp": (C.c_int, [vp, vp, u64p, C.POINTER(C.c_uint32)]),
        "bt_peel": (C.c_int, [vp, vp, u64p, C.POINTER(C.c_uint32)]),
        "bt_batch_upload": (C.c_int, [vp, C.POINTER(BtHostGraph), C.c_size_t, C.POINTER(vp)]),
        "bt_batch_free": (C.c_int, [vp]),
        "bt_batch_chunks": (C.c_int, [vp, C.POINTER(C.c_size_t)]),
        "bt_batch_run": (C.c_int, [vp, vp, C.c_int, C.POINTER(u64p), C.POINTER(BtBatchResult)]),
        "bt_serve_start": (C.c_int, [C.POINTER(BtServeConfig), C.POINTER(vp)]),
        "bt_serve_submit": (C.c_int, [vp, C.POINTER(BtHostGraph), C.c_int, C.POINTER(C.c_uint64)]),
        "bt_serve_poll": (C.c_int, [vp, C.c_uint64, C.POINTER(C.c_int)]),
        "bt_serve_wait": (C.c_int, [vp, C.c_uint64, u64p, C.POINTER(BtBatchResult), C.POINTER(C.c_double)]),
        "bt_serve_stats_get": (C.c_int, [vp, C.POINTER(BtServeStats)]),
        "bt_serve_stop": (C.c_int, [vp]),
    }
    for name, (res, args) in {**sig, **gpu_sig}.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def status_text(status: int) -> str:
    return lib().bt_status_str(status).decode()


def check(func: str, status: int) -> None:
    if status != 0:
        raise BitrussError(func, status, status_text(status))
