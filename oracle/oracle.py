"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py — never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(_HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "liboracle.so")
CLI_PATH = os.path.join(BUILD_DIR, "bt_oracle_cli")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class OrGraph(C.Structure):
    _fields_ = [
        ("u_num", C.c_uint32), ("l_num", C.c_uint32), ("n", C.c_uint32), ("m", C.c_uint32),
        ("neighbors", u32p), ("offsets", u32p), ("degrees", u32p), ("edges", u32p),
        ("edge_ids", u32p),
    ]


class OrStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in (
        "w_cnt", "w_peel", "i_peel", "b_peel", "d_peel", "sum_curr", "sub_rounds",
        "w_h", "i_h", "b_h")]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def build() -> None:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_ebfc.argtypes = [C.POINTER(OrGraph), C.c_int, u64p, u64p, u64p, C.POINTER(OrStats)]
        L.or_bfc.argtypes = [C.POINTER(OrGraph), C.c_int, u64p, C.POINTER(OrStats)]
        L.or_hidx.argtypes = [C.POINTER(OrGraph), C.c_int, u64p, C.POINTER(C.c_uint32), C.POINTER(OrStats)]
        L.or_msp.argtypes = [C.POINTER(OrGraph), C.c_int, u64p, C.POINTER(C.c_uint32), C.POINTER(OrStats)]
        L.or_msp_sample.argtypes = [C.POINTER(OrGraph), C.c_int, C.c_double, u64p, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_int), C.POINTER(OrStats)]
        L.or_msp_levels.argtypes = [C.POINTER(OrGraph), C.c_int, C.c_uint64, u64p, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_int), C.POINTER(OrStats)]
        L.or_peel_seq.argtypes = [C.POINTER(OrGraph), u64p]
        L.or_msp_bloom_ck.argtypes = [C.POINTER(OrGraph), C.c_int, C.c_uint64, C.c_int, u64p, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_int), C.POINTER(OrStats), u64p, C.c_uint32, CK_FN, C.c_void_p]
        L.or_msp_bloom_ck.restype = C.c_int
        L.or_mspp.argtypes = [C.POINTER(OrGraph), C.c_int, u64p, C.POINTER(C.c_uint32), C.POINTER(OrStats)]
        L.or_msp_bloom.argtypes = [C.POINTER(OrGraph), C.c_int, C.c_uint64, C.c_int, u64p, C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_int), C.POINTER(OrStats)]
        for f in (L.or_ebfc, L.or_bfc, L.or_hidx, L.or_msp, L.or_msp_sample, L.or_msp_levels, L.or_peel_seq,
                  L.or_msp_bloom, L.or_mspp):
            f.restype = C.c_int
        _lib = L
    return _lib


class _View:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, g):
        self.arrays = [np.ascontiguousarray(getattr(g, k), dtype=np.uint32)
                       for k in ("neighbors", "offsets", "degrees", "edges", "edge_ids")]
        p = [a.ctypes.data_as(u32p) for a in self.arrays]
        self.s = OrGraph(g.u_num, g.l_num, g.n, g.m, *p)


def _run(name, rc):
    if rc != 0:
        raise RuntimeError(f"oracle {name} failed with status {rc}")


def ebfc(g, threads: int = 1):
    """(support[m] uint64, total, support_max, stats) — cpu_bfc.cpp:24-122."""
    v = _View(g)
    sup = np.zeros(max(g.m, 1), np.uint64)
    tot, smax, st = C.c_uint64(), C.c_uint64(), OrStats()
    _run("ebfc", lib().or_ebfc(C.byref(v.s), threads, sup.ctypes.data_as(u64p), C.byref(tot),
                               C.byref(smax), C.byref(st)))
    return sup[: g.m], tot.value, smax.value, st.as_dict()


def bfc(g, threads: int = 1):
    """(total, stats) — cpu_bfc.cpp:137-203."""
    v = _View(g)
    tot, st = C.c_uint64(), OrStats()
    _run("bfc", lib().or_bfc(C.byref(v.s), threads, C.byref(tot), C.byref(st)))
    return tot.value, st.as_dict()


def hidx(g, threads: int = 1):
    """(bitruss[m] uint64, rounds, stats) — cpu_bitruss.cpp:125-239."""
    v = _View(g)
    out = np.zeros(max(g.m, 1), np.uint64)
    rounds, st = C.c_uint32(), OrStats()
    _run("hidx", lib().or_hidx(C.byref(v.s), threads, out.ctypes.data_as(u64p), C.byref(rounds), C.byref(st)))
    return out[: g.m], rounds.value, st.as_dict()


def msp(g, threads: int = 1):
    """(bitruss[m] uint64, levels, stats) — cpu_bitruss.cpp:241-597."""
    v = _View(g)
    out = np.zeros(max(g.m, 1), np.uint64)
    levels, st = C.c_uint32(), OrStats()
    _run("msp", lib().or_msp(C.byref(v.s), threads, out.ctypes.data_as(u64p), C.byref(levels), C.byref(st)))
    return out[: g.m], levels.value, st.as_dict()


def mspp(g, threads: int = 1):
    """(bitruss[m] uint64, levels, stats) — bitruss_mspp, bitruss_msp.cu:600-722."""
    v = _View(g)
    out = np.zeros(max(g.m, 1), np.uint64)
    levels, st = C.c_uint32(), OrStats()
    _run("mspp", lib().or_mspp(C.byref(v.s), threads, out.ctypes.data_as(u64p), C.byref(levels), C.byref(st)))
    return out[: g.m], levels.value, st.as_dict()


def msp_sample(g, threads: int, max_seconds: float):
    """msp cut short for timing: support counting, then peel levels until
    the first one that ends past max_seconds. Returns (levels_done,
    complete, stats) with the work counters of what was done."""
    v = _View(g)
    work = np.zeros(max(g.m, 1), np.uint64)
    levels, done, st = C.c_uint32(), C.c_int(), OrStats()
    _run("msp_sample", lib().or_msp_sample(C.byref(v.s), threads, max_seconds, work.ctypes.data_as(u64p),
                                           C.byref(levels), C.byref(done), C.byref(st)))
    return levels.value, bool(done.value), st.as_dict()


def msp_levels(g, threads: int, max_level: int):
    """msp through level max_level only: (state[m] uint64 — the result of a
    peeled edge, the current support of the others —, complete, stats)."""
    v = _View(g)
    out = np.zeros(max(g.m, 1), np.uint64)
    levels, done, st = C.c_uint32(), C.c_int(), OrStats()
    _run("msp_levels", lib().or_msp_levels(C.byref(v.s), threads, max_level, out.ctypes.data_as(u64p),
                                           C.byref(levels), C.byref(done), C.byref(st)))
    return out[: g.m], bool(done.value), st.as_dict()


def peel_seq(g):
    """Textbook sequential peel (cpu_bitruss.cpp:34-122): true bitruss numbers."""
    v = _View(g)
    out = np.zeros(max(g.m, 1), np.uint64)
    _run("peel_seq", lib().or_peel_seq(C.byref(v.s), out.ctypes.data_as(u64p)))
    return out[: g.m]


BLOOM_SORTED, BLOOM_EXACT, BLOOM_PLANT_SWAP = 1, 2, 4
CK_FN = C.CFUNCTYPE(None, C.c_uint64, u64p, C.c_uint32, C.c_void_p)


def msp_bloom_checkpoints(g, threads: int, levels, on_state, flags: int = 0):
    """or_msp_bloom's whole decomposition, calling on_state(level, state)
    with the prefix state after every level up to each of `levels`
    (ascending; state is a numpy view valid during the call only). Returns
    (result[m], levels, complete, stats)."""
    v = _View(g)
    out = np.zeros(max(g.m, 1), np.uint64)
    lv, done, st = C.c_uint32(), C.c_int(), OrStats()
    cks = np.ascontiguousarray(sorted(levels), dtype=np.uint64)

    def cb(level, ptr, m, _user):
        on_state(int(level), np.ctypeslib.as_array(ptr, shape=(max(int(m), 1),))[: int(m)])

    fn = CK_FN(cb)
    _run("msp_bloom_ck", lib().or_msp_bloom_ck(C.byref(v.s), threads, 0, int(flags), out.ctypes.data_as(u64p),
                                               C.byref(lv), C.byref(done), C.byref(st), cks.ctypes.data_as(u64p),
                                               len(cks), fn, None))
    return out[: g.m], lv.value, bool(done.value), st.as_dict()


def msp_bloom(g, threads: int = 1, max_level: int = 0, flags: int = 0):
    """msp over butterfly blooms (bt_oracle_bloom.c): (state[m] uint64,
    levels, complete, stats). With max_level = 0 the state is the bitruss
    result of or_msp; otherwise the prefix state of or_msp_levels. flags:
    BLOOM_SORTED forces the sorted-list form of the rule; BLOOM_EXACT gives
    the textbook peel (true bitruss numbers, as peel_seq and hidx);
    BLOOM_PLANT_SWAP puts a known slip into the rule (tests only)."""
    v = _View(g)
    out = np.zeros(max(g.m, 1), np.uint64)
    levels, done, st = C.c_uint32(), C.c_int(), OrStats()
    _run("msp_bloom", lib().or_msp_bloom(C.byref(v.s), threads, max_level, int(flags),
                                         out.ctypes.data_as(u64p), C.byref(levels), C.byref(done), C.byref(st)))
    return out[: g.m], levels.value, bool(done.value), st.as_dict()
