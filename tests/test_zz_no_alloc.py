"""The allocation-free entry point bt_ebfc_static (scratch in a fixed
static device array) against bt_ebfc, through the C-ABI: two calls in a
row, the second on a smaller graph, then a graph beyond the array's size,
which is refused."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from gpu_bitruss_amd import _lib
from tests import graphs as G

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[no-alloc] {msg}", flush=True)
    sys.stdout.flush()


def ebfc_static(dev, dg, m):
    out = np.zeros(max(m, 1), dtype=np.uint64)
    t, s = C.c_uint64(), C.c_uint64()
    st = _lib.lib().bt_ebfc_static(dev.handle, dg.handle, out.ctypes.data_as(_lib.u64p), C.byref(t), C.byref(s))
    return st, out[:m], t.value, s.value


@pytest.mark.gpu
def test_ebfc_static_equals_ebfc(dev):
    # largest first: the second call reuses the array with less data, so
    # anything left from the first call that it failed to clear shows up
    graphs = [("powerlaw_2m", gb.Graph.generate("powerlaw_2m")),
              ("hub", G.hub_graph(5, 6000, 20000, 400, 7000)),
              ("mix2", gb.Graph.generate("mix_graph", 2))]
    for name, g in graphs:
        say(f"{name}: m={g.m}")
        dg = dev.upload(g)
        try:
            want, wt, ws = dg.ebfc()
            for rep in range(2):
                st, got, t, s = ebfc_static(dev, dg, g.m)
                assert st == 0, (name, _lib.status_text(st))
                assert np.array_equal(got, want) and (t, s) == (wt, ws), (name, rep)
        finally:
            dg.free()


@pytest.mark.gpu
def test_ebfc_static_refuses_a_graph_beyond_the_scratch(dev):
    # 12 M distinct edges: the sort storage alone (about 24 B per edge) is
    # twice the static array
    rng = np.random.default_rng(3)
    u = rng.integers(0, 3000, size=13_000_000, dtype=np.uint64)
    v = rng.integers(0, 5000, size=13_000_000, dtype=np.uint64)
    key = np.unique(u * 5000 + v)[:12_000_000]
    g = gb.Graph.from_edges((key // 5000).astype(np.uint32), (key % 5000).astype(np.uint32))
    say(f"beyond the scratch: m={g.m}")
    dg = dev.upload(g)
    try:
        st, _, _, _ = ebfc_static(dev, dg, g.m)
        assert st == _lib.BT_ERR_SCRATCH, _lib.status_text(st)
        # the regular entry point still serves it
        _, t, _ = dg.ebfc(copy=False)
        assert t > 0
    finally:
        dg.free()
