"""Graph layer: process_graph semantics, .bin byte layout, reader
validation, text parsing and the seeded generators (CPU only)."""
from __future__ import annotations

import os
import struct

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from tests import graphs as G


def py_process_graph(pairs):
    """Independent restatement of process_graph (graph.cpp:74-242):
    first-appearance ids, lower shifted by u_num, sorted lists, edge ids in
    (upper, neighbour) order."""
    up, lo = {}, {}
    for u, v in pairs:
        up.setdefault(u, len(up))
        lo.setdefault(v, len(lo))
    nu = len(up)
    adj = {x: [] for x in range(nu + len(lo))}
    for u, v in pairs:
        a, b = up[u], nu + lo[v]
        adj[a].append(b)
        adj[b].append(a)
    for x in adj:
        adj[x].sort()
    edges = []
    for u in range(nu):
        for v in adj[u]:
            edges.append((u, v))
    eid = {e: i for i, e in enumerate(edges)}
    nbr, eids, off = [], [], [0]
    for x in range(len(adj)):
        for y in adj[x]:
            nbr.append(y)
            eids.append(eid[(x, y)] if x < nu else eid[(y, x)])
        off.append(len(nbr))
    return nu, len(lo), nbr, off, eids, edges


def test_aw_structure():
    g = G.from_pairs(G.AW_EDGES)
    assert (g.u_num, g.l_num, g.n, g.m) == (5, 3, 8, 10)
    assert g.edges.reshape(-1, 2).tolist() == [[0, 5], [0, 6], [1, 5], [1, 6], [2, 5], [2, 6], [2, 7],
                                               [3, 6], [4, 6], [4, 7]]
    assert g.offsets.tolist() == [0, 2, 4, 7, 8, 10, 13, 18, 20]
    assert g.degrees.tolist() == [2, 2, 3, 1, 2, 3, 5, 2]


@pytest.mark.parametrize("seed", range(10))
def test_process_graph_semantics(seed):
    rng = np.random.default_rng(seed)
    cells = rng.choice(40 * 40, size=int(rng.integers(1, 300)), replace=False)
    pairs = [(int(c // 40) * 7 + 3, int(c % 40) * 11 + 1) for c in cells]  # sparse raw ids
    g = G.from_pairs(pairs)
    nu, nl, nbr, off, eids, edges = py_process_graph(pairs)
    assert (g.u_num, g.l_num, g.m) == (nu, nl, len(pairs))
    assert g.neighbors.tolist() == nbr
    assert g.offsets.tolist() == off
    assert g.edge_ids.tolist() == eids
    assert [tuple(x) for x in g.edges.reshape(-1, 2).tolist()] == edges
    g.check_graph()


def test_bin_layout_and_round_trip(tmp_path):
    g = G.from_pairs(G.AW_EDGES)
    p = str(tmp_path / "aw.bin")
    g.graph_to_bin(p)
    raw = open(p, "rb").read()
    m, n = 10, 8
    assert len(raw) == 20 + 4 * (8 * m + 2 * n + 1)
    assert struct.unpack("<5I", raw[:20]) == (5, 3, 5, 10, 0)
    a = np.frombuffer(raw[20:], dtype="<u4")
    o = 0
    assert a[o:o + 2 * m].tolist() == g.neighbors.tolist(); o += 2 * m
    assert a[o:o + n + 1].tolist() == g.offsets.tolist(); o += n + 1
    assert a[o:o + n].tolist() == g.degrees.tolist(); o += n
    assert not a[o:o + m].any(); o += m                       # edge_support zeros
    assert a[o:o + 2 * m].tolist() == g.edges.tolist(); o += 2 * m
    assert a[o:o + 2 * m].tolist() == g.edge_ids.tolist(); o += 2 * m
    assert not a[o:o + m].any()                               # bitruss zeros
    h = gb.Graph(p, False)
    for k in ("neighbors", "offsets", "degrees", "edges", "edge_ids"):
        assert np.array_equal(getattr(h, k), getattr(g, k))
    assert (h.u_num, h.l_num, h.u_max, h.m, h.n) == (5, 3, 5, 10, 8)


def test_bin_reader_rejects_bad_files(tmp_path):
    g = G.from_pairs(G.AW_EDGES)
    p = str(tmp_path / "aw.bin")
    g.graph_to_bin(p)
    raw = bytearray(open(p, "rb").read())
    trunc = str(tmp_path / "trunc.bin")
    open(trunc, "wb").write(raw[:-4])
    with pytest.raises(gb.BitrussError):
        gb.Graph(trunc, False)
    bad = bytearray(raw)
    struct.pack_into("<I", bad, 20, 999)  # neighbour id out of range
    badp = str(tmp_path / "bad.bin")
    open(badp, "wb").write(bad)
    with pytest.raises(gb.BitrussError):
        gb.Graph(badp, False)
    with pytest.raises(gb.BitrussError):
        gb.Graph(str(tmp_path / "missing.bin"), False)
    with pytest.raises(ValueError):
        gb.Graph(str(tmp_path / "x.txt"), False)  # graph.cpp:271-274


def test_text_parsing(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("% bip unweighted\n% 10 5 3\n0 5 1 123\n\n0 6\n1 5\n1 6\t7\n")
    g = gb.Graph(str(p), True)
    assert (g.u_num, g.l_num, g.m) == (2, 2, 4)
    q = tmp_path / "bad.txt"
    q.write_text("0 5\nnot an edge\n")
    with pytest.raises(gb.BitrussError):
        gb.Graph(str(q), True)
    r = tmp_path / "neg.txt"
    r.write_text("0 -5\n")
    with pytest.raises(gb.BitrussError):
        gb.Graph(str(r), True)


def test_empty_graph(tmp_path):
    g = G.from_pairs(np.zeros((0, 2), dtype=np.uint32))
    assert g.m == 0 and g.n == 0
    p = str(tmp_path / "e.bin")
    g.graph_to_bin(p)
    assert os.path.getsize(p) == 24
    h = gb.Graph(p, False)
    assert h.m == 0


def test_uniform_generator():
    g = gb.Graph.generate("small_uniform")
    assert g.m == 10000 and g.u_num <= 1000 and g.l_num <= 1500
    pairs = set(map(tuple, g.edges.reshape(-1, 2).tolist()))
    assert len(pairs) == g.m  # distinct
    g2 = gb.Graph.generate("small_uniform")
    assert np.array_equal(g.neighbors, g2.neighbors) and np.array_equal(g.edge_ids, g2.edge_ids)


def test_chung_lu_generator_properties():
    g = gb.Graph.chung_lu(20000, 10000, 2.3, 2.3, 2.0e5, 500.0, 500.0, 5)
    assert 1.8e5 < g.m < 2.05e5
    pairs = set(map(tuple, g.edges.reshape(-1, 2).tolist()))
    assert len(pairs) == g.m
    deg = g.degrees.astype(np.int64)
    assert deg.max() <= 700  # expected degree capped at 500
    h = gb.Graph.chung_lu(20000, 10000, 2.3, 2.3, 2.0e5, 500.0, 500.0, 5)
    assert np.array_equal(g.neighbors, h.neighbors)


def test_mix_graph_ranges():
    for i in (0, 1, 2, 3, 4095):
        g = gb.Graph.generate("mix_graph", i)
        assert 1 <= g.m and g.u_num <= 5000 and g.l_num <= 5000
        assert len(set(map(tuple, g.edges.reshape(-1, 2).tolist()))) == g.m
    with pytest.raises(gb.BitrussError):
        gb.Graph.generate("mix_graph", 4096)


@pytest.mark.slow
def test_powerlaw_2m_workload():
    g = gb.Graph.generate("powerlaw_2m")
    assert 1.9e6 < g.m < 2.1e6
    assert g.degrees.max() <= 5000
