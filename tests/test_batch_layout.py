"""The batch layout (disjoint union of many graphs, bt_graph_union) on the
CPU: the union is a valid CSR, and the oracle's results on the union,
sliced per graph, equal its results on each graph alone — for ebfc, hidx
and msp (whose attribution rule compares edge ids and sides, so this checks
that the union keeps every such comparison inside a graph)."""
from __future__ import annotations

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from oracle import oracle
from tests import graphs as G


def batch_graphs():
    gs = [G.from_pairs(G.AW_EDGES), G.from_pairs(G.QUIRK7_EDGES), G.complete(4, 5),
          gb.Graph.from_edges([], [])]
    gs += [G.quirk_graph(s) for s in G.QUIRK_SEEDS + G.ORDER_SEEDS]
    gs += [G.random_bipartite(900 + i, 30 + 7 * i, 25 + 5 * i, 0.15) for i in range(4)]
    gs += [gb.Graph.generate("mix_graph", i) for i in (0, 1)]
    return gs


def test_union_is_a_valid_csr():
    gs = batch_graphs()
    u, e_off = gb.Graph.union(gs)
    u.check_graph()
    assert u.m == sum(g.m for g in gs) and u.n == sum(g.n for g in gs)
    assert u.u_num == sum(g.u_num for g in gs)
    assert list(np.diff(e_off)) == [g.m for g in gs]


def test_union_slices_equal_plain_results():
    gs = batch_graphs()
    u, e_off = gb.Graph.union(gs)
    us, ut, _, _ = oracle.ebfc(u, 4)
    uh, _, _ = oracle.hidx(u, 4)
    um, ul, _ = oracle.msp(u, 4)
    tot = 0
    for i, g in enumerate(gs):
        a, b = int(e_off[i]), int(e_off[i + 1])
        s, t, smax, _ = oracle.ebfc(g, 2)
        assert np.array_equal(us[a:b], s), i
        tot += t
        h, _, _ = oracle.hidx(g, 2)
        assert np.array_equal(uh[a:b], h), i
        r, lv, _ = oracle.msp(g, 2)
        assert np.array_equal(um[a:b], r), i
        assert lv == (int(r.max()) if g.m else 0)
    assert ut == tot


def test_union_rejects_overflow_and_bad_graphs():
    g = G.from_pairs(G.AW_EDGES)
    g.degrees = g.degrees.copy()
    g.degrees[1] += 1  # inconsistent with the offsets
    with pytest.raises(gb.BitrussError):
        gb.Graph.union([G.from_pairs(G.AW_EDGES), g])


def test_priority_wedges_match_the_oracle_count():
    """bt_graph_wedges (which sizes the unions of a batch) equals the wedge
    count of the oracle's butterfly counting loops."""
    for g in batch_graphs():
        _, st = oracle.bfc(g, 2)
        assert g.priority_wedges() == st["w_cnt"]
