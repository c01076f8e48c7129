"""Pins the CPU oracle (oracle/bt_oracle.c) against known answers, closed
forms and the independent restatements in tests/refimpl.py."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from tests import graphs as G
from tests import refimpl


def test_aw_known_answers():
    g = G.from_pairs(G.AW_EDGES)
    sup, total, smax, st = oracle.ebfc(g)
    assert total == 4 and smax == 3
    assert sup.tolist() == G.AW_SUPPORT
    assert oracle.bfc(g)[0] == 4
    h, rounds, _ = oracle.hidx(g)
    assert h.tolist() == G.AW_BITRUSS
    r, levels, _ = oracle.msp(g)
    assert r.tolist() == G.AW_BITRUSS and levels == 2
    assert oracle.peel_seq(g).tolist() == G.AW_BITRUSS


def test_quirk7_msp_differs_from_bitruss():
    g = G.from_pairs(G.QUIRK7_EDGES)
    sup, total, _, _ = oracle.ebfc(g)
    assert total == 2 and sup.tolist() == G.QUIRK7_SUPPORT
    assert oracle.hidx(g)[0].tolist() == G.QUIRK7_HIDX
    assert oracle.peel_seq(g).tolist() == G.QUIRK7_HIDX
    r, levels, _ = oracle.msp(g)
    assert r.tolist() == G.QUIRK7_MSP and levels == 2


@pytest.mark.parametrize("a,b", [(2, 2), (2, 5), (3, 4), (5, 5), (6, 3), (8, 7)])
def test_complete_bipartite_closed_forms(a, b):
    g = G.complete(a, b)
    k = (a - 1) * (b - 1)
    sup, total, smax, _ = oracle.ebfc(g)
    assert total == (a * (a - 1) // 2) * (b * (b - 1) // 2)
    assert set(sup.tolist()) == {k} and smax == k
    assert set(oracle.hidx(g)[0].tolist()) == {k}
    r, levels, _ = oracle.msp(g)
    assert set(r.tolist()) == {k} and levels == k


def test_zero_butterfly_graphs():
    star = G.from_pairs([(0, j) for j in range(20)])
    path = G.from_pairs([(i, i) for i in range(10)] + [(i, i + 1) for i in range(9)])
    for g in (star, path):
        sup, total, smax, _ = oracle.ebfc(g)
        assert total == 0 and smax == 0 and not sup.any()
        r, levels, _ = oracle.msp(g)
        assert not r.any() and levels == 0
        assert not oracle.hidx(g)[0].any()


@pytest.mark.parametrize("seed", range(12))
def test_counting_matches_pair_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    g = G.random_bipartite(seed, int(rng.integers(4, 14)), int(rng.integers(4, 14)), float(rng.uniform(0.2, 0.8)))
    total, sup = refimpl.brute_support(g)
    osup, ototal, osmax, _ = oracle.ebfc(g, threads=3)
    assert ototal == total
    assert osup.tolist() == sup
    assert osmax == (max(sup) if sup else 0)
    assert oracle.bfc(g, threads=2)[0] == total


@pytest.mark.parametrize("seed", range(12))
def test_hidx_equals_textbook_peel(seed):
    rng = np.random.default_rng(200 + seed)
    g = G.random_bipartite(50 + seed, int(rng.integers(4, 12)), int(rng.integers(4, 12)),
                           float(rng.uniform(0.3, 0.8)))
    truth = refimpl.textbook_peel(g)
    assert oracle.hidx(g, threads=3)[0].tolist() == truth
    assert oracle.peel_seq(g).tolist() == truth


@pytest.mark.parametrize("seed", range(16))
def test_msp_matches_butterfly_centric_rule(seed):
    rng = np.random.default_rng(300 + seed)
    g = G.random_bipartite(80 + seed, int(rng.integers(4, 12)), int(rng.integers(4, 12)),
                           float(rng.uniform(0.3, 0.85)))
    want, wlev = refimpl.msp_by_butterflies(g)
    got, levels, _ = oracle.msp(g, threads=3)
    assert got.tolist() == want
    assert levels == wlev


@pytest.mark.parametrize("seed", G.QUIRK_SEEDS + G.ORDER_SEEDS)
def test_msp_rule_on_quirk_graphs(seed):
    """Graphs where the attribution rule changes the answer: the oracle must
    follow the rule, not the true bitruss numbers."""
    g = G.quirk_graph(seed)
    want, wlev = refimpl.msp_by_butterflies(g)
    got, levels, _ = oracle.msp(g, threads=2)
    assert got.tolist() == want and levels == wlev
    if seed in G.QUIRK_SEEDS:
        assert got.tolist() != refimpl.textbook_peel(g)


@pytest.mark.parametrize("seed", G.QUIRK_SEEDS[:4] + [0, 1])
def test_msp_level_prefix_matches_butterfly_centric_rule(seed):
    """or_msp_levels (the state after levels 1..L) against the
    butterfly-centric restatement stopped at the same level, and the whole
    decomposition once L reaches the last level."""
    g = G.quirk_graph(seed) if seed in G.QUIRK_SEEDS else G.random_bipartite(
        500 + seed, 10, 9, 0.6)
    full, _ = refimpl.msp_by_butterflies(g)
    top = max(full) if full else 0
    for bound in sorted({1, 2, 3, max(1, top // 2), top, top + 3}):
        want, _ = refimpl.msp_by_butterflies(g, max_level=bound)
        got, complete, _ = oracle.msp_levels(g, 2, bound)
        assert got.tolist() == want, bound
        assert complete == (bound >= top)


@pytest.mark.parametrize("variant", ["always", "swap_pq"])
def test_planted_mistake_is_caught(variant):
    """A restatement with a plausible slip in the single-survivor rule must
    disagree with the oracle on some random case, so the msp pin can see
    such a mistake."""
    caught = False
    for seed in G.QUIRK_SEEDS + G.ORDER_SEEDS:
        g = G.quirk_graph(seed)
        if refimpl.msp_by_butterflies(g, variant)[0] != oracle.msp(g)[0].tolist():
            caught = True
            break
    assert caught


def test_results_independent_of_thread_count():
    g = G.random_bipartite(7, 30, 30, 0.4)
    a = oracle.msp(g, threads=1)
    b = oracle.msp(g, threads=6)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]
    h1 = oracle.hidx(g, threads=1)[0]
    h6 = oracle.hidx(g, threads=6)[0]
    assert np.array_equal(h1, h6)
    s1 = oracle.ebfc(g, threads=1)
    s6 = oracle.ebfc(g, threads=6)
    assert np.array_equal(s1[0], s6[0]) and s1[1] == s6[1]


def test_work_counters_on_aw():
    g = G.from_pairs(G.AW_EDGES)
    _, _, _, st = oracle.ebfc(g)
    # priority wedges of aw.graph, enumerated by hand: root 6 (degree 5)
    # has 5 (via 0, 1, 2, 2, 4), root 5 has 1 (5-2-7), root 2 has 1 (2-7-4)
    assert st["w_cnt"] == 7


def test_turn_graph_known_answers():
    """The hand-derived turn graph (tests/graphs.py): msp decrements e28 in
    {0,1}x{3,4}, bitruss_mspp does not (its pivot e0 is turned around), and
    the true bitruss numbers are msp's here."""
    g = G.from_pairs(G.TURN_EDGES)
    assert g.m == 32
    assert oracle.msp(g)[0].tolist() == G.TURN_MSP
    r, levels, _ = oracle.mspp(g)
    assert r.tolist() == G.TURN_MSPP and levels == 2
    assert oracle.peel_seq(g).tolist() == G.TURN_MSP
    assert refimpl.msp_by_butterflies(g, turned=refimpl.turned_edges(g))[0] == G.TURN_MSPP


@pytest.mark.parametrize("seed", list(range(6)))
def test_mspp_matches_butterfly_centric_rule(seed):
    """or_mspp against the butterfly-centric restatement with the turned
    pivots (derived, not retyped: only the pivot opposite the survivor sees
    its orientation) on graphs with a hub of turned edges, and quirk seeds."""
    for g in (G.turn_graph(seed), G.quirk_graph(G.QUIRK_SEEDS[seed % len(G.QUIRK_SEEDS)])):
        want, wl = refimpl.msp_by_butterflies(g, turned=refimpl.turned_edges(g))
        got, levels, _ = oracle.mspp(g, 2)
        assert got.tolist() == want and levels == wl


def test_mspp_equals_msp_without_turned_edges():
    for g in (G.from_pairs(G.AW_EDGES), G.from_pairs(G.QUIRK7_EDGES), G.complete(4, 6),
              G.random_bipartite(3, 20, 20, 0.3)):
        assert not refimpl.turned_edges(g)
        assert np.array_equal(oracle.mspp(g, 2)[0], oracle.msp(g, 2)[0])


def test_turn_slip_is_caught():
    """Ignoring the orientation (msp's rule) is a slip the turn graph shows."""
    g = G.from_pairs(G.TURN_EDGES)
    assert refimpl.msp_by_butterflies(g)[0] != oracle.mspp(g)[0].tolist()
