"""Batches on the GPU (bt_batch_*): every graph's slice of a batch run, and
its per-graph scalars, equal the plain single-graph calls on the same device
(and the oracle on the small graphs)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from oracle import oracle
from tests import graphs as G


def say(msg):
    print(f"[gpu-batch] {msg}", flush=True)
    sys.stdout.flush()


def batch_graphs():
    gs = [G.from_pairs(G.AW_EDGES), G.from_pairs(G.QUIRK7_EDGES), G.complete(9, 9), G.complete(40, 30),
          gb.Graph.from_edges([], [])]
    gs += [G.quirk_graph(s) for s in G.QUIRK_SEEDS + G.ORDER_SEEDS]
    gs += [G.random_bipartite(950 + i, 40 + 9 * i, 30 + 7 * i, 0.12) for i in range(6)]
    gs += [G.hub_graph(5, 700, 900, 60, 900)]
    gs += [gb.Graph.generate("mix_graph", i) for i in range(24)]
    return gs


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["bfc", "ebfc", "hidx", "msp"])
def test_batch_equals_plain(dev, algo):
    gs = batch_graphs()
    say(f"{algo}: {len(gs)} graphs, m = {sum(g.m for g in gs)}")
    b = dev.upload_batch(gs)
    arrays, scal = b.run(algo)
    b.free()
    for i, g in enumerate(gs):
        dg = dev.upload(g)
        if algo == "bfc":
            assert scal[i]["total"] == dg.bfc(), i
        elif algo == "ebfc":
            sup, total, smax = dg.ebfc()
            assert np.array_equal(arrays[i], sup), i
            assert scal[i]["total"] == total and scal[i]["support_max"] == smax, i
        elif algo == "hidx":
            h, rounds = dg.hidx()
            assert np.array_equal(arrays[i], h), i
            assert scal[i]["rounds"] == rounds, (i, scal[i]["rounds"], rounds)
        else:
            r, levels = dg.msp()
            assert np.array_equal(arrays[i], r), i
            assert scal[i]["levels"] == levels, i
        dg.free()


@pytest.mark.gpu
def test_batch_small_graphs_match_oracle(dev):
    gs = [G.from_pairs(G.AW_EDGES), G.from_pairs(G.QUIRK7_EDGES)] + [G.quirk_graph(s) for s in G.QUIRK_SEEDS]
    b = dev.upload_batch(gs)
    ms, _ = b.run("msp")
    hs, hsc = b.run("hidx")
    b.free()
    for i, g in enumerate(gs):
        r, _, _ = oracle.msp(g, 2)
        h, rounds, _ = oracle.hidx(g, 2)
        assert np.array_equal(ms[i], r), i
        assert np.array_equal(hs[i], h), i
        assert hsc[i]["rounds"] == rounds, i
    assert list(ms[0]) == G.AW_BITRUSS and list(ms[1]) == G.QUIRK7_MSP


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["ebfc", "hidx", "msp"])
def test_batch_in_many_unions(dev, algo):
    """A small per-union wedge bound splits the batch into many unions (a
    graph above the bound gets one of its own); results do not change."""
    gs = batch_graphs()
    dev.set_batch_lanes(1)
    try:
        whole = dev.upload_batch(gs)
    finally:
        dev.set_batch_lanes(2)
    assert whole.chunks() == 1
    a1, s1 = whole.run(algo)
    whole.free()
    dev.set_batch_wedges(20000)
    try:
        split = dev.upload_batch(gs)
    finally:
        dev.set_batch_wedges(0)
    assert split.chunks() > 5
    a2, s2 = split.run(algo)
    split.free()
    assert s1 == s2
    for x, y in zip(a1, a2):
        assert np.array_equal(x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["hidx", "msp"])
def test_batch_lanes_agree(dev, algo):
    """One lane (one union on all CUs), three and eight lanes running side
    by side on CU shares: the same per-graph results."""
    gs = batch_graphs()
    got = {}
    try:
        for lanes in (1, 3, 8):
            dev.set_batch_lanes(lanes)
            b = dev.upload_batch(gs)
            say(f"{algo}: {lanes} lanes, {b.chunks()} unions")
            assert b.chunks() >= lanes
            got[lanes] = b.run(algo)
            b.free()
    finally:
        dev.set_batch_lanes(2)
    a1, s1 = got[1]
    for lanes in (3, 8):
        a2, s2 = got[lanes]
        assert s1 == s2, lanes
        for x, y in zip(a1, a2):
            assert np.array_equal(x, y), lanes


@pytest.mark.gpu
def test_batch_level_modes_agree(dev):
    gs = batch_graphs()
    b = dev.upload_batch(gs)
    union, s2 = b.run("msp")
    dev.set_batch_levels("per_graph")
    try:
        per_graph, s1 = b.run("msp")
    finally:
        dev.set_batch_levels("union")
    b.free()
    assert s1 == s2
    for x, y in zip(per_graph, union):
        assert np.array_equal(x, y)


@pytest.mark.gpu
def test_batch_msp_paths_agree(dev):
    gs = [gb.Graph.generate("mix_graph", i) for i in range(8)]
    b = dev.upload_batch(gs)
    dev.set_msp_path("merge")
    try:
        merge, _ = b.run("msp")
    finally:
        dev.set_msp_path("auto")
    index, _ = b.run("msp")
    b.free()
    for a, c in zip(merge, index):
        assert np.array_equal(a, c)


@pytest.mark.gpu
def test_batch_full_mix_equals_plain(dev):
    """The batched configuration at full size: all 4,096 graphs of the mix
    (seeds 1000..5095, 86.7 M edges) decomposed as batches, every graph's
    msp results and level count against the plain call on that graph."""
    k = 4096
    gs = [gb.Graph.generate("mix_graph", i) for i in range(k)]
    b = dev.upload_batch(gs)
    try:
        say(f"full mix: {sum(g.m for g in gs)} edges in {b.chunks()} unions")
        arrays, scal = b.run("msp")
    finally:
        b.free()
    same = 0
    for i, g in enumerate(gs):
        dg = dev.upload(g)
        r, lv = dg.msp()
        dg.free()
        same += int(np.array_equal(r, arrays[i]) and lv == scal[i]["levels"])
        if i % 1024 == 1023:
            say(f"compared {i + 1} graphs, {same} identical")
    assert same == k
