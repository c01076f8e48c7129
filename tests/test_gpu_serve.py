"""The serving loop (bt_serve_*) on the GPU: requests submitted while
others run are batched, large ones are batched on a lane of their own, and
every result equals the plain call on the same graph (alone or in the
mix)."""
from __future__ import annotations

import sys
import threading
import time

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from tests import graphs as G


def say(msg):
    print(f"[gpu-serve] {msg}", flush=True)
    sys.stdout.flush()


def plain(dev, g, algo):
    dg = dev.upload(g)
    try:
        if algo == "msp":
            r, lv = dg.msp()
            return r, {"levels": lv}
        if algo == "hidx":
            r, rd = dg.hidx()
            return r, {"rounds": rd}
        if algo == "ebfc":
            r, t, smax = dg.ebfc()
            return r, {"total": t, "support_max": smax}
        return None, {"total": dg.bfc()}
    finally:
        dg.free()


def check(dev, g, algo, arr, scal):
    want, ws = plain(dev, g, algo)
    if want is not None:
        assert np.array_equal(arr, want)
    for k, v in ws.items():
        assert scal[k] == v, (k, scal[k], v)


@pytest.mark.gpu
def test_serve_mixed_stream(dev):
    """Small requests arrive from another thread while batches run; three
    requests are large (a low large_edges routes them to the large lane),
    with two algorithms, so that the large lane runs at least two batches."""
    mix = [gb.Graph.generate("mix_graph", i) for i in range(48)]
    cut = 50000
    gs = [g for g in mix if g.m < cut]
    gs += [G.from_pairs(G.AW_EDGES), G.from_pairs(G.QUIRK7_EDGES), gb.Graph.from_edges([], [])]
    large = [g for g in mix if g.m >= cut][:3]
    assert len(large) == 3
    algos = ["msp", "hidx", "ebfc", "bfc"]
    with gb.Server(large_edges=cut, max_batch=16) as srv:
        tickets = []

        def feed():
            for i, g in enumerate(gs):
                tickets.append((srv.submit(g, algos[i % 4]), g, algos[i % 4]))
                if i % 7 == 0:
                    time.sleep(0.002)

        th = threading.Thread(target=feed)
        th.start()
        big = [(srv.submit(g, a), g, a) for g, a in zip(large, ["msp", "hidx", "msp"])]
        th.join()
        results = [(srv.wait(t), g, a) for t, g, a in tickets + big]
        st = srv.stats()
    say(f"stats {st}")
    assert st["completed"] == len(gs) + len(large) and st["failed"] == 0
    assert st["large_runs"] >= 2 and st["batches"] >= 2
    for (arr, scal, lat), g, algo in results:
        assert lat >= 0
        check(dev, g, algo, arr, scal)


@pytest.mark.gpu
def test_serve_alone_equals_in_mix(dev):
    g = G.quirk_graph(G.QUIRK_SEEDS[0])
    with gb.Server() as srv:
        alone = srv.wait(srv.submit(g, "msp"))
    others = [gb.Graph.generate("mix_graph", i) for i in range(12)]
    with gb.Server(max_batch=64) as srv:
        ts = [srv.submit(o, "msp") for o in others[:6]]
        t = srv.submit(g, "msp")
        ts += [srv.submit(o, "msp") for o in others[6:]]
        mixed = srv.wait(t)
        for x in ts:
            srv.wait(x)
    assert np.array_equal(alone[0], mixed[0]) and alone[1] == mixed[1]


@pytest.mark.gpu
def test_serve_budget(dev):
    g = gb.Graph.generate("mix_graph", 7)
    with gb.Server(budget_bytes=1 << 20) as srv:  # below any request's estimate
        with pytest.raises(gb.BitrussError):
            srv.submit(g, "msp")
    # a budget that fits one request at a time still completes them all
    with gb.Server(budget_bytes=1 << 30, max_batch=8) as srv:
        ts = [srv.submit(gb.Graph.generate("mix_graph", i), "msp") for i in range(6)]
        outs = [srv.wait(t) for t in ts]
        assert srv.stats()["peak_bytes"] <= 1 << 30
    for i, (arr, scal, _) in enumerate(outs):
        check(dev, gb.Graph.generate("mix_graph", i), "msp", arr, scal)
