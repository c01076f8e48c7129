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


@pytest.mark.gpu
def test_serve_poisson_stream_as_configured(dev):
    """The serving configuration: a Poisson stream (seed 7, 50 requests/s)
    of mix graphs with 2% large Chung-Lu requests (m in [1M, 5M]), here its
    first 40 requests, which hold three large ones (the first 1.56M edges).
    Every result equals the plain call on the same graph, and the first
    large request's equals the oracle's digest."""
    import hashlib
    import json
    import os

    rng = np.random.default_rng(7)
    n_all = 400
    large = rng.random(n_all) < 0.02
    idx = rng.integers(0, 4096, n_all)
    n = 40
    reqs = [gb.Graph.generate("large_request", int(i)) if big else gb.Graph.generate("mix_graph", int(i))
            for i, big in zip(idx[:n], large[:n])]
    assert int(large[:n].sum()) == 3 and all(reqs[p].m >= 1_000_000 for p in np.nonzero(large[:n])[0])
    gaps = rng.exponential(1.0 / 50.0, n)
    with gb.Server(large_edges=1_000_000) as srv:
        tickets = [None] * n

        def feed():
            t = time.perf_counter()
            for i, g in enumerate(reqs):
                t += gaps[i]
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i] = srv.submit(g, "msp")

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        results = [srv.wait(t) for t in tickets]
        st = srv.stats()
    lat = np.array([r[2] for r in results])
    say(f"stream of {n}: p50 {np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms, {st}")
    assert st["completed"] == n and st["failed"] == 0 and st["large_runs"] >= 1
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "large_request_1204.json")))
    first = int(np.nonzero(large[:n])[0][0])
    assert int(idx[first]) == gold["index"] and reqs[first].m == gold["m"]
    arr = results[first][0]
    assert hashlib.sha256(np.ascontiguousarray(arr, dtype="<u8").tobytes()).hexdigest() == gold["msp_sha256"]
    assert results[first][1]["levels"] == gold["msp_levels"]
    for (arr, scal, _), g in zip(results, reqs):
        check(dev, g, "msp", arr, scal)
