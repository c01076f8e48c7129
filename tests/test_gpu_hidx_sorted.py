"""h-index rounds with the larger blooms' values kept in sorted chunks
(BT_OPT_HIDX_SORTED), through the C-ABI: the results and round counts equal
those of the rounds that stream every value and the CPU oracle's
(compute_bitruss_hindex, cpu_bitruss.cpp:200-239; the reference's GPU loop,
bitruss_hidx.cu:51-256), on graphs whose blooms span one chunk of 2,048
values and several and with edges in thousands of blooms (evaluated by a
whole workgroup), with the rounds one by one and as captured graphs, in
batches, and on the 2M-edge golden."""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from oracle import oracle
from tests import graphs as G

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg: str) -> None:
    print(f"[gpu-hidx-sorted] {msg}", flush=True)
    sys.stdout.flush()


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def twin_hubs(seed: int, p_hub, nv: int, nu: int, p: float) -> gb.Graph:
    """Upper hubs 0..len(p_hub)-1, hub h adjacent to each of nv lower
    vertices with probability p_hub[h] (two 0.9-hubs over 3,000 share ~2,430
    middles: blooms of two chunks), plus nu upper vertices with random edges
    of density p (small blooms and varied values)."""
    rng = np.random.default_rng(seed)
    pairs = [(h, j) for h, ph in enumerate(p_hub) for j in range(nv) if rng.random() < ph]
    mask = rng.random((nu, nv)) < p
    pairs += [(len(p_hub) + int(i), int(j)) for i, j in np.argwhere(mask)]
    return G.from_pairs(pairs)


def two_sided_hubs(seed: int, nu: int, nv: int, hu: int, pu: float, hv: int, pv: float, p: float) -> gb.Graph:
    """hu upper hubs adjacent to each lower vertex with probability pu, hv
    lower hubs adjacent to each upper vertex with probability pv, and random
    edges of density p: the edges between two hubs lie in thousands of
    blooms (the heavy edges, a workgroup each)."""
    rng = np.random.default_rng(seed)
    pairs = set()
    for h in range(hu):
        pairs |= {(h, int(j)) for j in np.nonzero(rng.random(nv) < pu)[0]}
    for h in range(hv):
        pairs |= {(int(i), h) for i in np.nonzero(rng.random(nu) < pv)[0]}
    pairs |= {(int(i), int(j)) for i, j in np.argwhere(rng.random((nu, nv)) < p)}
    return G.from_pairs(sorted(pairs))


def _graphs():
    gs = [("two_sided_hubs", two_sided_hubs(7, 3000, 3000, 3, 0.85, 3, 0.8, 0.002)),
          ("twin_hubs1", twin_hubs(1, [0.9, 0.9, 0.85, 0.3], 3000, 200, 0.01)),
          ("twin_hubs2", twin_hubs(2, [0.95, 0.7, 0.5, 0.5, 0.1], 2600, 300, 0.02)),
          ("hub", G.hub_graph(5, 6000, 20000, 400, 7000)),
          ("K3x2100", G.complete(3, 2100)),
          ("K40x40", G.complete(40, 40)),
          ("aw", G.from_pairs(G.AW_EDGES))]
    gs += [(f"quirk{s}", G.quirk_graph(s)) for s in G.QUIRK_SEEDS[:2]]
    gs += [(f"rand{s}", G.random_bipartite(s, 80, 120, 0.2)) for s in range(2)]
    gs += [(f"mix{i}", gb.Graph.generate("mix_graph", i)) for i in (0, 17)]
    return gs


@pytest.mark.gpu
def test_sorted_equals_stream_and_oracle(dev):
    with gb.Device(dev.device) as d:
        for name, g in _graphs():
            want, wrounds, _ = oracle.hidx(g, 8)
            dg = d.upload(g)
            for srt in (1, 0):
                for k in (0, 8):
                    d.set_hidx_sorted(srt)
                    d.set_hidx_graph_rounds(k)
                    say(f"{name} m={g.m}: sorted {srt}, {k} rounds per graph")
                    res, rounds = dg.hidx()
                    assert np.array_equal(res, want) and rounds == wrounds, (name, srt, k, rounds, wrounds)
            dg.free()


@pytest.mark.gpu
def test_sorted_in_batches(dev):
    """A batch union (twin hubs among mix graphs) with sorted chunks: each
    graph's results and round count equal its plain run that streams."""
    graphs = [gb.Graph.generate("mix_graph", i) for i in range(6)]
    graphs.insert(2, twin_hubs(3, [0.9, 0.9, 0.4], 2500, 150, 0.02))
    graphs.insert(4, two_sided_hubs(8, 2500, 2500, 2, 0.9, 2, 0.8, 0.002))
    with gb.Device(dev.device) as d:
        d.set_hidx_sorted(1)
        b = d.upload_batch(graphs)
        say(f"batch of {len(graphs)} graphs, sorted chunks")
        outs, results = b.run("hidx")
        b.free()
        d.set_hidx_sorted(0)
        for i, g in enumerate(graphs):
            dg = d.upload(g)
            want, wrounds = dg.hidx()
            dg.free()
            assert np.array_equal(outs[i], want) and results[i]["rounds"] == wrounds, i


@pytest.mark.gpu
def test_sorted_powerlaw_2m_golden(dev):
    """The 2M-edge graph with and without sorted chunks: the golden digest
    and the oracle's Jacobi round count both ways."""
    gold = json.load(open(os.path.join(GOLDEN, "powerlaw_2m.json")))
    g = gb.Graph.generate("powerlaw_2m")
    with gb.Device(dev.device) as d:
        dg = d.upload(g)
        for srt in (1, 0):
            d.set_hidx_sorted(srt)
            t0 = time.perf_counter()
            res, rounds = dg.hidx()
            say(f"powerlaw_2m hidx sorted {srt}: {rounds} rounds, {time.perf_counter() - t0:.2f} s")
            assert digest(res) == gold["hidx_sha256"] and rounds == gold["hidx_rounds_jacobi"], srt
        dg.free()
