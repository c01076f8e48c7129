"""The h-index rounds as captured HIP graphs (BT_OPT_HIDX_GRAPH_ROUNDS),
through the C-ABI: the results and the round count equal the rounds run one
by one with a read-back after each (the reference's loop,
bitruss_hidx.cu:219-231) and the CPU oracle (cpu_bitruss.cpp:200-239), for
graph lengths that end the loop inside a graph, on its last round and on
its first; each call replays its graph several times, with host
synchronisation between the replays."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from oracle import oracle
from tests import graphs as G

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu-hidx-graph] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
def test_graph_rounds_equal_rounds_one_by_one(dev):
    graphs = [("aw", G.from_pairs(G.AW_EDGES)), ("quirk7", G.from_pairs(G.QUIRK7_EDGES)),
              ("K9x9", G.complete(9, 9)), ("star", G.from_pairs([(0, i) for i in range(5)])),
              ("hub", G.hub_graph(5, 6000, 20000, 400, 7000))]
    graphs += [(f"rand{s}", G.random_bipartite(s, 60, 80, 0.15)) for s in range(3)]
    graphs += [(f"mix{i}", gb.Graph.generate("mix_graph", i)) for i in (0, 2, 17)]
    with gb.Device(dev.device) as d:
        for name, g in graphs:
            want, wrounds, _ = oracle.hidx(g, 8)
            dg = d.upload(g)
            for k in (0, 2, 8, 64):
                d.set_hidx_graph_rounds(k)
                say(f"{name} m={g.m}: {k} rounds per graph")
                res, rounds = dg.hidx()
                assert np.array_equal(res, want) and rounds == wrounds, (name, k, rounds, wrounds)
            dg.free()


@pytest.mark.gpu
def test_graph_rounds_in_batches(dev):
    """A batch union with graph rounds: every graph's results and round
    count equal its plain run with rounds one by one (the plain runs are
    checked against the oracle above and in the parity tests), and the
    oracle's on the first three graphs."""
    graphs = [gb.Graph.generate("mix_graph", i) for i in range(12)]
    with gb.Device(dev.device) as d:
        d.set_hidx_graph_rounds(4)
        b = d.upload_batch(graphs)
        say(f"batch of {len(graphs)} mix graphs, 4 rounds per graph")
        outs, results = b.run("hidx")
        b.free()
        d.set_hidx_graph_rounds(0)
        for i, g in enumerate(graphs):
            dg = d.upload(g)
            want, wrounds = dg.hidx()
            dg.free()
            assert np.array_equal(outs[i], want) and results[i]["rounds"] == wrounds, i
            if i < 3:
                w2, r2, _ = oracle.hidx(g, 8)
                assert np.array_equal(w2, want) and r2 == wrounds, i
