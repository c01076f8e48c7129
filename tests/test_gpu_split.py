"""Large blooms walked in chunks by every workgroup of the peel (a shared
large-bloom list fed by device-wide chunk counters).

SKIPPED: the first run of this path on MI355X hung in its first case (the
shared list taking every bloom above one middle) and its cause was not found
by reading the code; the path is taken out of the peel kernel, and these
tests stay as the specification for when it returns: with the shared list
taking every bloom above 1 or 64 middles, bt_msp, bt_mspp and bt_peel equal
the CPU oracle (cpu_bitruss.cpp:301-351, bitruss_msp.cu:600-722,
cpu_bitruss.cpp:34-122), and the 2M-edge decomposition and a batch union
equal their goldens and plain runs."""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from oracle import oracle
from tests import graphs as G

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="shared large-bloom walks hung on MI355X at their first run (cause not "
                                      "found); the path is out of the peel kernel")]

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg: str) -> None:
    print(f"[gpu-split] {msg}", flush=True)
    sys.stdout.flush()


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def graphs():
    out = [("aw", G.from_pairs(G.AW_EDGES)), ("quirk7", G.from_pairs(G.QUIRK7_EDGES)),
           ("K40x30", G.complete(40, 30)), ("K9x9", G.complete(9, 9))]
    out += [(f"quirk{s}", G.quirk_graph(s)) for s in G.QUIRK_SEEDS + G.ORDER_SEEDS]
    out += [(f"rand{s}", G.random_bipartite(s, 80, 90, 0.2)) for s in range(3)]
    out += [(f"mix{i}", gb.Graph.generate("mix_graph", i)) for i in (0, 2, 17)]
    out += [("hub", G.hub_graph(5, 6000, 20000, 400, 7000))]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("split", [1, 64])
def test_split_msp_matches_oracle(dev, split):
    with gb.Device(dev.device) as d:
        d.set_peel_split_middles(split)
        for name, g in graphs():
            say(f"msp {name} m={g.m} (split above {split})")
            dg = d.upload(g)
            res, levels = dg.msp()
            want, wlevels, _ = oracle.msp(g, 8)
            assert np.array_equal(res, want) and levels == wlevels, name
            r2, l2 = dg.mspp()
            w2, wl2, _ = oracle.mspp(g, 8)
            assert np.array_equal(r2, w2) and l2 == wl2, name
            r3, l3 = dg.peel()
            assert np.array_equal(r3, oracle.peel_seq(g)), name
            dg.free()


@pytest.mark.gpu
@pytest.mark.parametrize("split", [256, 4096])
def test_split_powerlaw_2m_golden(dev, split):
    with open(os.path.join(GOLDEN, "powerlaw_2m.json")) as f:
        gold = json.load(f)
    g = gb.Graph.generate("powerlaw_2m")
    with gb.Device(dev.device) as d:
        d.set_peel_split_middles(split)
        dg = d.upload(g)
        res, levels = dg.msp()
        assert digest(res) == gold["msp_sha256"] and levels == gold["msp_levels"]
        dg.free()


@pytest.mark.gpu
def test_split_batch_equals_plain(dev):
    gs = [gb.Graph.generate("mix_graph", i) for i in range(24)]
    with gb.Device(dev.device) as d:
        d.set_peel_split_middles(64)
        b = d.upload_batch(gs)
        outs, scal = b.run("msp")
        b.free()
        for i, g in enumerate(gs):
            want, wlevels, _ = oracle.msp(g, 8)
            assert np.array_equal(outs[i], want) and scal[i]["levels"] == wlevels, i
