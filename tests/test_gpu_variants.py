"""The reference's two further decompositions on the GPU, through the C-ABI:
bt_mspp (bitruss_mspp, bitruss_msp.cu:600-722, whose process_edge turns some
pivots around) against the CPU oracle's or_mspp, and bt_peel (the results of
bitruss_peeling_cpu, cpu_bitruss.cpp:34-122: true bitruss numbers) against
the textbook peel, the h-index fixpoint and the 2M-edge golden."""
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

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg: str) -> None:
    print(f"[gpu-variants] {msg}", flush=True)
    sys.stdout.flush()


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def small_graphs():
    out = [("aw", G.from_pairs(G.AW_EDGES)), ("quirk7", G.from_pairs(G.QUIRK7_EDGES)),
           ("turn", G.from_pairs(G.TURN_EDGES)), ("K6x9", G.complete(6, 9))]
    out += [(f"quirk{s}", G.quirk_graph(s)) for s in G.QUIRK_SEEDS + G.ORDER_SEEDS]
    out += [(f"turn{s}", G.turn_graph(s)) for s in range(4)]
    out += [(f"mix{i}", gb.Graph.generate("mix_graph", i)) for i in (0, 2, 17)]
    out += [("hub", G.hub_graph(5, 6000, 20000, 400, 7000))]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("c2cap", [128, 0])
def test_mspp_matches_oracle(dev, c2cap):
    with gb.Device(dev.device) as d:
        d.set_peel_c2cap(c2cap)
        for name, g in small_graphs():
            say(f"mspp {name} m={g.m} (c2cap {c2cap})")
            dg = d.upload(g)
            res, levels = dg.mspp()
            dg.free()
            want, wlevels, _ = oracle.mspp(g, 8)
            assert np.array_equal(res, want) and levels == wlevels, name
            if name == "turn":
                assert res.tolist() == G.TURN_MSPP


@pytest.mark.gpu
@pytest.mark.parametrize("c2cap", [128, 0])
def test_peel_gives_true_bitruss_numbers(dev, c2cap):
    with gb.Device(dev.device) as d:
        d.set_peel_c2cap(c2cap)
        for name, g in small_graphs():
            say(f"peel {name} m={g.m} (c2cap {c2cap})")
            dg = d.upload(g)
            res, kmax = dg.peel()
            dg.free()
            want = oracle.peel_seq(g)
            assert np.array_equal(res, want) and kmax == (int(want.max()) if g.m else 0), name


@pytest.mark.gpu
def test_peel_and_mspp_on_powerlaw_2m(dev):
    """The 2M-edge graph: bt_peel equals the golden h-index digest (the true
    bitruss numbers), and bt_mspp the oracle's or_mspp."""
    gold = json.load(open(os.path.join(GOLDEN, "powerlaw_2m.json")))
    g = gb.Graph.generate("powerlaw_2m")
    dg = dev.upload(g)
    try:
        say("peel powerlaw_2m")
        res, kmax = dg.peel()
        assert digest(res) == gold["hidx_sha256"] and kmax == gold["hidx_kmax"]
        say("mspp powerlaw_2m")
        res, levels = dg.mspp()
    finally:
        dg.free()
    if "mspp_sha256" in gold:
        assert digest(res) == gold["mspp_sha256"] and levels == gold["mspp_levels"]
    else:
        want, wlevels, _ = oracle.mspp(g, 16)
        assert np.array_equal(res, want) and levels == wlevels
