"""Counting launch settings per graph-size class (bt_tune_*, bt_autotune)
on the GPU, through the C-ABI: every candidate setting counts the same
supports and total as the CPU oracle (cpu_bfc.cpp:24-203) on graphs that
use every root bucket (wave, workgroup and hub roots), and an autotuning
run on the 2M-edge graph keeps the golden supports."""
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
    print(f"[gpu-tune] {msg}", flush=True)
    sys.stdout.flush()


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


@pytest.mark.gpu
def test_every_count_setting_gives_the_same_supports(dev):
    graphs = [("hub", G.hub_graph(5, 6000, 20000, 400, 7000)), ("mix17", gb.Graph.generate("mix_graph", 17)),
              ("dense", G.random_bipartite(3, 300, 400, 0.2)), ("K40x30", G.complete(40, 30))]
    cands = gb.tune_candidates()
    assert len(cands) == 48
    with gb.Device(dev.device) as d:
        for name, g in graphs:
            want, wtotal, wmax, _ = oracle.ebfc(g, 8)
            cls = gb.tune_class(g.m)
            dg = d.upload(g)
            say(f"{name}: m={g.m}, class {cls}, {len(cands)} settings")
            for cfg in cands:
                d.tune_set(cls, cfg)
                sup, total, smax = dg.ebfc()
                assert np.array_equal(sup, want) and total == wtotal and smax == wmax, (name, cfg)
                assert dg.bfc() == wtotal, (name, cfg)
            dg.free()


@pytest.mark.gpu
def test_autotune_keeps_the_golden_supports(dev):
    with open(os.path.join(GOLDEN, "powerlaw_2m.json")) as f:
        gold = json.load(f)
    g = gb.Graph.generate("powerlaw_2m")
    cls = gb.tune_class(g.m)
    with gb.Device(dev.device) as d:
        assert d.tune_get(cls) == gb.tune_default(cls)
        dg = d.upload(g)
        best, best_ms, times = dg.dev.autotune(dg, reps=1)
        say(f"powerlaw_2m: class {cls}, best {best} at {best_ms:.2f} ms; default "
            f"{gb.tune_default(cls)} at {times[gb.tune_default(cls)]:.2f} ms")
        assert d.tune_get(cls) == best and best in times and best_ms == min(times.values())
        sup, total, _ = dg.ebfc()
        assert digest(sup) == gold["ebfc_sha256"] and total == gold["butterflies"]
        dg.free()
