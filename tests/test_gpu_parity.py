"""Parity of the HIP path (through the C-ABI) with the CPU oracle:
bit-exact uint64 totals, supports and bitruss numbers."""
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
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def check_all(dev, g, name, threads=8, algos=("bfc", "ebfc", "msp", "hidx")):
    say(f"{name}: m={g.m} n={g.n}")
    dg = dev.upload(g)
    try:
        osup, ototal, osmax, _ = oracle.ebfc(g, threads)
        if "bfc" in algos:
            assert dg.bfc() == ototal, f"{name}: bfc total"
        if "ebfc" in algos:
            sup, total, smax = dg.ebfc()
            assert total == ototal and smax == osmax, f"{name}: ebfc totals"
            assert np.array_equal(sup, osup), f"{name}: ebfc supports"
        if "msp" in algos:
            res, levels = dg.msp()
            ores, olevels, _ = oracle.msp(g, threads)
            assert np.array_equal(res, ores), f"{name}: msp results"
            assert levels == olevels, f"{name}: msp levels"
        if "hidx" in algos:
            h, _ = dg.hidx()
            oh, _, _ = oracle.hidx(g, threads)
            assert np.array_equal(h, oh), f"{name}: hidx results"
    finally:
        dg.free()


def test_aw_known_answers(dev):
    g = G.from_pairs(G.AW_EDGES)
    dg = dev.upload(g)
    assert dg.bfc() == 4
    sup, total, smax = dg.ebfc()
    assert sup.tolist() == G.AW_SUPPORT and total == 4 and smax == 3
    res, levels = dg.msp()
    assert res.tolist() == G.AW_BITRUSS and levels == 2
    h, _ = dg.hidx()
    assert h.tolist() == G.AW_BITRUSS
    dg.free()


def test_quirk7(dev):
    g = G.from_pairs(G.QUIRK7_EDGES)
    dg = dev.upload(g)
    res, levels = dg.msp()
    assert res.tolist() == G.QUIRK7_MSP and levels == 2
    h, _ = dg.hidx()
    assert h.tolist() == G.QUIRK7_HIDX
    dg.free()


@pytest.mark.parametrize("a,b", [(2, 2), (3, 7), (9, 9), (40, 30)])
def test_complete_bipartite(dev, a, b):
    g = G.complete(a, b)
    k = (a - 1) * (b - 1)
    dg = dev.upload(g)
    assert dg.bfc() == (a * (a - 1) // 2) * (b * (b - 1) // 2)
    sup, _, _ = dg.ebfc()
    assert set(sup.tolist()) == {k}
    res, levels = dg.msp()
    assert set(res.tolist()) == {k} and levels == k
    h, _ = dg.hidx()
    assert set(h.tolist()) == {k}
    dg.free()


def test_empty_and_butterfly_free(dev):
    for pairs in ([(0, j) for j in range(50)], [(i, i) for i in range(30)] + [(i, i + 1) for i in range(29)]):
        g = G.from_pairs(pairs)
        dg = dev.upload(g)
        assert dg.bfc() == 0
        sup, total, smax = dg.ebfc()
        assert total == 0 and smax == 0 and not sup.any()
        res, levels = dg.msp()
        assert not res.any() and levels == 0
        h, rounds = dg.hidx()
        assert not h.any() and rounds == 1  # one round that changes nothing
        dg.free()
    g = G.from_pairs(np.zeros((0, 2), dtype=np.uint32))
    dg = dev.upload(g)
    assert dg.bfc() == 0
    assert dg.msp()[1] == 0
    assert dg.hidx()[1] == 1 == oracle.hidx(g, 1)[1]
    dg.free()


@pytest.mark.parametrize("seed", range(8))
def test_random_small(dev, seed):
    rng = np.random.default_rng(300 + seed)
    g = G.random_bipartite(80 + seed, int(rng.integers(4, 40)), int(rng.integers(4, 40)),
                           float(rng.uniform(0.2, 0.8)))
    check_all(dev, g, f"random{seed}")


@pytest.mark.parametrize("seed", G.QUIRK_SEEDS + G.ORDER_SEEDS)
def test_quirk_graphs(dev, seed):
    check_all(dev, G.quirk_graph(seed), f"quirk{seed}")


def test_hub_graph(dev):
    # hub degree above the LDS staging limit and hub roots in the dense path
    check_all(dev, G.hub_graph(5, 6000, 20000, 400, 7000), "hub6000")


@pytest.mark.parametrize("index", [0, 1, 2, 17, 1234])
def test_mix_graphs(dev, index):
    g = gb.Graph.generate("mix_graph", index)
    check_all(dev, g, f"mix{index}", algos=("bfc", "ebfc", "msp") if g.m > 200000 else
              ("bfc", "ebfc", "msp", "hidx"))


@pytest.mark.parametrize("path", ["merge", "index"])
def test_msp_paths_agree(dev, path):
    """Both peel paths against the oracle on graphs with many sub-rounds."""
    try:
        dev.set_msp_path(path)
        for seed in G.QUIRK_SEEDS[:3] + G.ORDER_SEEDS[:2]:
            check_all(dev, G.quirk_graph(seed), f"{path}-quirk{seed}", algos=("msp",))
        check_all(dev, gb.Graph.generate("mix_graph", 2), f"{path}-mix2", algos=("msp",))
        check_all(dev, G.hub_graph(5, 6000, 20000, 400, 7000), f"{path}-hub", algos=("msp",))
    finally:
        dev.set_msp_path("auto")


def test_small_uniform_golden(dev):
    gold = json.load(open(os.path.join(GOLDEN, "small_uniform.json")))
    g = gb.Graph.generate("small_uniform")
    assert g.m == gold["m"]
    dg = dev.upload(g)
    sup, total, smax = dg.ebfc()
    assert total == gold["butterflies"] and smax == gold["support_max"]
    assert digest(sup) == gold["ebfc_sha256"]
    res, levels = dg.msp()
    assert digest(res) == gold["msp_sha256"] and levels == gold["msp_levels"]
    h, _ = dg.hidx()
    assert digest(h) == gold["hidx_sha256"]
    dg.free()


def test_powerlaw_2m_golden(dev):
    path = os.path.join(GOLDEN, "powerlaw_2m.json")
    gold = json.load(open(path))
    g = gb.Graph.generate("powerlaw_2m")
    assert g.m == gold["m"]
    assert hashlib.sha256(g.neighbors.tobytes()).hexdigest() == gold["neighbors_sha256"]
    dg = dev.upload(g)
    say("powerlaw_2m ebfc")
    sup, total, smax = dg.ebfc()
    assert total == gold["butterflies"] and smax == gold["support_max"]
    assert digest(sup) == gold["ebfc_sha256"]
    say("powerlaw_2m msp")
    res, levels = dg.msp()
    assert digest(res) == gold["msp_sha256"] and levels == gold["msp_levels"]
    dg.free()


@pytest.mark.gpu
def test_powerlaw_2m_golden_hidx(dev):
    """h-index bitruss numbers and Jacobi round count of the 2M-edge graph
    against the oracle's (its CPU run took about two hours)."""
    gold = json.load(open(os.path.join(GOLDEN, "powerlaw_2m.json")))
    g = gb.Graph.generate("powerlaw_2m")
    dg = dev.upload(g)
    say("powerlaw_2m hidx")
    h, rounds = dg.hidx()
    assert digest(h) == gold["hidx_sha256"]
    assert int(h.max()) == gold["hidx_kmax"]
    assert rounds == gold["hidx_rounds_jacobi"]
    dg.free()


@pytest.mark.gpu
def test_msp_time_slices(dev):
    """The persistent peel cut into 1 ms slices (each launch resumes at the
    level boundary where the previous one returned) gives the golden
    results of the 2M-edge graph."""
    gold = json.load(open(os.path.join(GOLDEN, "powerlaw_2m.json")))
    g = gb.Graph.generate("powerlaw_2m")
    d = gb.Device(dev.device)
    try:
        d.set_peel_slice_ms(1)
        dg = d.upload(g)
        say("powerlaw_2m msp in 1 ms slices")
        res, levels = dg.msp()
        t = d.last_timings()
        dg.free()
    finally:
        d.close()
    say(f"slices: {t['peel_kernel_launches']}")
    assert t["peel_kernel_launches"] > 1
    assert digest(res) == gold["msp_sha256"] and levels == gold["msp_levels"]


@pytest.mark.gpu
def test_sorted_rule_lists(dev):
    """With the LDS list of both-frontier middles cut to nothing, every bloom
    that needs the attribution rule sorts its both-frontier middles into id
    lists and sums the rule by binary searches: results stay bit-exact
    against the oracle and the 2M-edge golden."""
    with gb.Device(dev.device) as d:
        d.set_peel_c2cap(0)
        _sorted_rule_lists(d)


def _sorted_rule_lists(dev):
    sorted_visits = 0
    graphs = [(f"quirk{s}", G.quirk_graph(s)) for s in G.QUIRK_SEEDS[:4] + G.ORDER_SEEDS[:2]]
    graphs += [("mix2", gb.Graph.generate("mix_graph", 2)), ("hub", G.hub_graph(5, 6000, 20000, 400, 7000))]
    for name, g in graphs:
        say(f"sorted lists: {name} m={g.m}")
        dg = dev.upload(g)
        res, levels = dg.msp()
        sorted_visits += dev.last_timings()["sorted_blooms"]
        dg.free()
        ores, olevels, _ = oracle.msp(g, 8)
        assert np.array_equal(res, ores) and levels == olevels, name
    gold = json.load(open(os.path.join(GOLDEN, "powerlaw_2m.json")))
    dg = dev.upload(gb.Graph.generate("powerlaw_2m"))
    say("sorted lists: powerlaw_2m")
    res, levels = dg.msp()
    sorted_visits += dev.last_timings()["sorted_blooms"]
    dg.free()
    say(f"bloom visits with sorted lists: {sorted_visits}")
    assert digest(res) == gold["msp_sha256"] and levels == gold["msp_levels"]
    assert sorted_visits > 0


@pytest.mark.gpu
def test_msp_collect_window(dev):
    """The collect window (a list of the live edges with support <= W, the
    others above it, split again when the window empties), which the peel
    uses on graphs with millions of edges, forced onto smaller graphs: the
    oracle's results and the 2M-edge golden."""
    with gb.Device(dev.device) as d:
        d.set_peel_window_edges(0)
        _msp_collect_window(d)


def _msp_collect_window(dev):
    for name, g in [("quirk%d" % s, G.quirk_graph(s)) for s in G.QUIRK_SEEDS[:3]] + \
            [("mix2", gb.Graph.generate("mix_graph", 2)), ("hub", G.hub_graph(5, 6000, 20000, 400, 7000))]:
        say(f"window: {name} m={g.m}")
        dg = dev.upload(g)
        res, levels = dg.msp()
        dg.free()
        ores, olevels, _ = oracle.msp(g, 8)
        assert np.array_equal(res, ores) and levels == olevels, name
    gold = json.load(open(os.path.join(GOLDEN, "powerlaw_2m.json")))
    d = gb.Device(dev.device)
    try:
        d.set_peel_window_edges(0)
        d.set_peel_slice_ms(5)  # resumes must keep the window's state
        dg = d.upload(gb.Graph.generate("powerlaw_2m"))
        say("window: powerlaw_2m in 5 ms slices")
        res, levels = dg.msp()
        dg.free()
    finally:
        d.close()
    assert digest(res) == gold["msp_sha256"] and levels == gold["msp_levels"]


@pytest.mark.gpu
def test_msp_level_prefix(dev):
    """A prefix of the decomposition (levels up to a bound; the peeled edges
    hold their result, the others their current support) equals the
    oracle's state after the same levels, with and without the collect
    window; a cut-short call reports BT_PARTIAL, a bound beyond the last
    level does not."""
    graphs = [("quirk%d" % s, G.quirk_graph(s)) for s in G.QUIRK_SEEDS[:3]]
    graphs += [("mix2", gb.Graph.generate("mix_graph", 2)), ("hub", G.hub_graph(5, 6000, 20000, 400, 7000))]
    for window in (0, 100000000):
        with gb.Device(dev.device) as d:
            d.set_peel_window_edges(window)
            for bound in (1, 3, 40):
                d.set_peel_max_level(bound)
                for name, g in graphs:
                    dg = d.upload(g)
                    res, _ = dg.msp()
                    partial = dg.partial
                    dg.free()
                    st, _, _ = oracle.msp_levels(g, 8, bound)
                    _, olevels, _ = oracle.msp(g, 8)
                    assert np.array_equal(res, st), (name, bound, window)
                    assert partial == (olevels > bound), (name, bound, window)
        say(f"level prefixes equal the oracle's (window threshold {window})")


@pytest.mark.gpu
def test_peel_on_part_of_the_device(dev):
    """The persistent peel on a subset of the CUs gives the same results,
    also with two contexts peeling side by side on disjoint CU counts."""
    import threading

    gs = [gb.Graph.generate("mix_graph", i) for i in (5, 6, 2)]
    want = []
    for g in gs:
        dg = dev.upload(g)
        want.append(dg.msp())
        dg.free()
    ndev = dev.device
    a, b = gb.Device(ndev), gb.Device(ndev)
    a.set_peel_blocks(64)
    b.set_peel_blocks(128)
    got = {}

    def run(d, name, graphs):
        out = []
        for g in graphs:
            dg = d.upload(g)
            out.append(dg.msp())
            dg.free()
        got[name] = out

    ta = threading.Thread(target=run, args=(a, "a", gs))
    tb = threading.Thread(target=run, args=(b, "b", gs[::-1]))
    ta.start()
    tb.start()
    ta.join()
    tb.join()
    a.close()
    b.close()
    for (r, lv), (wr, wl) in zip(got["a"], want):
        assert np.array_equal(r, wr) and lv == wl
    for (r, lv), (wr, wl) in zip(got["b"], want[::-1]):
        assert np.array_equal(r, wr) and lv == wl


@pytest.mark.gpu
def test_skewed_60m_golden(dev):
    """Butterflies and per-edge supports of the 60M-edge skewed graph
    (4.67e9 priority wedges, 3.3e9 bloom records) against the oracle's
    digests."""
    gold = json.load(open(os.path.join(GOLDEN, "skewed_60m.json")))
    g = gb.Graph.generate("skewed_60m")
    assert g.m == gold["m"]
    assert hashlib.sha256(g.neighbors.tobytes()).hexdigest() == gold["neighbors_sha256"]
    dg = dev.upload(g)
    say("skewed_60m bfc")
    assert dg.bfc() == gold["butterflies"]
    say("skewed_60m ebfc")
    sup, total, smax = dg.ebfc()
    assert total == gold["butterflies"] and smax == gold["support_max"]
    assert digest(sup) == gold["ebfc_sha256"]
    assert dev.last_timings()["wedges_count"] == gold["w_cnt"]
    dg.free()


@pytest.mark.gpu
def test_skewed_60m_msp_prefix(dev):
    """The 60M-edge graph's msp decomposition through its first levels (the
    oracles take hours for the whole one): the state after them — the peeled
    edges' results and the others' current supports — against the per-pivot
    oracle's (800 levels) and the bloom oracle's deeper checkpoints, on one
    context (repeated calls reuse its index buffers)."""
    gold = json.load(open(os.path.join(GOLDEN, "skewed_60m.json")))
    if "msp_prefix_sha256" not in gold:
        pytest.skip("the oracle's prefix state of skewed_60m is not recorded (make_golden.py --msp-prefix)")
    g = gb.Graph.generate("skewed_60m")
    # the per-pivot oracle's 800-level state, then the bloom oracle's deeper
    # checkpoints (its 800-level state is the per-pivot one)
    want = {gold["msp_prefix_levels"]: gold["msp_prefix_sha256"]}
    ck = {int(k): v for k, v in gold.get("msp_prefix_checkpoints", {}).items()}
    # every checkpoint up to 3,200 levels, and the two deepest while the
    # whole decomposition's digest is not recorded (each run peels from the
    # start; the whole run is test_skewed_60m_msp_whole)
    deep = [] if "msp_sha256" in gold else sorted(ck)[-2:]
    for level in sorted(ck):
        if level <= 3200 or level in deep:
            want[level] = ck[level]
    with gb.Device(dev.device) as d:
        # the bloom-index peel or a failure: never the merge path, which
        # would take hours here
        d.set_msp_path("index")
        dg = d.upload(g)
        for level in sorted(want):
            d.set_peel_max_level(level)
            say(f"skewed_60m msp through level {level}")
            res, _ = dg.msp()
            assert dg.partial
            assert digest(res) == want[level], level
        dg.free()


def _coh_skewed_60m(algo: str, tmp_path):
    """Runs `coh --gpu --algo <algo> --dump --json` (main.cu:96-129) on the
    60M-edge graph and returns its results and JSON summary. The decomposition
    reports its progress (each msp time slice, each graph of h-index rounds)
    to a log under gpurun_out/, so a long run shows it is alive while the
    test's own output is held by pytest."""
    import subprocess

    from gpu_bitruss_amd import _lib

    out_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpurun_out")
    os.makedirs(out_dir, exist_ok=True)
    log_path = os.path.join(out_dir, f"skewed_60m_{algo}_progress.log")
    binf, dump = str(tmp_path / "skewed_60m.bin"), str(tmp_path / f"skewed_60m_{algo}.u64")
    with open(log_path, "a") as log:
        log.write("[skewed_60m] generating the graph\n")
        log.flush()
        gb.Graph.generate("skewed_60m").graph_to_bin(binf)
        log.write(f"[skewed_60m] coh --gpu --algo {algo}\n")
        log.flush()
        env = dict(os.environ, BT_PEEL_PROGRESS="1")
        r = subprocess.run([_lib.COH_PATH, "--gpu", "--algo", algo, "--graph", binf, "--dump", dump, "--json"],
                           stdout=subprocess.PIPE, stderr=log, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stdout
    summary = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][0])
    return np.fromfile(dump, dtype="<u8"), summary


@pytest.mark.gpu
@pytest.mark.timeout(1000)
def test_skewed_60m_msp_whole(dev, tmp_path):
    """The 60M-edge graph's whole msp decomposition through the drop-in
    command line against the bloom oracle's digest and level count
    (or_msp_bloom, cpu_bitruss.cpp:536-597)."""
    gold = json.load(open(os.path.join(GOLDEN, "skewed_60m.json")))
    if "msp_sha256" not in gold:
        pytest.skip("the oracle's whole msp of skewed_60m is not recorded (make_golden.py --msp-bloom)")
    res, summary = _coh_skewed_60m("msp", tmp_path)
    assert res.size == gold["m"]
    assert digest(res) == gold["msp_sha256"] and summary["iters"] == gold["msp_levels"]


@pytest.mark.gpu
@pytest.mark.timeout(1000)
def test_skewed_60m_hidx_whole(dev, tmp_path):
    """The 60M-edge graph's h-index decomposition to its fixed point
    (`coh --gpu --algo hidx`, bitruss_hidx.cu:179-256; about 200 s with the
    sorted chunks) against the bitruss numbers of the oracle's exact bloom
    peel (or_msp_bloom with BLOOM_EXACT, which equals the textbook peel
    cpu_bitruss.cpp:34-122 on every graph the CPU tests hold): the Jacobi
    rounds' fixed point is the bitruss decomposition, as the 2M-edge golden
    shows for the oracle's own hidx. The round count (135 on MI355X, the same
    with and without sorted chunks on every graph tested) has no oracle
    here: the oracle's Jacobi rounds would stream 4.4e12 values each."""
    gold = json.load(open(os.path.join(GOLDEN, "skewed_60m.json")))
    if "peel_sha256" not in gold:
        pytest.skip("the oracle's exact peel of skewed_60m is not recorded")
    res, summary = _coh_skewed_60m("hidx", tmp_path)
    assert res.size == gold["m"]
    assert digest(res) == gold["peel_sha256"] and int(res.max()) == gold["peel_kmax"]
