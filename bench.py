"""Benchmark: GBiD (msp) decomposition throughput on one MI355X.

One step = one full GBiD decomposition (per-edge butterfly support, bloom
index, level-synchronous peeling with the reference's attribution rule) of
the Chung-Lu power-law graph 200,000 x 100,000, gamma 2.3, m ~ 2,000,000,
expected degree cap 5,000, seed 1, with the CSR already resident in HBM and
the results left on the device.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched] [--no-serve]
    python bench.py --workload mix [--graphs 4096] [--check]   # batched vs plain only
    python bench.py --workload serve [--requests 400] [--qps 25,100,400]

Prints one JSON line. `value` is GBiD wedges/s = (W_cnt + W_peel) / step
time, where W_cnt (priority wedges of the support count) and W_peel (wedges
the reference's peel loops visit) are the units of work counted by the CPU
oracle's instrumented loops; they are fixed numbers of the workload, kept
with its golden digests in tests/golden/powerlaw_2m.json.

The same line carries the other two parts of the metric, measured in the
same run after the timed steps: `batched` (the 4,096-graph mix decomposed as
batches against the plain path graph by graph: graphs/s of both, their
ratio, and how many sampled graphs gave identical results) and `serve` (a
Poisson stream of 400 requests, seed 7, 2% large: p50/p99 latency and
served graphs/s against the plain path one request after the other).

The 60M-edge graph is not the headline workload: one decomposition takes
about 270 s, so the 25 steps of the default run do not fit; its figures come
from `--workload skewed_60m` (one step) and are kept in profiles/.

`roofline` is for the dominant kernel, the persistent peel k_msp_persist:
its duration comes from HIP events recorded on the context's stream around
the launch, and its algorithmic bytes are the reference peel's per-unit
bytes times the units of the workload. `traffic` is the HBM bytes one launch
moves as measured by rocprofv3 counters (committed summary in profiles/).
The CPU baseline times the oracle on the host on a bounded sample: support
counting plus the first peel levels, cut after --cpu-seconds.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = "powerlaw_2m"
HBM_PEAK_GBS = 8000.0
KERNEL = "k_msp_persist"
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_msp_persist_powerlaw_2m_traffic.md")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def peel_bytes(st: dict) -> int:
    """Algorithmic HBM bytes of the reference's peel loops: per wedge 4 B w
    + 4 B wv id + 2 B flags, 4 B per intersection read, per butterfly 8 B
    ids + 2 B flags, 8 B per support decrement, 8 B per frontier edge."""
    return (10 * st["w_peel"] + 4 * st["i_peel"] + 10 * st["b_peel"] + 8 * st["d_peel"]
            + 8 * st["sum_curr"])


def measured_traffic():
    """(HBM bytes, L2 hit rate, atomics) per launch of the peel kernel from
    the committed rocprofv3 counter summary (None where it is not there)."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            text = f.read()
    except OSError:
        return None, None, None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9]+)", text)
    h = re.search(r"([0-9,]+) hits and\s+([0-9,]+) misses", text)
    a = re.search(r"TCC_ATOMIC_sum = TCC_EA0_ATOMIC_sum = ([0-9,]+)", text)
    hit = None
    if h:
        hits, misses = (int(x.replace(",", "")) for x in h.groups())
        hit = hits / (hits + misses)
    return (int(m.group(1)) if m else None), hit, (int(a.group(1).replace(",", "")) if a else None)


def cpu_baseline(g, threads: int, seconds: float):
    """The oracle's msp (support count, then peel levels) on the host, cut
    after the first level that ends past `seconds`."""
    from oracle import oracle

    t0 = time.perf_counter()
    levels, complete, st = oracle.msp_sample(g, threads, seconds)
    secs = time.perf_counter() - t0
    return secs, levels, complete, st


def batched_vs_plain(k: int, steps: int, warmup: int, lanes: int, plain_sample: int, check: bool):
    """The mix's first k graphs (seeds 1000..999+k): msp on all of them as
    batches (unions) against the plain path graph by graph, both with the
    graphs resident and the results left on the device. The plain path is
    timed on `plain_sample` graphs drawn at random (seed 0; all of them when
    0), whose results are also compared with the batch's."""
    import numpy as np

    import gpu_bitruss_amd as gb

    log(f"generating {k} mix graphs")
    gs = [gb.Graph.generate("mix_graph", i) for i in range(k)]
    m_total = sum(g.m for g in gs)
    dev = gb.Device(0)
    if lanes:
        dev.set_batch_lanes(lanes)
    b = dev.upload_batch(gs)
    log(f"batch: {m_total} edges in {b.chunks()} unions")
    for i in range(warmup):
        b.run("msp", copy=False)
    tb = []
    for i in range(steps):
        t0 = time.perf_counter()
        b.run("msp", copy=False)
        tb.append(time.perf_counter() - t0)
        log(f"batch step {i}: {tb[-1]:.2f} s")
    idx = list(range(k)) if not plain_sample or plain_sample >= k else \
        sorted(np.random.default_rng(0).choice(k, plain_sample, replace=False).tolist())
    dgs = [dev.upload(gs[i]) for i in idx]
    for dg in dgs[:16]:
        dg.msp(copy=False)
    t0 = time.perf_counter()
    for j, dg in enumerate(dgs):
        dg.msp(copy=False)
        if j % 512 == 511:
            log(f"plain: {j + 1} graphs, {time.perf_counter() - t0:.1f} s")
    tp = time.perf_counter() - t0
    same = None
    if check:
        arrays, scal = b.run("msp")
        same = 0
        for j, dg in enumerate(dgs):
            r, lv = dg.msp()
            same += int(np.array_equal(r, arrays[idx[j]]) and lv == scal[idx[j]]["levels"])
        log(f"identical: {same} of {len(idx)}")
    for dg in dgs:
        dg.free()
    b.free()
    dev.close()
    t = float(np.mean(tb))
    return gs, {"graphs": k, "edges": m_total, "batched_s": t, "batched_graphs_per_s": k / t,
                "plain_graphs_per_s": len(idx) / tp, "plain_sample": len(idx), "plain_s": tp,
                "speedup": (len(idx) / tp) and (k / t) / (len(idx) / tp), "identical": same,
                "identical_of": len(idx) if check else None, "lanes": lanes or 2}


def bench_mix(args) -> int:
    """The 4,096-graph mix (seeds 1000..5095): msp on all graphs as batches
    (unions) against the plain path graph by graph."""
    import numpy as np

    k = args.graphs
    gs, r = batched_vs_plain(k, args.steps, args.warmup, args.lanes, 0, args.check)
    m_total, t, tp, same = r["edges"], r["batched_s"], r["plain_s"], r["identical"]
    cpu = None
    if not args.no_cpu:
        # the oracle's whole msp on the mix graphs in order, until the
        # time bound: graphs/s on the host
        from oracle import oracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle msp on the mix graphs in order, {threads} threads, {args.cpu_seconds:.0f} s")
        t0 = time.perf_counter()
        done = 0
        while done < k and time.perf_counter() - t0 < args.cpu_seconds:
            oracle.msp(gs[done], threads)
            done += 1
        secs = time.perf_counter() - t0
        cpu = {"value": done / secs, "unit": "graphs/s", "cores": threads, "kind": "port",
               "sample": f"mix graphs 0..{done - 1} ({sum(g.m for g in gs[:done])} edges), whole msp each, "
                         f"{secs:.1f} s on {threads} threads"}
    print(json.dumps({
        "metric": "graphs/s, batched msp vs plain (4,096-graph mix)", "value": k / t, "unit": "graphs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"mix: {k} bipartite graphs, seeds 1000..{999 + k}, |U|,|V| log-uniform in "
                               "[50, 5000], density log-uniform in [0.5%, 5%], even seeds uniform, odd Chung-Lu 2.2",
                   "m": m_total, "algo": "msp", "lanes": args.lanes or 2},
        "plain_graphs_per_s": k / tp, "plain_s": tp, "speedup": tp / t, "identical": same,
        "cpu_baseline": cpu,
    }), flush=True)
    return 0


def serve_levels(n: int, qps_list, serve_cus: str, batch_wedges: int, light_workers: int = 0):
    """Poisson request stream (seed 7): 98% graphs of the mix distribution,
    2% large Chung-Lu requests; p50/p99 latency and graphs/s per offered QPS,
    against the plain path on the same requests one after the other (upload
    included, as a request would be served)."""
    import threading

    import numpy as np

    import gpu_bitruss_amd as gb

    rng = np.random.default_rng(7)
    large = rng.random(n) < 0.02
    idx = rng.integers(0, 4096, n)
    log(f"generating {n} requests ({int(large.sum())} large)")
    reqs = [gb.Graph.generate("large_request", int(i)) if big else gb.Graph.generate("mix_graph", int(i))
            for i, big in zip(idx, large)]
    dev = gb.Device(0)
    t0 = time.perf_counter()
    for g in reqs:
        dg = dev.upload(g)
        dg.msp(copy=False)
        dg.free()
    t_plain = time.perf_counter() - t0
    log(f"plain path, one graph after the other (upload included): {t_plain:.1f} s")
    dev.close()
    cus = tuple(int(x) for x in serve_cus.split(","))
    levels = []
    for qps in qps_list:
        gaps = rng.exponential(1.0 / qps, n)
        with gb.Server(large_edges=1_000_000, batch_wedges=batch_wedges, cus=cus, light_workers=light_workers) as srv:
            tickets = [None] * n
            start = time.perf_counter()

            def feed():
                t = start
                for i, g in enumerate(reqs):
                    t += gaps[i]
                    d = t - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    tickets[i] = srv.submit(g, "msp")

            th = threading.Thread(target=feed)
            th.start()
            th.join()
            lat = np.array([srv.wait(t)[2] for t in tickets])
            wall = time.perf_counter() - start
            st = srv.stats()
        p50, p99 = np.percentile(lat, 50), np.percentile(lat, 99)
        log(f"qps {qps:g}: p50 {p50:.1f} ms, p99 {p99:.1f} ms, {n / wall:.1f} graphs/s, {st['batches']} batches")
        levels.append({"qps": qps, "p50_ms": p50, "p99_ms": p99, "graphs_per_s": n / wall,
                       "batches": st["batches"], "large_runs": st["large_runs"]})
    return {"requests": n, "large": int(large.sum()), "plain_graphs_per_s": n / t_plain, "levels": levels,
            "lane_cus": list(cus), "batch_wedges": batch_wedges or 100_000_000,
            "light_workers": light_workers or 2}


def bench_serve(args) -> int:
    n = args.requests
    sv = serve_levels(n, [float(q) for q in args.qps.split(",")], args.serve_cus, args.serve_batch_wedges,
                      args.serve_light_workers)
    levels, cus = sv["levels"], sv["lane_cus"]
    best = max(levels, key=lambda x: x["graphs_per_s"])
    print(json.dumps({
        "metric": "served graphs/s and p50/p99 latency (Poisson stream, msp)", "value": best["graphs_per_s"],
        "unit": "graphs/s", "n_gpus": 1, "steps": len(levels), "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"serve: {n} requests, 98% mix graphs, 2% Chung-Lu 2.2 with m in [1M, 5M], "
                               "budget 64 GiB", "algo": "msp", "lane_cus": list(cus),
                   "batch_wedges": args.serve_batch_wedges or 100_000_000},
        "plain_graphs_per_s": sv["plain_graphs_per_s"], "levels": levels,
    }), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--workload", default=WORKLOAD, choices=[WORKLOAD, "skewed_60m", "mix", "serve"],
                    help="powerlaw_2m (the headline line), skewed_60m (the 60M-edge graph, one step is "
                         "~270 s), mix (batched vs plain) or serve")
    ap.add_argument("--no-batched", action="store_true", help="headline line without the batched part")
    ap.add_argument("--no-serve", action="store_true", help="headline line without the serving part")
    ap.add_argument("--plain-sample", type=int, default=512,
                    help="headline line: mix graphs the plain path is timed on (drawn at random, seed 0)")
    ap.add_argument("--serve-qps", type=float, default=50.0, help="headline line: offered requests/s")
    ap.add_argument("--graphs", type=int, default=4096, help="mix: number of graphs")
    ap.add_argument("--check", action="store_true", help="mix: compare every graph with the plain path")
    ap.add_argument("--lanes", type=int, default=0, help="mix: batch lanes (0: the library default, 2)")
    ap.add_argument("--requests", type=int, default=400, help="serve: requests per offered rate")
    ap.add_argument("--qps", default="25,100,400", help="serve: offered request rates")
    ap.add_argument("--serve-light-workers", type=int, default=0,
                    help="workers of the serving light lane (0: the library's default, 2)")
    ap.add_argument("--serve-cus", default="0,0,0",
                    help="serve: CUs of the light, heavy and large lanes (0: the library defaults)")
    ap.add_argument("--serve-batch-wedges", type=int, default=0,
                    help="serve: priority wedges per heavy batch at most (0: the library default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the bitruss path runs on a single MI355X)", file=sys.stderr)
        return 2
    if args.workload == "mix":
        return bench_mix(args)
    if args.workload == "serve":
        return bench_serve(args)

    import numpy as np

    import gpu_bitruss_amd as gb

    wl = args.workload
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", f"{wl}.json")))
    log(f"generating {wl}")
    g = gb.Graph.generate(wl)
    log(f"graph m={g.m} n={g.n}")
    if g.m != gold["m"]:
        raise RuntimeError("generated graph differs from the golden one")
    dev = gb.Device(0)
    dg = dev.upload(g)

    for i in range(args.warmup):
        log(f"warmup {i}")
        dg.msp(copy=False)
    kernel_ms, peel_ms, count_ms = [], [], []
    levels = 0
    sub_rounds = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, levels = dg.msp(copy=False)
        t = dev.last_timings()
        kernel_ms.append(t["peel_kernel_ms"])
        peel_ms.append(t["peel_ms"])
        count_ms.append(t["count_ms"] + t["prep_ms"])
        sub_rounds = t["sub_rounds"]
        log(f"step {i}: support {count_ms[-1]:.1f} ms, peel {peel_ms[-1]:.1f} ms "
            f"(kernel {kernel_ms[-1]:.1f} ms, {sub_rounds} sub-rounds)")
    wall = time.perf_counter() - t0
    ms = 1e3 * wall / max(args.steps, 1)
    if "msp_levels" in gold and levels != gold["msp_levels"]:
        raise RuntimeError(f"msp levels {levels} != {gold['msp_levels']}")
    wcount = dev.last_timings()["wedges_count"]
    if "msp_stats" not in gold:
        # the oracle's peel counters of this graph are not recorded: the
        # line carries the counting wedges and the time only
        print(json.dumps({"metric": "msp decomposition time", "value": ms / 1e3, "unit": "s",
                          "higher_is_better": False, "workload": wl, "m": g.m, "levels": levels,
                          "sub_rounds": sub_rounds, "w_cnt": wcount,
                          "phase_ms": {"support": float(np.mean(count_ms)), "peel": float(np.mean(peel_ms)),
                                       "peel_kernel": float(np.mean(kernel_ms))}}), flush=True)
        return 0
    stats = gold["msp_stats"]
    if stats["w_cnt"] != wcount:
        raise RuntimeError(f"wedge count {wcount} != {stats['w_cnt']}")
    wedges = stats["w_cnt"] + stats["w_peel"]
    value = wedges / (ms / 1e3)

    kms = float(np.mean(kernel_ms))
    nbytes = peel_bytes(stats)
    achieved = nbytes / (kms / 1e3) / 1e9
    traffic, l2_hit, atomics = measured_traffic() if wl == WORKLOAD else (None, None, None)
    out = {
        "metric": "GBiD wedges/s (W_cnt + W_peel per second), full msp decomposition",
        "value": value,
        "unit": "wedges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": ("powerlaw_2m: bipartite Chung-Lu 200,000 x 100,000, gamma 2.3/2.3, "
                         "expected degree cap 5,000, seed 1") if wl == WORKLOAD else
                        ("skewed_60m: bipartite Chung-Lu 4,000,000 (gamma 2.5) x 1,000,000 (gamma 2.0), "
                         "m ~ 60M, seed 2"),
            "m": g.m, "n": g.n, "algo": "msp",
        },
        "edges_peeled_per_s": g.m / (ms / 1e3),
        "wedges": {"count": stats["w_cnt"], "peel": stats["w_peel"]},
        "msp_levels": levels,
        "sub_rounds": sub_rounds,
        "phase_ms": {"support": float(np.mean(count_ms)), "peel": float(np.mean(peel_ms)),
                     "peel_kernel": kms},
        "roofline": {
            # no MFMA work: HBM is the only roof; `achieved` is the reference
            # peel's algorithmic bytes per launch over this kernel's time,
            # `traffic` the bytes the counters saw (committed summary), which
            # the bloom index keeps below the algorithmic bytes
            "bound": "hbm",
            "kernel": KERNEL,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_gbps": traffic / (kms / 1e3) / 1e9 if traffic else None,
            "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, ROOT) if traffic else None,
            "l2_hit_rate": l2_hit,
            "atomics_per_launch": atomics,
            "limiter": "latency of the dependent phases (sub-rounds and level collects), each ending in a "
                       "grid barrier; not bandwidth",
            "algorithmic_bytes": nbytes,
            "kernel_ms": kms,
        },
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle msp sample, {threads} threads, {args.cpu_seconds:.0f} s")
        secs, clev, complete, st = cpu_baseline(g, threads, args.cpu_seconds)
        if st["w_cnt"] != stats["w_cnt"]:
            raise RuntimeError("oracle wedge count differs from the golden one")
        done = "all" if complete else f"the first {clev}"
        out["cpu_baseline"] = {
            "value": (st["w_cnt"] + st["w_peel"]) / secs,
            "unit": "wedges/s",
            "cores": threads,
            "kind": "port",
            "sample": f"powerlaw_2m: support count (W_cnt {st['w_cnt']}) plus {done} of "
                      f"{gold['msp_levels']} peel levels (W_peel {st['w_peel']} of {stats['w_peel']}), "
                      f"{secs:.1f} s on {threads} threads",
            # the host's CPUs as the OS reports them; a run uses at most 16
            # threads (the share of the host that goes with one GPU)
            "host_cpus": os.cpu_count(),
        }
        # the same oracle on one thread, over a shorter bound
        log(f"cpu baseline: oracle msp sample, 1 thread, {args.cpu_seconds / 2:.0f} s")
        secs1, clev1, complete1, st1 = cpu_baseline(g, 1, args.cpu_seconds / 2)
        done1 = "all" if complete1 else f"the first {clev1}"
        out["cpu_baseline"]["single_thread"] = {
            "value": (st1["w_cnt"] + st1["w_peel"]) / secs1,
            "unit": "wedges/s",
            "cores": 1,
            "sample": f"powerlaw_2m: support count plus {done1} of {gold['msp_levels']} peel levels "
                      f"(W_peel {st1['w_peel']}), {secs1:.1f} s on 1 thread",
        }
    dg.free()
    dev.close()
    if wl == WORKLOAD and not args.no_batched:
        # the batched part of the metric: the whole 4,096-graph mix as
        # batches against the plain path on a random sample of its graphs
        _, bt = batched_vs_plain(4096, 1, 1, 0, args.plain_sample, True)
        out["batched"] = bt
        log(f"batched {bt['batched_graphs_per_s']:.1f} graphs/s, plain {bt['plain_graphs_per_s']:.1f}, "
            f"x{bt['speedup']:.2f}, identical {bt['identical']} of {bt['identical_of']}")
    if wl == WORKLOAD and not args.no_serve:
        out["serve"] = serve_levels(400, [args.serve_qps], args.serve_cus, args.serve_batch_wedges,
                                    args.serve_light_workers)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
