"""Benchmark: GBiD (msp) decomposition throughput on one MI355X.

One step = one full GBiD decomposition (per-edge butterfly support, then
level-synchronous peeling with the reference's attribution rule) of the
Chung-Lu power-law graph 200,000 x 100,000, gamma 2.3, m ~ 2,000,000,
expected degree cap 5,000, seed 1, with the CSR already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line. `value` is GBiD wedges/s = (W_cnt + W_peel) / step
time, where W_cnt (priority wedges of the support count) and W_peel (peel
wedges) are the units counted by the CPU oracle with the reference's loops.
The CPU baseline leg times the oracle on the host on the same graph.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = "powerlaw_2m"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(g, threads: int):
    """Oracle msp (ebfc + peel) on the host: returns (seconds, stats)."""
    from oracle import oracle

    t0 = time.perf_counter()
    _, levels, st = oracle.msp(g, threads)
    return time.perf_counter() - t0, st, levels


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the bitruss path runs on a single MI355X)", file=sys.stderr)
        return 2

    import numpy as np

    import gpu_bitruss_amd as gb

    log(f"generating {WORKLOAD}")
    g = gb.Graph.generate(WORKLOAD)
    log(f"graph m={g.m} n={g.n}")
    dev = gb.Device(0)
    dg = dev.upload(g)

    for i in range(args.warmup):
        log(f"warmup {i}")
        dg.msp(copy=False)
    times, peel, count = [], [], []
    sub_rounds = 0
    for i in range(args.steps):
        t0 = time.perf_counter()
        _, levels = dg.msp(copy=False)
        times.append(time.perf_counter() - t0)
        t = dev.last_timings()
        peel.append(t["peel_ms"])
        count.append(t["count_ms"] + t["prep_ms"])
        sub_rounds = t["sub_rounds"]
        log(f"step {i}: {times[-1] * 1e3:.1f} ms (support {count[-1]:.1f} ms, peel {peel[-1]:.1f} ms, "
            f"{sub_rounds} sub-rounds)")
    ms = 1e3 * float(np.mean(times))
    wcount = dev.last_timings()["wedges_count"]

    # units of work from the oracle (and the CPU baseline itself)
    cpu = None
    stats = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle msp with {threads} threads")
        secs, stats, clev = cpu_baseline(g, threads)
        assert clev == levels
        cpu = {"seconds": secs, "threads": threads}
    else:
        gold = json.load(open(os.path.join(ROOT, "tests", "golden", f"{WORKLOAD}.json")))
        stats = dict(gold["msp_stats"])
    assert stats["w_cnt"] == wcount, (stats["w_cnt"], wcount)
    wedges = stats["w_cnt"] + stats["w_peel"]
    value = wedges / (ms / 1e3)

    # roofline of the peel phase (algorithmic bytes of the reference's peel
    # loops: per wedge 4 B w + 4 B wv id + 2 B flags, 4 B per intersection
    # read, per butterfly 8 B ids + 2 B flags, 8 B per support decrement,
    # 8 B per frontier edge)
    peel_bytes = (10 * stats["w_peel"] + 4 * stats["i_peel"] + 10 * stats["b_peel"]
                  + 8 * stats["d_peel"] + 8 * stats["sum_curr"])
    peel_ms = float(np.mean(peel))
    achieved = peel_bytes / (peel_ms / 1e3) / 1e9
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
            "workload": "powerlaw_2m: bipartite Chung-Lu 200,000 x 100,000, gamma 2.3/2.3, "
                        "expected degree cap 5,000, seed 1",
            "m": g.m, "n": g.n, "algo": "msp",
        },
        "edges_peeled_per_s": g.m / (ms / 1e3),
        "wedges": {"count": stats["w_cnt"], "peel": stats["w_peel"]},
        "msp_levels": levels,
        "sub_rounds": sub_rounds,
        "phase_ms": {"support": float(np.mean(count)), "peel": peel_ms},
        "roofline": {
            "bound": "hbm",
            "kernel": "msp peel phase",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    if cpu:
        out["cpu_baseline"] = {
            "value": wedges / cpu["seconds"],
            "unit": "wedges/s",
            "cores": cpu["threads"],
            "kind": "port",
            "sample": f"full powerlaw_2m GBiD (ebfc + msp) by the CPU oracle, {cpu['seconds']:.1f} s",
        }
    print(json.dumps(out), flush=True)
    dg.free()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
