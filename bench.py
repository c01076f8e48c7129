"""Benchmark: GBiD (msp) decomposition throughput on one MI355X.

One step = one full GBiD decomposition (per-edge butterfly support, bloom
index, level-synchronous peeling with the reference's attribution rule) of
the Chung-Lu power-law graph 200,000 x 100,000, gamma 2.3, m ~ 2,000,000,
expected degree cap 5,000, seed 1, with the CSR already resident in HBM and
the results left on the device.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line. `value` is GBiD wedges/s = (W_cnt + W_peel) / step
time, where W_cnt (priority wedges of the support count) and W_peel (wedges
the reference's peel loops visit) are the units of work counted by the CPU
oracle's instrumented loops; they are fixed numbers of the workload, kept
with its golden digests in tests/golden/powerlaw_2m.json.

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
    """HBM bytes per launch of the peel kernel from the committed rocprofv3
    counter summary, or None when it is not there."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            m = re.search(r"hbm_bytes_per_launch:\s*([0-9]+)", f.read())
        return int(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(g, threads: int, seconds: float):
    """The oracle's msp (support count, then peel levels) on the host, cut
    after the first level that ends past `seconds`."""
    from oracle import oracle

    t0 = time.perf_counter()
    levels, complete, st = oracle.msp_sample(g, threads, seconds)
    secs = time.perf_counter() - t0
    return secs, levels, complete, st


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the bitruss path runs on a single MI355X)", file=sys.stderr)
        return 2

    import numpy as np

    import gpu_bitruss_amd as gb

    gold = json.load(open(os.path.join(ROOT, "tests", "golden", f"{WORKLOAD}.json")))
    log(f"generating {WORKLOAD}")
    g = gb.Graph.generate(WORKLOAD)
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
    if levels != gold["msp_levels"]:
        raise RuntimeError(f"msp levels {levels} != {gold['msp_levels']}")
    wcount = dev.last_timings()["wedges_count"]
    stats = gold["msp_stats"]
    if stats["w_cnt"] != wcount:
        raise RuntimeError(f"wedge count {wcount} != {stats['w_cnt']}")
    wedges = stats["w_cnt"] + stats["w_peel"]
    value = wedges / (ms / 1e3)

    kms = float(np.mean(kernel_ms))
    nbytes = peel_bytes(stats)
    achieved = nbytes / (kms / 1e3) / 1e9
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
        "phase_ms": {"support": float(np.mean(count_ms)), "peel": float(np.mean(peel_ms)),
                     "peel_kernel": kms},
        "roofline": {
            "bound": "hbm",
            "kernel": KERNEL,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(),
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
        }
    print(json.dumps(out), flush=True)
    dg.free()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
