"""Writes tests/golden/<workload>.json: digests of the CPU oracle's results
on the named seeded workloads, so GPU tests can check full-size outputs
without rerunning the slow CPU algorithms.

    python tests/golden/make_golden.py small_uniform powerlaw_2m

Each entry holds sha256 of the little-endian uint64 arrays (ebfc supports,
msp results, hidx results), plus the totals and the oracle work counters.
The hidx oracle is slow on powerlaw_2m (its Jacobi rounds took about two
hours on 5 threads); pass --skip-hidx to leave it out, and merge a run of
the oracle CLI made separately with --hidx-from:

    oracle/_build/bt_oracle_cli --graph powerlaw_2m.bin --algo hidx \
        --threads 5 --dump c2_hidx.u64 > c2_hidx.json
    python tests/golden/make_golden.py powerlaw_2m --skip-hidx \
        --hidx-from c2_hidx.json c2_hidx.u64

(the .bin is the generated graph written with Graph.graph_to_bin). The
60M-edge skewed graph is recorded the same way, msp from a separate CLI run
(--skip-msp leaves msp out, --msp-from merges the run):

    python tests/golden/make_golden.py skewed_60m --skip-msp --skip-hidx \
        --msp-from c3_msp.json c3_msp.u64

A decomposition too long for the oracle is recorded by a prefix instead:
--msp-prefix L stores the state after levels 1..L (peeled edges their
result, the others their current support):

    python tests/golden/make_golden.py skewed_60m --skip-msp --skip-hidx --msp-prefix 1000

--msp-bloom records msp with the bloom restatement (or_msp_bloom, pinned
against or_msp on the graphs the latter finishes and on the prefix above)
instead of the per-pivot oracle, whose merges take hours on the 60M-edge
graph:

    python tests/golden/make_golden.py skewed_60m --skip-hidx --msp-bloom \
        --msp-checkpoints 800,1600,3200,6400,12800,25600,51200,102400,150000,204800,250000

(3 h 18 min on 8 threads; --msp-checkpoints also records the state after
each listed level on the way). --peel-bloom records the bitruss numbers of
the exact bloom peel (BLOOM_EXACT, pinned against the textbook peel), the
fixed point of the h-index rounds, for graphs whose hidx oracle would take
too long:

    python tests/golden/make_golden.py skewed_60m --skip-hidx --skip-msp --peel-bloom

Fields of an existing file that a run does not recompute are kept.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import gpu_bitruss_amd as gb  # noqa: E402
from oracle import oracle  # noqa: E402


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="+")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 4)
    ap.add_argument("--skip-hidx", action="store_true")
    ap.add_argument("--skip-msp", action="store_true")
    ap.add_argument("--msp-bloom", action="store_true",
                    help="msp by the bloom restatement (or_msp_bloom) instead of or_msp")
    ap.add_argument("--msp-checkpoints", default="", metavar="L1,L2,...",
                    help="with --msp-bloom: also record the state after each listed level")
    ap.add_argument("--peel-bloom", action="store_true",
                    help="record the exact bloom peel's bitruss numbers (peel_sha256)")
    ap.add_argument("--msp-prefix", type=int, default=0, metavar="L",
                    help="also record the msp state after levels 1..L")
    ap.add_argument("--msp-from", nargs=2, metavar=("JSON", "U64"),
                    help="oracle CLI output (JSON line, --dump array) of msp on the one workload")
    ap.add_argument("--hidx-from", nargs=2, metavar=("JSON", "U64"),
                    help="oracle CLI output (JSON line, --dump array) of hidx on the one workload")
    args = ap.parse_args()
    if (args.hidx_from or args.msp_from) and len(args.workloads) != 1:
        ap.error("--hidx-from and --msp-from take one workload")
    for wl in args.workloads:
        g = gb.Graph.generate(wl)
        sup, total, smax, st = oracle.ebfc(g, args.threads)
        path = os.path.join(HERE, f"{wl}.json")
        out = {}
        if os.path.exists(path):
            with open(path) as f:
                out = json.load(f)
        out.update({
            "workload": wl, "m": g.m, "n": g.n, "u_num": g.u_num, "l_num": g.l_num,
            "neighbors_sha256": hashlib.sha256(g.neighbors.tobytes()).hexdigest(),
            "butterflies": total, "support_max": smax, "ebfc_sha256": digest(sup),
            "w_cnt": st["w_cnt"],
        })
        if not args.skip_msp and args.msp_bloom:
            cks = [int(x) for x in args.msp_checkpoints.split(",") if x]
            ck = dict(out.get("msp_prefix_checkpoints", {}))

            def on_state(level, state):
                d = digest(state)
                assert ck.get(str(level), d) == d, level
                ck[str(level)] = d

            res, levels, complete, st2 = oracle.msp_bloom_checkpoints(g, args.threads, cks, on_state, 0)
            assert complete
            out.update({"msp_levels": levels, "msp_sha256": digest(res), "msp_oracle": "or_msp_bloom",
                        "msp_stats": {k: st2[k] for k in ("w_peel", "i_peel", "b_peel", "d_peel", "sum_curr",
                                                          "sub_rounds")}})
            if ck:
                out["msp_prefix_checkpoints"] = ck
        if args.peel_bloom:
            res, levels, complete, st5 = oracle.msp_bloom(g, args.threads, 0, oracle.BLOOM_EXACT)
            assert complete
            out.update({"peel_sha256": digest(res), "peel_kmax": int(res.max()) if g.m else 0,
                        "peel_levels": levels, "peel_oracle": "or_msp_bloom (BLOOM_EXACT)",
                        "peel_stats": {k: st5[k] for k in ("w_peel", "b_peel", "sum_curr", "sub_rounds")}})
        elif not args.skip_msp:
            res, levels, st2 = oracle.msp(g, args.threads)
            out.update({"msp_levels": levels, "msp_sha256": digest(res), "msp_stats": st2, "msp_oracle": "or_msp"})
        if args.msp_prefix:
            st, _, st4 = oracle.msp_levels(g, args.threads, args.msp_prefix)
            out.update({"msp_prefix_levels": args.msp_prefix, "msp_prefix_sha256": digest(st),
                        "msp_prefix_stats": {k: st4[k] for k in ("w_peel", "b_peel", "sum_curr", "sub_rounds")}})
        if args.msp_from:
            with open(args.msp_from[0]) as f:
                run = json.loads(f.read().strip().splitlines()[-1])
            r = np.fromfile(args.msp_from[1], dtype="<u8")
            assert run["algo"] == "msp" and run["m"] == g.m and r.size == g.m, "msp run of another graph"
            out.update({"msp_levels": run["iters"], "msp_sha256": digest(r),
                        "msp_stats": {k: run[k] for k in ("w_peel", "i_peel", "b_peel", "d_peel",
                                                          "sum_curr", "sub_rounds")}})
        if not args.skip_hidx:
            h, rounds, st3 = oracle.hidx(g, args.threads)
            out.update({"hidx_sha256": digest(h), "hidx_kmax": int(h.max()) if g.m else 0,
                        "hidx_rounds_jacobi": rounds, "hidx_stats": st3})
        if args.hidx_from:
            with open(args.hidx_from[0]) as f:
                run = json.loads(f.read().strip().splitlines()[-1])
            h = np.fromfile(args.hidx_from[1], dtype="<u8")
            assert run["algo"] == "hidx" and run["m"] == g.m and h.size == g.m, "hidx run of another graph"
            out.update({"hidx_sha256": digest(h), "hidx_kmax": int(h.max()) if g.m else 0,
                        "hidx_rounds_jacobi": run["iters"],
                        "hidx_stats": {k: run[k] for k in ("w_h", "i_h", "b_h")}})
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
        print("wrote", path)


if __name__ == "__main__":
    main()
