"""Measures the counting launch settings per graph-size class on the GPU and
writes the table the library is built with (csrc/bt_tune_table.h).

    python -m gpu_bitruss_amd.tune [--reps 3] [--out FILE] [--no-60m] [--json FILE]

A size class c holds graphs of 4^c to 4^(c+1) - 1 edges (batch unions: edges
per graph). Each class with a representative graph is tuned on it: graphs of
the 4,096-graph mix for the small classes, the 2M-edge power-law graph and
the 60M-edge skewed graph for the large ones. bt_autotune times every
candidate (best of --reps runs of the per-edge support count), checks that
all candidates give the same supports, and the fastest is written for the
class; classes without a representative take the nearest tuned class's
settings. Rebuild the library afterwards (__graft_entry__.build()).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import gpu_bitruss_amd as gb  # noqa: E402

CLASSES = 16


def representatives(with_60m: bool):
    """(class, name, graph) per class that has one: mix graphs first by
    class, then the named large workloads."""
    picked = {}
    for i in range(0, 4096, 7):
        if len(picked) >= 7:
            break
        g = gb.Graph.generate("mix_graph", i)
        c = gb.tune_class(g.m)
        if 3 <= c <= 9 and c not in picked:
            picked[c] = (f"mix_graph {i}", g)
    out = [(c, n, g) for c, (n, g) in sorted(picked.items())]
    g = gb.Graph.generate("powerlaw_2m")
    out.append((gb.tune_class(g.m), "powerlaw_2m", g))
    if with_60m:
        g = gb.Graph.generate("skewed_60m")
        out.append((gb.tune_class(g.m), "skewed_60m", g))
    return out


def write_table(path: str, table: dict, notes: dict) -> None:
    lines = [
        "// Counting launch settings per graph-size class, measured on MI355X by",
        "// gpu_bitruss_amd/tune.py (which rewrites this file). Class c holds graphs",
        "// of 4^c to 4^(c+1) - 1 edges (batch unions: edges per graph). Fields:",
        "// wave_wedges, block_wedges, block_threads, hub_threads.",
        "#pragma once",
        "static const bt_count_config kTuneTable[BT_TUNE_CLASSES] = {",
    ]
    for c in range(CLASSES):
        a, b, t, h = table[c]
        lines.append(f"    {{{a}, {b}, {t}, {h}}},  // class {c}: {notes[c]}")
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "bt_tune_table.h"))
    ap.add_argument("--no-60m", action="store_true", help="leave out the 60M-edge graph")
    ap.add_argument("--json", default="", help="also write every candidate's time here")
    args = ap.parse_args()
    report = {}
    tuned = {}
    with gb.Device(0) as d:
        for c, name, g in representatives(not args.no_60m):
            t0 = time.time()
            dg = d.upload(g)
            default = gb.tune_default(c)
            best, best_ms, times = d.autotune(dg, reps=args.reps)
            dg.free()
            d.trim()
            tuned[c] = best
            report[c] = {"graph": name, "m": g.m, "best": best, "best_ms": best_ms,
                         "default": default, "default_ms": times[default],
                         "times": {",".join(map(str, k)): v for k, v in times.items()}}
            print(f"[tune] class {c:2d} {name} (m={g.m}): best {best} {best_ms:.3f} ms, "
                  f"previous {default} {times[default]:.3f} ms ({time.time() - t0:.1f} s)", flush=True)
    table, notes = {}, {}
    for c in range(CLASSES):
        if c in tuned:
            r = report[c]
            table[c] = tuned[c]
            notes[c] = f"{r['graph']}, m {r['m']}: {r['best_ms']:.3f} ms (previous {r['default_ms']:.3f} ms)"
        else:
            near = min(tuned, key=lambda k: (abs(k - c), k))
            table[c] = tuned[near]
            notes[c] = f"as class {near}"
    write_table(args.out, table, notes)
    print(f"[tune] wrote {args.out}", flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
