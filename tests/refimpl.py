"""Independent pure-Python restatements for tiny graphs, used to pin the
oracle. They are derived from the mathematics, not retyped from the
reference loops:

* butterflies by upper-vertex pairs: total = sum over pairs {a,b} of
  C(|N(a) & N(b)|, 2); support(a,x) = sum over b with x in N(a)&N(b) of
  (|N(a)&N(b)| - 1);
* true bitruss numbers by textbook peeling (always remove an edge of
  minimum support, decrementing its live butterflies' other edges while
  they exceed it);
* msp by butterflies: each sub-round looks at every live butterfly with a
  frontier edge; survivors are decremented once when 2 or 3 edges survive,
  and the single survivor f of a butterfly with three frontier edges is
  decremented iff (o<q<p) or (p<o<q) or (q<o<p), where o is the frontier
  edge opposite f, p the one sharing f's upper vertex and q the one
  sharing f's lower vertex (this is what the per-pivot rule of
  cpu_bitruss.cpp:301-316 amounts to, derived by placing f in each of the
  three pivot roles).
"""
from __future__ import annotations

from collections import defaultdict
from itertools import combinations


def structure(g):
    """(edges list [(u, v)], adjacency sets, edge-id dict) of a Graph."""
    m = g.m
    edges = [(int(g.edges[2 * e]), int(g.edges[2 * e + 1])) for e in range(m)]
    adj = defaultdict(set)
    eid = {}
    for e, (u, v) in enumerate(edges):
        adj[u].add(v)
        adj[v].add(u)
        eid[(u, v)] = e
    return edges, adj, eid


def butterflies(g):
    """All butterflies as (ua, ub, vx, vy) with their 4 edge ids
    (ua-vx, ua-vy, ub-vx, ub-vy)."""
    edges, adj, eid = structure(g)
    out = []
    uppers = sorted({u for u, _ in edges})
    for a, b in combinations(uppers, 2):
        common = sorted(adj[a] & adj[b])
        for x, y in combinations(common, 2):
            out.append((a, b, x, y, eid[(a, x)], eid[(a, y)], eid[(b, x)], eid[(b, y)]))
    return out


def brute_support(g):
    """(total, support list) by upper-vertex pairs."""
    edges, adj, eid = structure(g)
    sup = [0] * g.m
    total = 0
    uppers = sorted({u for u, _ in edges})
    for a, b in combinations(uppers, 2):
        common = adj[a] & adj[b]
        c = len(common)
        total += c * (c - 1) // 2
        if c >= 2:
            for x in common:
                sup[eid[(a, x)]] += c - 1
                sup[eid[(b, x)]] += c - 1
    return total, sup


def textbook_peel(g):
    """True bitruss numbers."""
    _, sup = brute_support(g)
    sup = list(sup)
    bfs = butterflies(g)
    by_edge = defaultdict(list)
    for i, bf in enumerate(bfs):
        for e in bf[4:]:
            by_edge[e].append(i)
    dead = [False] * len(bfs)
    done = [False] * g.m
    out = [0] * g.m
    for _ in range(g.m):
        e = min((x for x in range(g.m) if not done[x]), key=lambda x: sup[x])
        out[e] = sup[e]
        for i in by_edge[e]:
            if dead[i]:
                continue
            dead[i] = True
            for f in bfs[i][4:]:
                if f != e and sup[f] > sup[e]:
                    sup[f] -= 1
        done[e] = True
    return out


def msp_by_butterflies(g, variant: str = "", max_level: int = 0, turned=frozenset()):
    """(result list, levels) of the reference msp, butterfly-centric.

    `variant` plants a deliberate mistake for the tests that show the pins
    can see one: "always" decrements every single survivor, "swap_pq"
    exchanges the roles of the two frontier edges adjacent to it.
    `max_level` > 0 stops before the first level above it: the list then
    holds the peeled edges' results and the other edges' supports.
    `turned` holds the edges that bitruss_mspp takes as pivots the other way
    round (process_edge, bitruss_msp.cu:366-378). Only the pivot opposite
    the survivor sees its orientation: its literal condition names the
    survivor's neighbours by the pivot's own (u, v), so turning it around
    exchanges p and q in (o<q<p); the adjacent pivots' conditions read
    pivot < other adjacent < opposite either way."""
    _, sup = brute_support(g)
    sup = list(sup)
    bfs = butterflies(g)
    processed = [s == 0 for s in sup]
    level_max = 0
    while True:
        alive = [e for e in range(g.m) if not processed[e]]
        if not alive:
            break
        level = min(sup[e] for e in alive)
        if max_level and level > max_level:
            break
        level_max = level
        curr = {e for e in alive if sup[e] == level}
        while curr:
            dec = defaultdict(int)
            for (_, _, _, _, ax, ay, bx, by) in bfs:
                four = (ax, ay, bx, by)
                if any(processed[e] for e in four):
                    continue
                inc = [e for e in four if e in curr]
                if not inc:
                    continue
                surv = [e for e in four if e not in curr]
                if len(surv) in (2, 3):
                    for f in surv:
                        dec[f] += 1
                elif len(surv) == 1:
                    f = surv[0]
                    # roles relative to f: opposite, same upper, same lower
                    opposite = {ax: by, ay: bx, bx: ay, by: ax}
                    same_upper = {ax: ay, ay: ax, bx: by, by: bx}
                    same_lower = {ax: bx, bx: ax, ay: by, by: ay}
                    o, p, q = opposite[f], same_upper[f], same_lower[f]
                    if variant == "swap_pq":
                        p, q = q, p
                    first = (o < p < q) if o in turned else (o < q < p)
                    if variant == "always" or first or p < o < q or q < o < p:
                        dec[f] += 1
            nxt = set()
            for f, k in dec.items():
                s = sup[f]
                if s > level and s - k <= level:
                    nxt.add(f)
                sup[f] = max(level, s - k)
            for e in curr:
                processed[e] = True
            curr = nxt
    return sup, level_max


def turned_edges(g):
    """Edges bitruss_mspp takes the other way round: deg(upper) > 10 deg(lower)."""
    deg = [int(x) for x in g.degrees]
    return frozenset(e for e in range(g.m) if deg[int(g.edges[2 * e])] > 10 * deg[int(g.edges[2 * e + 1])])
