"""Small test graphs shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

import gpu_bitruss_amd as gb

# data/aw.graph of the reference (10 edges)
AW_EDGES = [(0, 5), (0, 6), (1, 5), (1, 6), (2, 5), (2, 6), (2, 7), (3, 6), (4, 6), (4, 7)]
# hand-derived answers for aw.graph
AW_SUPPORT = [2, 2, 2, 2, 2, 3, 1, 0, 1, 1]
AW_BITRUSS = [2, 2, 2, 2, 2, 2, 1, 0, 1, 1]

# 7-edge graph where msp's attribution rule differs from true bitruss numbers
QUIRK7_EDGES = [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (2, 0), (2, 2)]
QUIRK7_SUPPORT = [2, 1, 1, 1, 1, 1, 1]
QUIRK7_MSP = [2, 1, 1, 1, 1, 1, 1]
QUIRK7_HIDX = [1, 1, 1, 1, 1, 1, 1]


def from_pairs(pairs) -> gb.Graph:
    a = np.array(pairs, dtype=np.uint32).reshape(-1, 2)
    return gb.Graph.from_edges(a[:, 0], a[:, 1])


def complete(a: int, b: int) -> gb.Graph:
    return from_pairs([(i, j) for i in range(a) for j in range(b)])


def random_bipartite(seed: int, nu: int, nv: int, density: float) -> gb.Graph:
    rng = np.random.default_rng(seed)
    mask = rng.random((nu, nv)) < density
    pairs = np.argwhere(mask)
    rng.shuffle(pairs)
    if len(pairs) == 0:
        pairs = np.array([[0, 0]])
    return from_pairs(pairs)


def hub_graph(seed: int, hub_deg: int, extra: int, nu: int, nv: int) -> gb.Graph:
    """One upper hub adjacent to hub_deg lower vertices plus random edges
    (exercises the hub-root counting path and large N(u))."""
    rng = np.random.default_rng(seed)
    pairs = {(0, j) for j in range(hub_deg)}
    while len(pairs) < hub_deg + extra:
        pairs.add((int(rng.integers(1, nu)), int(rng.integers(0, nv))))
    p = np.array(sorted(pairs), dtype=np.uint32)
    rng.shuffle(p)
    return from_pairs(p)


# random sparse graphs on which msp differs from the true bitruss numbers
# (found by scanning seeds 0..199 of quirk_graph), plus seeds on which the
# single-survivor order condition matters with the opposite edge smallest
# (found by scanning 200..1499)
QUIRK_SEEDS = [24, 46, 56, 139, 177]
ORDER_SEEDS = [211, 216, 244]


def quirk_graph(seed: int) -> gb.Graph:
    rng = np.random.default_rng(seed)
    nu, nv = int(rng.integers(6, 20)), int(rng.integers(6, 20))
    d = float(rng.uniform(0.15, 0.45))
    return random_bipartite(1000 + seed, nu, nv, d)
