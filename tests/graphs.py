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


# A graph on which bitruss_mspp's turned-around pivots change the result.
# After the first-appearance relabel: upper 0 (degree 27: lower 3, 4 and 25
# leaves), upper 1 (lower 3, 4, 5), upper 2 (lower 4, 5); edge ids e0=(0,3),
# e1=(0,4), e2..e26 the leaves, e27=(1,3), e28=(1,4), e29=(1,5), e30=(2,4),
# e31=(2,5). Butterflies {0,1}x{3,4} and {1,2}x{4,5}; supports 1 except
# e28 = 2. Level 1 peels everything but e28 at once. In {0,1}x{3,4} the
# survivor e28 has o = e0, p = e27, q = e1: msp's o<q<p (0<1<27) decrements
# it; e0 has deg 27 > 10 * deg(lower 3) = 20, so mspp takes it as a pivot
# turned around, whose condition reads o<p<q (0<27<1): no decrement. In
# {1,2}x{4,5} (o = e31, p = e29, q = e30) no ordering holds. So msp gives
# e28 = 1 and mspp e28 = 2; the textbook peel (true bitruss) gives 1.
TURN_EDGES = [(1, 1), (0, 0), (0, 1), (1, 0), (2, 2), (0, 2), (2, 0)] + [(1, 3 + i) for i in range(25)]
TURN_MSP = [1, 1] + [0] * 25 + [1, 1, 1, 1, 1]
TURN_MSPP = [1, 1] + [0] * 25 + [1, 2, 1, 1, 1]


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


def turn_graph(seed: int) -> gb.Graph:
    """One upper vertex adjacent to every lower vertex, the others sparse:
    its edges have an upper endpoint of more than ten times the degree of
    the lower one, which bitruss_mspp takes as pivots the other way round."""
    rng = np.random.default_rng(seed)
    nu, nv = int(rng.integers(6, 12)), int(rng.integers(25, 40))
    d = float(rng.uniform(0.05, 0.12))
    pairs = {(0, j) for j in range(nv)}
    for i in range(1, nu):
        for j in range(nv):
            if rng.random() < d:
                pairs.add((i, j))
    p = np.array(sorted(pairs), dtype=np.uint32)
    rng.shuffle(p)
    return from_pairs(p)
