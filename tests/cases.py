"""Small seeded graphs shared by the parity tests (sizes the oracle finishes
in well under a second)."""
from __future__ import annotations

import numpy as np

from cumetis_amd import graphs
from cumetis_amd.engine import Graph


def hub_graph(seed: int = 5, side: int = 90, hub_deg: int = 6000, mids=(90, 700, 3000),
              offsets=np.int32) -> Graph:
    """Grid plus one hub row above 4096 entries and medium rows of the given
    degrees, so that every sort bin of the contraction is exercised."""
    rng = np.random.default_rng(seed)
    n0 = side * side
    idx = np.arange(n0).reshape(side, side)
    u = [idx[:, :-1].ravel(), idx[:-1, :].ravel()]
    v = [idx[:, 1:].ravel(), idx[1:, :].ravel()]
    extra = [hub_deg, *mids]
    n = n0 + len(extra)
    for k, d in enumerate(extra):
        nb = rng.choice(n0, size=d, replace=False)
        u.append(np.full(d, n0 + k))
        v.append(nb)
    u = np.concatenate(u)
    v = np.concatenate(v)
    w = rng.integers(1, 50, size=len(u))
    return graphs.from_edges(n, u, v, w, offsets)


def isolated(n: int = 10) -> Graph:
    return Graph(np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32), None)


def named_cases():
    return {
        "delaunay4720": graphs.delaunay2d(4720, seed=1),
        "grid32w100": graphs.grid2d(32, 32, wmax=100, seed=2),
        "rmat10": graphs.rmat(10, 16, seed=3, offsets=np.int32),
        "rgg3000w16": graphs.rgg2d(3000, 8.0, wmax=16, seed=4),
        "hub": hub_graph(),
        "path2": graphs.from_edges(2, np.array([0]), np.array([1]), np.array([3])),
        "isolated10": isolated(10),
    }


def with_weights(g: Graph) -> Graph:
    """Materialise unit weights so that the oracle and engine see the same arrays."""
    aw = np.ones(g.m, np.int32) if g.adjwgt is None else g.adjwgt
    vw = np.ones(g.n, np.int32) if g.vwgt is None else g.vwgt
    return Graph(g.xadj, g.adjncy, aw, vw)
