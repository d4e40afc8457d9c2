"""Seeded synthetic graphs of the benchmark configurations and METIS I/O.

All generators return symmetric CSR graphs without self loops or duplicate
edges; edge weights are drawn once per undirected edge and mirrored, vertex
weights are 1.  Rows are sorted by neighbour id.
"""
from __future__ import annotations

import numpy as np

from .engine import Graph


def from_edges(n: int, u: np.ndarray, v: np.ndarray, w: np.ndarray | None = None,
               offsets=np.int32) -> Graph:
    """CSR from undirected edges (u, v); drops self loops and duplicates
    (keeping the first weight of a duplicate)."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    keep = u != v
    u, v = u[keep], v[keep]
    if w is not None:
        w = np.asarray(w, np.int64)[keep]
    a, b = np.minimum(u, v), np.maximum(u, v)
    key = a * n + b
    key, first = np.unique(key, return_index=True)
    a, b = key // n, key % n
    if w is not None:
        w = w[first]
    src = np.concatenate([a, b])
    dst = np.concatenate([b, a])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    xadj = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=xadj[1:])
    adjwgt = None
    if w is not None:
        adjwgt = np.concatenate([w, w])[order].astype(np.int32)
    return Graph(xadj.astype(offsets), dst.astype(np.int32), adjwgt, None)


def grid2d(nx: int, ny: int, wmax: int = 0, seed: int = 1, offsets=np.int32) -> Graph:
    """5-point stencil grid; wmax > 0 draws edge weights uniform in [1, wmax]."""
    idx = np.arange(nx * ny, dtype=np.int64).reshape(ny, nx)
    u = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    v = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    w = None
    if wmax > 0:
        w = np.random.default_rng(seed).integers(1, wmax + 1, size=len(u))
    return from_edges(nx * ny, u, v, w, offsets)


def grid3d(nx: int, ny: int, nz: int, wmax: int = 0, seed: int = 1, offsets=np.int32) -> Graph:
    """7-point stencil grid."""
    idx = np.arange(nx * ny * nz, dtype=np.int64).reshape(nz, ny, nx)
    u = np.concatenate([idx[:, :, :-1].ravel(), idx[:, :-1, :].ravel(), idx[:-1, :, :].ravel()])
    v = np.concatenate([idx[:, :, 1:].ravel(), idx[:, 1:, :].ravel(), idx[1:, :, :].ravel()])
    w = None
    if wmax > 0:
        w = np.random.default_rng(seed).integers(1, wmax + 1, size=len(u))
    return from_edges(nx * ny * nz, u, v, w, offsets)


def rgg2d(n: int, mean_degree: float = 8.0, wmax: int = 0, seed: int = 1,
          offsets=np.int32) -> Graph:
    """2D random geometric graph in the unit square, radius for the given
    mean degree (r = sqrt(d / (pi n)))."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    r = np.sqrt(mean_degree / (np.pi * n))
    pairs = cKDTree(pts).query_pairs(r, output_type="ndarray")
    w = None
    if wmax > 0:
        w = rng.integers(1, wmax + 1, size=len(pairs))
    return from_edges(n, pairs[:, 0], pairs[:, 1], w, offsets)


def delaunay2d(n: int, wmax: int = 0, seed: int = 1, offsets=np.int32) -> Graph:
    """Planar Delaunay triangulation of n uniform points in the unit square."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    tri = Delaunay(pts).simplices.astype(np.int64)
    u = np.concatenate([tri[:, 0], tri[:, 1], tri[:, 2]])
    v = np.concatenate([tri[:, 1], tri[:, 2], tri[:, 0]])
    w = None
    if wmax > 0:
        w = rng.integers(1, wmax + 1, size=len(u))
    return from_edges(n, u, v, w, offsets)


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1, offsets=np.int64, chunk: int = 1 << 24) -> Graph:
    """R-MAT (Chakrabarti, Zhan, Faloutsos 2004), symmetrised, self loops and
    duplicates removed, unit weights."""
    n = 1 << scale
    E = edge_factor * n
    rng = np.random.default_rng(seed)
    keys = []
    ab, abc = a + b, a + b + c
    for s in range(0, E, chunk):
        e = min(chunk, E - s)
        u = np.zeros(e, np.int64)
        v = np.zeros(e, np.int64)
        for bit in range(scale):
            r = rng.random(e)
            down = r >= ab                       # quadrants c, d
            right = (r >= a) & (r < ab) | (r >= abc)  # quadrants b, d
            u |= down.astype(np.int64) << bit
            v |= right.astype(np.int64) << bit
        keep = u != v
        u, v = u[keep], v[keep]
        keys.append(np.minimum(u, v) * n + np.maximum(u, v))
    key = np.unique(np.concatenate(keys))
    del keys
    lo, hi = key // n, key % n
    del key
    src = np.concatenate([lo, hi])
    dst = np.concatenate([hi, lo])
    del lo, hi
    order = np.lexsort((dst, src))
    dst = dst[order].astype(np.int32)
    cnt = np.bincount(src, minlength=n)
    del src, order
    xadj = np.zeros(n + 1, np.int64)
    np.cumsum(cnt, out=xadj[1:])
    return Graph(xadj.astype(offsets), dst, None, None)


def _take_csr(lib, n, px, pa, pw, offsets, weighted):
    import ctypes as C
    m = C.cast(px.value + 8 * n, C.POINTER(C.c_int64))[0]
    xadj = np.ctypeslib.as_array(C.cast(px.value, C.POINTER(C.c_int64)), shape=(n + 1,)).copy()
    adj = np.ctypeslib.as_array(C.cast(pa.value, C.POINTER(C.c_int32)), shape=(max(m, 1),))[:m].copy()
    wgt = np.ctypeslib.as_array(C.cast(pw.value, C.POINTER(C.c_int32)), shape=(max(m, 1),))[:m].copy()
    for p in (px, pa, pw):
        lib.cm_free_host(p)
    return Graph(xadj.astype(offsets), adj, wgt if weighted else None, None)


def rmat_fast(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
              seed: int = 1, offsets=np.int64) -> Graph:
    """R-MAT from the engine library's multithreaded generator (a different
    seeded instance than ``rmat``, same distribution)."""
    import ctypes as C
    from . import _lib
    lib = _lib.load()
    n = C.c_int64()
    px, pa, pw = C.c_void_p(), C.c_void_p(), C.c_void_p()
    rc = lib.cm_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(n), C.byref(px), C.byref(pa),
                         C.byref(pw))
    if rc != 0:
        raise RuntimeError(f"cm_gen_rmat failed ({rc})")
    return _take_csr(lib, n.value, px, pa, pw, offsets, False)


def rgg2d_fast(n: int, mean_degree: float = 8.0, wmax: int = 0, seed: int = 1,
               offsets=np.int32) -> Graph:
    """2D random geometric graph from the engine library's generator."""
    import ctypes as C
    from . import _lib
    lib = _lib.load()
    px, pa, pw = C.c_void_p(), C.c_void_p(), C.c_void_p()
    rc = lib.cm_gen_rgg2d(n, mean_degree, wmax, seed, C.byref(px), C.byref(pa), C.byref(pw))
    if rc != 0:
        raise RuntimeError(f"cm_gen_rgg2d failed ({rc})")
    return _take_csr(lib, n, px, pa, pw, offsets, wmax > 0)


# ------------------------------------------------------------- METIS files
def read_metis(path: str) -> Graph:
    """METIS .graph reader (format of cu:3712-3851: header n m [fmt], 1-based
    neighbour ids, optional vertex/edge weights, '%' comment lines)."""
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith("%")]
    head = lines[0].split()
    n, m = int(head[0]), int(head[1])
    fmt = int(head[2]) if len(head) > 2 else 0
    fs = f"{fmt % 1000:03d}"
    vw, ew = fs[1] == "1", fs[2] == "1"
    xadj = np.zeros(n + 1, np.int64)
    adj, wts, vwgt = [], [], np.ones(n, np.int32)
    for i in range(n):
        tok = lines[1 + i].split() if 1 + i < len(lines) else []
        k = 0
        if vw:
            vwgt[i] = int(tok[0])
            k = 1
        step = 2 if ew else 1
        row = tok[k:]
        for t in range(0, len(row) - (step - 1), step):
            adj.append(int(row[t]) - 1)
            wts.append(int(row[t + 1]) if ew else 1)
        xadj[i + 1] = len(adj)
    if xadj[-1] != 2 * m:
        raise ValueError(f"header says {m} edges, file has {xadj[-1] // 2}")
    return Graph(xadj.astype(np.int32), np.array(adj, np.int32), np.array(wts, np.int32),
                 vwgt if vw else None)


def write_metis(path: str, g: Graph):
    ew = g.adjwgt is not None
    vw = g.vwgt is not None
    fmt = f"{int(vw)}{int(ew)}" if (vw or ew) else ""
    with open(path, "w") as f:
        f.write(f"{g.n} {g.m // 2}" + (f" {int(fmt):03d}" if fmt else "") + "\n")
        for i in range(g.n):
            parts = []
            if vw:
                parts.append(str(int(g.vwgt[i])))
            for j in range(int(g.xadj[i]), int(g.xadj[i + 1])):
                parts.append(str(int(g.adjncy[j]) + 1))
                if ew:
                    parts.append(str(int(g.adjwgt[j])))
            f.write(" ".join(parts) + "\n")
