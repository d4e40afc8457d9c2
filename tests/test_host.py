"""Host parts of the product (no GPU): the C++ initial partition against the
oracle's restatement, and the METIS reader/writer of the C ABI."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from cases import named_cases, with_weights
from cumetis_amd import _lib, graphs

GOLD = Path(__file__).resolve().parent / "golden"


def product_initpart(g, k, seed=1, parallel=False):
    lib = _lib.load()
    import oracle.oracle as O
    O.srand(seed)  # both draw from the same glibc stream
    x = np.ascontiguousarray(g.xadj, np.int64)
    a = np.ascontiguousarray(g.adjncy, np.int32)
    w = np.ascontiguousarray(g.adjwgt, np.int32)
    v = np.ascontiguousarray(g.vwgt, np.int32)
    where = np.zeros(max(g.n, 1), np.int32)
    obj = C.c_int64()
    fn = lib.cm_initial_partition_par if parallel else lib.cm_initial_partition
    rc = fn(g.n, x.ctypes.data, a.ctypes.data, w.ctypes.data, v.ctypes.data, k, where.ctypes.data, C.byref(obj))
    assert rc == 0
    return where[:g.n], obj.value


@pytest.mark.parametrize("name,k", [("delaunay4720", 8), ("grid32w100", 5), ("rgg3000w16", 16),
                                    ("rmat10", 4), ("hub", 3), ("path2", 2)])
def test_initial_partition_matches_oracle(orc, name, k):
    g = with_weights(named_cases()[name])
    wp, op = product_initpart(g, k)
    orc.srand(1)
    wo, oo = orc.initpart(g.xadj, g.adjncy, g.adjwgt, g.vwgt, k)
    assert np.array_equal(wp, wo)
    assert op == oo
    assert op == orc.edgecut(g.xadj, g.adjncy, g.adjwgt, wp)  # the bisection cuts sum to the k-way cut


@pytest.mark.parametrize("k", [2, 8, 64])
def test_initial_partition_many_components(orc, k):
    """R-MAT scale 13 keeps a third of its vertices isolated, so every BFS of
    the bisection restarts thousands of times from the k-th unseen vertex
    (cu:2585-2595); the product finds it with a Fenwick tree instead of the
    reference's scan, and must pick the same vertex every time."""
    g = with_weights(graphs.rmat(13, 8, seed=6, offsets=np.int64))
    assert (np.diff(g.xadj) == 0).sum() > g.n // 5
    wp, op = product_initpart(g, k)
    orc.srand(1)
    wo, oo = orc.initpart(g.xadj, g.adjncy, g.adjwgt, g.vwgt, k)
    assert np.array_equal(wp, wo) and op == oo


def test_initial_partition_weighted_vertices(orc):
    g = with_weights(graphs.grid2d(25, 25, wmax=30, seed=4))
    g.vwgt[:] = np.random.default_rng(1).integers(1, 6, size=g.n)
    wp, op = product_initpart(g, 6)
    orc.srand(1)
    wo, oo = orc.initpart(g.xadj, g.adjncy, g.adjwgt, g.vwgt, 6)
    assert np.array_equal(wp, wo) and op == oo


def test_read_metis_3elt():
    lib = _lib.load()
    n, m = C.c_int64(), C.c_int64()
    px, pa, pw, pv = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    err = C.create_string_buffer(256)
    rc = lib.cm_read_metis(str(GOLD / "3elt.graph").encode(), C.byref(n), C.byref(m), C.byref(px),
                           C.byref(pa), C.byref(pw), C.byref(pv), err, 256)
    assert rc == 0, err.value
    ref = graphs.read_metis(str(GOLD / "3elt.graph"))
    xadj = np.ctypeslib.as_array((C.c_int64 * (n.value + 1)).from_address(px.value)).copy()
    adj = np.ctypeslib.as_array((C.c_int32 * m.value).from_address(pa.value)).copy()
    for p in (px, pa, pw, pv):
        lib.cm_free_host(p)
    assert n.value == 4720 and m.value == 27444
    assert np.array_equal(xadj, ref.xadj.astype(np.int64)) and np.array_equal(adj, ref.adjncy)


@pytest.mark.parametrize("text,msg", [
    ("3\n2\n1\n", b"number of vertices and edges"),
    ("0 1\n", b"must be positive"),
    ("2 1 112\n2\n1\n", b"Cannot read this type"),
    ("2 1\n3\n1\n", b"out of bounds"),
    ("2 1\n1\n1\n", b"self loop"),
    ("2 1 1\n2 0\n1 0\n", b"must be positive"),
    ("3 1\n2\n1 3\n2\n", b"more edges"),
    ("3 2\n2\n1\n\n", b"only found"),
    ("3 1\n2\n1\n", b"Premature end"),
])
def test_read_metis_errors(tmp_path, text, msg):
    lib = _lib.load()
    f = tmp_path / "bad.graph"
    f.write_text(text)
    n, m = C.c_int64(), C.c_int64()
    px, pa, pw, pv = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    err = C.create_string_buffer(512)
    rc = lib.cm_read_metis(str(f).encode(), C.byref(n), C.byref(m), C.byref(px), C.byref(pa),
                           C.byref(pw), C.byref(pv), err, 512)
    assert rc == _lib.CM_ERR_INPUT
    assert msg in err.value


def _read(path, threads):
    import os
    lib = _lib.load()
    os.environ["CM_READ_THREADS"] = str(threads)
    try:
        n, m = C.c_int64(), C.c_int64()
        px, pa, pw, pv = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib.cm_read_metis(str(path).encode(), C.byref(n), C.byref(m), C.byref(px), C.byref(pa),
                               C.byref(pw), C.byref(pv), err, 512)
    finally:
        del os.environ["CM_READ_THREADS"]
    if rc != 0:
        return rc, err.value
    get = lambda p, t, k: np.ctypeslib.as_array((t * k).from_address(p.value)).copy() if k else np.zeros(0)
    out = (get(px, C.c_int64, n.value + 1), get(pa, C.c_int32, m.value), get(pw, C.c_int32, m.value),
           get(pv, C.c_int32, n.value))
    for p_ in (px, pa, pw, pv):
        lib.cm_free_host(p_)
    return rc, out


@pytest.mark.parametrize("weights", ["none", "edges", "both"])
def test_read_metis_parallel_equals_sequential(tmp_path, weights):
    """The chunked reader gives the same arrays with 1, 7 and 16 threads, and
    the arrays of an independent line-by-line reader, on a file with comment
    lines and empty rows (isolated vertices) spread through it."""
    g = graphs.rgg2d(20000, 6.0, wmax=16, seed=8)
    vw = np.random.default_rng(2).integers(0, 9, size=g.n).astype(np.int32)
    gg = graphs.Graph(g.xadj, g.adjncy, g.adjwgt if weights != "none" else None, vw if weights == "both" else None)
    f = tmp_path / "g.graph"
    graphs.write_metis(str(f), gg)
    lines = f.read_text().split("\n")
    rng = np.random.default_rng(3)
    for at in sorted(rng.integers(1, len(lines) - 1, size=40), reverse=True):
        lines.insert(int(at), "% a comment line")
    f.write_text("\n".join(lines))
    ref = graphs.read_metis(str(f))
    assert (np.diff(ref.xadj) == 0).sum() > 0
    outs = [_read(f, t) for t in (1, 7, 16)]
    for rc, (x, a, w, v) in outs:
        assert rc == 0
        np.testing.assert_array_equal(x, ref.xadj.astype(np.int64))
        np.testing.assert_array_equal(a, ref.adjncy)
        np.testing.assert_array_equal(w, ref.adjwgt)
        np.testing.assert_array_equal(v, vw if weights == "both" else np.ones(g.n, np.int32))


@pytest.mark.parametrize("threads", [1, 5, 16])
def test_read_metis_parallel_first_error_wins(tmp_path, threads):
    """Errors in several chunks: the one first in file order is reported,
    as the sequential reader would; a line that passes the declared edge
    count before its own bad token reports the overflow."""
    g = graphs.grid2d(60, 60, seed=1)
    f = tmp_path / "g.graph"
    graphs.write_metis(str(f), g)
    lines = f.read_text().split("\n")
    n = g.n
    bad = lines.copy()
    bad[1 + 3000] = bad[1 + 3000] + " 9999"   # out of bounds, late
    bad[1 + 1200] = bad[1 + 1200] + " 1201"   # self loop of vertex 1201, earlier
    f.write_text("\n".join(bad))
    rc, msg = _read(f, threads)
    assert rc == _lib.CM_ERR_INPUT and msg == b"Vertex 1201 has a self loop"
    # one edge too many on the last line, followed by an out-of-range id
    over = lines.copy()
    over[n] = over[n] + f" {n - 60} 99999"
    f.write_text("\n".join(over))
    rc, msg = _read(f, threads)
    assert rc == _lib.CM_ERR_INPUT and b"more edges in the file than" in msg
    # fewer vertex lines than declared
    f.write_text("\n".join(lines[:n - 5]))
    rc, msg = _read(f, threads)
    assert rc == _lib.CM_ERR_INPUT and msg == f"Premature end of input file while reading vertex {n - 5}.".encode()


def test_write_partition(tmp_path):
    lib = _lib.load()
    part = np.array([3, 0, 2, 1], np.int32)
    f = tmp_path / "g.graph.part.4"
    assert lib.cm_write_partition(str(f).encode(), part.ctypes.data, 4) == 0
    assert f.read_text() == "3\n0\n2\n1\n"


@pytest.mark.parametrize("k", [2, 8, 64])
def test_parallel_initial_partition(orc, k):
    """The parallel form that the pipeline takes on coarsest graphs of 2^14
    vertices or more, here on an R-MAT graph a fifth of whose vertices are
    isolated: the same result on every call, its objval the oracle's edge cut
    of its result, and the isolated vertices, in index order, each on the
    lightest part (the lowest id among equals) after the rest is split."""
    g = with_weights(graphs.rmat(13, 8, seed=6, offsets=np.int64))
    deg = np.diff(g.xadj)
    iso = np.nonzero(deg == 0)[0]
    assert len(iso) > g.n // 5
    w1, o1 = product_initpart(g, k, parallel=True)
    w2, o2 = product_initpart(g, k, parallel=True)
    assert np.array_equal(w1, w2) and o1 == o2
    assert 0 <= w1.min() and w1.max() < k
    assert o1 == orc.edgecut(g.xadj, g.adjncy, g.adjwgt, w1)
    vw = np.asarray(g.vwgt, np.int64)
    pw = np.bincount(w1[deg > 0], weights=vw[deg > 0], minlength=k).astype(np.int64)
    for v in iso:
        p = int(np.argmin(pw))
        assert w1[v] == p, v
        pw[p] += vw[v]
    assert pw.max() * k <= 1.03 * vw.sum() + k * vw.max()


def test_parallel_initial_partition_components(orc):
    """Two disjoint cliques of 40 vertices and 30 isolated vertices, k = 2:
    the parallel form separates the cliques (cut 0) and splits the isolated
    vertices evenly, while the serial form keeps matching the oracle."""
    n0 = 40
    src, dst = np.triu_indices(n0, 1)
    e = np.concatenate([np.stack([src, dst], 1), np.stack([src, dst], 1) + n0])
    g = with_weights(graphs.from_edges(2 * n0 + 30, e[:, 0], e[:, 1]))
    wp, op = product_initpart(g, 2, parallel=True)
    assert op == 0 and orc.edgecut(g.xadj, g.adjncy, g.adjwgt, wp) == 0
    assert len(set(wp[:n0])) == 1 and len(set(wp[n0:2 * n0])) == 1 and wp[0] != wp[n0]
    pw = np.bincount(wp, weights=g.vwgt, minlength=2)
    assert abs(pw[0] - pw[1]) <= g.vwgt.max()
    we, oe = product_initpart(g, 2)
    orc.srand(1)
    wo, oo = orc.initpart(g.xadj, g.adjncy, g.adjwgt, g.vwgt, 2)
    assert np.array_equal(we, wo) and oe == oo


def test_parallel_initial_partition_stalled_coarsening(orc):
    """Four stars of 150 K leaves whose centres form a ring: a leaf can only
    match its centre, so coarsening stops after one level on a graph of more
    than 2^20 entries and the five region-growing tries run on threads (the
    first on the step's own arrays, the others on copies).  Same checks:
    one result on every call, objval the oracle's cut of it."""
    leaves, stars = 150_000, 4
    n = stars * (leaves + 1)
    centre = np.arange(stars) * (leaves + 1)
    src = np.repeat(centre, leaves)
    dst = (centre[:, None] + 1 + np.arange(leaves)[None, :]).ravel()
    ring = np.stack([centre, np.roll(centre, -1)], 1)
    g = with_weights(graphs.from_edges(n, np.concatenate([src, ring[:, 0]]), np.concatenate([dst, ring[:, 1]])))
    assert g.m >= 1 << 20
    w1, o1 = product_initpart(g, 4, parallel=True)
    w2, o2 = product_initpart(g, 4, parallel=True)
    assert np.array_equal(w1, w2) and o1 == o2
    assert 0 <= w1.min() and w1.max() < 4
    assert o1 == orc.edgecut(g.xadj, g.adjncy, g.adjwgt, w1)
