"""GPU parity at the sizes the benchmark quotes, bit for bit against the oracle.

- the 725x725 5-point grid (adjwgt U[1,100], unit vertex weights, k=64,
  int32 offsets): every level of the hierarchy, rounds included;
- R-MAT scale 22, edge factor 16 (.57/.19/.19, seed 1), symmetrised and
  deduplicated, int64 offsets, given WITHOUT weights (the engine's level 0
  then reads the shared unit array, the oracle gets explicit ones), k=256:
  level 0's matching, round count and numbering, and the coarse CSR
  (xadj, adjncy, adjwgt, vwgt).  The hierarchy stops after this level
  (nc/n = 0.88 >= 0.75, cu:1843-1845), so this is the whole coarsening.

- the serving mix's large job, a 160^3 7-point grid (4,096,000 vertices,
  24,422,400 entries) given WITHOUT weights, k=128: every level of the
  hierarchy (11 levels down to 26,855 vertices), level 0's rounds included.

The inputs are the benchmark's own (same generators, seeds and parameters),
given through host views, so the engine validates them as well.
"""
import numpy as np
import pytest

from cases import with_weights
from conftest import say
from cumetis_amd import graphs
from cumetis_amd.engine import Graph

pytestmark = pytest.mark.gpu


def test_hierarchy_grid725_k64(engine, orc):
    g = graphs.grid2d(725, 725, wmax=100, seed=1)
    gw = with_weights(g)
    say(f"[grid725 hierarchy] n={g.n} m={g.m} k=64")
    h = engine.coarsen(g, 64, seed=1)
    ref = orc.coarsen(gw.xadj, gw.adjncy, gw.adjwgt, gw.vwgt, 64, matcher="handshake", seed=1)
    T = orc.threshold(g.n, 64)
    mv = orc.maxvwgt(g.n, T)
    assert (T, mv) == (4380, 180)
    assert h.num_levels == len(ref) == 11
    for i, R in enumerate(ref):
        L = h.level(i)
        assert (L.n, L.m) == (R["n"], len(R["adjncy"])), i
        np.testing.assert_array_equal(L.xadj.astype(np.int64), R["xadj"])
        np.testing.assert_array_equal(L.adjncy, R["adjncy"])
        np.testing.assert_array_equal(L.adjwgt, R["adjwgt"])
        np.testing.assert_array_equal(L.vwgt, R["vwgt"])
        if i < len(ref) - 1:
            np.testing.assert_array_equal(L.match, R["match"])
            np.testing.assert_array_equal(L.cmap, R["cmap"])
            _, rounds = orc.handshake_match(R["xadj"], R["adjncy"], R["adjwgt"], R["vwgt"], mv, seed=1)
            assert L.rounds == rounds, i
    h.release()


def test_rmat22_int64_unit_weights_k256(engine, orc):
    say("[rmat22 level 0] generating")
    g = graphs.rmat_fast(22, 16, seed=1, offsets=np.int64)
    assert g.adjwgt is None and g.vwgt is None and g.xadj.dtype == np.int64
    say(f"[rmat22 level 0] n={g.n} m={g.m} k=256 int64 offsets, no weights")
    T = orc.threshold(g.n, 256)
    mv = orc.maxvwgt(g.n, T)
    assert (T, mv) == (26214, 240)
    h = engine.coarsen(g, 256, seed=1)
    assert h.num_levels == 2
    L0 = h.level(0)
    L1 = h.level(1)
    assert L0.xadj.dtype == np.int64 and L1.xadj.dtype == np.int64
    h.release()
    say("[rmat22 level 0] oracle")
    match, rounds = orc.handshake_match(g.xadj, g.adjncy, None, None, mv, seed=1)
    assert L0.rounds == rounds == 11
    np.testing.assert_array_equal(L0.match, match)
    cmap, nc = orc.number(match)
    np.testing.assert_array_equal(L0.cmap, cmap)
    assert L1.n == nc == 3701554
    cx, ca, cw, cv = orc.cpu_contract(g.xadj, g.adjncy, None, None, match, cmap, nc)
    np.testing.assert_array_equal(L1.xadj, cx)
    np.testing.assert_array_equal(L1.vwgt, cv)
    np.testing.assert_array_equal(L1.adjncy, ca)
    np.testing.assert_array_equal(L1.adjwgt, cw)
    assert not orc.coarsen_continue(nc, g.n, int(cx[-1]), T)


@pytest.mark.parametrize("unit", [False, True])
def test_rmat20_int64_hierarchy(engine, orc, unit):
    """A smaller R-MAT with int64 offsets, every level against the oracle:
    hub rows, all row classes and the frontier on level 0, with and without
    edge weights."""
    g = graphs.rmat_fast(20, 16, seed=3, offsets=np.int64)
    if not unit:
        src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.xadj))
        a = np.minimum(src, g.adjncy)
        b = np.maximum(src, g.adjncy)
        g = Graph(g.xadj, g.adjncy, (1 + (a * 7919 + b * 104729) % 13).astype(np.int32), None)
    gw = with_weights(g)
    say(f"[rmat20 int64 hierarchy] unit={unit} n={g.n} m={g.m}")
    h = engine.coarsen(g, 64, seed=4)
    ref = orc.coarsen(gw.xadj, gw.adjncy, gw.adjwgt, gw.vwgt, 64, matcher="handshake", seed=4)
    assert h.num_levels == len(ref)
    mv = orc.maxvwgt(g.n, orc.threshold(g.n, 64))
    for i, R in enumerate(ref):
        L = h.level(i)
        assert L.xadj.dtype == np.int64
        np.testing.assert_array_equal(L.xadj, R["xadj"])
        np.testing.assert_array_equal(L.adjncy, R["adjncy"])
        np.testing.assert_array_equal(L.adjwgt, R["adjwgt"])
        np.testing.assert_array_equal(L.vwgt, R["vwgt"])
        if i < len(ref) - 1:
            np.testing.assert_array_equal(L.match, R["match"])
            np.testing.assert_array_equal(L.cmap, R["cmap"])
            _, rounds = orc.handshake_match(R["xadj"], R["adjncy"], R["adjwgt"], R["vwgt"], mv, seed=4)
            assert L.rounds == rounds, i
    h.release()


def test_hierarchy_grid160cubed_k128(engine, orc):
    g = graphs.grid3d(160, 160, 160)
    assert g.adjwgt is None and g.vwgt is None
    say(f"[grid160^3 hierarchy] n={g.n} m={g.m} k=128, no weights")
    h = engine.coarsen(g, 128, seed=1)
    ref = orc.coarsen(g.xadj, g.adjncy, None, None, 128, matcher="handshake", seed=1)
    T = orc.threshold(g.n, 128)
    mv = orc.maxvwgt(g.n, T)
    assert (T, mv) == (29257, 210)
    assert h.num_levels == len(ref) == 11
    for i, R in enumerate(ref):
        L = h.level(i)
        assert (L.n, L.m) == (R["n"], len(R["adjncy"])), i
        np.testing.assert_array_equal(L.xadj.astype(np.int64), R["xadj"])
        np.testing.assert_array_equal(L.adjncy, R["adjncy"])
        np.testing.assert_array_equal(L.adjwgt, R["adjwgt"])
        np.testing.assert_array_equal(L.vwgt, R["vwgt"])
        if i < len(ref) - 1:
            np.testing.assert_array_equal(L.match, R["match"])
            np.testing.assert_array_equal(L.cmap, R["cmap"])
    _, rounds = orc.handshake_match(ref[0]["xadj"], ref[0]["adjncy"], ref[0]["adjwgt"], ref[0]["vwgt"], mv, seed=1)
    assert h.level(0).rounds == rounds
    h.release()
