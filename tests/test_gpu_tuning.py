"""GPU: the launch-shape parameters of a context (cm_tuning) change no
result.  Extreme values give hierarchies bit-exact with the oracle, and
round batches replayed from cached hipGraphs give the hierarchies of
direct launches on every replay, with other eager work and
synchronisations between the replays."""
import numpy as np
import pytest

from cases import hub_graph, with_weights
from conftest import say
from cumetis_amd import graphs
from cumetis_amd._lib import CM_ERR_INVALID, CM_ROW_ASCENDING, CM_ROW_FIRST_SEEN, EngineError
from cumetis_amd.engine import Graph

pytestmark = pytest.mark.gpu

EXTREME = dict(first_round_batch=2, level_batch_margin=1, large_level_batch=2, compact_round=2,
               compact_min_n=1, speculate_edges=0, number_one_tile_max=0, seg_wave_bins=1)


def _graphs():
    return {
        "grid200w100": graphs.grid2d(200, 200, wmax=100, seed=3),
        "rmat14": graphs.rmat(14, 16, seed=2, offsets=np.int64),
        "hub": hub_graph(seed=7, side=150, hub_deg=9000, mids=(2048, 2049, 300, 100)),
        "delaunay3000": graphs.delaunay2d(3000, seed=4),
    }


def _levels(h):
    out = []
    for i in range(h.num_levels):
        L = h.level(i)
        out.append((L.rounds, L.xadj.astype(np.int64), L.adjncy, L.adjwgt, L.vwgt, L.match, L.cmap))
    return out


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x[0] == y[0]
        for u, v in zip(x[1:], y[1:]):
            assert (u is None) == (v is None)
            if u is not None:
                np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("order", [CM_ROW_FIRST_SEEN, CM_ROW_ASCENDING])
@pytest.mark.parametrize("name", ["grid200w100", "rmat14", "hub", "delaunay3000"])
def test_extreme_tuning_bit_exact_with_oracle(engine, orc, name, order):
    g = with_weights(_graphs()[name])
    k = 16
    say(f"[tuning extreme] {name} order={order}")
    old = engine.set_tuning(**EXTREME)
    try:
        h = engine.coarsen(g, k, seed=3, row_order=order)
        got = _levels(h)
        h.release()
    finally:
        engine.set_tuning(**old)
    ref = orc.coarsen(g.xadj, g.adjncy, g.adjwgt, g.vwgt, k, matcher="handshake", seed=3)
    assert len(got) == len(ref)
    mv = orc.maxvwgt(int(g.vwgt.sum()), orc.threshold(g.n, k))
    for i, R in enumerate(ref):
        np.testing.assert_array_equal(got[i][1], R["xadj"])
        np.testing.assert_array_equal(got[i][4], R["vwgt"])
        a, w = got[i][2], got[i][3]
        if order == CM_ROW_ASCENDING and i > 0:
            ra, rw = R["adjncy"], R["adjwgt"]
            ro = np.lexsort((ra, np.repeat(np.arange(R["n"]), np.diff(R["xadj"]))))
            np.testing.assert_array_equal(a, ra[ro])
            np.testing.assert_array_equal(w, rw[ro])
        else:
            np.testing.assert_array_equal(a, R["adjncy"])
            np.testing.assert_array_equal(w, R["adjwgt"])
        if i < len(ref) - 1:
            np.testing.assert_array_equal(got[i][5], R["match"])
            _, rounds = orc.handshake_match(R["xadj"], R["adjncy"], R["adjwgt"], R["vwgt"], mv, seed=3)
            assert got[i][0] == rounds


@pytest.mark.parametrize("name", ["grid200w100", "rmat14", "hub"])
def test_graph_replays_equal_direct_launches(engine, name):
    g = with_weights(_graphs()[name])
    say(f"[graph replays] {name}")
    h = engine.coarsen(g, 16, seed=5)
    direct = _levels(h)
    h.release()
    old = engine.set_tuning(graph_rounds=1)
    try:
        for rep in range(4):
            h = engine.coarsen(g, 16, seed=5)  # the second and later runs replay
            _same(_levels(h), direct)
            h.release()
            # eager work between the replays on the same stream
            engine.match(g, 40, seed=rep)
            engine.sync()
    finally:
        engine.set_tuning(**old)


def test_graph_replays_batch(engine):
    rng = np.random.default_rng(3)
    gs = [with_weights(graphs.grid2d(int(s), int(s), wmax=16, seed=int(s))) for s in rng.integers(30, 120, size=6)]
    gs += [with_weights(graphs.rgg2d(int(n), 8.0, wmax=16, seed=int(n))) for n in rng.integers(1000, 20000, size=6)]
    ks = [int(k) for k in rng.integers(1, 32, size=len(gs)) * 2]
    say(f"[graph replays batch] graphs={len(gs)}")
    hs = engine.coarsen_batch(gs, ks, seed=9)
    direct = [_levels(h) for h in hs]
    for h in hs:
        h.release()
    old = engine.set_tuning(graph_rounds=1)
    try:
        for rep in range(3):
            hs = engine.coarsen_batch(gs, ks, seed=9)
            for h, d in zip(hs, direct):
                _same(_levels(h), d)
                h.release()
            engine.coarsen(gs[0], 8, seed=1).release()
            engine.sync()
    finally:
        engine.set_tuning(**old)


def test_tuning_rejects_out_of_range(engine):
    for bad in (dict(first_round_batch=1), dict(seg_wave_bins=7), dict(number_one_tile_max=9000),
                dict(graph_rounds=2), dict(compact_round=1)):
        with pytest.raises(EngineError) as e:
            engine.set_tuning(**bad)
        assert e.value.code == CM_ERR_INVALID
    assert engine.tuning()["first_round_batch"] >= 2


@pytest.mark.parametrize("name", ["grid200w100", "rmat14"])
def test_run_timing_counts_every_launch(engine, name):
    """Kernel timing by runs (mode 2: one event pair around consecutive
    launches of one kernel) counts the same launches per kernel as timing
    every launch (mode 1), and neither changes the hierarchy."""
    g = with_weights(_graphs()[name])
    say(f"[run timing] {name}")
    h = engine.coarsen(g, 16, seed=5)
    plain = _levels(h)
    h.release()
    counts = {}
    try:
        for mode in (1, 2):
            engine.set_timing(mode)
            engine.reset_stats()
            h = engine.coarsen(g, 16, seed=5)
            _same(_levels(h), plain)
            h.release()
            counts[mode] = engine.kernel_stats()
    finally:
        engine.set_timing(0)
    assert {k: v["count"] for k, v in counts[1].items()} == {k: v["count"] for k, v in counts[2].items()}
    assert counts[2]["match_round"]["ms"] > 0 and counts[2]["contract_small"]["ms"] > 0
    with pytest.raises(EngineError) as e:
        engine.set_timing(3)
    assert e.value.code == CM_ERR_INVALID
