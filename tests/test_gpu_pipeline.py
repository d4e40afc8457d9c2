"""GPU: batched coarsening against one-by-one, and the full partition
pipeline against the oracle pipeline."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cases import hub_graph, named_cases, with_weights
from conftest import say
from cumetis_amd import graphs
from cumetis_amd._lib import CM_ROW_ASCENDING, CM_ROW_FIRST_SEEN
from cumetis_amd.engine import Graph, concat_graphs

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def batch_graphs():
    rng = np.random.default_rng(17)
    gs = [graphs.grid2d(40, 30, wmax=16, seed=1), graphs.rgg2d(2500, 8.0, wmax=16, seed=2),
          Graph(np.zeros(1, np.int32), np.zeros(0, np.int32), None, None),          # empty graph
          graphs.grid2d(7, 5, wmax=16, seed=3), named_cases()["isolated10"],
          graphs.rmat(11, 8, seed=5, offsets=np.int32), graphs.delaunay2d(3000, wmax=16, seed=6),
          graphs.rgg2d(600, 8.0, wmax=16, seed=7), hub_graph(seed=9, side=70, hub_deg=4500, mids=(80, 300)),
          graphs.grid2d(64, 64, wmax=16, seed=8)]
    out = []
    for g in gs:
        g = with_weights(g)
        g.vwgt[:] = rng.integers(1, 3, size=g.n)
        out.append(g)
    return out


def levels_equal(a, b):
    assert a.num_levels == b.num_levels
    for i in range(a.num_levels):
        x, y = a.level(i), b.level(i)
        assert x.n == y.n and x.m == y.m, i
        for f in ("xadj", "adjncy", "adjwgt", "vwgt", "match", "cmap"):
            u, v = getattr(x, f), getattr(y, f)
            if u is None or v is None:
                assert u is None and v is None
            else:
                assert np.array_equal(u.astype(np.int64), v.astype(np.int64)), (i, f)


@pytest.mark.parametrize("order", [CM_ROW_FIRST_SEEN, CM_ROW_ASCENDING])
def test_batch_equals_alone(engine, order):
    gs = batch_graphs()
    ks = [8, 4, 2, 2, 2, 16, 8, 4, 4, 32]
    seeds = list(range(3, 3 + len(gs)))
    say(f"[batch vs alone] graphs={len(gs)} order={order}")
    hs = engine.coarsen_batch(gs, ks, seed=seeds, row_order=order)
    for b, g in enumerate(gs):
        alone = engine.coarsen(g, ks[b], seed=seeds[b], row_order=order)
        levels_equal(hs[b], alone)
        alone.release()
    for h in hs:
        h.release()


def test_batch_device_input_and_int64(engine):
    gs = batch_graphs()[:6]
    cg, voff, eoff = concat_graphs(gs)
    cg64 = Graph(cg.xadj.astype(np.int64), cg.adjncy, cg.adjwgt, cg.vwgt)
    say("[batch device/int64]")
    a = engine.coarsen_batch(gs, 8, concat=(cg, voff, eoff))
    b = engine.coarsen_batch(gs, 8, concat=(cg64, voff, eoff), on_device=True)
    for x, y in zip(a, b):
        levels_equal(x, y)


@pytest.mark.parametrize("balance", [True, False])
@pytest.mark.parametrize("name,k", [("delaunay4720", 8), ("grid100", 16), ("rgg", 4), ("hub", 8)])
def test_partition_bit_exact_with_oracle_pipeline(engine, orc, name, k, balance):
    if name == "grid100":
        g = graphs.grid2d(100, 100, wmax=100, seed=2)
    elif name == "rgg":
        g = graphs.rgg2d(8000, 8.0, wmax=16, seed=3)
    elif name == "hub":
        g = hub_graph()
    else:
        g = named_cases()[name]
    g = with_weights(g)
    say(f"[partition vs oracle pipeline] {name} k={k} balance={balance}")
    part, rep = engine.partition(g, k, seed=1, balance=balance)
    where, levels = orc.partition(g.xadj, g.adjncy, g.adjwgt, g.vwgt, k, matcher="handshake", seed=1,
                                  balance=balance)
    assert rep.levels == len(levels)
    assert np.array_equal(part, where)
    assert rep.edgecut == orc.edgecut(g.xadj, g.adjncy, g.adjwgt, part)


def _imbalance(vw, part, k):
    pw = np.bincount(part, weights=vw, minlength=k)
    return pw.max() * k / vw.sum()


@pytest.mark.parametrize("name,k", [("delaunay4720", 8), ("grid200w100", 64), ("rgg20k", 16)])
def test_partition_quality_vs_reference_pipeline(engine, orc, name, k):
    """Edge cut of the GPU pipeline against the reference pipeline (its CPU
    HEM coarsening + the same initial partition and refinement)."""
    if name == "grid200w100":
        g = graphs.grid2d(200, 200, wmax=100, seed=1)
    elif name == "rgg20k":
        g = graphs.rgg2d(20000, 8.0, wmax=16, seed=1)
    else:
        g = named_cases()[name]
    g = with_weights(g)
    say(f"[quality] {name} k={k}")
    part, rep = engine.partition(g, k, seed=1)
    ref, _ = orc.partition(g.xadj, g.adjncy, g.adjwgt, g.vwgt, k, matcher="cpu")
    cut = rep.edgecut
    rcut = orc.edgecut(g.xadj, g.adjncy, g.adjwgt, ref)
    imb = _imbalance(g.vwgt, part, k)
    print(f"cut={cut} ref_cut={rcut} ratio={cut / rcut:.4f} imb={imb:.4f} "
          f"ref_imb={_imbalance(g.vwgt, ref, k):.4f}", flush=True)
    assert part.min() >= 0 and part.max() < k
    assert imb <= 1.03 + 1e-6


def test_cli_3elt(tmp_path):
    g = tmp_path / "3elt.graph"
    g.write_text((ROOT / "tests" / "golden" / "3elt.graph").read_text())
    say("[cli] 3elt k=8")
    r = subprocess.run([str(ROOT / "cumetis_amd" / "cuMetis"), str(g), "8"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Edge-cut=" in r.stdout and "cuMetis_Partition_time=" in r.stdout
    part = np.array([int(x) for x in (tmp_path / "3elt.graph.part.8").read_text().split()])
    assert len(part) == 4720 and part.min() == 0 and part.max() == 7
    cut = int(r.stdout.split("Edge-cut=")[1].split()[0])
    gg = graphs.read_metis(str(g))
    src = np.repeat(np.arange(gg.n), np.diff(gg.xadj))
    assert cut == int((part[src] != part[gg.adjncy]).sum()) // 2
