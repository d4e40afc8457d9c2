"""The reference's own graph (mtx/3elt.mtx, converted with mtx_to_graph.c
semantics into tests/golden/3elt.graph) through the oracle and the engine."""
from pathlib import Path

import numpy as np
import pytest

from cumetis_amd import graphs
from cumetis_amd.engine import Graph

GOLD = Path(__file__).resolve().parent / "golden"


def load3elt():
    g = graphs.read_metis(str(GOLD / "3elt.graph"))
    return Graph(g.xadj, g.adjncy, np.ones(g.m, np.int32), np.ones(g.n, np.int32))


def test_3elt_shape():
    g = load3elt()
    assert g.n == 4720 and g.m == 2 * 13722       # header of mtx/3elt.mtx
    src = np.repeat(np.arange(g.n), np.diff(g.xadj))
    assert not np.any(src == g.adjncy)


def test_3elt_oracle_fixture(orc):
    g = load3elt()
    f = np.load(GOLD / "3elt_k8.npz")
    assert int(f["T"]) == orc.threshold(g.n, 8)
    orc.srand(1)
    match, cmap, nc, perm = orc.cpu_match(g.xadj, g.adjncy, g.adjwgt, None, int(f["maxvwgt"]))
    assert np.array_equal(match, f["match"]) and np.array_equal(perm, f["perm"])
    cx, ca, cw, cv = orc.cpu_contract(g.xadj, g.adjncy, g.adjwgt, None, match, cmap, nc)
    for k, a in (("cxadj", cx), ("cadjncy", ca), ("cadjwgt", cw), ("cvwgt", cv)):
        assert np.array_equal(a, f[k]), k


@pytest.mark.gpu
def test_3elt_gpu_contract_matches_fixture(engine):
    g = load3elt()
    f = np.load(GOLD / "3elt_k8.npz")
    nc = len(f["cvwgt"])
    cx, ca, cw, cv = engine.contract(g, f["match"], f["cmap"], nc)
    assert np.array_equal(cx.astype(np.int64), f["cxadj"])
    assert np.array_equal(ca, f["cadjncy"]) and np.array_equal(cw, f["cadjwgt"])
    assert np.array_equal(cv, f["cvwgt"])
