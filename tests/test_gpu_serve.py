"""GPU: the continuous-batching server.  Every job's partition must be the
same alone (cm_partition) and in the mix, and its reported edge cut the
oracle's cut of that partition.  A request's trials (4 by default) share one
device copy of its graph."""
import threading
import time

import numpy as np
import pytest

from cases import hub_graph, named_cases, with_weights
from conftest import say
from cumetis_amd import graphs
from cumetis_amd._lib import EngineError
from cumetis_amd.engine import Server

pytestmark = pytest.mark.gpu


def mix():
    gs = [(graphs.grid2d(60, 50, wmax=16, seed=1), 8), (graphs.rgg2d(4000, 8.0, wmax=16, seed=2), 4),
          (named_cases()["delaunay4720"], 8), (graphs.grid2d(120, 90, wmax=100, seed=3), 16),
          (graphs.rmat(11, 8, seed=5, offsets=np.int32), 4), (hub_graph(seed=9), 2),
          (graphs.rgg2d(900, 8.0, wmax=16, seed=7), 2), (graphs.grid3d(20, 20, 20, seed=4), 32),
          (named_cases()["path2"], 2), (named_cases()["isolated10"], 2), (graphs.grid2d(33, 33, seed=8), 3)]
    return [(with_weights(g), k) for g, k in gs]


def test_server_results_equal_alone(engine, orc):
    jobs = mix()
    say(f"[serve] {len(jobs)} jobs, staggered submits")
    alone = [engine.partition(g, k, seed=1) for g, k in jobs]
    with Server(0, hbm_budget=256 << 20, max_slots=4) as srv:
        tickets = []
        for i, (g, k) in enumerate(jobs):
            tickets.append(srv.submit(g, k, seed=1))
            if i % 3 == 2:
                time.sleep(0.005)  # later requests arrive while earlier ones run
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for (part, rep), (apart, arep), (g, k) in zip(res, alone, jobs):
        assert np.array_equal(part, apart)
        assert rep.edgecut == arep.edgecut and rep.levels == arep.levels
        assert rep.edgecut == orc.edgecut(g.xadj, g.adjncy, g.adjwgt, part)
    assert st["completed"] == len(jobs)
    assert st["max_active"] >= 2, st


def test_server_concurrent_clients(engine):
    jobs = mix()[:6]
    say("[serve] two client threads")
    alone = [engine.partition(g, k, seed=1)[0] for g, k in jobs]
    out = [None] * len(jobs)
    with Server(0, max_slots=8) as srv:
        def client(idx):
            for i in idx:
                t = srv.submit(jobs[i][0], jobs[i][1], seed=1)
                out[i] = srv.wait(t)[0]
        th = [threading.Thread(target=client, args=(list(range(s, len(jobs), 2)),)) for s in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    for a, b in zip(out, alone):
        assert np.array_equal(a, b)


def test_server_rejects_bad_input(engine):
    g = graphs.from_edges(3, np.array([0, 1]), np.array([1, 2]))
    bad = graphs.from_edges(3, np.array([0, 1]), np.array([1, 2]))
    bad.adjncy = bad.adjncy.copy()
    bad.adjncy[0] = 0  # self loop on vertex 0
    with Server(0, max_slots=2) as srv:
        with pytest.raises(EngineError):
            srv.submit(bad, 2)
        t = srv.submit(g, 2)
        part, rep = srv.wait(t)
        assert len(part) == 3


def test_server_mix_with_large_job(engine, orc):
    """The serving configuration's large job, a 160^3 7-point grid (4,096,000
    vertices, k=128), submitted among small jobs of the batch distribution:
    every result equals the same job partitioned alone."""
    big = graphs.grid3d(160, 160, 160)
    jobs = [(with_weights(graphs.rgg2d(3000, 8.0, wmax=16, seed=21)), 8), (big, 128),
            (with_weights(graphs.grid2d(90, 70, wmax=16, seed=22)), 16),
            (with_weights(graphs.rgg2d(12000, 8.0, wmax=16, seed=23)), 32)]
    say(f"[serve] large job n={big.n} k=128 among {len(jobs) - 1} small ones")
    alone = [engine.partition(g, k, seed=1) for g, k in jobs]
    with Server(0, max_slots=16) as srv:
        tickets = [srv.submit(g, k, seed=1) for g, k in jobs]
        res = [srv.wait(t) for t in tickets]
    for (part, rep), (apart, arep), (g, k) in zip(res, alone, jobs):
        assert np.array_equal(part, apart)
        assert rep.edgecut == arep.edgecut
        assert rep.edgecut == orc.edgecut(g.xadj, g.adjncy, g.adjwgt, part)
    pw = np.bincount(res[1][0], minlength=128)
    assert pw.max() * 128 / big.n <= 1.03 + 1e-6
