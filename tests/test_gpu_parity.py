"""GPU parity: every operator of the HIP engine against the CPU oracle, bit
for bit, through the C ABI."""
import numpy as np
import pytest

from cases import named_cases, with_weights
from conftest import say
from cumetis_amd import graphs
from cumetis_amd._lib import CM_ROW_ASCENDING, CM_ROW_FIRST_SEEN
from cumetis_amd.engine import Graph

pytestmark = pytest.mark.gpu

CASES = named_cases()


def sort_rows(xadj, adjncy, adjwgt):
    a = adjncy.copy()
    w = adjwgt.copy()
    for i in range(len(xadj) - 1):
        s, e = int(xadj[i]), int(xadj[i + 1])
        o = np.argsort(a[s:e], kind="stable")
        a[s:e] = a[s:e][o]
        w[s:e] = w[s:e][o]
    return a, w


def maxv_for(g, k=8, orc=None):
    T = orc.threshold(g.n, k)
    return orc.maxvwgt(int(g.n if g.vwgt is None else g.vwgt.sum()), T)


@pytest.mark.parametrize("name", list(CASES))
def test_number(engine, orc, name):
    g = with_weights(CASES[name])
    say(f"[number] case={name} n={g.n} m={g.m}")
    orc.srand(1)
    match, _, _, _ = orc.cpu_match(g.xadj, g.adjncy, g.adjwgt, g.vwgt, maxv_for(g, orc=orc))
    cmap, nc = engine.number(g.xadj, match)
    ecmap, enc = orc.number(match)
    assert nc == enc
    np.testing.assert_array_equal(cmap, ecmap)


def _random_matching(rng, n, frac):
    """A symmetric matching of n vertices: random disjoint pairs over a
    fraction of them, the rest matched to themselves."""
    match = np.arange(n, dtype=np.int32)
    perm = rng.permutation(n)
    k = int(n * frac) // 2
    a, b = perm[:k], perm[k:2 * k]
    match[a], match[b] = b, a
    return match


@pytest.mark.parametrize("one_tile", [0, 4096])
@pytest.mark.parametrize("n", [1, 63, 1023, 1024, 1025, 2047, 2048, 2049, 4095, 4096, 4097, 6143, 8193, 70001])
def test_number_tile_edges(engine, orc, n, one_tile):
    """Numbering across the edges of its tiles (2048 vertices, and 4096 for
    a level numbered by one tile): cmap and nc equal the oracle's."""
    rng = np.random.default_rng(n)
    match = _random_matching(rng, n, 0.8)
    deg = rng.integers(0, 9, size=n)
    xadj = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    say(f"[number tiles] n={n} one_tile={one_tile}")
    old = engine.set_tuning(number_one_tile_max=one_tile)
    try:
        cmap, nc = engine.number(xadj, match)
    finally:
        engine.set_tuning(**old)
    ecmap, enc = orc.number(match)
    assert nc == enc
    np.testing.assert_array_equal(cmap, ecmap)


@pytest.mark.parametrize("order", [CM_ROW_FIRST_SEEN, CM_ROW_ASCENDING])
@pytest.mark.parametrize("matcher", ["cpu", "handshake"])
@pytest.mark.parametrize("name", list(CASES))
def test_contract(engine, orc, name, matcher, order):
    g = with_weights(CASES[name])
    say(f"[contract] case={name} matcher={matcher} order={order} n={g.n} m={g.m}")
    mv = maxv_for(g, orc=orc)
    if matcher == "cpu":
        orc.srand(1)
        match, cmap, nc, _ = orc.cpu_match(g.xadj, g.adjncy, g.adjwgt, g.vwgt, mv)
    else:
        match, _ = orc.handshake_match(g.xadj, g.adjncy, g.adjwgt, g.vwgt, mv, seed=11)
        cmap, nc = orc.number(match)
    cx, ca, cw, cv = engine.contract(g, match, cmap, nc, order)
    ex, ea, ew, ev = orc.cpu_contract(g.xadj, g.adjncy, g.adjwgt, g.vwgt, match, cmap, nc)
    np.testing.assert_array_equal(cx.astype(np.int64), ex)
    np.testing.assert_array_equal(cv, ev)
    if order == CM_ROW_ASCENDING:
        ea, ew = sort_rows(ex, ea, ew)
    np.testing.assert_array_equal(ca, ea)
    np.testing.assert_array_equal(cw, ew)


@pytest.mark.parametrize("name", list(CASES))
def test_match_bit_exact(engine, orc, name):
    g = with_weights(CASES[name])
    say(f"[match] case={name} n={g.n} m={g.m}")
    mv = maxv_for(g, orc=orc)
    for seed in (1, 12345):
        m_gpu, r_gpu = engine.match(g, mv, seed=seed)
        m_orc, r_orc = orc.handshake_match(g.xadj, g.adjncy, g.adjwgt, g.vwgt, mv, seed=seed)
        np.testing.assert_array_equal(m_gpu, m_orc)
        assert r_gpu == r_orc


@pytest.mark.parametrize("offsets", [np.int32, np.int64])
@pytest.mark.parametrize("unit", [False, True])
def test_match_split_rows(engine, orc, unit, offsets):
    """Rows cut into many chunks: a hub of 140000 entries (more chunks than a
    wave has lanes) and rows exactly at and just past the chunk boundary,
    with 32- and 64-bit offsets."""
    from cases import hub_graph
    g = hub_graph(seed=21, side=400, hub_deg=140000, mids=(2048, 2049, 4096, 4097), offsets=offsets)
    if unit:
        g = Graph(g.xadj, g.adjncy, None, None)
    g = with_weights(g)
    say(f"[match split rows] unit={unit} offsets={np.dtype(offsets).name} n={g.n} m={g.m}")
    mv = maxv_for(g, orc=orc)
    for seed in (1, 99):
        m_gpu, r_gpu = engine.match(g, mv, seed=seed)
        m_orc, r_orc = orc.handshake_match(g.xadj, g.adjncy, g.adjwgt, g.vwgt, mv, seed=seed)
        np.testing.assert_array_equal(m_gpu, m_orc)
        assert r_gpu == r_orc


def _symmetric_weights(g, wmax, seed):
    """Edge weights that depend only on the unordered pair, vertex weights 1..3."""
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.xadj.astype(np.int64)))
    a = np.minimum(src, g.adjncy.astype(np.int64))
    b = np.maximum(src, g.adjncy.astype(np.int64))
    aw = (1 + (a * 7919 + b * 104729 + seed) % wmax).astype(np.int32)
    vw = np.random.default_rng(seed).integers(1, 4, size=g.n).astype(np.int32)
    return Graph(g.xadj, g.adjncy, aw, vw)


@pytest.mark.parametrize("wmax", [1, 7, 1000])
def test_match_small_mid_rows(engine, orc, wmax):
    """A skewed graph (rows above 64 entries exist), so rows of 17..64
    entries go to the 16-lane group path and rows of 65..128 to the 32-lane
    one, with edge weights that prune by weight (wmax 7: many ties broken by
    the hash; 1000: few ties) and vertex weights that make some pairs too
    heavy."""
    g = _symmetric_weights(graphs.rmat(12, 16, seed=5, offsets=np.int32), wmax, seed=wmax)
    deg = np.diff(g.xadj)
    assert deg.max() > 128 and ((deg > 16) & (deg <= 64)).sum() > 100 and ((deg > 64) & (deg <= 128)).sum() > 50
    say(f"[match small mid rows] wmax={wmax} n={g.n} m={g.m}")
    mv = maxv_for(g, orc=orc)
    for seed in (2, 31):
        m_gpu, r_gpu = engine.match(g, mv, seed=seed)
        m_orc, r_orc = orc.handshake_match(g.xadj, g.adjncy, g.adjwgt, g.vwgt, mv, seed=seed)
        np.testing.assert_array_equal(m_gpu, m_orc)
        assert r_gpu == r_orc


@pytest.mark.parametrize("name", ["rmat13", "hub", "rmat20", "rmat20_off64"])
def test_match_unit_weights(engine, orc, name):
    """Graphs given without weights use the shared array of ones, and the row
    scans then read its first word instead of streaming it: the matching must
    be the oracle's with explicit unit weights (rmat20 also runs the frontier
    rounds)."""
    from cases import hub_graph
    if name == "rmat13":
        g = graphs.rmat(13, 16, seed=8, offsets=np.int32)
    elif name == "hub":
        g = hub_graph(seed=22, side=300, hub_deg=50000, mids=(2048, 2049, 300, 100))
    else:
        g = graphs.rmat_fast(20, 8, seed=10, offsets=np.int64 if name.endswith("off64") else np.int32)
    g = Graph(g.xadj, g.adjncy, None, None)
    say(f"[match unit weights] case={name} n={g.n} m={g.m} offsets={g.xadj.dtype}")
    gw = with_weights(g)
    mv = maxv_for(gw, orc=orc)
    for seed in (3, 41):
        m_gpu, r_gpu = engine.match(g, mv, seed=seed)
        m_orc, r_orc = orc.handshake_match(gw.xadj, gw.adjncy, gw.adjwgt, gw.vwgt, mv, seed=seed)
        np.testing.assert_array_equal(m_gpu, m_orc)
        assert r_gpu == r_orc


@pytest.mark.parametrize("order", [CM_ROW_FIRST_SEEN, CM_ROW_ASCENDING])
@pytest.mark.parametrize("name", ["rmat13", "hub"])
def test_contract_unit_weights(engine, orc, name, order):
    """A graph given without weights: the contraction kernels get no weight
    array and count every edge as 1.  The coarse graph must be the oracle's
    with explicit unit weights, on the register, hash (segments of all
    classes, a hub segment) and sorted paths."""
    from cases import hub_graph
    if name == "rmat13":
        g = graphs.rmat(13, 16, seed=12, offsets=np.int32)
    else:
        g = hub_graph(seed=24, side=300, hub_deg=50000, mids=(2048, 2049, 300, 100))
    g = Graph(g.xadj, g.adjncy, None, None)
    gw = with_weights(g)
    say(f"[contract unit weights] case={name} order={order} n={g.n} m={g.m}")
    mv = maxv_for(gw, orc=orc)
    match, _ = orc.handshake_match(gw.xadj, gw.adjncy, gw.adjwgt, gw.vwgt, mv, seed=13)
    cmap, nc = orc.number(match)
    cx, ca, cw, cv = engine.contract(g, match, cmap, nc, order)
    ex, ea, ew, ev = orc.cpu_contract(gw.xadj, gw.adjncy, gw.adjwgt, gw.vwgt, match, cmap, nc)
    np.testing.assert_array_equal(cx.astype(np.int64), ex)
    np.testing.assert_array_equal(cv, ev)
    if order == CM_ROW_ASCENDING:
        ea, ew = sort_rows(ex, ea, ew)
    np.testing.assert_array_equal(ca, ea)
    np.testing.assert_array_equal(cw, ew)
    assert cw.sum() == g.m - int((cmap[np.repeat(np.arange(g.n), np.diff(g.xadj))] == cmap[g.adjncy]).sum())


def test_contract_unit_weights_repeated_edges(engine, orc):
    """Unit weights stage hashed rows packed (a coarse edge weight of at most
    4 beside the id), which holds only without repeated edges.  A graph
    whose rows repeat some neighbours 5 times gives coarse weights above 4:
    the packed pass flags it and the level is redone on the sorted path,
    still equal to the oracle."""
    from cases import hub_graph
    g0 = hub_graph(seed=31, side=120, hub_deg=6000, mids=(300, 100))
    n = g0.n
    rows = [list(g0.adjncy[g0.xadj[i]:g0.xadj[i + 1]]) for i in range(n)]
    for a in (n - 1, n - 2):  # the two mid rows (hashed segments)
        for b in rows[a][:5]:
            rows[a] += [b] * 4
            rows[b] += [a] * 4
    xadj = np.zeros(n + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=xadj[1:])
    adj = np.concatenate([np.asarray(r, np.int32) for r in rows])
    g = Graph(xadj.astype(np.int32), adj, None, None)
    gw = with_weights(g)
    say(f"[contract repeated edges] n={g.n} m={g.m}")
    mv = maxv_for(gw, orc=orc)
    match, _ = orc.handshake_match(gw.xadj, gw.adjncy, gw.adjwgt, gw.vwgt, mv, seed=5)
    cmap, nc = orc.number(match)
    engine.set_timing(True)
    engine.reset_stats()
    try:
        cx, ca, cw, cv = engine.contract(g, match, cmap, nc, CM_ROW_FIRST_SEEN)
        stats = engine.kernel_stats()
    finally:
        engine.set_timing(False)
    assert "hc_dedup_block" in stats and "contract_gather" in stats  # packed pass, then the sorted path
    ex, ea, ew, ev = orc.cpu_contract(gw.xadj, gw.adjncy, gw.adjwgt, gw.vwgt, match, cmap, nc)
    assert ew.max() > 4
    np.testing.assert_array_equal(cx.astype(np.int64), ex)
    np.testing.assert_array_equal(cv, ev)
    np.testing.assert_array_equal(ca, ea)
    np.testing.assert_array_equal(cw, ew)


@pytest.mark.parametrize("name", ["rmat20", "rmat20_off64", "grid1024"])
def test_match_frontier(engine, orc, name):
    """Levels of 2^20 vertices or more list what is still unmatched after
    round 5 (vertices, class rows, hub chunks) and run the later rounds over
    those lists: the matching and the round count must be those of the
    oracle.  R-MAT scale 20 has every row class and hubs; the 1024^2 grid
    has only thread-per-row rows."""
    if name.startswith("rmat20"):
        off = np.int64 if name.endswith("off64") else np.int32
        g = _symmetric_weights(graphs.rmat_fast(20, 8, seed=9, offsets=off), 7, seed=3)
    else:
        g = with_weights(graphs.grid2d(1024, 1024, wmax=100, seed=2))
    assert g.n >= 1 << 20
    say(f"[match frontier] case={name} n={g.n} m={g.m}")
    mv = maxv_for(g, orc=orc)
    for seed in (5, 77):
        engine.set_timing(True)
        engine.reset_stats()
        try:
            m_gpu, r_gpu = engine.match(g, mv, seed=seed)
            stats = engine.kernel_stats()
        finally:
            engine.set_timing(False)
        m_orc, r_orc = orc.handshake_match(g.xadj, g.adjncy, g.adjwgt, g.vwgt, mv, seed=seed)
        assert r_orc >= 7 and "match_compact" in stats  # list launches did accepts
        np.testing.assert_array_equal(m_gpu, m_orc)
        assert r_gpu == r_orc


@pytest.mark.parametrize("offsets", [np.int32, np.int64])
@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("hub_deg,side", [(140000, 400), (600000, 800), (1100000, 1100)])
def test_contract_hub_partitions(engine, orc, hub_deg, side, unit, offsets):
    """Hub segments cut into many key partitions: 35 and 147 partitions
    (entries bucketed by partition in each chunk), the second past the length
    whose entries can be listed in LDS (2^19), so the second pass reads the
    bucket slices again; and 269 partitions, more than are bucketed, so every
    partition block reads the whole segment.  Unit: the graph is given
    without weights, so the hub pass keeps no weight copies."""
    from cases import hub_graph
    g0 = hub_graph(seed=23, side=side, hub_deg=hub_deg, mids=(5000, 2100, 1500, 600), offsets=offsets)
    if unit:
        g0 = Graph(g0.xadj, g0.adjncy, None, None)
    g = with_weights(g0)
    say(f"[contract hub partitions] hub_deg={hub_deg} unit={unit} offsets={np.dtype(offsets).name} "
        f"n={g.n} m={g.m}")
    mv = maxv_for(g, orc=orc)
    match, _ = orc.handshake_match(g.xadj, g.adjncy, g.adjwgt, g.vwgt, mv, seed=3)
    cmap, nc = orc.number(match)
    cx, ca, cw, cv = engine.contract(g0, match, cmap, nc, CM_ROW_FIRST_SEEN)
    ex, ea, ew, ev = orc.cpu_contract(g.xadj, g.adjncy, g.adjwgt, g.vwgt, match, cmap, nc)
    np.testing.assert_array_equal(cx.astype(np.int64), ex)
    np.testing.assert_array_equal(cv, ev)
    np.testing.assert_array_equal(ca, ea)
    np.testing.assert_array_equal(cw, ew)


def test_contract_hub_table_overflow(engine, orc):
    """A hub whose coarse neighbour ids all fall into one key partition
    (9000 distinct ids against an 8192-slot table): the partition's table
    fills, and the level is redone on the sorted path, still bit-exact.
    Isolated vertices keep cmap[v] = v below the hub, so the ids are chosen
    through the partition hash of the contraction (a multiplicative hash of
    the id, scaled to the partition count)."""
    P = 3  # partitions of a 9001-entry segment of 4096-entry partitions
    ids = np.arange(30000, dtype=np.uint64)
    part = ((((ids * np.uint64(0x85EBCA77)) & np.uint64(0xFFFFFFFF)) * np.uint64(P)) >> np.uint64(32))
    S = np.nonzero(part == 0)[0][:9000]
    assert len(S) == 9000
    hub = 30000
    g = graphs.from_edges(hub + 1, np.full(len(S), hub), S, np.random.default_rng(4).integers(1, 50, size=len(S)))
    g = with_weights(g)
    say(f"[contract hub overflow] n={g.n} m={g.m}")
    mv = maxv_for(g, orc=orc)
    match, _ = orc.handshake_match(g.xadj, g.adjncy, g.adjwgt, g.vwgt, mv, seed=3)
    cmap, nc = orc.number(match)
    assert np.array_equal(cmap[:hub], np.arange(hub))  # the ids above are the coarse ids
    engine.set_timing(True)
    engine.reset_stats()
    try:
        cx, ca, cw, cv = engine.contract(g, match, cmap, nc, CM_ROW_FIRST_SEEN)
        stats = engine.kernel_stats()
    finally:
        engine.set_timing(False)
    assert "hc_dedup_hub" in stats and "contract_gather" in stats  # the hub pass ran, then the sorted path
    ex, ea, ew, ev = orc.cpu_contract(g.xadj, g.adjncy, g.adjwgt, g.vwgt, match, cmap, nc)
    np.testing.assert_array_equal(cx.astype(np.int64), ex)
    np.testing.assert_array_equal(cv, ev)
    np.testing.assert_array_equal(ca, ea)
    np.testing.assert_array_equal(cw, ew)


def test_match_repeatable(engine, orc):
    g = with_weights(CASES["rgg3000w16"])
    say("[match repeat] rgg3000w16")
    mv = maxv_for(g, orc=orc)
    a, _ = engine.match(g, mv, seed=99)
    b, _ = engine.match(g, mv, seed=99)
    c, _ = engine.match(g, mv, seed=100)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)


def _segsort_case(rng, sizes):
    segs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    n = int(np.sum(sizes))
    keys = rng.integers(-1000, 1000, size=n).astype(np.int32)
    vals = np.arange(n, dtype=np.int32)
    return keys, vals, segs


@pytest.mark.parametrize("sizes", [
    [0, 1, 2, 3, 4, 5, 7, 8, 9, 16, 17, 31, 32, 33, 63, 64],
    [65, 100, 128, 129, 255, 256, 257, 511, 512, 513, 1000, 1024, 1025, 2047, 2048, 2049, 4095, 4096],
    [4097, 5000, 8192, 8193, 20000, 70001],
    list(np.random.default_rng(3).integers(0, 300, size=2000)),
])
def test_segsort_i32(engine, sizes):
    rng = np.random.default_rng(len(sizes))
    keys, vals, segs = _segsort_case(rng, sizes)
    say(f"[segsort] nseg={len(segs)} n={len(keys)} max={max(sizes)}")
    k, v = engine.segsort(keys, vals, segs)
    bounds = list(segs) + [len(keys)]
    for s in range(len(segs)):
        a, b = bounds[s], bounds[s + 1]
        o = np.argsort(keys[a:b], kind="stable")
        np.testing.assert_array_equal(k[a:b], keys[a:b][o])
        np.testing.assert_array_equal(v[a:b], vals[a:b][o])


def test_project(engine, orc):
    rng = np.random.default_rng(7)
    nc, n = 5000, 12000
    cmap = rng.integers(0, nc, size=n).astype(np.int32)
    wc = rng.integers(0, 64, size=nc).astype(np.int32)
    say("[project] n=12000")
    np.testing.assert_array_equal(engine.project(cmap, wc), orc.project(cmap, wc))


@pytest.mark.parametrize("order", [CM_ROW_FIRST_SEEN, CM_ROW_ASCENDING])
@pytest.mark.parametrize("name,k", [("grid100w100", 8), ("delaunay4720", 8), ("rmat12", 16), ("hub", 4),
                                    ("path_increasing", 2), ("path_increasing_off64", 2), ("rmat12_off64", 16),
                                    ("hub_off64", 4)])
def test_hierarchy(engine, orc, name, k, order):
    if name == "grid100w100":
        g = graphs.grid2d(100, 100, wmax=100, seed=9)
    elif name.startswith("rmat12"):
        g = graphs.rmat(12, 16, seed=4, offsets=np.int64 if name.endswith("off64") else np.int32)
    elif name == "hub_off64":
        from cases import hub_graph
        g = hub_graph(offsets=np.int64)
    elif name.startswith("path_increasing"):
        # weights rise along the path: each round matches only the heaviest
        # remaining edge, so level 0 needs 200 rounds, far more than one batch
        # of launches or the 32 rounds of the single-workgroup matching of
        # small levels (the ordinary launches continue from its state, and
        # the contraction is redone after the rounds resume)
        n = 400
        g = graphs.from_edges(n, np.arange(n - 1), np.arange(1, n), np.arange(1, n))
        if name.endswith("off64"):
            g = Graph(g.xadj.astype(np.int64), g.adjncy, g.adjwgt, g.vwgt)
    else:
        g = CASES[name]
    g = with_weights(g)
    say(f"[hierarchy] case={name} k={k} order={order}")
    h = engine.coarsen(g, k, seed=5, row_order=order)
    ref = orc.coarsen(g.xadj, g.adjncy, g.adjwgt, g.vwgt, k, matcher="handshake", seed=5)
    assert h.num_levels == len(ref)
    for i, R in enumerate(ref):
        L = h.level(i)
        assert L.n == R["n"]
        assert L.xadj.dtype == g.xadj.dtype
        np.testing.assert_array_equal(L.xadj.astype(np.int64), R["xadj"])
        np.testing.assert_array_equal(L.vwgt, R["vwgt"])
        ea, ew = R["adjncy"], R["adjwgt"]
        if order == CM_ROW_ASCENDING and i > 0:
            ea, ew = sort_rows(R["xadj"], ea, ew)
        np.testing.assert_array_equal(L.adjncy, ea)
        np.testing.assert_array_equal(L.adjwgt, ew)
        if i < len(ref) - 1:
            np.testing.assert_array_equal(L.match, R["match"])
            np.testing.assert_array_equal(L.cmap, R["cmap"])
            _, rounds = orc.handshake_match(R["xadj"], R["adjncy"], R["adjwgt"], R["vwgt"],
                                            maxv_for(g, k, orc), seed=5)
            assert L.rounds == rounds, i
    h.release()


def test_hierarchy_int64_offsets(engine, orc):
    g = with_weights(graphs.grid2d(60, 50, wmax=20, seed=1))
    g64 = Graph(g.xadj.astype(np.int64), g.adjncy, g.adjwgt, g.vwgt)
    say("[hierarchy int64] grid60x50")
    a = engine.coarsen(g, 8, seed=3)
    b = engine.coarsen(g64, 8, seed=3)
    assert a.num_levels == b.num_levels
    for i in range(a.num_levels):
        la, lb = a.level(i), b.level(i)
        assert lb.xadj.dtype == np.int64
        np.testing.assert_array_equal(la.xadj.astype(np.int64), lb.xadj)
        np.testing.assert_array_equal(la.adjncy, lb.adjncy)
        np.testing.assert_array_equal(la.adjwgt, lb.adjwgt)


def test_hierarchy_device_input(engine, orc):
    g = with_weights(graphs.grid2d(40, 40, wmax=9, seed=2))
    say("[hierarchy device view] grid40x40")
    a = engine.coarsen(g, 8, seed=3)
    b = engine.coarsen(g, 8, seed=3, on_device=True)
    assert a.num_levels == b.num_levels
    for i in range(a.num_levels):
        np.testing.assert_array_equal(a.level(i).adjncy, b.level(i).adjncy)
