#!/usr/bin/env python3
"""Coarsening benchmark on one MI355X.

A step is one full multilevel coarsening (match -> number -> contract per
level until the reference stop rule ends it) of synthetic input already
resident in HBM.  Legs:
  grid725   2D 5-point grid 725x725, adjwgt U[1,100], unit vwgt, k=64, int32
            CSR (the main line);
  rmat22    R-MAT scale 22, edge factor 16 (.57/.19/.19), symmetrised, unit
            weights, k=256, int64 offsets;
  batch512  512 graphs (60% 2D RGG with n log-uniform in [1k,200k], mean
            degree 8; 40% grids with side U{32..400}; adjwgt U[1,16];
            k U{2,4,..,64}): one batched coarsening against the same graphs
            coarsened one by one;
  ascending  grid725 and rmat22 with CM_ROW_ASCENDING (the reference GPU
            contraction's row order: the binned segmented sort);
  delaunay4720, serve (a Poisson stream of partition requests through the
            server), e2e (cm_partition of the grid, k=64) and e2e_rmat22
            (cm_partition of the R-MAT graph, k=256, one trial: the host
            initial partition of its 3.7 M-vertex coarsest graph takes most
            of it).
Prints ONE JSON line: the main leg's metric, its dominant kernel's roofline
(algorithmic bytes / hipEvent-timed launch duration on the engine's stream)
and the reference CPU coarsening path timed on this host; the other legs are
reported under "legs".
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12


def level_bytes(n, m, nc, mc, w):
    """Compulsory bytes of one level: matching w(n+1)+8m+8n, numbering 8n,
    contraction w(n+1)+8m+12n+w(nc+1)+8mc+4nc."""
    return 2 * w * (n + 1) + 16 * m + 28 * n + w * (nc + 1) + 8 * mc + 4 * nc


def hier_shape(eng, h):
    from cumetis_amd import _lib as L
    out = []
    for i in range(eng.lib.cm_hier_num_levels(h)):
        li = L.LevelInfo()
        eng._check(eng.lib.cm_hier_level(h, i, C.byref(li)))
        out.append((li.n, li.m, li.rounds))
    return out


def shape_metrics(shape, w):
    edges = sum(m // 2 for (n, m, r) in shape[:-1])
    algo = sum(level_bytes(shape[i][0], shape[i][1], shape[i + 1][0], shape[i + 1][1], w)
               for i in range(len(shape) - 1))
    return edges, algo


def timed_steps(eng, step, warmup, steps):
    for _ in range(warmup):
        eng.lib.cm_hier_release(eng.ctx, step())
    eng.sync()
    times, last = [], None
    for i in range(steps):
        t = time.perf_counter()
        h = step()
        eng.sync()
        times.append(time.perf_counter() - t)
        if last is not None:
            eng.lib.cm_hier_release(eng.ctx, last)
        last = h
    return times, last


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated on a copy
    kernel), or None when no counter run covers it."""
    p = Path(__file__).resolve().parent / "profiles" / f"{workload}_traffic.json"
    if not p.exists():
        return None
    b = json.loads(p.read_text()).get("bytes_per_launch", {}).get(kernel)
    return None if b is None else int(b)


def kernel_profile(eng, step, reps=3, workload=None, plain_s=None):
    """Per-kernel hipEvent timing on the engine stream (separate pass).
    Events go around every run of consecutive launches of one kernel (the
    rounds of a matching batch are one run), so no event sits between the
    launches of a run; a run's time includes the gaps between its launches.
    An event pair still adds a gap around every run: the step time with
    events minus the plain step time (plain_s, from the timed steps), spread
    over the step's launches, is reported beside the roofline."""
    eng.set_timing(2)
    eng.reset_stats()
    wall = []
    for _ in range(reps):
        t = time.perf_counter()
        h = step()
        eng.sync()
        wall.append(time.perf_counter() - t)
        eng.lib.cm_hier_release(eng.ctx, h)
    eng.sync()
    ks = eng.kernel_stats()
    eng.set_timing(False)
    launches = sum(s["count"] for s in ks.values()) / reps
    overhead_ms = None
    if plain_s is not None:
        overhead_ms = max(0.0, (float(np.median(wall)) - plain_s) * 1e3)
    priced = {n: s for n, s in ks.items() if s["bytes"] > 0}
    roof = None
    if priced:
        name, s = max(priced.items(), key=lambda kv: kv[1]["ms"])
        avg_s = s["ms"] / s["count"] / 1e3
        bpl = s["bytes"] / s["count"]
        ach = bpl / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": name, "achieved": round(ach, 2), "peak": 8000.0,
                "unit": "GB/s", "frac": round(ach / 8000.0, 4),
                "traffic": measured_traffic(workload, name) if workload else None,
                "avg_launch_us": round(avg_s * 1e6, 2), "algo_bytes_per_launch": int(bpl)}
        roof["launches_per_step"] = round(launches, 1)
        roof["kernel_launches_per_step"] = round(s["count"] / reps, 1)
        if overhead_ms is not None:
            roof["timing_overhead_ms_per_step"] = round(overhead_ms, 3)
    per = {n: round(s["ms"] / reps, 4) for n, s in sorted(ks.items(), key=lambda kv: -kv[1]["ms"])[:14]}
    return roof, per


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def coarsen_leg(eng, g, k, order, warmup, steps, profile=True, workload=None):
    from cumetis_amd import _lib as L
    progress(f"coarsen leg n={g.n} m={g.m} k={k}")
    view, keep = eng.device_graph(g)  # input resident in HBM before timing
    eng.sync()
    p = L.CoarsenParams(k, 0, 0, order, 0, 0, 1, L.CM_FLAG_TRUSTED_INPUT, 0)

    def step():
        h = C.c_void_p()
        eng._check(eng.lib.cm_coarsen(eng.ctx, C.byref(view), C.byref(p), C.byref(h)))
        return h

    times, h = timed_steps(eng, step, warmup, steps)
    shape = hier_shape(eng, h)
    eng.lib.cm_hier_release(eng.ctx, h)
    w = 8 if g.xadj.dtype == np.int64 else 4
    edges, algo = shape_metrics(shape, w)
    med = float(np.median(times))
    out = {"medges_per_s": round(edges / med / 1e6, 2), "ms": round(med * 1e3, 3),
           "p10_ms": round(float(np.percentile(times, 10)) * 1e3, 3),
           "p90_ms": round(float(np.percentile(times, 90)) * 1e3, 3),
           "levels": len(shape), "coarsest_n": int(shape[-1][0]),
           "rounds": [r for (_, _, r) in shape[:-1]],
           "hbm_frac_whole_coarsening": round(algo / med / HBM_PEAK, 5), "algo_bytes": int(algo),
           "input_medges_per_s": round((g.m // 2) / med / 1e6, 2), "edges": int(edges)}
    roof, per = kernel_profile(eng, step, workload=workload, plain_s=med) if profile else (None, {})
    # the same coarsening from host arrays (the cuMetis_PartGraph boundary):
    # includes the host-to-device copies of the graph; reported beside the
    # device-resident value, never as it
    hview, hkeep = eng.host_view(g)

    def hstep():
        h = C.c_void_p()
        eng._check(eng.lib.cm_coarsen(eng.ctx, C.byref(hview), C.byref(p), C.byref(h)))
        return h

    ht, hh = timed_steps(eng, hstep, 1, max(2, min(steps, 5)))
    eng.lib.cm_hier_release(eng.ctx, hh)
    hmed = float(np.median(ht))
    out["pcie_inclusive"] = {"ms": round(hmed * 1e3, 3), "medges_per_s": round(edges / hmed / 1e6, 2)}
    del keep, hkeep
    return out, roof, per


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(g, k, budget_s=12.0):
    """Reference CPU coarsening path (cuMetis_cpu_coarsen restated in C,
    built with -O3, single thread, pinned to one core of this host),
    repeated within a time budget (at least one full run)."""
    import os
    from oracle import oracle as O
    x = g.xadj.astype(np.int64)
    w = g.adjwgt if g.adjwgt is not None else np.ones(g.m, np.int32)
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        reps, edges, t0 = 0, 0, time.perf_counter()
        while True:
            O.srand(1)
            t = time.perf_counter()
            lv = O.coarsen(x, g.adjncy, w, None, k, matcher="cpu")
            dt = time.perf_counter() - t
            edges += sum(int(L["xadj"][-1]) // 2 for L in lv[:-1])
            reps += 1
            if time.perf_counter() - t0 > budget_s or dt * (reps + 1) > budget_s * 1.5:
                break
        el = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    return {"value": round(edges / el / 1e6, 2), "unit": "Medges/s", "cores": 1, "kind": "port",
            "sample": f"{reps} full CPU coarsening run(s) of the same graph (cpu_match + "
                      f"cpu_create_cgraph per level, k-way threshold), {el:.1f} s, pinned to core {core}",
            "cpu_model": cpu_model(), "host_cpus": len(old)}


def stream_copy_gbs(eng, reps=20, nbytes=1 << 30):
    """Measured device-to-device copy bandwidth on this GPU (read + write
    bytes per second of a 1 GiB copy through the engine's stream), the
    on-box ceiling for streaming kernels."""
    a, b = C.c_void_p(), C.c_void_p()
    eng._check(eng.lib.cm_dmalloc(eng.ctx, nbytes, C.byref(a)))
    eng._check(eng.lib.cm_dmalloc(eng.ctx, nbytes, C.byref(b)))
    try:
        eng._check(eng.lib.cm_memcpy_d2d(eng.ctx, b, a, nbytes))
        eng.sync()
        t = time.perf_counter()
        for _ in range(reps):
            eng._check(eng.lib.cm_memcpy_d2d(eng.ctx, b, a, nbytes))
        eng.sync()
        s = (time.perf_counter() - t) / reps
    finally:
        eng.lib.cm_dfree(eng.ctx, a)
        eng.lib.cm_dfree(eng.ctx, b)
    return round(2 * nbytes / s / 1e9, 1)


def batch_graphs(count=512, seed=2024):
    from cumetis_amd import graphs
    rng = np.random.default_rng(seed)
    gs, ks = [], []
    for i in range(count):
        if rng.random() < 0.6:
            n = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
            g = graphs.rgg2d_fast(n, 8.0, wmax=16, seed=seed + i)
        else:
            s = int(rng.integers(32, 401))
            g = graphs.grid2d(s, s, wmax=16, seed=seed + i)
        gs.append(g)
        ks.append(int(2 * rng.integers(1, 33)))
    return gs, ks


def batch_leg(eng, count, order, steps, cpu=True):
    from cumetis_amd import _lib as L
    from cumetis_amd.engine import concat_graphs
    progress(f"batch leg: generating {count} graphs")
    gs, ks = batch_graphs(count)
    cg, voff, eoff = concat_graphs(gs)
    progress(f"batch leg: sum n={cg.n} sum m={cg.m}")
    view, keep = eng.device_graph(cg)
    nb = len(gs)
    P = (L.CoarsenParams * nb)()
    for b in range(nb):
        P[b] = L.CoarsenParams(ks[b], 0, 0, order, 0, 0, 1000 + b, L.CM_FLAG_TRUSTED_INPUT, 0)
    vo, eo = np.ascontiguousarray(voff, np.int64), np.ascontiguousarray(eoff, np.int64)
    H = (C.c_void_p * nb)()

    def batched():
        eng._check(eng.lib.cm_coarsen_batch(eng.ctx, nb, C.byref(view),
                                            vo.ctypes.data_as(C.POINTER(C.c_int64)),
                                            eo.ctypes.data_as(C.POINTER(C.c_int64)), P, H))
        eng.sync()
        return [H[b] for b in range(nb)]

    def release(hs):
        for h in hs:
            eng.lib.cm_hier_release(eng.ctx, h)

    release(batched())
    # the same batch with every round batch replayed from a cached hipGraph
    old_tuning = eng.set_tuning(graph_rounds=1)
    release(batched())  # captures
    gt = []
    for _ in range(steps):
        t = time.perf_counter()
        release(batched())
        gt.append(time.perf_counter() - t)
    eng.set_tuning(**old_tuning)
    bt = []
    for _ in range(steps):
        t = time.perf_counter()
        hs = batched()
        bt.append(time.perf_counter() - t)
        if _ < steps - 1:
            release(hs)
    bshape = [hier_shape(eng, h) for h in hs]
    release(hs)
    progress("batch leg: one by one")
    # one by one: each graph resident on its own before timing
    views = [eng.device_graph(g) for g in gs]
    eng.sync()

    def one_by_one():
        out = []
        for b, (v, _) in enumerate(views):
            h = C.c_void_p()
            eng._check(eng.lib.cm_coarsen(eng.ctx, C.byref(v), C.byref(P[b]), C.byref(h)))
            out.append(h)
        eng.sync()
        return out

    release(one_by_one())
    t = time.perf_counter()
    hs = one_by_one()
    ot = time.perf_counter() - t
    oshape = [hier_shape(eng, h) for h in hs]
    release(hs)
    same_shape = all([(n, m) for (n, m, _) in a] == [(n, m) for (n, m, _) in b] for a, b in zip(bshape, oshape))
    # every level array of a sample of graphs, batched against alone
    progress("batch leg: level arrays of a sample")
    sample = list(range(0, nb, max(1, nb // 16)))[:16]
    hb = batched()
    arrays_equal = True
    for b in sample:
        h = C.c_void_p()
        eng._check(eng.lib.cm_coarsen(eng.ctx, C.byref(views[b][0]), C.byref(P[b]), C.byref(h)))
        arrays_equal &= hier_arrays_equal(eng, C.c_void_p(hb[b]), h)
        eng.lib.cm_hier_release(eng.ctx, h)
    release(hb)
    del views
    cuts = batch_cuts(gs, ks, sample[:8])
    med = float(np.median(bt))
    out = {"graphs": nb, "sum_n": int(voff[-1]), "sum_m": int(eoff[-1]),
           "batched_graphs_per_s": round(nb / med, 1), "batched_ms": round(med * 1e3, 2),
           "one_by_one_graphs_per_s": round(nb / ot, 1), "one_by_one_ms": round(ot * 1e3, 2),
           "speedup": round(ot / med, 2),
           "batched_ms_graph_replay": round(float(np.median(gt)) * 1e3, 2),
           "tuning": eng.tuning(), "level_shapes_identical": bool(same_shape),
           "level_arrays_identical_sample": {"graphs": len(sample), "identical": bool(arrays_equal)},
           "cuts": cuts}
    if cpu:
        out["cpu_all_cores"] = batch_cpu_all_cores(gs, ks)
    return out


def hier_arrays_equal(eng, ha, hb):
    """Every level's xadj/adjncy/adjwgt/vwgt/match/cmap of two hierarchies,
    local values (bases subtracted by cm_hier_copy_level)."""
    from cumetis_amd.engine import Hierarchy
    a, b = Hierarchy(eng, ha), Hierarchy(eng, hb)
    try:
        if a.num_levels != b.num_levels:
            return False
        for i in range(a.num_levels):
            x, y = a.level(i), b.level(i)
            for f in ("xadj", "adjncy", "adjwgt", "vwgt", "match", "cmap"):
                u, v = getattr(x, f), getattr(y, f)
                if (u is None) != (v is None):
                    return False
                if u is not None and not np.array_equal(u.astype(np.int64), v.astype(np.int64)):
                    return False
        return True
    finally:
        a.h = b.h = None  # released by the caller


def batch_cuts(gs, ks, sample):
    """Edge cuts of a sample of the batch's graphs partitioned alone
    (cm_partition) and through the server, whose coarsening runs every
    admitted job's level as one batched union per step."""
    from cumetis_amd.engine import Engine, Server
    alone = []
    with Engine(0) as e2:
        for b in sample:
            alone.append(int(e2.partition(gs[b], ks[b], seed=1)[1].edgecut))
    with Server(0, max_slots=64) as srv:
        tickets = [srv.submit(gs[b], ks[b], seed=1) for b in sample]
        served = [int(srv.wait(t)[1].edgecut) for t in tickets]
    return {"graphs": len(sample), "alone": alone, "batched_union": served, "equal": alone == served}


def _cpu_coarsen_edges(g, k):
    from oracle import oracle as O
    x = g.xadj.astype(np.int64)
    w = g.adjwgt if g.adjwgt is not None else np.ones(g.m, np.int32)
    lv = O.coarsen(x, g.adjncy, w, None, k, matcher="cpu")
    return sum(int(L["xadj"][-1]) // 2 for L in lv[:-1])


def batch_cpu_all_cores(gs, ks):
    """The reference CPU coarsening of the whole batch with one graph per
    host core (threads calling the C restatement, which runs without the
    GIL); its rand() stream is shared, so the matchings differ from a
    one-at-a-time run, the work does not."""
    import concurrent.futures as cf
    import os
    cores = len(os.sched_getaffinity(0))
    t = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        edges = sum(ex.map(lambda b: _cpu_coarsen_edges(gs[b], ks[b]), range(len(gs))))
    el = time.perf_counter() - t
    return {"graphs_per_s": round(len(gs) / el, 1), "medges_per_s": round(edges / el / 1e6, 2),
            "cores": cores, "seconds": round(el, 2), "kind": "port",
            "sample": f"all {len(gs)} graphs, one per core"}


def e2e_leg(eng, g, k, trials=0, reps=3):
    """The whole pipeline (cm_partition: coarsening, host initial partition,
    uncoarsening with balancing and refinement) from host arrays, with the
    phase times of the reference's report lines (cu:3901-3919)."""
    reps_out = []
    for _ in range(reps):
        part, rep = eng.partition(g, k, seed=1, trials=trials)
        reps_out.append(rep)
        progress(f"e2e k={k}: {rep.total_ms:.1f} ms (coarsest {rep.coarsest_n})")
    r = min(reps_out, key=lambda x: x.total_ms)
    vw = np.ones(g.n, np.int64) if g.vwgt is None else np.asarray(g.vwgt, np.int64)
    pw = np.bincount(part, weights=vw, minlength=k)
    return {"total_ms": round(r.total_ms, 2), "coarsen_ms": round(r.coarsen_ms, 2),
            "init_ms": round(r.init_ms, 2), "uncoarsen_ms": round(r.uncoarsen_ms, 2),
            "trials": int(r.trials), "edgecut": int(r.edgecut), "levels": int(r.levels),
            "coarsest_n": int(r.coarsest_n), "imbalance": round(float(pw.max() * k / vw.sum()), 4),
            "runs": reps}


def serve_leg(count=120, rate=20.0, seed=7, trials=0):
    """Poisson stream of partition requests: the batch distribution plus 5%
    3D 7-point grids 160^3 (k=128), at a fixed offered load."""
    from cumetis_amd import graphs
    from cumetis_amd import _lib as L
    from cumetis_amd.engine import Server
    progress(f"serve leg: {count} requests at {rate}/s, trials {trials or 'default'}")
    rng = np.random.default_rng(seed)
    small, ks = batch_graphs(max(8, count), seed=seed)
    big = graphs.grid3d(160, 160, 160)
    reqs = []
    for i in range(count):
        if rng.random() < 0.05:
            reqs.append((big, 128))
        else:
            reqs.append((small[i], ks[i]))
    gaps = rng.exponential(1.0 / rate, size=count)
    with Server(0, max_slots=64) as srv:
        t0 = time.perf_counter()
        tickets = []
        due = t0
        for (g, k), gap in zip(reqs, gaps):
            due += gap
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(srv.submit(g, k, seed=1, trials=trials))
        progress("serve leg: all submitted")
        reps = [srv.wait(t)[1] for t in tickets]
        wall = time.perf_counter() - t0
        st = srv.stats()
    return {"requests": count, "offered_per_s": rate, "trials_per_request": trials or L.CM_DEFAULT_TRIALS,
            "large_jobs": sum(1 for g, k in reqs if k == 128),
            "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
            "graphs_per_s": round(count / wall, 2), "max_concurrent_jobs": st["max_active"],
            "mean_coarsen_levels": round(float(np.mean([r.levels for r in reps])), 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--order", default="first_seen", choices=["first_seen", "ascending"])
    ap.add_argument("--legs", default="grid725,delaunay4720,rmat22,ascending,e2e_rmat22,batch512,serve,e2e",
                    help="comma list of: grid725 (always run), delaunay4720, rmat22, ascending (grid725 and "
                         "rmat22 with ascending rows: the binned segmented sort), batch512, serve, "
                         "e2e (grid 725^2 k=64 through cm_partition), e2e_rmat22 (R-MAT s22 k=256, one "
                         "trial, about 25 s: the host initial partition of a 3.7 M-vertex coarsest graph)")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(every box here has a single MI355X)", file=sys.stderr)
        sys.exit(2)

    from cumetis_amd import _lib as L
    from cumetis_amd import graphs
    from cumetis_amd.engine import Engine

    order = L.CM_ROW_FIRST_SEEN if args.order == "first_seen" else L.CM_ROW_ASCENDING
    legs = [x for x in args.legs.split(",") if x]
    eng = Engine(0)
    g = graphs.grid2d(725, 725, wmax=100, seed=1)
    main_out, roof, per = coarsen_leg(eng, g, 64, order, args.warmup, args.steps,
                                      profile=not args.no_profile,
                                      workload="grid725" if args.order == "first_seen" else None)
    extra = {}
    if "delaunay4720" in legs:
        progress("delaunay4720")
        gd = graphs.delaunay2d(4720, seed=1)
        o, rd, pd = coarsen_leg(eng, gd, 8, order, args.warmup, args.steps, profile=not args.no_profile)
        o["roofline"] = rd
        o["kernels_ms_per_step"] = pd
        o["config"] = "2D Delaunay triangulation of 4720 uniform points (seed 1), unit weights, k=8, int32 CSR"
        if not args.no_cpu:
            o["cpu_baseline"] = cpu_baseline(gd, 8, budget_s=5.0)
        extra["delaunay4720"] = o
    if "rmat22" in legs:
        progress("rmat22: generating")
        gr = graphs.rmat_fast(22, 16, seed=1, offsets=np.int64)
        o, rr, pr = coarsen_leg(eng, gr, 256, order, args.warmup, args.steps, profile=not args.no_profile,
                                workload="rmat22" if args.order == "first_seen" else None)
        o["roofline"] = rr
        o["kernels_ms_per_step"] = pr
        o["config"] = "R-MAT s22 ef16 (.57/.19/.19), symmetrised, unit weights, k=256, int64 offsets"
        # actually required bytes: with no edge weights given, the level-0
        # kernels read no adjwgt (8m of the 16m per level are unit weights)
        w0 = 4 * gr.m  # one int32 per entry, read twice (matching, contraction)
        o["algo_bytes_without_unit_adjwgt"] = int(o["algo_bytes"] - 2 * w0)
        o["hbm_frac_without_unit_adjwgt"] = round(o["algo_bytes_without_unit_adjwgt"] / (o["ms"] / 1e3) / HBM_PEAK, 5)
        if not args.no_cpu:
            progress("rmat22: cpu baseline")
            o["cpu_baseline"] = cpu_baseline(gr, 256, budget_s=10.0)
        extra["rmat22"] = o
        if "ascending" in legs and args.order == "first_seen":
            progress("rmat22: ascending rows")
            oa, _, pa = coarsen_leg(eng, gr, 256, L.CM_ROW_ASCENDING, 2, max(5, args.steps // 4), profile=True)
            extra["rmat22_ascending"] = {"ms": oa["ms"], "medges_per_s": oa["medges_per_s"],
                                         "kernels_ms_per_step": pa}
        if "e2e_rmat22" in legs:
            progress("rmat22: end to end (one trial)")
            extra["e2e_rmat22"] = e2e_leg(eng, gr, 256, trials=1, reps=1)
        del gr
    if "ascending" in legs and args.order == "first_seen":
        # the reference GPU contraction's row order (cu:1332): the binned
        # segmented sort instead of the first-seen hash passes
        progress("grid725: ascending rows")
        oa, _, pa = coarsen_leg(eng, g, 64, L.CM_ROW_ASCENDING, args.warmup, args.steps, profile=True)
        extra["grid725_ascending"] = {"ms": oa["ms"], "medges_per_s": oa["medges_per_s"],
                                      "p10_ms": oa["p10_ms"], "p90_ms": oa["p90_ms"], "kernels_ms_per_step": pa}
    if "batch512" in legs:
        extra["batch512"] = batch_leg(eng, args.batch_count, order, 3, cpu=not args.no_cpu)
    if "serve" in legs:
        extra["serve"] = serve_leg()
        extra["serve_one_trial"] = serve_leg(trials=1)
    if "e2e" in legs:
        progress("e2e grid725")
        extra["e2e_grid725"] = e2e_leg(eng, g, 64)
    progress("stream copy")
    copy_gbs = stream_copy_gbs(eng)
    if roof is not None:
        roof["measured_copy_gbs"] = copy_gbs
        roof["frac_of_copy"] = round(roof["achieved"] / copy_gbs, 4)
    progress("cpu baseline")
    cpu = None if args.no_cpu else cpu_baseline(g, 64)
    out = {
        "metric": "coarsening Medges/s & %HBM peak; batched vs one-by-one graphs/s at equal cut",
        "value": main_out["medges_per_s"], "unit": "Medges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": main_out["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "grid2d 725x725 5-point, adjwgt U[1,100], unit vwgt, k=64, int32 CSR",
                   "row_order": args.order, "n": int(g.n), "m": int(g.m),
                   "levels": main_out["levels"], "coarsest_n": main_out["coarsest_n"],
                   "rounds": main_out["rounds"]},
        "hbm_frac_whole_coarsening": main_out["hbm_frac_whole_coarsening"],
        "copy_frac_whole_coarsening": round(main_out["hbm_frac_whole_coarsening"] * 8000.0 / copy_gbs, 5),
        "algo_bytes": main_out["algo_bytes"], "input_medges_per_s": main_out["input_medges_per_s"],
        "p10_ms": main_out["p10_ms"], "p90_ms": main_out["p90_ms"],
        "pcie_inclusive": main_out["pcie_inclusive"],
        "roofline": roof, "cpu_baseline": cpu, "kernels_ms_per_step": per, "legs": extra,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
