#!/usr/bin/env python3
"""Coarsening benchmark on one MI355X.

A step is one full multilevel coarsening (match -> number -> contract per
level, until the reference stop rule ends it) of one synthetic graph that is
already resident in HBM.  Default workload: 725x725 5-point grid, int32 edge
weights uniform in [1,100], unit vertex weights, k=64, int32 CSR.

Prints ONE JSON line (see the metric/config fields) with the dominant
kernel's roofline (algorithmic bytes / hipEvent-timed launch duration on the
engine's stream) and the reference CPU coarsening path timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
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


def make_graph(cfg: str):
    from cumetis_amd import graphs
    if cfg == "grid725":
        return graphs.grid2d(725, 725, wmax=100, seed=1), 64, "grid2d 725x725 5-point, adjwgt U[1,100], k=64, int32 CSR"
    if cfg == "rmat22":
        return graphs.rmat(22, 16, seed=1, offsets=np.int64), 256, "R-MAT s22 ef16 (.57/.19/.19), symmetrised, unit weights, k=256, int64 offsets"
    if cfg == "delaunay4720":
        return graphs.delaunay2d(4720, seed=1), 8, "Delaunay 4720 uniform points, unit weights, k=8"
    raise SystemExit(f"unknown config {cfg}")


def cpu_baseline(g, k, budget_s=12.0):
    """Reference CPU coarsening path (cuMetis_cpu_coarsen restated in C,
    single thread), repeated within a time budget."""
    from oracle import oracle as O
    x = g.xadj.astype(np.int64)
    w = g.adjwgt if g.adjwgt is not None else np.ones(g.m, np.int32)
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
    return {"value": edges / el / 1e6, "unit": "Medges/s", "cores": 1, "kind": "port",
            "sample": f"{reps} full CPU coarsening run(s) of the same graph (cpu_match + "
                      f"cpu_create_cgraph per level, k-way threshold), {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="grid725")
    ap.add_argument("--order", default="first_seen", choices=["first_seen", "ascending"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(every box here has a single MI355X)", file=sys.stderr)
        sys.exit(2)

    from cumetis_amd import _lib as L
    from cumetis_amd.engine import Engine

    g, k, desc = make_graph(args.config)
    order = L.CM_ROW_FIRST_SEEN if args.order == "first_seen" else L.CM_ROW_ASCENDING
    eng = Engine(0)
    view, keep = eng.device_graph(g)      # input resident in HBM before timing
    eng.sync()

    import ctypes as C
    p = L.CoarsenParams(k, 0, 0, order, 0, 0, 1, L.CM_FLAG_TRUSTED_INPUT, 0)

    def step():
        h = C.c_void_p()
        eng._check(eng.lib.cm_coarsen(eng.ctx, C.byref(view), C.byref(p), C.byref(h)))
        return h

    for _ in range(args.warmup):
        eng.lib.cm_hier_release(eng.ctx, step())
    eng.sync()
    times = []
    hs = None
    for i in range(args.steps):
        t = time.perf_counter()
        h = step()
        eng.sync()
        times.append(time.perf_counter() - t)
        if i == args.steps - 1:
            hs = h
        else:
            eng.lib.cm_hier_release(eng.ctx, h)
    eng.sync()
    total = float(np.sum(times))
    med = float(np.median(times))

    # hierarchy shape -> edges and algorithmic bytes
    nlev = eng.lib.cm_hier_num_levels(hs)
    infos = []
    for i in range(nlev):
        li = L.LevelInfo()
        eng._check(eng.lib.cm_hier_level(hs, i, C.byref(li)))
        infos.append((li.n, li.m, li.rounds))
    w = 8 if g.xadj.dtype == np.int64 else 4
    edges = sum(m // 2 for (n, m, r) in infos[:-1])
    algo = sum(level_bytes(infos[i][0], infos[i][1], infos[i + 1][0], infos[i + 1][1], w)
               for i in range(nlev - 1))
    eng.lib.cm_hier_release(eng.ctx, hs)

    # per-kernel hipEvent timing pass (separate from the timed steps)
    roof = None
    kstats = {}
    if not args.no_profile:
        eng.set_timing(True)
        eng.reset_stats()
        for _ in range(3):
            eng.lib.cm_hier_release(eng.ctx, step())
        eng.sync()
        kstats = eng.kernel_stats()
        eng.set_timing(False)
        priced = {n: s for n, s in kstats.items() if s["bytes"] > 0}
        if priced:
            name, s = max(priced.items(), key=lambda kv: kv[1]["ms"])
            avg_s = s["ms"] / s["count"] / 1e3
            bpl = s["bytes"] / s["count"]
            ach = bpl / avg_s / 1e9
            roof = {"bound": "hbm", "kernel": name, "achieved": round(ach, 2), "peak": 8000.0,
                    "unit": "GB/s", "frac": round(ach / 8000.0, 4), "traffic": None,
                    "avg_launch_us": round(avg_s * 1e6, 2), "algo_bytes_per_launch": int(bpl)}

    cpu = None if args.no_cpu else cpu_baseline(g, k)
    value = edges / med / 1e6
    out = {
        "metric": "coarsening Medges/s & %HBM peak; batched vs one-by-one graphs/s at equal cut",
        "value": round(value, 2), "unit": "Medges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(med * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": desc, "levels": nlev, "row_order": args.order,
                   "n": int(g.n), "m": int(g.m), "coarsest_n": int(infos[-1][0]),
                   "rounds": [r for (_, _, r) in infos[:-1]]},
        "hbm_frac_whole_coarsening": round(algo / med / HBM_PEAK, 5),
        "algo_bytes": int(algo), "input_edges_per_s_M": round((g.m // 2) / med / 1e6, 2),
        "p10_ms": round(float(np.percentile(times, 10)) * 1e3, 3),
        "p90_ms": round(float(np.percentile(times, 90)) * 1e3, 3),
        "roofline": roof, "cpu_baseline": cpu,
        "kernels_ms_per_step": {n: round(s["ms"] / 3, 4) for n, s in
                                sorted(kstats.items(), key=lambda kv: -kv[1]["ms"])[:12]},
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
