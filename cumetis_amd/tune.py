"""Autotuner of the engine's launch-shape parameters on this GPU.

``python -m cumetis_amd.tune [--out DIR] [--record DIR]`` times the coarsening of four
workloads under candidate values of each cm_tuning parameter, one parameter
at a time (coordinate descent from the current defaults, candidates
interleaved and repeated so that drift on the box spreads over all of them),
and keeps a candidate only when it lowers the summed normalised time by more
than the noise margin.  It writes

  OUT/cm_tuned.hpp         the defaults every new context takes (the build
                           compiles cumetis_amd/csrc/cm_tuned.hpp, the default
                           OUT);
  RECORD/tuned_gfx950.json every measurement behind the choice (default
                           cumetis_amd/).

Workloads (the benchmark's shapes): the 725x725 grid (k=64) in first-seen and
in ascending row order (the latter runs the segmented sort), R-MAT scale 22
(k=256, int64 offsets, no weights) and 128 graphs of the batch distribution
coarsened as one batch.  None of the parameters changes a result
(tests/test_gpu_tuning.py); the tuner checks that the hierarchy shapes of
every candidate equal those of the defaults.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
sys.path.insert(0, str(PKG.parent))

CANDIDATES = {
    "graph_rounds": [0, 1],
    "first_round_batch": [7, 8, 9, 10, 11],
    "level_batch_margin": [1, 2, 3],
    "large_level_batch": [11, 12, 13, 14, 16],
    "compact_round": [4, 5, 6],
    "compact_min_n": [1 << 18, 1 << 19, 1 << 20, 1 << 21],
    "speculate_edges": [1 << 22, 1 << 23, 1 << 24, 1 << 25],
    "number_one_tile_max": [0, 1024, 2048, 4096],
    "seg_wave_bins": [4, 5, 6],
}
# which workloads each parameter can affect
AFFECTS = {
    "graph_rounds": ["grid725", "rmat22", "batch128"],
    "first_round_batch": ["grid725", "batch128"],
    "level_batch_margin": ["grid725", "batch128"],
    "large_level_batch": ["rmat22"],
    "compact_round": ["rmat22"],
    "compact_min_n": ["rmat22", "grid725"],
    "speculate_edges": ["rmat22"],
    "number_one_tile_max": ["grid725", "batch128"],
    "seg_wave_bins": ["grid725_asc"],
}
NOISE = 0.01  # a candidate must win by more than 1%


def progress(msg):
    print(f"[tune] {msg}", file=sys.stderr, flush=True)


class Workloads:
    def __init__(self, eng):
        from cumetis_amd import _lib as L
        from cumetis_amd import graphs
        from cumetis_amd.engine import concat_graphs
        self.eng, self.L = eng, L
        progress("generating workloads")
        grid = graphs.grid2d(725, 725, wmax=100, seed=1)
        rmat = graphs.rmat_fast(22, 16, seed=1, offsets=np.int64)
        rng = np.random.default_rng(2024)
        gs, ks = [], []
        for i in range(128):
            if rng.random() < 0.6:
                n = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
                gs.append(graphs.rgg2d_fast(n, 8.0, wmax=16, seed=2024 + i))
            else:
                s = int(rng.integers(32, 401))
                gs.append(graphs.grid2d(s, s, wmax=16, seed=2024 + i))
            ks.append(int(2 * rng.integers(1, 33)))
        cg, voff, eoff = concat_graphs(gs)
        self.keep = []
        self.views = {}
        for name, g in (("grid", grid), ("rmat", rmat), ("batch", cg)):
            v, k = eng.device_graph(g)
            self.views[name] = v
            self.keep.append(k)
        self.voff = np.ascontiguousarray(voff, np.int64)
        self.eoff = np.ascontiguousarray(eoff, np.int64)
        self.nb = len(gs)
        self.P = (L.CoarsenParams * self.nb)()
        for b in range(self.nb):
            self.P[b] = L.CoarsenParams(ks[b], 0, 0, L.CM_ROW_FIRST_SEEN, 0, 0, 1000 + b,
                                        L.CM_FLAG_TRUSTED_INPUT, 0)
        eng.sync()

    def _coarsen(self, view, k, order):
        L, eng = self.L, self.eng
        p = L.CoarsenParams(k, 0, 0, order, 0, 0, 1, L.CM_FLAG_TRUSTED_INPUT, 0)
        h = C.c_void_p()
        eng._check(eng.lib.cm_coarsen(eng.ctx, C.byref(view), C.byref(p), C.byref(h)))
        return [h]

    def _batch(self):
        eng = self.eng
        H = (C.c_void_p * self.nb)()
        eng._check(eng.lib.cm_coarsen_batch(eng.ctx, self.nb, C.byref(self.views["batch"]),
                                            self.voff.ctypes.data_as(C.POINTER(C.c_int64)),
                                            self.eoff.ctypes.data_as(C.POINTER(C.c_int64)), self.P, H))
        return [H[b] for b in range(self.nb)]

    def run(self, name):
        L = self.L
        if name == "grid725":
            return self._coarsen(self.views["grid"], 64, L.CM_ROW_FIRST_SEEN)
        if name == "grid725_asc":
            return self._coarsen(self.views["grid"], 64, L.CM_ROW_ASCENDING)
        if name == "rmat22":
            return self._coarsen(self.views["rmat"], 256, L.CM_ROW_FIRST_SEEN)
        return self._batch()

    def shape(self, hs):
        from cumetis_amd import _lib as L
        out = []
        for h in hs:
            for i in range(self.eng.lib.cm_hier_num_levels(h)):
                li = L.LevelInfo()
                self.eng._check(self.eng.lib.cm_hier_level(h, i, C.byref(li)))
                out.append((li.n, li.m, li.rounds))
        return out

    def time(self, name, steps):
        eng = self.eng
        for h in self.run(name):  # warm-up (and, with graph_rounds, the capture)
            eng.lib.cm_hier_release(eng.ctx, h)
        eng.sync()
        ts, shape = [], None
        for _ in range(steps):
            t = time.perf_counter()
            hs = self.run(name)
            eng.sync()
            ts.append(time.perf_counter() - t)
            if shape is None:
                shape = self.shape(hs)
            for h in hs:
                eng.lib.cm_hier_release(eng.ctx, h)
        return float(np.median(ts)), shape


STEPS = {"grid725": 15, "grid725_asc": 9, "rmat22": 7, "batch128": 3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(PKG / "csrc"), help="directory for cm_tuned.hpp")
    ap.add_argument("--record", default=None, help="directory for tuned_gfx950.json (default: the package)")
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args()
    from cumetis_amd.engine import Engine
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    eng = Engine(0)
    W = Workloads(eng)
    cur = eng.tuning()
    base_shapes = {}
    record = {"start": dict(cur), "params": {}}
    for param, cands in CANDIDATES.items():
        wl = AFFECTS[param]
        progress(f"{param}: {cands} on {wl}")
        times = {c: {w: [] for w in wl} for c in cands}
        for rep in range(args.repeats):
            for c in cands:
                eng.set_tuning(**{**cur, param: c})
                for w in wl:
                    t, shape = W.time(w, STEPS[w])
                    base_shapes.setdefault(w, shape)
                    if shape != base_shapes[w]:
                        raise RuntimeError(f"{param}={c} changed the hierarchy of {w}")
                    times[c][w].append(t)
        med = {c: {w: float(np.median(v)) for w, v in tw.items()} for c, tw in times.items()}
        ref = cur[param] if cur[param] in cands else cands[0]
        cost = {c: sum(med[c][w] / med[ref][w] for w in wl) / len(wl) for c in cands}
        best = min(cands, key=lambda c: cost[c])
        keep = best if cost[best] < cost[ref] * (1 - NOISE) else ref
        record["params"][param] = {"candidates": cands, "workloads": wl,
                                   "median_ms": {str(c): {w: round(v * 1e3, 4) for w, v in med[c].items()}
                                                 for c in cands},
                                   "cost": {str(c): round(v, 4) for c, v in cost.items()}, "chosen": keep}
        progress(f"{param}: cost {json.dumps(record['params'][param]['cost'])} -> {keep}")
        cur[param] = keep
        eng.set_tuning(**cur)
    record["chosen"] = dict(cur)
    hdr = ["// cm_tuned.hpp -- launch-shape defaults of a new context on gfx950, written",
           "// by cumetis_amd/tune.py from its measurements (see that script for the",
           "// workloads and candidates; the record is cumetis_amd/tuned_gfx950.json).",
           "// Every value changes launch shapes only, never a result.",
           "#pragma once", "#include <stdint.h>", "", "namespace cm {"]
    names = {"first_round_batch": "int32_t", "level_batch_margin": "int32_t", "large_level_batch": "int32_t",
             "compact_round": "int32_t", "compact_min_n": "int64_t", "speculate_edges": "int64_t",
             "number_one_tile_max": "int32_t", "seg_wave_bins": "int32_t", "graph_rounds": "int32_t"}
    for k, t in names.items():
        v = int(cur[k])
        val = f"int64_t({v})" if t == "int64_t" else str(v)
        hdr.append(f"constexpr {t} TUNED_{k.upper()} = {val};")
    hdr += ["}  // namespace cm", ""]
    (out / "cm_tuned.hpp").write_text("\n".join(hdr))
    rec = Path(args.record) if args.record else PKG
    rec.mkdir(parents=True, exist_ok=True)
    (rec / "tuned_gfx950.json").write_text(json.dumps(record, indent=1) + "\n")
    progress(f"wrote {out / 'cm_tuned.hpp'}")
    print(json.dumps(record["chosen"]), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
