"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def _build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            _build()
        _lib = C.CDLL(str(LIB))
        P = C.c_void_p
        sig = {
            "orc_srand": (None, [C.c_uint]),
            "orc_rand": (C.c_int, []),
            "orc_randarrayofp": (None, [C.c_int, P, C.c_int, C.c_int]),
            "orc_log2": (C.c_int, [C.c_int]),
            "orc_threshold": (C.c_int, [C.c_int, C.c_int]),
            "orc_maxvwgt": (C.c_int, [C.c_longlong, C.c_int]),
            "orc_cpu_match": (C.c_int, [C.c_int, P, P, P, P, C.c_int, P, P, P]),
            "orc_number": (C.c_int, [C.c_int, P, P]),
            "orc_cpu_contract": (C.c_int64, [C.c_int, P, P, P, P, P, P, C.c_int, P, P, P, P]),
            "orc_coarsen_continue": (C.c_int, [C.c_int, C.c_int, C.c_int64, C.c_int]),
            "orc_edge_hash": (C.c_uint32, [C.c_uint64, C.c_uint32, C.c_uint32]),
            "orc_handshake_match": (C.c_int, [C.c_int, P, P, P, P, C.c_int, C.c_uint64, C.c_int, P]),
            "orc_project": (None, [C.c_int, P, P, P]),
            "orc_edgecut": (C.c_int64, [C.c_int, P, P, P, P]),
            "orc_pwgts": (None, [C.c_int, P, P, C.c_int, P]),
            "orc_segsort_u64": (None, [C.c_int64, P, P, P]),
            "orc_initpart": (C.c_int, [C.c_int, P, P, P, P, C.c_int, P]),
            "orc_kway_refine": (None, [C.c_int, P, P, P, P, C.c_int, P]),
            "orc_kway_balance": (C.c_int, [C.c_int, P, P, P, P, C.c_int, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(_lib, k)
            f.restype = r
            f.argtypes = a
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


def _csr(xadj, adjncy, adjwgt, vwgt):
    n = len(xadj) - 1
    x = np.ascontiguousarray(xadj, np.int64)
    a = np.ascontiguousarray(adjncy, np.int32)
    w = np.ones(len(a), np.int32) if adjwgt is None else np.ascontiguousarray(adjwgt, np.int32)
    v = np.ones(n, np.int32) if vwgt is None else np.ascontiguousarray(vwgt, np.int32)
    return n, x, a, w, v


def srand(seed: int = 1):
    lib().orc_srand(seed)


def rand() -> int:
    return lib().orc_rand()


def randarrayofp(n: int, m: int, flag: int = 1, p=None):
    out = np.zeros(max(n, 1), np.int32) if p is None else np.ascontiguousarray(p, np.int32).copy()
    lib().orc_randarrayofp(n, _p(out), m, flag)
    return out[:n]


def threshold(n: int, k: int) -> int:
    return lib().orc_threshold(n, k)


def maxvwgt(tvwgt: int, T: int) -> int:
    return lib().orc_maxvwgt(tvwgt, T)


def cpu_match(xadj, adjncy, adjwgt, vwgt, maxv):
    n, x, a, w, v = _csr(xadj, adjncy, adjwgt, vwgt)
    match = np.empty(max(n, 1), np.int32)
    cmap = np.empty(max(n, 1), np.int32)
    perm = np.empty(max(n, 1), np.int32)
    nc = lib().orc_cpu_match(n, _p(x), _p(a), _p(w), _p(v), maxv, _p(match), _p(cmap), _p(perm))
    return match[:n], cmap[:n], nc, perm[:n]


def number(match):
    m = np.ascontiguousarray(match, np.int32)
    cmap = np.empty(max(len(m), 1), np.int32)
    nc = lib().orc_number(len(m), _p(m), _p(cmap))
    return cmap[:len(m)], nc


def cpu_contract(xadj, adjncy, adjwgt, vwgt, match, cmap, nc):
    n, x, a, w, v = _csr(xadj, adjncy, adjwgt, vwgt)
    mt = np.ascontiguousarray(match, np.int32)
    cm = np.ascontiguousarray(cmap, np.int32)
    cx = np.empty(nc + 1, np.int64)
    ca = np.empty(max(len(a), 1), np.int32)
    cw = np.empty(max(len(a), 1), np.int32)
    cv = np.empty(max(nc, 1), np.int32)
    mc = lib().orc_cpu_contract(n, _p(x), _p(a), _p(w), _p(v), _p(mt), _p(cm), nc, _p(cx), _p(ca),
                                _p(cw), _p(cv))
    return cx, ca[:mc], cw[:mc], cv[:nc]


def coarsen_continue(nc, nf, mc, T) -> bool:
    return bool(lib().orc_coarsen_continue(nc, nf, mc, T))


def edge_hash(seed, a, b) -> int:
    return lib().orc_edge_hash(seed, a, b)


def handshake_match(xadj, adjncy, adjwgt, vwgt, maxv, seed=1, max_rounds=4096):
    n, x, a, w, v = _csr(xadj, adjncy, adjwgt, vwgt)
    match = np.empty(max(n, 1), np.int32)
    r = lib().orc_handshake_match(n, _p(x), _p(a), _p(w), _p(v), maxv, seed, max_rounds, _p(match))
    return match[:n], r


def project(cmap, where_c):
    cm = np.ascontiguousarray(cmap, np.int32)
    wc = np.ascontiguousarray(where_c, np.int32)
    wf = np.empty(max(len(cm), 1), np.int32)
    lib().orc_project(len(cm), _p(cm), _p(wc), _p(wf))
    return wf[:len(cm)]


def edgecut(xadj, adjncy, adjwgt, where) -> int:
    n, x, a, w, _ = _csr(xadj, adjncy, adjwgt, None)
    wh = np.ascontiguousarray(where, np.int32)
    return lib().orc_edgecut(n, _p(x), _p(a), _p(w), _p(wh))


def pwgts(vwgt, where, nparts, n=None):
    wh = np.ascontiguousarray(where, np.int32)
    n = len(wh) if n is None else n
    v = np.ones(n, np.int32) if vwgt is None else np.ascontiguousarray(vwgt, np.int32)
    out = np.empty(nparts, np.int64)
    lib().orc_pwgts(n, _p(v), _p(wh), nparts, _p(out))
    return out


def segsort_u64(segoff, keys, vals):
    so = np.ascontiguousarray(segoff, np.int64)
    k = np.ascontiguousarray(keys, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().orc_segsort_u64(len(so) - 1, _p(so), _p(k), _p(v))
    return k, v


def initpart(xadj, adjncy, adjwgt, vwgt, nparts):
    n, x, a, w, v = _csr(xadj, adjncy, adjwgt, vwgt)
    where = np.zeros(max(n, 1), np.int32)
    obj = lib().orc_initpart(n, _p(x), _p(a), _p(w), _p(v), nparts, _p(where))
    return where[:n], obj


def kway_refine(xadj, adjncy, adjwgt, vwgt, nparts, where):
    n, x, a, w, v = _csr(xadj, adjncy, adjwgt, vwgt)
    wh = np.ascontiguousarray(where, np.int32).copy()
    lib().orc_kway_refine(n, _p(x), _p(a), _p(w), _p(v), nparts, _p(wh))
    return wh


def kway_balance(xadj, adjncy, adjwgt, vwgt, nparts, where):
    """The engine's balancing pass (no reference counterpart): moves vertices
    out of parts above int(w/k * 1.03).  Returns (where, moves)."""
    n, x, a, w, v = _csr(xadj, adjncy, adjwgt, vwgt)
    wh = np.ascontiguousarray(where, np.int32).copy()
    moved = lib().orc_kway_balance(n, _p(x), _p(a), _p(w), _p(v), nparts, _p(wh))
    return wh, moved


def coarsen(xadj, adjncy, adjwgt, vwgt, nparts, matcher="cpu", seed=1, T=None, maxv=None):
    """Reference level loop (cu:1824-1850 / 1854-1875) with the k-way threshold
    (cu:3634); matcher 'cpu' = cuMetis_cpu_match, 'handshake' = the engine's
    matching model.  Returns a list of levels (dicts)."""
    n = len(xadj) - 1
    v0 = np.ones(n, np.int32) if vwgt is None else np.asarray(vwgt, np.int32)
    w0 = np.ones(int(xadj[-1]), np.int32) if adjwgt is None else np.asarray(adjwgt, np.int32)
    T = threshold(n, nparts) if T is None else T
    maxv = maxvwgt(int(v0.sum()), T) if maxv is None else maxv
    levels = [dict(n=n, xadj=np.asarray(xadj, np.int64), adjncy=np.asarray(adjncy, np.int32),
                   adjwgt=w0, vwgt=v0)]
    while True:
        L = levels[-1]
        if matcher == "cpu":
            match, cmap, nc, _ = cpu_match(L["xadj"], L["adjncy"], L["adjwgt"], L["vwgt"], maxv)
        else:
            match, _ = handshake_match(L["xadj"], L["adjncy"], L["adjwgt"], L["vwgt"], maxv, seed)
            cmap, nc = number(match)
        cx, ca, cw, cv = cpu_contract(L["xadj"], L["adjncy"], L["adjwgt"], L["vwgt"], match, cmap, nc)
        L["match"], L["cmap"] = match, cmap
        levels.append(dict(n=nc, xadj=cx, adjncy=ca, adjwgt=cw, vwgt=cv))
        if not coarsen_continue(nc, L["n"], int(cx[-1]), T):
            break
    return levels


def partition(xadj, adjncy, adjwgt, vwgt, nparts, matcher="handshake", seed=1, balance=False, trials=1):
    """The multilevel pipeline (cu:3544-3616) on the CPU: coarsen (matcher
    'cpu' = the reference's CPU HEM, 'handshake' = the engine's matching
    model), initial partition of the coarsest level (glibc stream from seed
    1), then refine the coarsest level and, for every finer level, project and
    refine.  balance=True runs the engine's balancing pass before each
    refinement (the engine's default; the reference has none).  trials > 1
    (handshake only) repeats the pipeline with seeds seed .. seed+trials-1
    and keeps the lowest edge cut, the earlier trial on ties.
    Returns (where, levels) of the kept trial."""
    n = len(xadj) - 1
    if nparts == 1 or n == 0:
        return np.zeros(n, np.int32), []
    assert trials == 1 or matcher == "handshake"
    best = None
    for t in range(trials):
        if matcher == "cpu":
            srand(1)
        levels = coarsen(xadj, adjncy, adjwgt, vwgt, nparts, matcher=matcher, seed=seed + t)
        C = levels[-1]
        srand(1)
        where, _ = initpart(C["xadj"], C["adjncy"], C["adjwgt"], C["vwgt"], nparts)
        for L in [C] + list(reversed(levels[:-1])):
            if L is not C:
                where = project(L["cmap"], where)
            if balance:
                where, _ = kway_balance(L["xadj"], L["adjncy"], L["adjwgt"], L["vwgt"], nparts, where)
            where = kway_refine(L["xadj"], L["adjncy"], L["adjwgt"], L["vwgt"], nparts, where)
        cut = edgecut(xadj, adjncy, adjwgt, where)
        if best is None or cut < best[0]:
            best = (cut, where, levels)
    return best[1], best[2]
