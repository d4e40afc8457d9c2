"""Host-side mirror of the engine's operator interface.

Mirrors the reference's coarsening operators (cuMetis_gpu_match cu:1699,
cuMetis_gpu_create_cgraph cu:1291, bb_segsort include/bb_segsort.h:35,
cuMetis_kway_project cu:3079, cuMetis_coarsen cu:1824, cuMetis_PartGraph
cu:3622) on top of the C ABI.  Arrays cross as numpy; device buffers are
``DeviceArray`` objects from the engine's own pool.  Every operator runs on
the GPU through ``libcumetis_amd.so``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


class DeviceArray:
    """Device buffer owned by an Engine's pool."""

    def __init__(self, eng: "Engine", nbytes: int, dtype, count: int):
        self.eng = eng
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(int(nbytes), 1)
        p = C.c_void_p()
        eng._check(eng.lib.cm_dmalloc(eng.ctx, self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, eng: "Engine", a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(eng, a.nbytes, a.dtype, a.size)
        if a.nbytes:
            eng._check(eng.lib.cm_memcpy_h2d(eng.ctx, d.ptr, a.ctypes.data, a.nbytes))
        return d

    @classmethod
    def empty(cls, eng: "Engine", count: int, dtype) -> "DeviceArray":
        dt = np.dtype(dtype)
        return cls(eng, dt.itemsize * int(count), dt, count)

    def to_host(self, count: int | None = None) -> np.ndarray:
        n = self.count if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        if n:
            self.eng._check(self.eng.lib.cm_memcpy_d2h(self.eng.ctx, out.ctypes.data, self.ptr,
                                                       out.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.eng.ctx:
            self.eng.lib.cm_dfree(self.eng.ctx, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class Graph:
    """Host CSR graph (METIS conventions, 0-based)."""
    xadj: np.ndarray
    adjncy: np.ndarray
    adjwgt: np.ndarray | None = None
    vwgt: np.ndarray | None = None

    @property
    def n(self) -> int:
        return len(self.xadj) - 1

    @property
    def m(self) -> int:
        return int(self.xadj[-1])


@dataclass
class Level:
    n: int
    m: int
    rounds: int
    xadj: np.ndarray
    adjncy: np.ndarray
    adjwgt: np.ndarray
    vwgt: np.ndarray
    match: np.ndarray | None
    cmap: np.ndarray | None


class Hierarchy:
    """Device-resident level hierarchy returned by ``Engine.coarsen``."""

    def __init__(self, eng: "Engine", h: C.c_void_p):
        self.eng = eng
        self.h = h

    @property
    def num_levels(self) -> int:
        return self.eng.lib.cm_hier_num_levels(self.h)

    def info(self, i: int) -> L.LevelInfo:
        li = L.LevelInfo()
        self.eng._check(self.eng.lib.cm_hier_level(self.h, i, C.byref(li)))
        return li

    def level(self, i: int) -> Level:
        li = self.info(i)
        n, m = li.n, li.m
        odt = np.int64 if li.offset_type == L.CM_OFF64 else np.int32
        xadj = np.empty(n + 1, odt)
        adjncy = np.empty(m, np.int32)
        adjwgt = np.empty(m, np.int32)
        vwgt = np.empty(n, np.int32)
        last = i == self.num_levels - 1
        match = None if last else np.empty(n, np.int32)
        cmap = None if last else np.empty(n, np.int32)
        ptr = lambda a: a.ctypes.data if a is not None else None
        self.eng._check(self.eng.lib.cm_hier_copy_level(
            self.eng.ctx, self.h, i, ptr(xadj), ptr(adjncy), ptr(adjwgt), ptr(vwgt), ptr(match),
            ptr(cmap)))
        return Level(n, m, li.rounds, xadj, adjncy, adjwgt, vwgt, match, cmap)

    def project(self, level: int, where_c: DeviceArray, where_f: DeviceArray):
        self.eng._check(self.eng.lib.cm_project(self.eng.ctx, self.h, level, where_c.ptr,
                                                where_f.ptr))

    def release(self):
        if self.h:
            self.eng.lib.cm_hier_release(self.eng.ctx, self.h)
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Engine:
    """One device context of the coarsening engine."""

    def __init__(self, device: int = 0, lib_path=None):
        self.lib = L.load(lib_path)
        self.ctx = C.c_void_p()
        rc = self.lib.cm_ctx_create(device, None, C.byref(self.ctx))
        if rc != L.CM_OK:
            raise L.EngineError(rc, f"cm_ctx_create(device={device}) failed")

    # ------------------------------------------------------------ plumbing
    def _check(self, rc: int):
        if rc != L.CM_OK:
            msg = self.lib.cm_ctx_last_error(self.ctx)
            raise L.EngineError(rc, msg.decode() if msg else "")

    def close(self):
        if self.ctx:
            self.lib.cm_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        self._check(self.lib.cm_ctx_sync(self.ctx))

    def set_timing(self, on):
        """Kernel timing: False/0 off, True/1 events around every launch, 2
        events around every run of consecutive launches of one kernel."""
        self._check(self.lib.cm_ctx_set_timing(self.ctx, int(on)))

    def reset_stats(self):
        self._check(self.lib.cm_ctx_reset_stats(self.ctx))

    def kernel_stats(self) -> dict:
        out = {}
        i = 0
        buf = C.create_string_buffer(128)
        cnt, ms, by = C.c_int64(), C.c_double(), C.c_double()
        while self.lib.cm_ctx_kernel_stat(self.ctx, i, buf, 128, C.byref(cnt), C.byref(ms),
                                          C.byref(by)) == L.CM_OK:
            out[buf.value.decode()] = {"count": cnt.value, "ms": ms.value, "bytes": by.value}
            i += 1
        return out

    def tuning(self) -> dict:
        t = L.Tuning()
        self._check(self.lib.cm_ctx_get_tuning(self.ctx, C.byref(t)))
        return {f: getattr(t, f) for f, _ in L.Tuning._fields_ if f != "reserved"}

    def set_tuning(self, **kw):
        """Change launch-shape parameters (cm_tuning); returns the old values."""
        old = self.tuning()
        t = L.Tuning()
        self._check(self.lib.cm_ctx_get_tuning(self.ctx, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        self._check(self.lib.cm_ctx_set_tuning(self.ctx, C.byref(t)))
        return old

    def pool_bytes(self):
        r, u = C.c_int64(), C.c_int64()
        self._check(self.lib.cm_ctx_pool_bytes(self.ctx, C.byref(r), C.byref(u)))
        return r.value, u.value

    def to_device(self, a: np.ndarray) -> DeviceArray:
        return DeviceArray.from_host(self, a)

    def empty(self, count: int, dtype) -> DeviceArray:
        return DeviceArray.empty(self, count, dtype)

    # ----------------------------------------------------------- views
    @staticmethod
    def _offset_type(xadj: np.ndarray) -> int:
        return L.CM_OFF64 if xadj.dtype == np.int64 else L.CM_OFF32

    def device_graph(self, g: Graph):
        """Copy a host graph to device; returns (view, keepalive)."""
        odt = np.int64 if g.xadj.dtype == np.int64 else np.int32
        xa = self.to_device(g.xadj.astype(odt))
        aj = self.to_device(g.adjncy.astype(np.int32))
        aw = self.to_device(g.adjwgt.astype(np.int32)) if g.adjwgt is not None else None
        vw = self.to_device(g.vwgt.astype(np.int32)) if g.vwgt is not None else None
        v = L.GraphView(g.n, g.m, self._offset_type(g.xadj.astype(odt)), 1, xa.ptr, aj.ptr,
                        aw.ptr if aw else None, vw.ptr if vw else None)
        return v, (xa, aj, aw, vw)

    def host_view(self, g: Graph):
        odt = np.int64 if g.xadj.dtype == np.int64 else np.int32
        keep = [np.ascontiguousarray(g.xadj, dtype=odt),
                np.ascontiguousarray(g.adjncy, dtype=np.int32),
                None if g.adjwgt is None else np.ascontiguousarray(g.adjwgt, dtype=np.int32),
                None if g.vwgt is None else np.ascontiguousarray(g.vwgt, dtype=np.int32)]
        v = L.GraphView(g.n, g.m, self._offset_type(keep[0]), 0, keep[0].ctypes.data,
                        keep[1].ctypes.data,
                        None if keep[2] is None else keep[2].ctypes.data,
                        None if keep[3] is None else keep[3].ctypes.data)
        return v, keep

    # ----------------------------------------------------------- operators
    def match(self, g: Graph, maxvwgt: int, seed: int = 1, max_rounds: int = 0):
        """Seeded handshake matching; returns (match, rounds)."""
        v, keep = self.device_graph(g)
        d = self.empty(max(g.n, 1), np.int32)
        r = C.c_int32()
        self._check(self.lib.cm_match(self.ctx, C.byref(v), int(maxvwgt), C.c_uint64(seed),
                                      int(max_rounds), d.ptr, C.byref(r)))
        return d.to_host(g.n), r.value

    def number(self, xadj: np.ndarray, match: np.ndarray):
        """Coarse numbering; returns (cmap, nc)."""
        n = len(xadj) - 1
        odt = np.int64 if xadj.dtype == np.int64 else np.int32
        dx = self.to_device(xadj.astype(odt))
        dm = self.to_device(match.astype(np.int32))
        dc = self.empty(max(n, 1), np.int32)
        nc = C.c_int64()
        self._check(self.lib.cm_number(self.ctx, n, dx.ptr, self._offset_type(xadj.astype(odt)),
                                       dm.ptr, dc.ptr, C.byref(nc)))
        return dc.to_host(n), nc.value

    def contract(self, g: Graph, match: np.ndarray, cmap: np.ndarray, nc: int,
                 row_order: int = L.CM_ROW_FIRST_SEEN):
        """Contraction; returns (cxadj, cadjncy, cadjwgt, cvwgt)."""
        v, keep = self.device_graph(g)
        odt = np.int64 if g.xadj.dtype == np.int64 else np.int32
        dm = self.to_device(match.astype(np.int32))
        dc = self.to_device(cmap.astype(np.int32))
        cx = self.empty(nc + 1, odt)
        ca = self.empty(max(g.m, 1), np.int32)
        cw = self.empty(max(g.m, 1), np.int32)
        cv = self.empty(max(nc, 1), np.int32)
        mc = C.c_int64()
        self._check(self.lib.cm_contract(self.ctx, C.byref(v), dm.ptr, dc.ptr, int(nc),
                                         int(row_order), cx.ptr, ca.ptr, cw.ptr, cv.ptr,
                                         C.byref(mc)))
        return cx.to_host(nc + 1), ca.to_host(mc.value), cw.to_host(mc.value), cv.to_host(nc)

    def segsort(self, keys: np.ndarray, vals: np.ndarray, segs: np.ndarray):
        """bb_segsort<int,int> semantics; returns sorted (keys, vals)."""
        n = len(keys)
        dk = self.to_device(keys.astype(np.int32))
        dv = self.to_device(vals.astype(np.int32))
        ds = self.to_device(segs.astype(np.int32))
        self._check(self.lib.cm_segsort_i32(self.ctx, dk.ptr, dv.ptr, n, ds.ptr, len(segs)))
        return dk.to_host(n), dv.to_host(n)

    def project(self, cmap: np.ndarray, where_c: np.ndarray) -> np.ndarray:
        n = len(cmap)
        dc = self.to_device(cmap.astype(np.int32))
        dw = self.to_device(where_c.astype(np.int32))
        df = self.empty(max(n, 1), np.int32)
        self._check(self.lib.cm_project_raw(self.ctx, n, dc.ptr, dw.ptr, df.ptr))
        return df.to_host(n)

    def coarsen(self, g: Graph, nparts: int, seed: int = 1, row_order: int = L.CM_ROW_FIRST_SEEN,
                threshold: int = 0, maxvwgt: int = 0, max_levels: int = 0, max_rounds: int = 0,
                on_device: bool = False, trusted: bool = False) -> Hierarchy:
        if on_device:
            v, keep = self.device_graph(g)
        else:
            v, keep = self.host_view(g)
        p = L.CoarsenParams(nparts, threshold, maxvwgt, row_order, max_levels, max_rounds, seed,
                            L.CM_FLAG_TRUSTED_INPUT if trusted else 0, 0)
        h = C.c_void_p()
        self._check(self.lib.cm_coarsen(self.ctx, C.byref(v), C.byref(p), C.byref(h)))
        hier = Hierarchy(self, h)
        hier._keep = keep
        return hier

    def partition(self, g: Graph, nparts: int, seed: int = 1,
                  row_order: int = L.CM_ROW_FIRST_SEEN, balance: bool = True, trials: int = 0,
                  exact_initpart: bool = False, host_init_coarsen: bool = False):
        """cm_partition; balance=False leaves out the balancing pass (the
        reference's pipeline has none); trials 0 = CM_DEFAULT_TRIALS;
        exact_initpart keeps the serial initial partition on coarsest graphs
        of 2^14 vertices or more (CM_FLAG_EXACT_INITPART); host_init_coarsen
        keeps the parallel form's large recursion steps coarsening on the host
        (CM_FLAG_HOST_INIT_COARSEN)."""
        v, keep = self.host_view(g)
        flags = ((0 if balance else L.CM_FLAG_NO_BALANCE) | (L.CM_FLAG_EXACT_INITPART if exact_initpart else 0)
                 | (L.CM_FLAG_HOST_INIT_COARSEN if host_init_coarsen else 0))
        p = L.CoarsenParams(nparts, 0, 0, row_order, 0, 0, seed, flags, trials)
        part = np.empty(max(g.n, 1), np.int32)
        rep = L.PartReport()
        self._check(self.lib.cm_partition(self.ctx, C.byref(v), C.byref(p), part.ctypes.data,
                                          C.byref(rep)))
        return part[:g.n], rep


def concat_graphs(gs):
    """Concatenate graphs for Engine.coarsen_batch: global offsets, local ids."""
    n = [g.n for g in gs]
    m = [g.m for g in gs]
    voff = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    eoff = np.concatenate([[0], np.cumsum(m)]).astype(np.int64)
    use64 = any(g.xadj.dtype == np.int64 for g in gs)
    odt = np.int64 if use64 else np.int32
    xadj = np.concatenate([np.asarray(g.xadj[:-1], np.int64) + eoff[b] for b, g in enumerate(gs)]
                          + [[eoff[-1]]]).astype(odt)
    adjncy = np.concatenate([np.asarray(g.adjncy, np.int32) for g in gs] or [np.zeros(0, np.int32)])
    adjwgt = np.concatenate([np.ones(g.m, np.int32) if g.adjwgt is None else np.asarray(g.adjwgt, np.int32)
                             for g in gs] or [np.zeros(0, np.int32)])
    vwgt = np.concatenate([np.ones(g.n, np.int32) if g.vwgt is None else np.asarray(g.vwgt, np.int32)
                           for g in gs] or [np.zeros(0, np.int32)])
    return Graph(xadj, adjncy, adjwgt, vwgt), voff, eoff


def _coarsen_batch(self, gs, nparts, seed=1, row_order=L.CM_ROW_FIRST_SEEN, max_rounds=0,
                   on_device=False, concat=None):
    """Batched coarsening; returns one Hierarchy per graph."""
    cg, voff, eoff = concat if concat is not None else concat_graphs(gs)
    nb = len(voff) - 1
    ks = list(nparts) if hasattr(nparts, "__len__") else [nparts] * nb
    seeds = list(seed) if hasattr(seed, "__len__") else [seed] * nb
    if on_device:
        v, keep = self.device_graph(cg)
    else:
        v, keep = self.host_view(cg)
    P = (L.CoarsenParams * max(nb, 1))()
    for b in range(nb):
        P[b] = L.CoarsenParams(int(ks[b]), 0, 0, row_order, 0, max_rounds, int(seeds[b]), 0, 0)
    H = (C.c_void_p * max(nb, 1))()
    vo = np.ascontiguousarray(voff, np.int64)
    eo = np.ascontiguousarray(eoff, np.int64)
    self._check(self.lib.cm_coarsen_batch(self.ctx, nb, C.byref(v), vo.ctypes.data_as(C.POINTER(C.c_int64)),
                                          eo.ctypes.data_as(C.POINTER(C.c_int64)), P, H))
    out = []
    for b in range(nb):
        h = Hierarchy(self, C.c_void_p(H[b]))
        h._keep = keep
        out.append(h)
    return out


Engine.coarsen_batch = _coarsen_batch


class Server:
    """Continuous-batching partition server (cm_server_* of the C ABI)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_slots: int = 64, lib_path=None):
        self.lib = L.load(lib_path)
        self.s = C.c_void_p()
        rc = self.lib.cm_server_create(device, int(hbm_budget), int(max_slots), C.byref(self.s))
        if rc != L.CM_OK:
            raise L.EngineError(rc, "cm_server_create failed")
        self._keep = {}
        self._n = {}

    def submit(self, g: Graph, nparts: int, seed: int = 1, balance: bool = True, trials: int = 0) -> int:
        odt = np.int64 if g.xadj.dtype == np.int64 else np.int32
        keep = [np.ascontiguousarray(g.xadj, odt), np.ascontiguousarray(g.adjncy, np.int32),
                None if g.adjwgt is None else np.ascontiguousarray(g.adjwgt, np.int32),
                None if g.vwgt is None else np.ascontiguousarray(g.vwgt, np.int32)]
        v = L.GraphView(g.n, g.m, L.CM_OFF64 if odt == np.int64 else L.CM_OFF32, 0, keep[0].ctypes.data,
                        keep[1].ctypes.data, None if keep[2] is None else keep[2].ctypes.data,
                        None if keep[3] is None else keep[3].ctypes.data)
        p = L.CoarsenParams(int(nparts), 0, 0, L.CM_ROW_FIRST_SEEN, 0, 0, int(seed),
                            0 if balance else L.CM_FLAG_NO_BALANCE, int(trials))
        t = C.c_uint64()
        rc = self.lib.cm_server_submit(self.s, C.byref(v), C.byref(p), C.byref(t))
        if rc != L.CM_OK:
            raise L.EngineError(rc, "cm_server_submit failed")
        self._n[t.value] = g.n
        return t.value

    def wait(self, ticket: int):
        n = self._n.pop(ticket)
        part = np.empty(max(n, 1), np.int32)
        rep = L.PartReport()
        rc = self.lib.cm_server_wait(self.s, C.c_uint64(ticket), part.ctypes.data, C.byref(rep))
        if rc != L.CM_OK:
            raise L.EngineError(rc, "job failed")
        return part[:n], rep

    def stats(self) -> dict:
        c, p50, p99, gps, ma = C.c_int64(), C.c_double(), C.c_double(), C.c_double(), C.c_int32()
        self.lib.cm_server_stats(self.s, C.byref(c), C.byref(p50), C.byref(p99), C.byref(gps), C.byref(ma))
        return {"completed": c.value, "p50_ms": p50.value, "p99_ms": p99.value,
                "graphs_per_s": gps.value, "max_active": ma.value}

    def close(self):
        if self.s:
            self.lib.cm_server_destroy(self.s)
            self.s = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
