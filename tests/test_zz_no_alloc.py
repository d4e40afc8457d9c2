"""GPU: the allocation-free segmented sort (cm_segsort_i32_noalloc) against
cm_segsort_i32, through the C ABI: two calls in a row (the second smaller,
over scratch the first left behind), every bin of the sort (register, LDS
and hub rows with merge passes), no growth of the context's pool, and the
size limit rejected before anything runs."""
import numpy as np
import pytest

from conftest import say
from cumetis_amd import _lib as L

pytestmark = pytest.mark.gpu


def _case(seed, sizes):
    rng = np.random.default_rng(seed)
    sizes = np.asarray(sizes, np.int64)
    segs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    n = int(sizes.sum())
    keys = rng.integers(-50000, 50000, size=n).astype(np.int32)
    vals = rng.integers(0, 1 << 30, size=n).astype(np.int32)
    return keys, vals, segs


def _run(engine, fn, keys, vals, segs):
    dk = engine.to_device(keys)
    dv = engine.to_device(vals)
    ds = engine.to_device(segs)
    rc = fn(engine.ctx, dk.ptr, dv.ptr, len(keys), ds.ptr, len(segs))
    out = (dk.to_host(len(keys)), dv.to_host(len(keys))) if rc == L.CM_OK else None
    for d in (dk, dv, ds):
        d.free()
    return rc, out


def _expected(keys, vals, segs):
    k, v = keys.copy(), vals.copy()
    bounds = list(segs) + [len(keys)]
    for s in range(len(segs)):
        a, b = bounds[s], bounds[s + 1]
        o = np.argsort(keys[a:b], kind="stable")
        k[a:b], v[a:b] = keys[a:b][o], vals[a:b][o]
    return k, v


def test_noalloc_segsort_matches_segsort(engine):
    rng = np.random.default_rng(11)
    big = np.concatenate([rng.integers(0, 70, size=4000), [100, 300, 700, 1500, 3000, 4096],
                          [5000, 20000, 70001], rng.integers(0, 5, size=3000)])
    small = np.concatenate([rng.integers(0, 40, size=900), [129, 2049, 9000]])
    for step, (seed, sizes) in enumerate([(1, big), (2, small)]):
        keys, vals, segs = _case(seed, sizes)
        say(f"[noalloc segsort] call {step + 1}: n={len(keys)} segments={len(segs)} max={int(max(sizes))}")
        rc, ref = _run(engine, engine.lib.cm_segsort_i32, keys, vals, segs)
        assert rc == L.CM_OK
        reserved0, in_use0 = engine.pool_bytes()
        rc, got = _run(engine, engine.lib.cm_segsort_i32_noalloc, keys, vals, segs)
        assert rc == L.CM_OK, engine.lib.cm_ctx_last_error(engine.ctx)
        # the arrays of _run come from the pool and go back to it; the call
        # itself takes nothing from it
        assert engine.pool_bytes() == (reserved0, in_use0)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])
        ek, ev = _expected(keys, vals, segs)
        np.testing.assert_array_equal(got[0], ek)
        np.testing.assert_array_equal(got[1], ev)


def test_noalloc_segsort_rejects_oversize(engine):
    n = 2_000_000  # about 48 MB of scratch, more than CM_NOALLOC_SCRATCH_BYTES (32 MiB)
    keys = np.zeros(n, np.int32)
    vals = np.zeros(n, np.int32)
    segs = np.arange(0, n, 1000, dtype=np.int32)
    say(f"[noalloc segsort] oversize n={n}")
    rc, _ = _run(engine, engine.lib.cm_segsort_i32_noalloc, keys, vals, segs)
    assert rc == L.CM_ERR_UNSUPPORTED
    assert b"scratch" in engine.lib.cm_ctx_last_error(engine.ctx)
