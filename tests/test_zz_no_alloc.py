"""ebi_hip_query_noalloc (fixed per-GPU scratch, no allocation) against
ebi_hip_query, through the C ABI: every op on an in-situ codec (BUFF), a
fused XOR codec (Gorilla), Chimp128 and a codec decoded to HBM
(Uncompressed); two calls in a row per case, the second on a smaller column,
so the second call runs on scratch the first one left behind.  Results must
be bit-identical (the two entry points run the same kernels).  A query whose
scratch exceeds the fixed size is refused with EBI_INVALID_ARGUMENT."""
import ctypes as C

import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G

pytestmark = pytest.mark.gpu

OPS = [E.OP_SUM, E.OP_MIN, E.OP_MAX, E.OP_FILTER, E.OP_FILTER_SUM, E.OP_MATERIALIZE, E.OP_FILTER_MATERIALIZE]


def _run(ctx, dec, op, pred, mask, noalloc):
    q = E.Query(op, dec, pred if op in (E.OP_FILTER, E.OP_FILTER_SUM, E.OP_FILTER_MATERIALIZE) else None, mask)
    fn = E.lib().ebi_hip_query_noalloc if noalloc else E.lib().ebi_hip_query
    rc = fn(ctx._h, C.byref(q.c))
    assert rc == E.OK, (rc, E.lib().ebi_hip_last_error(ctx._h))
    return q.result()


def _same(a, b):
    if isinstance(a, float):
        return np.array([a]).view(np.uint64)[0] == np.array([b]).view(np.uint64)[0]
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float64:
        return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


@pytest.mark.parametrize("codec", ["buff", "gorilla", "chimp128", "uncompressed"])
def test_noalloc_matches_query(ctx, codec):
    for n, chunk, seed in ((300_000, 65536, 1), (40_000, 4096, 2)):  # the second call smaller
        vals = G.random_walk(n, seed=seed)
        data = E.encode(vals, E.CompressorConfig.by_name(codec, scale=100), E.ChunkOption.record_count(chunk))
        dec = E.Decoder(ctx, data)
        rng = np.random.default_rng(seed)
        q = np.quantile(vals, [0.3, 0.8])
        pred = E.Predicate.range(E.RangeValue.inclusive(float(q[0])), E.RangeValue.exclusive(float(q[1])))
        for mask in (None, np.flatnonzero(rng.random(n) < 0.4)):
            for op in OPS:
                print(f"{codec} n={n} op={op} mask={'none' if mask is None else len(mask)}", flush=True)
                a = _run(ctx, dec, op, pred, mask, True)
                b = _run(ctx, dec, op, pred, mask, False)
                assert _same(a, b), (codec, n, op)
        dec.close()


def test_noalloc_refuses_oversized(ctx):
    # materializing 9 M decoded records needs 72 MiB of record scratch: refused before any device work
    n = 9_000_000
    data = E.encode(np.zeros(n), E.CompressorConfig.uncompressed(), E.ChunkOption.record_count(1 << 20))
    dec = E.Decoder(ctx, data)
    q = E.Query(E.OP_MATERIALIZE, dec)
    rc = E.lib().ebi_hip_query_noalloc(ctx._h, C.byref(q.c))
    assert rc == E.INVALID_ARGUMENT
    assert b"above the fixed" in E.lib().ebi_hip_last_error(ctx._h)
    # the same query through the allocating entry point still works
    assert E.run_query(ctx, E.Query(E.OP_SUM, dec)) == 0.0
    dec.close()
