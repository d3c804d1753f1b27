"""GPU tests of the serving loop (ebi_hip_serve_*): continuous batching of
concurrent queries over resident columns.

Every query's result in the mix must equal its result alone (the
synchronous entry point) bit for bit -- bitsets, materialized f64 patterns,
min/max and sums (the same plan and the same deterministic reduction order
run in both) -- and equal the CPU oracle (sums within 1e-12 relative of a
compensated sum).  Admission is by device memory: a query whose scratch
cannot fit the budget is rejected, queries that fit one at a time but not
all together wait for memory and complete.
"""
import threading

import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G
from oracle.oracle import OracleDecoder

pytestmark = pytest.mark.gpu


def same(a, b):
    if isinstance(a, float):
        return np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64) or (np.isnan(a) and np.isnan(b))
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float64:
        return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


COLS = [("walk2", "buff", 100, 300_000, 65536), ("walk2", "delta_sprintz", 100, 200_000, 32768),
        ("unif3", "gorilla", 1000, 150_000, 65536), ("sin4", "chimp128", 10000, 120_000, 16384),
        ("step2", "rle", 100, 250_000, 65536), ("walk2", "elf", 100, 100_000, 8192),
        ("unif3", "chimp", 1000, 90_000, 65536), ("sin4", "uncompressed", 1, 70_000, 4096)]


@pytest.fixture(scope="module")
def columns(ctx):
    out = []
    for gen, codec, scale, n, chunk in COLS:
        fn, _ = G.GENERATORS[gen]
        vals = fn(n)
        f = E.encode(vals, E.CompressorConfig.by_name(codec, scale=scale), E.ChunkOption.record_count(chunk))
        out.append((E.Decoder(ctx, f), OracleDecoder(f), vals))
    yield out
    for d, _, _ in out:
        d.close()


def make_queries(columns, seed):
    rng = np.random.default_rng(seed)
    qs = []
    for d, o, vals in columns:
        lo, hi = np.quantile(vals, [0.3, 0.6])
        rng_pred = E.Predicate.range(E.RangeValue.inclusive(float(lo)), E.RangeValue.exclusive(float(hi)))
        gt = E.Predicate.range(E.RangeValue.exclusive(float(np.quantile(vals, 0.9))), E.RangeValue.none())
        n = d.number_of_records
        mask = np.sort(rng.choice(n, size=n // 3, replace=False))
        qs += [
            (E.OP_FILTER, dict(predicate=rng_pred)),
            (E.OP_FILTER, dict(predicate=gt, bitmask=mask)),
            (E.OP_FILTER_MATERIALIZE, dict(predicate=rng_pred)),
            (E.OP_MATERIALIZE, dict()),
            (E.OP_MATERIALIZE, dict(bitmask=mask)),
            (E.OP_SUM, dict()),
            (E.OP_SUM, dict(bitmask=mask)),
            (E.OP_MIN, dict(bitmask=mask)),
            (E.OP_MAX, dict()),
            (E.OP_MAX, dict(chunk_id=0)),
        ]
        qs[-10:] = [(d, o) + q for q in qs[-10:]]
    order = rng.permutation(len(qs))
    return [qs[i] for i in order]


def alone(ctx, d, op, kw):
    return E.run_query(ctx, E.Query(op, d, **kw))


def oracle_check(o, op, kw, got):
    p, m, cid = kw.get("predicate"), kw.get("bitmask"), kw.get("chunk_id")
    if op == E.OP_FILTER:
        assert np.array_equal(got, o.filter(p, bitmask=m, chunk_id=cid))
    elif op == E.OP_FILTER_MATERIALIZE:
        assert same(got, o.filter_materialize(p, bitmask=m, chunk_id=cid))
    elif op == E.OP_MATERIALIZE:
        assert same(got, o.materialize(bitmask=m, chunk_id=cid))
    elif op == E.OP_SUM:
        w = o.sum(bitmask=m, chunk_id=cid)
        c = o.sum_compensated(bitmask=m, chunk_id=cid)
        assert got == w or abs(got - c) <= 1e-12 * abs(c)
    elif op == E.OP_MIN:
        assert same(got, o.min(bitmask=m, chunk_id=cid))
    else:
        assert same(got, o.max(bitmask=m, chunk_id=cid))


def test_mix_equals_alone_and_oracle(ctx, columns):
    qs = make_queries(columns, seed=1)
    ref = [alone(ctx, d, op, kw) for d, o, op, kw in qs]
    for (d, o, op, kw), r in zip(qs, ref):
        oracle_check(o, op, kw, r)
    with E.Server(ctx, device_budget=1 << 30, max_batch=16, n_streams=4) as srv:
        tickets = srv.submit([E.Query(op, d, **kw) for d, o, op, kw in qs])
        for t, r, (d, o, op, kw) in zip(tickets, ref, qs):
            got, lat = srv.wait(t)
            assert same(got, r), (d.scheme, op, kw.keys())
            assert lat >= 0.0
        st = srv.stats()
    assert st["submitted"] == len(qs) and st["completed"] == len(qs) and st["rejected"] == 0
    assert st["batches"] >= 1 and st["max_inflight"] >= 2


def test_arrivals_while_others_run(ctx, columns):
    # waves submitted while earlier waves are still on the device
    qs = make_queries(columns, seed=2)
    ref = [alone(ctx, d, op, kw) for d, o, op, kw in qs]
    with E.Server(ctx, device_budget=1 << 30, max_batch=8, n_streams=3) as srv:
        tickets = []
        for i in range(0, len(qs), 7):
            tickets += srv.submit([E.Query(op, d, **kw) for d, o, op, kw in qs[i:i + 7]])
        for t, r in zip(tickets, ref):
            got, _ = srv.wait(t)
            assert same(got, r)
        assert srv.stats()["batches"] >= 2


def test_memory_admission_waits_and_rejects(ctx, columns):
    d = columns[3][0]  # Chimp128, 120,000 records: decode scratch ~1 MB per query
    qs = [E.Query(E.OP_MATERIALIZE, d) for _ in range(12)]
    want = d.materialize()
    # room for about two such queries at once: the rest wait for memory
    with E.Server(ctx, device_budget=3 << 20, max_batch=64) as srv:
        tickets = srv.submit(qs)
        for t in tickets:
            got, _ = srv.wait(t)
            assert same(got, want)
        st = srv.stats()
        assert st["rejected"] == 0 and st["max_inflight"] < 12 and st["device_bytes_peak"] <= 3 << 20
    # a query larger than the whole budget is rejected
    with E.Server(ctx, device_budget=1 << 19) as srv:
        [t] = srv.submit([E.Query(E.OP_MATERIALIZE, d)])
        with pytest.raises(E.EbiError) as ei:
            srv.wait(t)
        assert ei.value.status == E.REJECTED
        assert srv.stats()["rejected"] == 1


def test_concurrent_submitters(ctx, columns):
    qs = make_queries(columns, seed=3)
    ref = [alone(ctx, d, op, kw) for d, o, op, kw in qs]
    errors = []
    with E.Server(ctx, device_budget=1 << 30, max_batch=16) as srv:
        def worker(part):
            try:
                idx = list(range(part, len(qs), 4))
                ts = srv.submit([E.Query(qs[i][2], qs[i][0], **qs[i][3]) for i in idx])
                for i, t in zip(idx, ts):
                    got, _ = srv.wait(t)
                    if not same(got, ref[i]):
                        errors.append(i)
            except Exception as e:  # reported below
                errors.append(repr(e))
        th = [threading.Thread(target=worker, args=(p,)) for p in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    assert not errors


def test_invalid_query_completes_with_error(ctx, columns):
    d = columns[0][0]
    q = E.Query(E.OP_FILTER, d, predicate=E.Predicate.eq(1.0))
    q.c.pred.kind = 7  # not a predicate kind
    with E.Server(ctx) as srv:
        [t] = srv.submit([q])
        with pytest.raises(E.EbiError) as ei:
            srv.wait(t)
        assert ei.value.status == E.INVALID_ARGUMENT


def test_query_entry_point_matches_methods(ctx, columns):
    d, o, vals = columns[0]
    p = E.Predicate.range(E.RangeValue.inclusive(99.0), E.RangeValue.inclusive(101.0))
    assert np.array_equal(E.run_query(ctx, E.Query(E.OP_FILTER, d, predicate=p)), d.filter(p))
    assert same(E.run_query(ctx, E.Query(E.OP_SUM, d)), d.sum())


def test_filter_sum_equals_filter_then_sum(ctx, columns):
    for d, o, vals in columns:
        lo, hi = np.quantile(vals, [0.2, 0.7])
        p = E.Predicate.range(E.RangeValue.inclusive(float(lo)), E.RangeValue.inclusive(float(hi)))
        ids = d.filter(p)
        want = d.sum(bitmask=ids) if ids.size else None
        got = E.run_query(ctx, E.Query(E.OP_FILTER_SUM, d, predicate=p))
        if want is not None:
            assert same(got, want), d.scheme
        # with an input bitmask and a chunk restriction too
        m = np.arange(0, d.number_of_records, 3)
        ids = d.filter(p, bitmask=m, chunk_id=0)
        got = E.run_query(ctx, E.Query(E.OP_FILTER_SUM, d, predicate=p, bitmask=m, chunk_id=0))
        if ids.size:
            assert same(got, d.sum(bitmask=ids, chunk_id=0)), d.scheme


def test_decode_op_writes_device_buffer(ctx, columns):
    with E.Server(ctx, device_budget=1 << 28) as srv:
        bufs, qs = [], []
        for d, o, vals in columns:
            b = ctx.alloc(8 * d.number_of_records)
            bufs.append(b)
            qs.append(E.Query(E.OP_DECODE, d, out=b))
        ts = srv.submit(qs)
        for t, b, (d, o, vals) in zip(ts, bufs, columns):
            srv.wait(t)
            assert same(b.to_host(count=d.number_of_records), o.materialize())
    for b in bufs:
        b.free()


def test_decodes_sharing_one_output_buffer(ctx, columns):
    """Many in-flight DECODE queries of one column into one shared device
    buffer on several streams all complete without error, and the buffer ends
    up holding the column: a decode only writes its output (for Chimp128 the
    parse's payloads live in the query's own scratch, not in the output)."""
    for d, o, vals in columns:
        if d.scheme not in (E.CompressorConfig.by_name("chimp128").scheme, E.CompressorConfig.by_name("elf").scheme,
                            E.CompressorConfig.by_name("gorilla").scheme):
            continue
        b = ctx.alloc(8 * d.number_of_records)
        with E.Server(ctx, device_budget=1 << 30, max_batch=4, n_streams=4) as srv:
            ts = srv.submit([E.Query(E.OP_DECODE, d, out=b) for _ in range(24)])
            for t in ts:
                srv.wait(t)  # raises on a device error
        assert same(b.to_host(count=d.number_of_records), o.materialize()), d.scheme
        b.free()
