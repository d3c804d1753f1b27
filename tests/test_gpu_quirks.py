"""GPU parity on the reference's quirks that ordinary data reaches:

* a canonical NaN (0x7ff8000000000000, numpy's nan) in a Gorilla / Chimp /
  Chimp128 / Elf / ElfOnChimp column ends its chunk there -- the reference's readers stop
  at the first NaN and return the records before it without an error
  (chunk_reader/gorilla.rs:149-152, chimp.rs:119, chimp_n.rs:219-265,
  elf.rs:105-258, consumed by decoder/query.rs:15-154);
* BUFF sum over a constant chunk (B = 0) returns value * n_records even
  under a bitmask (buff/sum.rs:80-84);
* DeltaSprintz chunks of bit width W = 0 (all records equal) queried with
  and without a bitmask (chunk_reader/sprintz.rs:91-446).

Every result is compared with the CPU oracle on the same file (bit-exact;
sums of the XOR codecs within 1e-12 relative of the compensated sum)."""
import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G
from oracle.oracle import OracleDecoder

pytestmark = pytest.mark.gpu

XOR = ["gorilla", "chimp", "chimp128", "elf", "elf_on_chimp"]


def same_bits(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _nan_column(n, at, seed=5):
    v = G.random_walk(n, seed=seed)
    v[list(at)] = np.nan
    return v


def _all_ops(d, o, n, vals, xor_sum=True):
    rng = np.random.default_rng(n)
    R, I, X, N = E.Predicate.range, E.RangeValue.inclusive, E.RangeValue.exclusive, E.RangeValue.none()
    fin = vals[np.isfinite(vals)]
    q = np.quantile(fin, [0.1, 0.6])
    preds = [R(I(float(q[0])), X(float(q[1]))), R(N, N), E.Predicate.ne(float(fin[3])), E.Predicate.eq(float(fin[7]))]
    masks = [None, np.flatnonzero(rng.random(n) < 0.5), [0, n - 1]]
    for cid in (None, 0, 1, 2):
        for m in masks:
            print(f"  chunk_id={cid} mask={'none' if m is None else len(m)}", flush=True)
            assert same_bits(d.materialize(m, cid), o.materialize(m, cid)), ("materialize", cid)
            for op in ("min", "max"):
                g, w = getattr(d, op)(m, cid), getattr(o, op)(m, cid)
                assert same_bits([g], [w]), (op, cid, g, w)
            g, w = d.sum(m, cid), o.sum(m, cid)
            c = o.sum_compensated(m, cid)
            assert g == w or abs(g - c) <= 1e-12 * max(abs(c), 1e-300), ("sum", cid, g, w, c)
            for p in preds:
                assert np.array_equal(d.filter(p, m, cid), o.filter(p, m, cid)), ("filter", p, cid)
                assert same_bits(d.filter_materialize(p, m, cid), o.filter_materialize(p, m, cid)), ("fm", p, cid)


@pytest.mark.parametrize("codec", XOR)
@pytest.mark.parametrize("chunk", [1000, 64, 4096])
def test_nan_ends_chunk(ctx, codec, chunk):
    """NaN at record 0 of the column, mid-chunk, in a chunk's last record,
    twice in one chunk and in the column's last record."""
    n = 9000
    at = [0, 1500, 1999, 3100, 3105, 8999]
    vals = _nan_column(n, at)
    data = E.encode(vals, E.CompressorConfig.by_name(codec), E.ChunkOption.record_count(chunk))
    print(f"{codec} chunk={chunk}", flush=True)
    d, o = E.Decoder(ctx, data), OracleDecoder(data)
    try:
        # closed form: each chunk's records before its first NaN, in order
        want = []
        for c0 in range(0, n, chunk):
            part = vals[c0:c0 + chunk]
            k = np.flatnonzero(np.isnan(part))
            want.append(part[: k[0]] if k.size else part)
        want = np.concatenate(want)
        got = d.materialize()
        assert same_bits(got, want)
        _all_ops(d, o, n, vals)
        # the device decode writes each chunk's records before its NaN
        buf = ctx.alloc(8 * n)
        try:
            d.decode_device(buf)
            dev = buf.to_host(np.float64, n)
        finally:
            buf.free()
        for c0 in range(0, n, chunk):
            part = vals[c0:c0 + chunk]
            k = np.flatnonzero(np.isnan(part))
            m = k[0] if k.size else part.size
            assert same_bits(dev[c0:c0 + m], part[:m]), c0
        # the same queries through the serving loop (fused query kernels)
        with E.Server(ctx) as srv:
            p = E.Predicate.range(E.RangeValue.inclusive(99.0), E.RangeValue.none())
            qs = [E.Query(E.OP_SUM, d), E.Query(E.OP_MAX, d), E.Query(E.OP_FILTER, d, p),
                  E.Query(E.OP_MATERIALIZE, d), E.Query(E.OP_FILTER_SUM, d, p)]
            res = [srv.wait(t)[0] for t in srv.submit(qs)]
        assert res[0] == d.sum() and same_bits([res[1]], [o.max()])
        assert np.array_equal(res[2], o.filter(p)) and same_bits(res[3], want)
        assert res[4] == d.sum(o.filter(p))
    finally:
        d.close()


@pytest.mark.parametrize("codec", XOR)
def test_nan_matmul_precondition(ctx, codec):
    """A data matrix cannot come from a chunk that ends early (the reference's
    slice of the decoded chunk panics, chunk_reader.rs:271)."""
    vals = G.uniform(64, -10, 10, 3)
    vals[10] = np.nan
    d = E.Decoder(ctx, E.encode(vals, E.CompressorConfig.by_name(codec), E.ChunkOption.record_count(64)))
    with pytest.raises(E.EbiError) as ei:
        d.matmul_cuda(np.eye(8), (8, 8), (8, 8))
    assert ei.value.status == E.PRECONDITION
    d.close()


@pytest.mark.parametrize("chunk", [1000, 512, 4096])
def test_buff_constant_chunk_sum_ignores_bitmask(ctx, chunk):
    """buff/sum.rs:80-84: B = 0 -> value * n_records, bitmask or not; min /
    max of a B = 0 chunk are its value; an in-chunk-empty bitmask prunes."""
    n = 5000
    vals = np.full(n, 12.5)
    vals[2000:2500] = 3.25  # one chunk (at 1000 per chunk) holds two levels
    data = E.encode(vals, E.CompressorConfig.buff(scale=100), E.ChunkOption.record_count(chunk))
    d, o = E.Decoder(ctx, data), OracleDecoder(data)
    rng = np.random.default_rng(chunk)
    for m in (None, np.flatnonzero(rng.random(n) < 0.3), [5, 7, 4999], list(range(1000, 1010))):
        for cid in (None, 0, 3):
            g, w = d.sum(m, cid), o.sum(m, cid)
            assert same_bits([g], [w]), ("sum", m if m is None else len(m), cid, g, w)
            for op in ("min", "max"):
                assert same_bits([getattr(d, op)(m, cid)], [getattr(o, op)(m, cid)]), op
            assert same_bits(d.materialize(m, cid), o.materialize(m, cid))
    # the quirk itself: a sparse bitmask inside one constant chunk sums the whole chunk
    if chunk == 1000:
        assert d.sum([5, 7]) == 12.5 * 1000
    d.close()


@pytest.mark.parametrize("chunk", [1000, 777])
def test_sprintz_zero_width_queries(ctx, chunk):
    """DeltaSprintz W = 0 chunks (every record equals q0 / scale), queried
    with and without a bitmask, beside chunks of nonzero width."""
    n = 4000
    vals = np.full(n, 7.25)
    vals[1500:2500] = G.random_walk(1000, seed=3)  # chunks with W > 0 in the middle
    data = E.encode(vals, E.CompressorConfig.delta_sprintz(scale=100), E.ChunkOption.record_count(chunk))
    d, o = E.Decoder(ctx, data), OracleDecoder(data)
    rng = np.random.default_rng(chunk)
    R, I, X, N = E.Predicate.range, E.RangeValue.inclusive, E.RangeValue.exclusive, E.RangeValue.none()
    preds = [E.Predicate.eq(7.25), E.Predicate.ne(7.25), R(I(7.0), X(7.5)), R(X(7.25), N), R(N, N),
             R(I(99.0), I(101.0))]
    for m in (None, np.flatnonzero(rng.random(n) < 0.4), [0, 1, 3999], list(range(100, 140))):
        for cid in (None, 0, 2):
            g, w = d.sum(m, cid), o.sum(m, cid)
            assert same_bits([g], [w]), ("sum", cid, g, w)
            for op in ("min", "max"):
                assert same_bits([getattr(d, op)(m, cid)], [getattr(o, op)(m, cid)]), op
            assert same_bits(d.materialize(m, cid), o.materialize(m, cid))
            for p in preds:
                assert np.array_equal(d.filter(p, m, cid), o.filter(p, m, cid)), (p, cid)
                assert same_bits(d.filter_materialize(p, m, cid), o.filter_materialize(p, m, cid)), (p, cid)
    d.close()
