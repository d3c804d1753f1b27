"""GPU query parity: filter / filter_materialize / materialize / sum / min /
max through the C ABI against the CPU oracle and the reference's known
answers.  Bitsets, materialized f64 patterns and min/max are bit-exact; BUFF
and DeltaSprintz sums are bit-exact (integer-exact per chunk, as the
reference computes them); other sums are within 1e-12 relative of an FP64
compensated sum of the selected records."""
import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G
from oracle.oracle import OracleDecoder
from tests.helpers import chunk_of, known_answers, pred_of, round_at

pytestmark = pytest.mark.gpu
KA = known_answers()
CODECS = G.ALL_CODECS


def enc(values, codec, scale, chunk):
    return E.encode(np.array(values, dtype=np.float64), E.CompressorConfig.by_name(codec, scale=scale), chunk)


def same_bits(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("codec", CODECS)
def test_known_answers(ctx, codec):
    for case in KA["filter"]:
        scale = KA["scales"]["filter"].get(codec, 1)
        d = E.Decoder(ctx, enc(case["values"], codec, scale, chunk_of(case["chunk"])))
        p = pred_of(case["pred"])
        got = d.filter(p, case["bitmask"], case["chunk_id"])
        assert got.tolist() == case["expected"], case["name"]
        fm = d.filter_materialize(p, case["bitmask"], case["chunk_id"])
        want = np.array(case["values"])[case["expected"]]
        if codec in ("buff", "delta_sprintz"):
            fm, want = round_at(fm, scale), round_at(want, scale)
        assert fm.tolist() == want.tolist(), case["name"]
    for case in KA["materialize"]:
        scale = KA["scales"]["materialize"].get(codec, 1)
        d = E.Decoder(ctx, enc(case["values"], codec, scale, chunk_of(case["chunk"])))
        got = d.materialize(case["bitmask"], case["chunk_id"])
        want = np.array(case["expected"])
        if codec in ("buff", "delta_sprintz"):
            got, want = round_at(got, scale), round_at(want, scale)
        assert got.tolist() == want.tolist(), case["name"]
    for case in KA["aggregate"]:
        scale = KA["scales"]["aggregate"].get(codec, 1)
        d = E.Decoder(ctx, enc(case["values"], codec, scale, chunk_of(case["chunk"])))
        assert getattr(d, case["op"])(case["bitmask"]) == case["expected"], case["name"]


def _preds(vals, rng):
    q = np.quantile(vals, [0.001, 0.01, 0.1, 0.5, 0.9])
    x = float(vals[rng.integers(len(vals))])
    R, I, X, N = E.Predicate.range, E.RangeValue.inclusive, E.RangeValue.exclusive, E.RangeValue.none()
    return [
        R(I(float(q[3])), I(float(q[4]))),
        R(X(float(q[2])), N),
        R(N, X(float(q[1]))),
        R(I(float(q[0])), X(float(q[0]) + 0.5)),
        R(N, N),
        E.Predicate.eq(x),
        E.Predicate.ne(x),
        E.Predicate.eq(x + 1e6),
        R(I(1e9), N),
        R(N, I(-1e9)),
    ]


CASES = [("walk2", 20000, 1003), ("unif3", 12000, 4096), ("step2", 20000, 2048), ("sin4", 9000, 777)]


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("gen,n,chunk", CASES)
def test_random_parity(ctx, codec, gen, n, chunk):
    fn, dec = G.GENERATORS[gen]
    vals = fn(n, seed=n)
    data = enc(vals, codec, 10 ** dec, E.ChunkOption.record_count(chunk))
    d = E.Decoder(ctx, data)
    o = OracleDecoder(data)
    rng = np.random.default_rng(n + chunk)
    masks = [None, np.flatnonzero(rng.random(n) < 0.5), np.flatnonzero(rng.random(n) < 0.003), []]
    for cid in (None, 1):
        for m in masks:
            assert same_bits(d.materialize(m, cid), o.materialize(m, cid))
            for op in ("min", "max"):
                g, w = getattr(d, op)(m, cid), getattr(o, op)(m, cid)
                assert same_bits([g], [w]), (op, g, w)
            g, w = d.sum(m, cid), o.sum(m, cid)
            if codec in ("buff", "delta_sprintz"):
                assert same_bits([g], [w]), ("sum", g, w)
            else:
                c = o.sum_compensated(m, cid)
                assert abs(g - c) <= 1e-12 * max(abs(c), 1e-300) or g == c, ("sum", g, c)
            for p in _preds(vals, rng):
                ids = o.filter(p, m, cid)
                assert np.array_equal(d.filter(p, m, cid), ids), (p, cid)
                assert same_bits(d.filter_materialize(p, m, cid), o.filter_materialize(p, m, cid)), (p, cid)
                # the fused filter + sum: sum(Some(filter(..))) of the reference's Decoder
                g = E.run_query(ctx, E.Query(E.OP_FILTER_SUM, d, predicate=p, bitmask=m, chunk_id=cid))
                w = o.sum(ids, cid)
                if codec in ("buff", "delta_sprintz"):
                    assert same_bits([g], [w]), ("filter_sum", p, g, w)
                else:
                    c = o.sum_compensated(ids, cid)
                    assert abs(g - c) <= 1e-12 * max(abs(c), 1e-300) or g == c, ("filter_sum", p, g, c)


@pytest.mark.parametrize("codec", ["uncompressed", "gorilla", "buff", "chimp128"])
def test_materialize_many_blocks(ctx, codec):
    """Materialize across many stream-compaction blocks (32,768 records
    each): every record, dense and sparse random bitmasks, one contiguous
    run crossing block edges, a single record per block and a range filter;
    bit-exact against the oracle."""
    n = 300_001
    vals = G.random_walk(n, seed=11)
    data = enc(vals, codec, 100, E.ChunkOption.record_count(65536))
    d = E.Decoder(ctx, data)
    o = OracleDecoder(data)
    rng = np.random.default_rng(5)
    masks = [None, np.flatnonzero(rng.random(n) < 0.6), np.flatnonzero(rng.random(n) < 0.001),
             np.arange(30_000, 140_000), np.arange(32_767, n, 32_768), [n - 1]]
    for m in masks:
        assert same_bits(d.materialize(m), o.materialize(m))
    q = np.quantile(vals, [0.2, 0.7])
    p = E.Predicate.range(E.RangeValue.inclusive(float(q[0])), E.RangeValue.exclusive(float(q[1])))
    assert same_bits(d.filter_materialize(p, None), o.filter_materialize(p, None))
    d.close()


@pytest.mark.parametrize("center,spread", [(5e11, 1e9), (1e9, 1e6)])
def test_buff_sum_past_exact_integers(ctx, center, spread):
    """BUFF sums whose chunk integers (first case) or whose total magnitude
    (second) pass 2^53 take the in-order f64 fold instead of the exact
    integer total; both must equal the reference's `result += v`, bit for
    bit, with and without a bitmask, through sum and the fused filter + sum.
    The first case also runs the wide-delta (B >= 24) path."""
    rng = np.random.default_rng(int(center) % 1000)
    vals = np.round(center + rng.uniform(-spread, spread, 60000), 3)
    data = enc(vals, "buff", 1000, E.ChunkOption.record_count(1000))
    d, o = E.Decoder(ctx, data), OracleDecoder(data)
    m = np.flatnonzero(rng.random(vals.size) < 0.4)
    p = E.Predicate.range(E.RangeValue.inclusive(center - spread / 2), E.RangeValue.none())
    for mask in (None, m):
        assert same_bits([d.sum(mask)], [o.sum(mask)])
        ids = o.filter(p, mask)
        got = E.run_query(ctx, E.Query(E.OP_FILTER_SUM, d, predicate=p, bitmask=mask))
        assert same_bits([got], [o.sum(ids)])
