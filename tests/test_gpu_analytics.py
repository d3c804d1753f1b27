"""GPU ML queries and multi-column aggregations against the reference's known
answers and the oracle: knn1, the CPU-layout matmul, the chunk readers'
distance_squared / dot_product (bit-exact: the same sequential folds) and
Aggregation / Expr (MAX / MIN and NONE bit-exact, SUM within 1e-12
relative of the compensated sum)."""
import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G
from oracle.oracle import OracleDecoder, aggregate
from tests.helpers import chunk_of, known_answers
from tests.test_oracle_analytics import KINDS, nodes_of

pytestmark = pytest.mark.gpu
KA = known_answers()


def enc(values, codec, chunk, scale=100):
    return E.encode(np.array(values, dtype=np.float64), E.CompressorConfig.by_name(codec, scale=scale), chunk)


def bits_eq(a, b):
    return np.array([a], dtype=np.float64).view(np.uint64)[0] == np.array([b], dtype=np.float64).view(np.uint64)[0]


@pytest.mark.parametrize("codec", KA["ml_codecs"])
def test_knn1_and_matmul_known_answers(ctx, codec):
    for case in KA["knn1"]:
        d = E.Decoder(ctx, enc(case["data"], codec, chunk_of(case["chunk"])))
        assert d.knn1(case["target"])[0] == case["expected"], case["name"]
        d.close()
    for case in KA["matmul_cpu"]:
        data = np.array(case["data"])
        b, h, w = data.shape
        target = np.array(case["target"])
        d = E.Decoder(ctx, enc(data.ravel(), codec, chunk_of(case["chunk"])))
        got = d.matmul(target.ravel(), (target.shape[1], target.shape[0]), (h, w))
        assert got.tolist() == np.array(case["expected"]).ravel().tolist(), case["name"]
        d.close()


@pytest.mark.parametrize("codec", ["gorilla", "buff", "chimp128", "uncompressed", "elf"])
@pytest.mark.parametrize("L,chunk", [(8, 4096), (7, 1000), (100, 777), (3, 5), (13, 13)])
def test_knn1_parity(ctx, codec, L, chunk):
    n = 12000
    vals = G.random_walk(n, seed=L + chunk)
    data = enc(vals, codec, E.ChunkOption.record_count(chunk))
    d, o = E.Decoder(ctx, data), OracleDecoder(data)
    rng = np.random.default_rng(L)
    for trial in range(3):
        start = int(rng.integers(0, n - L))
        target = vals[start:start + L] + (0.0 if trial == 0 else rng.normal(0, 0.05, L).round(2))
        gi, gd = d.knn1(target)
        wi, wd = o.knn1(target)
        assert gi == wi and bits_eq(gd, wd), (trial, gi, wi, gd, wd)
    # a target equal to many vectors: ties resolved as the reference does
    const = enc(np.full(n, 3.5), codec, E.ChunkOption.record_count(chunk))
    dc, oc = E.Decoder(ctx, const), OracleDecoder(const)
    assert dc.knn1(np.full(L, 3.5)) == oc.knn1(np.full(L, 3.5))
    d.close()
    dc.close()


@pytest.mark.parametrize("codec", ["gorilla", "chimp", "elf", "chimp128"])
def test_knn1_nan_truncated(ctx, codec):
    vals = G.random_walk(5000, seed=2)
    vals[[700, 2500]] = np.nan
    data = enc(vals, codec, E.ChunkOption.record_count(1000))
    d, o = E.Decoder(ctx, data), OracleDecoder(data)
    for L in (6, 9):
        t = vals[1200:1200 + L]
        gi, gd = d.knn1(t)
        wi, wd = o.knn1(t)
        assert gi == wi and bits_eq(gd, wd)
    d.close()


@pytest.mark.parametrize("codec", ["gorilla", "buff", "rle", "delta_sprintz"])
def test_matmul_and_folds_parity(ctx, codec):
    rng = np.random.default_rng(9)
    vals = np.round(rng.uniform(-10, 10, 6000), 2)
    for chunk, (dr, dc), tc in ((600, (4, 5), 3), (1000, (3, 8), 2), (777, (7, 3), 5)):
        data = enc(vals[: (6000 // (dr * dc)) * dr * dc], codec, E.ChunkOption.record_count(chunk))
        d, o = E.Decoder(ctx, data), OracleDecoder(data)
        target = np.round(rng.uniform(-3, 3, dc * tc), 3)
        try:
            want = o.matmul(target, (dc, tc), (dr, dc))
        except Exception:
            with pytest.raises(E.EbiError):
                d.matmul(target, (dc, tc), (dr, dc))
        else:
            got = d.matmul(target, (dc, tc), (dr, dc))
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (chunk, dr, dc)
        for cid, off, ln in ((0, 0, 50), (1, 100, 700), (2, 5, 3), (0, 10_000, 4)):
            t = np.round(rng.uniform(-2, 2, ln), 2)
            assert bits_eq(d.distance_squared(cid, off, t), o.distance_squared(cid, off, t))
            assert bits_eq(d.dot_product(cid, off, t), o.dot_product(cid, off, t))
        d.close()


def test_expr_known_answers(ctx):
    for case in KA["expr"]:
        decs = [E.Decoder(ctx, enc(c, "uncompressed", chunk_of(KA["expr_chunk"]))) for c in case["columns"]]
        aggs = [E.Aggregation(KINDS[k] if k != "none" else E.AGG_NONE, E.Expr(nodes_of(p)))
                for k, p in case["aggs"]]
        # the C ABI's kinds: 0 none, 1 max, 2 min, 3 sum (as KINDS)
        got = E.Aggregation.compute_multiple(aggs, decs, case["bitmask"])
        assert got == case["expected"], case["name"]
        for dd in decs:
            dd.close()


@pytest.mark.parametrize("codecs", [("gorilla", "buff"), ("uncompressed", "chimp128"), ("delta_sprintz", "elf"),
                                    ("rle", "rle")])
def test_expr_parity(ctx, codecs):
    n, chunk = 20000, 3000
    a = G.random_walk(n, seed=1)
    b = G.uniform(n, 1.0, 9.0, 2, seed=2)
    fa = bytes(enc(a, codecs[0], E.ChunkOption.record_count(chunk)))
    fb = bytes(enc(b, codecs[1], E.ChunkOption.record_count(chunk)))
    da, db = E.Decoder(ctx, fa), E.Decoder(ctx, fb)
    rng = np.random.default_rng(4)
    progs = [
        E.Expr.index(0) * E.Expr.index(1) + E.Expr.index(1),
        (E.Expr.index(0) - 100.0) / E.Expr.index(1),
        E.Expr.literal(2.0) * E.Expr.index(1),
        E.Expr.index(1),
    ]
    for mask in (None, np.flatnonzero(rng.random(n) < 0.3), []):
        for kind in (E.AGG_SUM, E.AGG_MAX, E.AGG_MIN, E.AGG_NONE):
            for p in progs:
                if kind == E.AGG_NONE and len(p.nodes) == 1:
                    continue
                got = E.Aggregation(kind, p).compute([da, db], mask)
                want, comp = aggregate([fa, fb], p.nodes, kind, mask)
                if kind == E.AGG_NONE:
                    assert np.array_equal(np.asarray(got).view(np.uint64), want.view(np.uint64))
                elif kind == E.AGG_SUM:
                    assert got == want or abs(got - comp) <= 1e-12 * max(abs(comp), 1e-300), (got, want, comp)
                else:
                    assert bits_eq(got, want), (kind, got, want)
    with pytest.raises(E.EbiError) as ei:
        E.Aggregation.sum(E.Expr.literal(1.0) + E.Expr.literal(2.0)).compute([da, db])
    assert ei.value.status == E.PRECONDITION
    da.close()
    db.close()
