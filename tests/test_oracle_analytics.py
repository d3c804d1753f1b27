"""The oracle's ML queries and aggregations against the reference's own
known answers (ebi/tests/ml_query.rs:266-450, ebi/tests/aggregation.rs:472-750,
transcribed in tests/golden/known_answers.json), plus the closed forms the
mathematics fixes: a sum of squares of identical vectors is 0, a dot
product with a unit vector picks one record."""
import numpy as np
import pytest

import ebi_amd as E
from oracle.oracle import OracleDecoder, aggregate
from tests.helpers import chunk_of, known_answers

KA = known_answers()
OPS = {"+": 2, "-": 3, "*": 4, "/": 5}
KINDS = {"none": 0, "max": 1, "min": 2, "sum": 3}


def nodes_of(prog):
    out = []
    for n in prog:
        if n[0] == "i":
            out.append((0, n[1], 0.0))
        elif n[0] == "l":
            out.append((1, 0, n[1]))
        else:
            out.append((OPS[n[0]], 0, 0.0))
    return out


def enc(values, codec, chunk):
    return bytes(E.encode(np.array(values, dtype=np.float64), E.CompressorConfig.by_name(codec), chunk))


@pytest.mark.parametrize("codec", KA["ml_codecs"])
def test_oracle_knn1_known_answers(codec):
    for case in KA["knn1"]:
        o = OracleDecoder(enc(case["data"], codec, chunk_of(case["chunk"])))
        idx, _ = o.knn1(case["target"])
        assert idx == case["expected"], case["name"]


@pytest.mark.parametrize("codec", KA["ml_codecs"])
def test_oracle_matmul_known_answers(codec):
    for case in KA["matmul_cpu"]:
        data = np.array(case["data"])
        b, h, w = data.shape
        target = np.array(case["target"])  # columns
        o = OracleDecoder(enc(data.ravel(), codec, chunk_of(case["chunk"])))
        got = o.matmul(target.ravel(), (target.shape[1], target.shape[0]), (h, w))
        assert got.tolist() == np.array(case["expected"]).ravel().tolist(), case["name"]


def test_oracle_expr_known_answers():
    for case in KA["expr"]:
        files = [enc(c, "uncompressed", chunk_of(KA["expr_chunk"])) for c in case["columns"]]
        for (kind, prog), want in zip(case["aggs"], case["expected"]):
            got, _ = aggregate(files, nodes_of(prog), KINDS[kind], case["bitmask"])
            assert got == want, case["name"]


def test_oracle_fold_closed_forms():
    v = np.round(np.random.default_rng(3).uniform(-5, 5, 200), 2)
    o = OracleDecoder(enc(v, "gorilla", E.ChunkOption.record_count(64)))
    assert o.distance_squared(1, 10, v[74:90]) == 0.0
    unit = np.zeros(5)
    unit[3] = 1.0
    assert o.dot_product(2, 7, unit) == v[128 + 7 + 3]
    # clamped offset and the zip's truncation at the chunk's end
    assert o.dot_product(3, 60, np.ones(10)) == sum(v[192 + 60:200])
    assert o.dot_product(0, 500, np.ones(3)) == 0.0
