"""Pins the CPU oracle (and the host encoder feeding it) against the
reference's own known-answer vectors (tests/golden/known_answers.json,
transcribed from ebi/tests/query.rs and ebi/tests/aggregation.rs)."""
import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G
from oracle.oracle import OracleDecoder
from tests.helpers import chunk_of, known_answers, pred_of, round_at

KA = known_answers()
CODECS = G.ALL_CODECS


def enc(values, codec, scale, chunk):
    return E.encode(np.array(values, dtype=np.float64), E.CompressorConfig.by_name(codec, scale=scale),
                    chunk_of(chunk))


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("case", KA["filter"], ids=lambda c: c["name"])
def test_filter_known_answers(codec, case):
    scale = KA["scales"]["filter"].get(codec, 1)
    d = OracleDecoder(enc(case["values"], codec, scale, case["chunk"]))
    got = d.filter(pred_of(case["pred"]), case["bitmask"], case["chunk_id"])
    assert got.tolist() == case["expected"]


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("case", KA["filter"], ids=lambda c: c["name"])
def test_filter_materialize_known_answers(codec, case):
    # query.rs:118-189: the filter_materialize output equals the input values
    # at the expected ids (compared after rounding at the codec's scale)
    scale = KA["scales"]["filter"].get(codec, 1)
    d = OracleDecoder(enc(case["values"], codec, scale, case["chunk"]))
    got = d.filter_materialize(pred_of(case["pred"]), case["bitmask"], case["chunk_id"])
    want = np.array(case["values"])[case["expected"]]
    if codec in ("buff", "delta_sprintz"):
        got, want = round_at(got, scale), round_at(want, scale)
    assert got.tolist() == want.tolist()


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("case", KA["materialize"], ids=lambda c: c["name"])
def test_materialize_known_answers(codec, case):
    scale = KA["scales"]["materialize"].get(codec, 1)
    d = OracleDecoder(enc(case["values"], codec, scale, case["chunk"]))
    got = d.materialize(case["bitmask"], case["chunk_id"])
    want = np.array(case["expected"])
    if codec in ("buff", "delta_sprintz"):
        got, want = round_at(got, scale), round_at(want, scale)
    assert got.tolist() == want.tolist()


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("case", KA["aggregate"], ids=lambda c: c["name"])
def test_aggregate_known_answers(codec, case):
    scale = KA["scales"]["aggregate"].get(codec, 1)
    d = OracleDecoder(enc(case["values"], codec, scale, case["chunk"]))
    got = getattr(d, case["op"])(case["bitmask"])
    assert got == case["expected"]
