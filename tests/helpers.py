"""Shared test helpers: fixtures loading and predicate construction."""
import json
from pathlib import Path

import numpy as np

import ebi_amd as E

GOLDEN = Path(__file__).resolve().parent / "golden"


def known_answers():
    return json.loads((GOLDEN / "known_answers.json").read_text())


def chunk_of(spec):
    kind, value = spec
    return E.ChunkOption(kind, value)


def pred_of(d):
    if d["kind"] == 0:
        return E.Predicate.eq(d["value"])
    if d["kind"] == 1:
        return E.Predicate.ne(d["value"])
    s, e = d["start"], d["end"]
    return E.Predicate.range(E.RangeValue(s[0], float(s[1])), E.RangeValue(e[0], float(e[1])))


def round_at(x, scale):
    """(x * scale).round() / scale with Rust's half-away rounding (aggregation.rs:14-16)."""
    y = np.asarray(x, dtype=np.float64) * scale
    return np.copysign(np.floor(np.abs(y) + 0.5), y) / scale


def scale_for(codec, table):
    return table.get(codec, 1)
