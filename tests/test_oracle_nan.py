"""The oracle's NaN end semantics, pinned by a closed form: a chunk of a
Gorilla / Chimp / Chimp128 / Elf / ElfOnChimp column decodes to its records before its
first canonical NaN (the reference's readers stop there without an error,
chunk_reader/gorilla.rs:149-152, chimp.rs:119, chimp_n.rs:219-265,
elf.rs:105-258), and the queries see exactly those records
(decoder/query.rs:15-154)."""
import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G
from oracle.oracle import OracleDecoder


@pytest.mark.parametrize("codec", ["gorilla", "chimp", "chimp128", "elf", "elf_on_chimp"])
def test_oracle_nan_truncates_chunks(codec):
    n, chunk = 3000, 700
    vals = G.random_walk(n, seed=9)
    vals[[0, 450, 699, 1400, 1401, 2999]] = np.nan
    data = E.encode(vals, E.CompressorConfig.by_name(codec), E.ChunkOption.record_count(chunk))
    o = OracleDecoder(data)
    kept, ids = [], []
    for c0 in range(0, n, chunk):
        part = vals[c0:c0 + chunk]
        k = np.flatnonzero(np.isnan(part))
        m = k[0] if k.size else part.size
        kept.append(part[:m])
        ids.append(np.arange(c0, c0 + m))
    kept, ids = np.concatenate(kept), np.concatenate(ids)
    got = o.materialize()
    assert np.array_equal(got.view(np.uint64), kept.view(np.uint64))
    allp = E.Predicate.range(E.RangeValue.none(), E.RangeValue.none())
    assert np.array_equal(o.filter(allp), ids)
    assert o.max() == kept.max() and o.min() == kept.min()
    s = 0.0
    for c0 in range(0, n, chunk):  # naive sum per chunk, chunk sums added in order
        part = vals[c0:c0 + chunk]
        k = np.flatnonzero(np.isnan(part))
        cs = 0.0
        for x in part[: (k[0] if k.size else part.size)]:
            cs += x
        s += cs
    assert o.sum() == s
    # the smallest case: [1, 2, nan, 3, 4] keeps [1, 2]
    small = E.encode(np.array([1.0, 2.0, np.nan, 3.0, 4.0]), E.CompressorConfig.by_name(codec),
                     E.ChunkOption.full())
    assert OracleDecoder(small).materialize().tolist() == [1.0, 2.0]
