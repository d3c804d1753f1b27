"""BUFF controlled-precision decode (Decoder::with_precision,
api/decoder.rs:159-161; decode_with_precision, decoder/chunk_reader/buff/
controlled_precision.rs:5-75) in the oracle, pinned by hand-derived cases.

Reading the reference's loop by hand: with U = frac - PRECISION_MAP[p]
(0 if negative), plane 0 is always read, plane k >= 1 while the B - 8k bits
below it exceed U, the LSB tail (r = B mod 8 bits) only after every plane and
if r > U; the bits not read are zero.  So the decoded value is
(base + (delta with its low `drop` bits cleared)) / 2^frac, with `drop`
worked out below per case, and delta = fixed(v) - base."""
import numpy as np
import pytest

import ebi_amd as E
from oracle.oracle import OracleDecoder, lib as orc


def _fixed(v, frac):
    return np.array([orc().orc_buff_fixed(float(x), frac) for x in v], dtype=np.int64)


# (scale, frac, values spanning B bits, {precision: drop bits}) -- derived by hand:
#  B = 20, frac 8 (P = 2 planes, r = 4): p0: U = 8, plane 1 (12 > 8) read, tail 4 > 8 no -> drop 4;
#    p1: U = 3, tail 4 > 3 -> drop 0; p2: U = 0 -> drop 0.
#  B = 27, frac 11 (P = 3, r = 3): p0: U = 11, plane 1 (19 > 11) yes, plane 2 (11 > 11) no -> drop 11;
#    p1: U = 6, plane 2 (11 > 6) yes, tail 3 > 6 no -> drop 3; p2: U = 3, tail 3 > 3 no -> drop 3;
#    p3: U = 0 -> drop 0.
#  B = 16, frac 15 (P = 2, r = 0): p0: U = 15, plane 1 (8 > 15) no -> drop 8; p1: U = 10 -> drop 8;
#    p2: U = 7, plane 1 (8 > 7) yes -> drop 0.
CASES = [
    (100, 8, (0.0, 4000.0), 20, {0: 4, 1: 0, 2: 0}),
    (1000, 11, (-30000.0, 35000.0), 27, {0: 11, 1: 3, 2: 3, 3: 0}),
    (10000, 15, (0.0, 1.9), 16, {0: 8, 1: 8, 2: 0, 4: 0}),
]


@pytest.mark.parametrize("scale,frac,span,bits,drops", CASES)
def test_precision_drops_hand_derived(scale, frac, span, bits, drops):
    rng = np.random.default_rng(scale)
    dec = int(round(np.log10(scale)))
    v = np.round(rng.uniform(span[0], span[1], 3000), dec)
    v[0], v[1] = span[0], span[1]
    data = E.encode(v, E.CompressorConfig.buff(scale=scale), E.ChunkOption.full())
    f = _fixed(v, frac)
    base = int(f.min())
    delta = (f - base).astype(np.uint64)
    assert int(delta.max()).bit_length() == bits  # the chunk's B
    o = OracleDecoder(data)
    full = o.materialize()
    for p, drop in drops.items():
        o.with_precision(p)
        got = o.materialize()
        kept = delta & ~np.uint64((1 << drop) - 1)
        want = (base + kept.astype(np.int64)).astype(np.float64) / 2.0 ** frac
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (p, drop)
        # never above the full decode, and below it by less than 2^(drop - frac)
        assert np.all(got <= full) and np.all(full - got < 2.0 ** (drop - frac))
    o.with_precision(12)
    assert np.array_equal(o.materialize().view(np.uint64), full.view(np.uint64))


def test_precision_small_widths_and_constant():
    """B < 8 and B = 0 chunks decode every bit whatever the precision."""
    for vals in (np.full(100, 3.5), np.round(np.linspace(1.0, 1.2, 100), 2)):
        data = E.encode(vals, E.CompressorConfig.buff(scale=100), E.ChunkOption.record_count(30))
        o = OracleDecoder(data)
        full = o.materialize()
        for p in (0, 1, 5):
            o.with_precision(p)
            assert np.array_equal(o.materialize().view(np.uint64), full.view(np.uint64))
