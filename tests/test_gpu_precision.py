"""GPU BUFF controlled-precision decode (Decoder::with_precision,
api/decoder.rs:159-161; chunk_reader/buff/controlled_precision.rs:5-75)
against the oracle: materialize, filter_materialize and the device decode
bit-exact at every precision 0..12 and back at full precision; in-situ
filters and aggregates unchanged by the precision, as in the reference."""
import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G
from oracle.oracle import OracleDecoder

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("gen", list(G.GENERATORS))
@pytest.mark.parametrize("chunk", [4096, 1003])
def test_precision_parity(ctx, gen, chunk):
    fn, dec = G.GENERATORS[gen]
    n = 20000
    vals = fn(n, seed=chunk)
    data = E.encode(vals, E.CompressorConfig.buff(scale=10 ** dec), E.ChunkOption.record_count(chunk))
    d, o = E.Decoder(ctx, data), OracleDecoder(data)
    rng = np.random.default_rng(chunk)
    mask = np.flatnonzero(rng.random(n) < 0.3)
    q = np.quantile(vals, [0.2, 0.7])
    p = E.Predicate.range(E.RangeValue.inclusive(float(q[0])), E.RangeValue.exclusive(float(q[1])))
    full_sum, full_max, full_ids = d.sum(mask), d.max(), d.filter(p)
    buf = ctx.alloc(8 * n)
    try:
        for prec in list(range(0, 13)) + [-1]:
            print(f"{gen} chunk={chunk} precision={prec}", flush=True)
            d.with_precision(prec)
            o.with_precision(prec)
            assert same_bits(d.materialize(), o.materialize()), prec
            assert same_bits(d.materialize(mask), o.materialize(mask)), prec
            assert same_bits(d.filter_materialize(p, mask), o.filter_materialize(p, mask)), prec
            d.decode_device(buf)
            assert same_bits(buf.to_host(np.float64, n), o.materialize()), prec
            # in-situ queries ignore the precision
            assert d.sum(mask) == full_sum and d.max() == full_max
            assert np.array_equal(d.filter(p), full_ids)
    finally:
        buf.free()
        d.close()


def test_precision_range_checked(ctx):
    d = E.Decoder(ctx, E.encode(np.arange(10.0), E.CompressorConfig.buff(scale=10), E.ChunkOption.full()))
    with pytest.raises(E.EbiError):
        d.with_precision(13)
    with pytest.raises(E.EbiError):
        d.with_precision(-2)
    d.close()
