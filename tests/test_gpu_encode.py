"""GPU encoders (BUFF, DeltaSprintz; ChunkWriter::write, encoder.rs:317-405):
the file the kernels write is byte-identical to the host encoder's, itself
the restatement of compressor/buff.rs:222-305 and compressor/sprintz.rs:151-
205 pinned by the golden vectors; a resident column re-encoded in HBM
answers queries as the oracle does on the host-encoded file."""
import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G
from oracle.oracle import OracleDecoder

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(1)
    yield "walk2", G.random_walk(50_003, seed=3), 100
    yield "unif3", G.uniform(40_000, -1000, 1000, 3, seed=4), 1000
    yield "sin4", G.sinusoid(30_000, seed=5), 10000
    yield "const", np.full(10_000, 7.25), 100
    yield "big", np.round(rng.uniform(-1e12, 1e12, 5000), 0), 1
    yield "tiny", np.array([1.5]), 10
    yield "step", G.step_series(20_000, seed=6), 100


@pytest.mark.parametrize("codec", ["buff", "delta_sprintz"])
@pytest.mark.parametrize("chunk", [E.ChunkOption.record_count(4096), E.ChunkOption.record_count(1003),
                                   E.ChunkOption.record_count(5), E.ChunkOption.full(),
                                   E.ChunkOption.byte_size_best_effort(65536)])
def test_gpu_encode_bytes_equal_host(ctx, codec, chunk):
    for name, vals, scale in _cases():
        cfg = E.CompressorConfig.by_name(codec, scale=scale)
        want = E.encode(vals, cfg, chunk)
        buf = ctx.upload(vals)
        try:
            got = E.encode_device(ctx, buf, vals.size, cfg, chunk)
        finally:
            buf.free()
        assert got.size == want.size and np.array_equal(got, want), (name, codec, got.size, want.size,
                                                                     np.flatnonzero(got[:min(got.size, want.size)] != want[:min(got.size, want.size)])[:5])


def test_gpu_encode_unsupported_scheme(ctx):
    buf = ctx.upload(np.arange(10.0))
    with pytest.raises(E.EbiError) as ei:
        E.encode_device(ctx, buf, 10, E.CompressorConfig.gorilla(), E.ChunkOption.full())
    assert ei.value.status == E.UNSUPPORTED
    buf.free()


def test_reencode_resident_column(ctx):
    """Decode a Gorilla column into HBM, re-encode it there as BUFF and as
    DeltaSprintz, and query the new resident columns."""
    n = 100_000
    vals = G.random_walk(n, seed=8)
    src = E.Decoder(ctx, E.encode(vals, E.CompressorConfig.gorilla(), E.ChunkOption.record_count(8192)))
    buf = ctx.alloc(8 * n)
    src.decode_device(buf)
    p = E.Predicate.range(E.RangeValue.inclusive(99.0), E.RangeValue.exclusive(101.0))
    for cfg in (E.CompressorConfig.buff(scale=100), E.CompressorConfig.delta_sprintz(scale=100)):
        d = E.Decoder.encode_on_device(ctx, buf, n, cfg, E.ChunkOption.record_count(16384))
        o = OracleDecoder(E.encode(vals, cfg, E.ChunkOption.record_count(16384)))
        assert np.array_equal(d.materialize().view(np.uint64), o.materialize().view(np.uint64))
        assert np.array_equal(d.filter(p), o.filter(p))
        assert d.sum() == o.sum() and d.max() == o.max()
        d.close()
    buf.free()
    src.close()
