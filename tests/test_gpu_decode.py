"""GPU decode parity: every codec's HIP decoder against the CPU oracle,
bit-exact on the decoded f64 patterns (ebi/tests/round_trip.rs:249-336 is the
reference's own round-trip strategy)."""
import numpy as np
import pytest

import ebi_amd as E
from ebi_amd import datagen as G
from oracle.oracle import OracleDecoder

pytestmark = pytest.mark.gpu

CODECS = G.ALL_CODECS


def _gpu_decode(ctx, data, per_chunk=False):
    dec = E.Decoder(ctx, data)
    buf = ctx.alloc(max(8 * dec.number_of_records, 8))
    try:
        (dec.decode_per_chunk if per_chunk else dec.decode_device)(buf)
        return buf.to_host(np.float64, dec.number_of_records)
    finally:
        buf.free()
        dec.close()


def _check(ctx, values, codec, chunk, scale=100, per_chunk=False):
    data = E.encode(values, E.CompressorConfig.by_name(codec, scale=scale), chunk)
    want = OracleDecoder(data).materialize()
    got = _gpu_decode(ctx, data, per_chunk)
    assert got.shape == want.shape
    bad = np.flatnonzero(got.view(np.uint64) != want.view(np.uint64))
    assert bad.size == 0, f"{codec} {chunk}: {bad.size} mismatches, first at {bad[:5]}: {got[bad[:3]]} vs {want[bad[:3]]}"


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("gen", list(G.GENERATORS))
def test_decode_generators(ctx, codec, gen):
    fn, dec = G.GENERATORS[gen]
    vals = fn(20000)
    _check(ctx, vals, codec, E.ChunkOption.record_count(4096), scale=10 ** dec)


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("chunk", [E.ChunkOption.record_count(5), E.ChunkOption.record_count(1003),
                                   E.ChunkOption.byte_size_best_effort(5274), E.ChunkOption.full()])
def test_decode_ragged_chunks(ctx, codec, chunk):
    vals = G.random_walk(10003, seed=7)
    _check(ctx, vals, codec, chunk)


@pytest.mark.parametrize("blocks_per_cu", [1, 2, 4, 8])
def test_decode_grid_shapes(blocks_per_cu, monkeypatch):
    """Every persistent grid size gives the same records, with small chunks so
    that the segment kernel's jobs spread over workgroups that share CUs (a
    spilling build of that kernel decoded Elf wrongly in every workgroup but
    the first on a CU)."""
    monkeypatch.setenv("EBI_HIP_BLOCKS_PER_CU", str(blocks_per_cu))  # read when a context is created
    c = E.Context(0)
    try:
        vals = G.random_walk(40000, seed=7)
        for codec in ("elf", "gorilla", "chimp", "chimp128"):
            for chunk in (3, 5, 64):
                _check(c, vals, codec, E.ChunkOption.record_count(chunk))
    finally:
        c.close()


@pytest.mark.parametrize("codec", [c for c in CODECS if c not in ("buff", "delta_sprintz")])
def test_decode_xor_adversarial(ctx, codec):
    _check(ctx, G.xor_values(), codec, E.ChunkOption.record_count(66))
    _check(ctx, G.xor_values2(), codec, E.ChunkOption.record_count(3))


@pytest.mark.parametrize("codec", CODECS)
def test_decode_per_chunk_comparator(ctx, codec):
    vals = G.random_walk(6000, seed=11)
    _check(ctx, vals, codec, E.ChunkOption.record_count(1000), per_chunk=True)
    # several segment jobs per chunk (XOR codecs), ragged last chunk
    _check(ctx, G.random_walk(50001, seed=12), codec, E.ChunkOption.record_count(8192), per_chunk=True)


@pytest.mark.parametrize("codec", CODECS)
def test_decode_single_record_and_constant(ctx, codec):
    _check(ctx, np.array([42.5]), codec, E.ChunkOption.record_count(8))
    _check(ctx, np.full(3000, 100.19), codec, E.ChunkOption.record_count(700))


def test_decode_batch_mixed_codecs(ctx):
    """One persistent launch over columns of every codec, uneven chunk sizes."""
    rng = np.random.default_rng(3)
    decs, outs, wants = [], [], []
    for i, codec in enumerate(CODECS * 2):
        n = int(rng.integers(1000, 30000))
        chunk = int(rng.integers(300, 9000))
        vals = G.random_walk(n, seed=100 + i)
        data = E.encode(vals, E.CompressorConfig.by_name(codec, scale=100), E.ChunkOption.record_count(chunk))
        decs.append(E.Decoder(ctx, data))
        outs.append(ctx.alloc(8 * n))
        wants.append(OracleDecoder(data).materialize())
    E.decode_batch(ctx, decs, outs)
    for d, o, w in zip(decs, outs, wants):
        got = o.to_host(np.float64, d.number_of_records)
        assert np.array_equal(got.view(np.uint64), w.view(np.uint64)), E.SCHEME_NAMES[d.scheme]
        o.free()
        d.close()


def test_prepared_batch_graph_replays_match_eager(ctx):
    """The tuned, graph-captured batch replays bit-exact, several times, with
    eager decodes of the same columns (and syncs) between the replays."""
    specs = [("walk2", "gorilla", 70001, 8192), ("unif3", "chimp128", 50000, 4096), ("step2", "rle", 90000, 16384),
             ("sin4", "buff", 60000, 65536), ("walk2", "delta_sprintz", 40000, 2048), ("unif3", "elf", 30000, 1000),
             ("sin4", "uncompressed", 20000, 333), ("walk2", "chimp", 45000, 5000)]
    decs, wants, outs = [], [], []
    for gen, codec, n, chunk in specs:
        fn, dec = G.GENERATORS[gen]
        f = E.encode(fn(n), E.CompressorConfig.by_name(codec, scale=10 ** dec), E.ChunkOption.record_count(chunk))
        decs.append(E.Decoder(ctx, f))
        wants.append(OracleDecoder(f).materialize())
        outs.append(ctx.alloc(8 * n))
    b = E.Batch(ctx, decs, outs, autotune=True)
    cfg = b.config
    assert cfg["blocks_per_cu"] in (4, 8, 16) and cfg["tuned_ms"] > 0
    eager_out = ctx.alloc(8 * max(s[2] for s in specs))
    for rep in range(3):
        for o in outs:  # cleared between replays: every replay must write every record
            o.zero()
        b.run(eager=(rep == 1))
        for d, w, o in zip(decs, wants, outs):
            got = o.to_host(count=d.number_of_records)
            assert np.array_equal(got.view(np.uint64), w.view(np.uint64))
        decs[rep].decode_device(eager_out)  # the program's own eager work between replays
    b.close()
    eager_out.free()
    for o in outs:
        o.free()
    for d in decs:
        d.close()


@pytest.mark.parametrize("mixed", [0, 1])
def test_prepared_batch_mixed_and_two_kernel(mixed, monkeypatch):
    """Untuned prepared batches of tile and segment codecs: one k_decode_mix
    launch over both job lists, or k_decode then k_decode_seg, at small and
    large grids (small chunks, so that one wave takes jobs of both kinds and
    of several columns); bit-exact against the oracle in both forms."""
    monkeypatch.setenv("EBI_HIP_BATCH_MIXED", str(mixed))  # read when a context is created
    for bpc in (1, 16):
        monkeypatch.setenv("EBI_HIP_BLOCKS_PER_CU", str(bpc))
        c = E.Context(0)
        try:
            specs = [("walk2", "elf", 30011, 97), ("unif3", "chimp128", 20000, 300), ("step2", "rle", 25000, 640),
                     ("sin4", "buff", 21000, 1500), ("walk2", "delta_sprintz", 17000, 700),
                     ("walk2", "gorilla", 26000, 129), ("sin4", "uncompressed", 9000, 333), ("unif3", "chimp", 15000, 64)]
            decs, wants, outs = [], [], []
            for gen, codec, n, chunk in specs:
                fn, dec = G.GENERATORS[gen]
                f = E.encode(fn(n), E.CompressorConfig.by_name(codec, scale=10 ** dec),
                             E.ChunkOption.record_count(chunk))
                decs.append(E.Decoder(c, f))
                wants.append(OracleDecoder(f).materialize())
                outs.append(c.alloc(8 * n))
            b = E.Batch(c, decs, outs, autotune=False)
            assert b.config["mixed"] == bool(mixed) and b.config["blocks_per_cu"] == bpc
            for eager in (True, False, False):
                for o in outs:
                    o.zero()
                b.run(eager=eager)
                if eager and mixed:  # one kernel: the segment phase is empty
                    assert c.last_phase_ms(0) > 0 and c.last_phase_ms(2) < 0.05
                for d, w, o in zip(decs, wants, outs):
                    got = o.to_host(count=d.number_of_records)
                    assert np.array_equal(got.view(np.uint64), w.view(np.uint64)), (E.SCHEME_NAMES[d.scheme], bpc)
            b.close()
            for o in outs:
                o.free()
            for d in decs:
                d.close()
        finally:
            c.close()


@pytest.mark.parametrize("seed", range(6))
def test_decode_random_shapes(ctx, seed):
    """Random record counts, chunk sizes and generators for the XOR codecs
    (segment kernel, Chimp128 resolve), one column at a time and all together
    in one batched launch, bit-exact against the oracle."""
    rng = np.random.default_rng(1000 + seed)
    files, wants = [], []
    for codec in ("gorilla", "chimp", "chimp128", "elf"):
        gen = list(G.GENERATORS)[int(rng.integers(len(G.GENERATORS)))]
        fn, dec = G.GENERATORS[gen]
        n = int(rng.integers(1, 120_000))
        chunk = int(rng.choice([1, 2, 7, 63, 64, 65, 129, 1000, 4095, 4097, 20000]))
        vals = fn(n, seed=int(rng.integers(1 << 30)))
        f = E.encode(vals, E.CompressorConfig.by_name(codec, scale=10 ** dec), E.ChunkOption.record_count(chunk))
        want = OracleDecoder(f).materialize()
        got = _gpu_decode(ctx, f)
        bad = np.flatnonzero(got.view(np.uint64) != want.view(np.uint64))
        assert bad.size == 0, f"{codec} {gen} n={n} chunk={chunk}: {bad.size} mismatches from {bad[:5]}"
        files.append(f)
        wants.append(want)
    decs = [E.Decoder(ctx, f) for f in files]
    outs = [ctx.alloc(max(8 * d.number_of_records, 8)) for d in decs]
    try:
        E.decode_batch(ctx, decs, outs)
        for d, o, w in zip(decs, outs, wants):
            assert np.array_equal(o.to_host(np.float64, d.number_of_records).view(np.uint64), w.view(np.uint64))
    finally:
        for o in outs:
            o.free()
        for d in decs:
            d.close()


@pytest.mark.parametrize("codec", ["gorilla", "chimp", "chimp128", "elf"])
@pytest.mark.parametrize("how", ["zeroed", "tail_zeroed"])
def test_damaged_chunk(ctx, codec, how):
    """A damaged XOR chunk (its body, or the second half of it, zeroed): the
    oracle reports an error for the whole column, and so does every decode
    and query on the GPU that selects that chunk -- UnexpectedEndOfChunk
    where the oracle says so; the intact chunks still decode and query
    bit-exactly (chunk_id 0)."""
    from oracle.oracle import OracleError
    vals = G.random_walk(5000, seed=3)
    data = bytearray(E.encode(vals, E.CompressorConfig.by_name(codec, scale=100), E.ChunkOption.record_count(1000)))
    off, size, _, _ = E.file_chunks(bytes(data))
    o, sz = int(off[2]), int(size[2])
    lo = o if how == "zeroed" else o + sz // 2
    data[lo:o + sz] = bytes(o + sz - lo)
    data = bytes(data)
    orc = OracleDecoder(data)
    with pytest.raises(OracleError) as oe:
        orc.materialize()
    d = E.Decoder(ctx, data)
    buf = ctx.alloc(8 * d.number_of_records)
    try:
        for run in (lambda: d.decode_device(buf), lambda: d.decode_per_chunk(buf), d.materialize, d.sum, d.max,
                    lambda: d.filter(E.Predicate.eq(100.0))):
            with pytest.raises(E.EbiError) as ge:
                run()
            if oe.value.status == 1:
                assert ge.value.status == E.UNEXPECTED_END
        p = E.Predicate.range(E.RangeValue.inclusive(99.0), E.RangeValue.inclusive(101.0))
        assert np.array_equal(d.materialize(None, 0).view(np.uint64), orc.materialize(None, 0).view(np.uint64))
        assert np.array_equal(d.filter(p, None, 0), orc.filter(p, None, 0))
        assert np.float64(d.max(None, 0)).view(np.uint64) == np.float64(orc.max(None, 0)).view(np.uint64)
    finally:
        buf.free()
        d.close()


@pytest.mark.parametrize("codec", CODECS)
def test_prepared_batch_shapes(ctx, codec):
    """Every launch shape the batch tuner may pick decodes bit-exactly: tiles
    per tile-codec job 1 / 2 / 4 / 8 (RLE: output steps per job), grids of 4
    and 16 blocks per CU, one mixed launch and the two kernels -- on two
    columns of the codec (random walk, step series: RLE runs up to 255) with
    chunk sizes that leave ragged last tiles and a short last chunk."""
    decs, outs, wants = [], [], []
    for i, (gen, chunk) in enumerate((("walk2", 5000), ("step2", 70000))):
        fn, dec = G.GENERATORS[gen]
        vals = fn(150_001, seed=40 + i)
        data = E.encode(vals, E.CompressorConfig.by_name(codec, scale=10 ** dec), E.ChunkOption.record_count(chunk))
        decs.append(E.Decoder(ctx, data))
        outs.append(ctx.alloc(8 * decs[-1].number_of_records))
        wants.append(OracleDecoder(data).materialize())
    b = E.Batch(ctx, decs, outs, autotune=False)
    try:
        for tpj, bpc, mixed in ((1, 4, False), (2, 16, True), (8, 4, True), (8, 16, False), (4, 8, False)):
            b.set_config(bpc, mixed, tpj)
            for o in outs:
                o.zero()
            b.run()
            for d, o, w in zip(decs, outs, wants):
                got = o.to_host(np.float64, d.number_of_records)
                assert np.array_equal(got.view(np.uint64), w.view(np.uint64)), (codec, tpj, bpc, mixed)
        with pytest.raises(E.EbiError):
            b.set_config(8, False, 3)  # not a tile-job shape
    finally:
        b.close()
        for o in outs:
            o.free()
        for d in decs:
            d.close()


def test_prepared_batch_tuning_cache(ctx):
    """A second batch of the same columns reuses the first one's tuned shape
    from the context's cache (no timing) and decodes the same records."""
    rng = np.random.default_rng(9)
    decs, outs = [], []
    for codec in ("buff", "gorilla", "rle"):
        vals = G.random_walk(200_000, seed=int(rng.integers(1000)))
        data = E.encode(vals, E.CompressorConfig.by_name(codec, scale=100), E.ChunkOption.record_count(20000))
        decs.append(E.Decoder(ctx, data))
        outs.append(ctx.alloc(8 * 200_000))
    b1 = E.Batch(ctx, decs, outs, autotune=True)
    c1 = b1.config
    b1.run()
    first = [o.to_host(np.float64, 200_000) for o in outs]
    b1.close()
    b2 = E.Batch(ctx, decs, outs, autotune=True)
    c2 = b2.config
    for o in outs:
        o.zero()
    b2.run()
    assert c2["cached"]
    assert (c1["blocks_per_cu"], c1["mixed"], c1["tiles_per_job"]) == (c2["blocks_per_cu"], c2["mixed"],
                                                                      c2["tiles_per_job"])
    for o, f in zip(outs, first):
        assert np.array_equal(o.to_host(np.float64, 200_000).view(np.uint64), f.view(np.uint64))
    b2.close()
    for o in outs:
        o.free()
    for d in decs:
        d.close()


def test_decode_hand_derived_vectors(ctx):
    """The hand-derived chunk bodies (one record per Chimp/Chimp128/Elf flag,
    BUFF widths 5/9/17, Sprintz W = 0) decode on the GPU to the same f64
    patterns as on the oracle."""
    from tests.test_golden_vectors import CASES
    for name, (codec, values, scale, _bits, *lossy) in sorted(CASES.items()):
        data = E.encode(np.array(values), E.CompressorConfig.by_name(codec, scale=scale),
                        E.ChunkOption.record_count(64))
        want = np.array(lossy[0] if lossy else values)
        got = _gpu_decode(ctx, data)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), name
    for vals, scale in (([0.0, 300.0, 1.5], 100), ([1.0, 1.25, 1.5], 10), ([1.0, 1.0051, 2.5], 100)):
        data = E.encode(np.array(vals), E.CompressorConfig.buff(scale=scale), E.ChunkOption.record_count(8))
        want = OracleDecoder(data).materialize()
        assert np.array_equal(_gpu_decode(ctx, data).view(np.uint64), want.view(np.uint64)), vals


@pytest.mark.parametrize("codec,scale", [("rle", 100), ("buff", 10000), ("gorilla", 100), ("uncompressed", 1)])
def test_upload_spans_several_windows(ctx, codec, scale):
    """Columns whose device image (RLE: 12 Mi fields of 9 bytes; BUFF ~3 B
    per record; Gorilla; Uncompressed 96 MiB) is repacked and copied over
    several 32 MiB windows of the pinned upload buffer: windows cut through
    RLE run and value arrays, BUFF planes and XOR streams."""
    vals = G.uniform(12 << 20, -1000.0, 1000.0, 4 if codec == "buff" else 2, seed=9)
    data = E.encode(vals, E.CompressorConfig.by_name(codec, scale=scale), E.ChunkOption.record_count(1 << 20))
    want = OracleDecoder(data).materialize()
    got = _gpu_decode(ctx, data)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
