"""GPU parity of the decompress-then-matmul path.

* Decoder.matmul_cuda on the reference's exact known answers
  (ebi/tests/ml_query.rs:455-576, on the codecs the reference runs them on).
* Random batched shapes (edges off the 128 x 128 tile grid, k off the
  16-wide K slab) against a float64 numpy product of the oracle-decoded data:
  decoded operands are bit-exact (the decode parity tests pin them), the
  product is within 1e-12 relative in the Frobenius norm.
* gemm_decoded over two row-chunked columns (one matrix row per chunk,
  Chimp128 / Elf / BUFF at scale 1000, uniform[-10, 10] at 3 decimals).
"""
import numpy as np
import pytest

import ebi_amd as E
from oracle.oracle import OracleDecoder
from tests.helpers import chunk_of, known_answers

pytestmark = pytest.mark.gpu
KA = known_answers()
TOL = 1e-12  # ||C - C_ref||_F / ||C_ref||_F


def rel_fro(c, ref):
    d = np.linalg.norm(np.asarray(c) - np.asarray(ref))
    r = np.linalg.norm(ref)
    return d / r if r else d


def uniform3(n, seed):
    return E.synth(1, n, seed, 3, -10.0, 10.0)


@pytest.mark.parametrize("codec", KA["matmul_codecs"])
@pytest.mark.parametrize("case", range(len(KA["matmul"])))
def test_matmul_known_answers(ctx, codec, case):
    t = KA["matmul"][case]
    data = np.array(t["data"], dtype=np.float64)  # batch x m x k
    target = np.array(t["target"], dtype=np.float64)
    expected = np.array(t["expected"], dtype=np.float64).ravel()
    b, m, k = data.shape
    n = target.shape[1]
    f = E.encode(data.ravel(), E.CompressorConfig.by_name(codec), chunk_of(t["chunk"]))
    d = E.Decoder(ctx, f)
    got = d.matmul_cuda(target, (k, n), (m, k))
    d.close()
    assert np.array_equal(got, expected), t["name"]


@pytest.mark.parametrize("codec,scale", [("uncompressed", 1), ("chimp128", 1), ("elf", 1), ("buff", 1000),
                                         ("gorilla", 1)])
@pytest.mark.parametrize("batch,m,k,n", [(3, 130, 37, 200), (1, 256, 256, 256), (5, 17, 129, 9), (2, 1, 300, 1)])
def test_matmul_random_shapes(ctx, codec, scale, batch, m, k, n):
    vals = uniform3(batch * m * k, seed=m * 1000 + k)
    target = uniform3(k * n, seed=n + 7).reshape(k, n)
    chunk = E.ChunkOption.record_count(max(m * k, 1000))
    f = E.encode(vals, E.CompressorConfig.by_name(codec, scale=scale), chunk)
    decoded = OracleDecoder(f).materialize().reshape(batch, m, k)
    ref = np.matmul(decoded, target).ravel()
    d = E.Decoder(ctx, f)
    got = d.matmul_cuda(target, (k, n), (m, k))
    d.close()
    assert got.shape == ref.shape
    assert rel_fro(got, ref) <= TOL


def test_matmul_preconditions(ctx):
    f = E.encode(np.arange(12, dtype=np.float64), E.CompressorConfig.by_name("uncompressed"),
                 E.ChunkOption.record_count(12))
    d = E.Decoder(ctx, f)
    with pytest.raises(E.EbiError) as ei:  # data_cols != target_rows
        d.matmul_cuda(np.zeros((4, 2)), (4, 2), (2, 3))
    assert ei.value.status == E.PRECONDITION
    with pytest.raises(E.EbiError) as ei:  # 12 records are not a multiple of 5 x 1
        d.matmul_cuda(np.zeros((1, 2)), (1, 2), (5, 1))
    assert ei.value.status == E.PRECONDITION
    d.close()


@pytest.mark.parametrize("codec,scale", [("chimp128", 1), ("elf", 1), ("buff", 1000)])
@pytest.mark.parametrize("m,k,n", [(384, 512, 256), (1000, 999, 1001)])
def test_gemm_decoded_row_chunked(ctx, codec, scale, m, k, n):
    a_vals = uniform3(m * k, seed=11)
    b_vals = uniform3(k * n, seed=12)
    cfg = E.CompressorConfig.by_name(codec, scale=scale)
    fa = E.encode(a_vals, cfg, E.ChunkOption.record_count(k))  # one row of A per chunk
    fb = E.encode(b_vals, cfg, E.ChunkOption.record_count(n))  # one row of B per chunk
    a_dec = OracleDecoder(fa).materialize().reshape(m, k)
    b_dec = OracleDecoder(fb).materialize().reshape(k, n)
    ref = a_dec @ b_dec
    da, db = E.Decoder(ctx, fa), E.Decoder(ctx, fb)
    c = E.gemm_decoded(ctx, da, db, m, n, k)
    assert rel_fro(c, ref) <= TOL
    # device-resident result equals the host-returned one
    buf = ctx.alloc(8 * m * n)
    E.gemm_decoded(ctx, da, db, m, n, k, out=buf)
    assert np.array_equal(buf.to_host(count=m * n).reshape(m, n), c)
    buf.free()
    da.close()
    db.close()


def test_gemm_decoded_shape_check(ctx):
    f = E.encode(np.ones(6), E.CompressorConfig.by_name("uncompressed"), E.ChunkOption.record_count(3))
    d = E.Decoder(ctx, f)
    with pytest.raises(E.EbiError) as ei:
        E.gemm_decoded(ctx, d, d, 2, 2, 2)  # 6 records are not 2 x 2
    assert ei.value.status == E.PRECONDITION
    d.close()


def test_dgemm_exact_small_integers(ctx):
    # integer operands with exact products: every entry must be exact
    rng = np.random.default_rng(5)
    m, n, k = 200, 300, 150
    a = rng.integers(-8, 9, size=(m, k)).astype(np.float64)
    b = rng.integers(-8, 9, size=(k, n)).astype(np.float64)
    fa = E.encode(a.ravel(), E.CompressorConfig.by_name("uncompressed"), E.ChunkOption.record_count(4096))
    fb = E.encode(b.ravel(), E.CompressorConfig.by_name("uncompressed"), E.ChunkOption.record_count(4096))
    da, db = E.Decoder(ctx, fa), E.Decoder(ctx, fb)
    c = E.gemm_decoded(ctx, da, db, m, n, k)
    assert np.array_equal(c, a @ b)
    da.close()
    db.close()
