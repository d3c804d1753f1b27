"""Hand-derived byte vectors, layout formulas and closed forms that pin the
host encoder and the CPU oracle independently of each other.

Each chunk body below is written out bit by bit from the reference encoder
rules (file:line cited per case), not produced by this repository's encoder;
the test checks that the encoder emits exactly these bytes and that the
oracle decodes them to the expected values.
"""
import struct

import numpy as np
import pytest

import ebi_amd as E
from oracle.oracle import OracleDecoder, lib as olib


def bits_to_bytes(bits: str) -> bytes:
    bits = bits.replace(" ", "")
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


def b64(x: float) -> str:
    return format(struct.unpack(">Q", struct.pack(">d", x))[0], "064b")


def chunk_body(data: bytes, k: int = 0) -> bytes:
    phys, size, _, _ = E.file_chunks(data)
    return bytes(data[int(phys[k]):int(phys[k]) + int(size[k])])


NAN = "0111111111111000" + "0" * 48

def f64(u: int) -> float:
    return struct.unpack(">d", struct.pack(">Q", u))[0]


ONE_P1 = f64(0x3FF0000000000001)  # 1.0 + 1 ulp
ONE_P3 = f64(0x3FF0000000000003)  # 1.0 + 3 ulp

# name -> (codec, values, scale, bits[, decoded values when lossy])
CASES = {
    # compressor/gorilla.rs:125-181: 12.0 then 24.0 (xor 0x0010.., L=11, T=52,
    # sig=1), then NaN (xor with 24.0 = 0x3FC0.., L=2, T=54, sig=8), then a 0.
    "gorilla": ("gorilla", [12.0, 24.0], 1,
                b64(12.0) + "1 1 01011 000000 1" + "1 1 00010 000111 11111111" + "0"),
    # compressor/chimp.rs:99-146: case 01 with lead code 1 (8), sig 4; then
    # NaN: case 01, lead code 0, sig 10.
    "chimp": ("chimp", [12.0, 24.0], 1,
              b64(12.0) + "01 001 000100 0001" + "01 000 001010 0011111111" + "0"),
    # compressor/chimp.rs:99-146, one record per flag:
    #   1.0 -> 1.0: xor 0 -> "00";
    #   1.0 -> 1.0+1ulp: xor 1 (lz 63 -> round 24, code 7, tz 0 <= 6), lead
    #     differs from the stored 65 -> "11" + "111" + 40 low bits;
    #   +1ulp -> +3ulp: xor 2 (lz 62 -> round 24 = stored lead, tz 1) ->
    #     "10" + 40 low bits;
    #   close: NaN ^ (1.0+3ulp) = 0x4008000000000003 (lz 1 -> round 0, tz 0)
    #     -> "11" + "000" + 64 bits, then a 0.
    "chimp_flags": ("chimp", [1.0, 1.0, ONE_P1, ONE_P3], 1,
                    b64(1.0) + "00" + "11 111" + "0" * 39 + "1" + "10" + "0" * 38 + "10"
                    + "11 000" + format(0x4008000000000003, "064b") + "0"),
    # compressor/chimp_n.rs:196-258: key match on slot 0 (tz 52 > 13), flag-01
    # field 512*(128+0) + 64*1 + 4 in 18 bits; NaN matches slot 1 (tz 54).
    "chimp128": ("chimp128", [12.0, 24.0], 1,
                 b64(12.0) + format(512 * 128 + 64 + 4, "018b") + "0001"
                 + format(512 * 129 + 0 + 10, "018b") + "0011111111" + "0"),
    # compressor/chimp_n.rs:196-258 with N=128 (9-bit flag-00 field, 18-bit
    # flag-01 field, threshold 13, key = low 14 bits):
    #   1.0 -> 1.0: key hit on slot 0, xor 0 -> 9-bit index 0;
    #   -> 1.0+1ulp: key 1 unseen -> previous slot, xor 1 (lead 24, code 7)
    #     -> 5-bit 24+7 and 40 low bits;
    #   -> 1.0+3ulp: key 3 unseen, xor 2, lead 24 = stored -> "10" + 40 bits;
    #   close: NaN key 0 was last seen at index 1 (slot 1 = 1.0), xor
    #     0x4008000000000000 has tz 51 > 13 -> 18-bit 512*(128+1) + 64*0 + 13
    #     then the 13 significant bits 0x4008 >> 3 = 2049, then a 0.
    "chimp128_flags": ("chimp128", [1.0, 1.0, ONE_P1, ONE_P3], 1,
                       b64(1.0) + "000000000" + "11111" + "0" * 39 + "1" + "10" + "0" * 38 + "10"
                       + format(512 * 129 + 13, "018b") + format(2049, "013b") + "0"),
    # compressor/elf.rs:91-160 + 186-277: 1.5 has beta*=2, alpha=1, but its
    # low 48 mantissa bits are zero -> not erased (prefix 10); first ElfXor
    # value: 7-bit tz (51) then 12 bits; close: prefix 10, NaN as case 10
    # (lead code 0, center 2, one payload bit), then 0.
    "elf": ("elf", [1.5], 1, "10 0110011 001111111111" + "10 100000010 0" + "0"),
    # compressor/elf.rs:91-160 (outer) and 186-277 (ElfXor), every prefix:
    #   1.1: beta* 2, alpha 1, g = 4 + 0, 48 bits erased (delta != 0), new
    #     beta* -> "11" + "0010"; v' = 0x3FF1000000000000 goes through
    #     write_first: tz 48 in 7 bits, v' >> 49 in 15 bits;
    #   1.1: same beta* -> "0"; xor 0 -> "01";
    #   1.2: "0"; v' = 0x3FF3.., xor 0x0002.. (lead 14 -> round 12, tz 49):
    #     case 10, 9-bit ((2<<3|2)<<4)|3 = 291, (xor >> 50) in 2 bits;
    #   1.3: "0"; v' = 0x3FF4.., xor 0x0007.. (round 12, tz 48 < 49): case
    #     10, header 292, (xor >> 49) = 3 in 3 bits;
    #   1.5: not erased -> "10"; xor 0x000C.. (round 12, tz 50 >= 48):
    #     case 00, 2 + 4 bits of xor >> 48 = 12;
    #   pi: beta* 16 -> "10"; xor 0x7FF121FB54442D18 (round 0, tz 3): case
    #     11, 11-bit (3<<3|0)<<6|61 = 1597, then xor >> 4 in 60 bits;
    #   close: "10"; NaN ^ pi = 0x3FF121FB54442D18 (round 0, tz 3 >= 3):
    #     case 00, 2 + 61 bits of xor >> 3; then a 0.
    "elf_flags": ("elf", [1.1, 1.1, 1.2, 1.3, 1.5, np.pi], 1,
                  "11 0010" + format(48, "07b") + format(0x3FF1000000000000 >> 49, "015b")
                  + "0 01"
                  + "0" + format(291, "09b") + "00"
                  + "0" + format(292, "09b") + "011"
                  + "10" + "001100"
                  + "10" + format(1597, "011b") + format(0x7FF121FB54442D18 >> 4, "060b")
                  + "10" + format(0x3FF121FB54442D18 >> 3, "063b") + "0"),
    # compressor/elf.rs:91-160 over compressor/chimp.rs:99-146:
    #   1.1: "11" + "0010", Chimp first value v' = 0x3FF1000000000000 in 64;
    #   1.1: "0", xor 0 -> "00";
    #   1.2: "0", xor 0x0002.. (round 12 -> code 2, tz 49 > 6): case 01,
    #     "010", sig 3 in 6 bits, xor >> 49 = 1 in 3 bits;
    #   close: "10", NaN ^ 0x3FF3.. = 0x400B.. (round 0, tz 48): case 01,
    #     "000", sig 16, 0x400B in 16 bits, then a 0.
    "elf_on_chimp": ("elf_on_chimp", [1.1, 1.1, 1.2], 1,
                     "11 0010" + format(0x3FF1000000000000, "064b")
                     + "0 00"
                     + "0 01 010 000011 001"
                     + "10 01 000 010000" + format(0x400B, "016b") + "0"),
    # compressor/sprintz.rs:151-205 at scale 100: q = 100, 150, 125; zigzag
    # deltas 0, 100, 49 -> W = 7.
    "delta_sprintz": ("delta_sprintz", [1.0, 1.5, 1.25], 100,
                      format(100, "064b") + format(100, "032b") + format(7, "08b")
                      + "0000000 1100100 0110001", [1.0, 1.5, 1.25]),
    # compressor/sprintz.rs:161-168: all zigzag deltas 0 -> W = 0, the header
    # (q0 = 25, scale 10, W 0) and no fields.
    "delta_sprintz_w0": ("delta_sprintz", [2.5, 2.5, 2.5], 10,
                         format(25, "064b") + format(10, "032b") + format(0, "08b"), [2.5, 2.5, 2.5]),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_msb_codecs_bytes(case):
    codec, values, scale, bits, *lossy = CASES[case]
    data = E.encode(np.array(values), E.CompressorConfig.by_name(codec, scale=scale),
                    E.ChunkOption.record_count(64))
    assert chunk_body(data) == bits_to_bytes(bits)
    got = OracleDecoder(data).materialize()
    want = np.array(lossy[0] if lossy else values)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_rle_bytes():
    # compressor/run_length.rs:64-92 and format/run_length.rs
    data = E.encode(np.array([1.0, 1.0, 1.0, 2.0]), E.CompressorConfig.rle(), E.ChunkOption.record_count(64))
    want = struct.pack("<Q", 2) + bytes([3]) + struct.pack("<d", 1.0) + bytes([1]) + struct.pack("<d", 2.0)
    assert chunk_body(data) == want
    assert OracleDecoder(data).materialize().tolist() == [1.0, 1.0, 1.0, 2.0]


def test_rle_run_split_at_255():
    data = E.encode(np.full(300, 7.0), E.CompressorConfig.rle(), E.ChunkOption.record_count(1000))
    body = chunk_body(data)
    assert struct.unpack("<Q", body[:8])[0] == 2 and body[8] == 255 and body[17] == 45


def test_buff_bytes():
    # compressor/buff.rs:222-305 at scale 100 -> frac 8: fixed(1.0)=256,
    # fixed(1.0051)=257, fixed(2.5)=640; base 256, deltas 0, 1, 384, B = 9:
    # one plane ((d >> 1) ^ 0x80) and a 1-bit LSB-first tail.
    data = E.encode(np.array([1.0, 1.0051, 2.5]), E.CompressorConfig.buff(scale=100), E.ChunkOption.record_count(8))
    hdr = struct.pack("<5I", 256, 0, 3, 9, 8)
    want = hdr + bytes([0x80, 0x80, 0x40]) + bytes([0b010])
    assert chunk_body(data) == want
    assert OracleDecoder(data).materialize().tolist() == [1.0, 257 / 256, 2.5]


def test_buff_wide_bytes():
    # compressor/buff.rs:262-294 at scale 100 -> frac 8: fixed 0, 76800, 384;
    # base 0, B = 17: planes (d >> 9) and (d >> 1), each byte ^ 0x80, then
    # the 1-bit LSB-first tail after finish_write_byte.
    data = E.encode(np.array([0.0, 300.0, 1.5]), E.CompressorConfig.buff(scale=100), E.ChunkOption.record_count(8))
    want = struct.pack("<5I", 0, 0, 3, 17, 8) + bytes([0x80, 0x16, 0x80]) + bytes([0x80, 0x80, 0x40]) + bytes([0])
    assert chunk_body(data) == want
    assert OracleDecoder(data).materialize().tolist() == [0.0, 300.0, 1.5]


def test_buff_narrow_bytes():
    # compressor/buff.rs:271-276 at scale 10 -> frac 5: fixed 32, 40, 48;
    # base 32, deltas 0, 8, 16, B = 5 < 8: 5-bit LSB-first fields only, no
    # byte planes: bit 8 (delta 8) and bit 14 (delta 16) set.
    data = E.encode(np.array([1.0, 1.25, 1.5]), E.CompressorConfig.buff(scale=10), E.ChunkOption.record_count(8))
    assert chunk_body(data) == struct.pack("<5I", 32, 0, 3, 5, 5) + bytes([0x00, 0x41])
    assert OracleDecoder(data).materialize().tolist() == [1.0, 1.25, 1.5]


def test_buff_single_value_has_zero_width():
    data = E.encode(np.array([3.25, 3.25]), E.CompressorConfig.buff(scale=10), E.ChunkOption.record_count(8))
    body = chunk_body(data)
    # scale 10 -> precision 1 -> frac 5; fixed = 104; B = 0: header only
    assert body == struct.pack("<5I", 104, 0, 2, 0, 5)


def test_buff_fixed_point_closed_form():
    # into_fixed_representation_with_fractional_part_bits_length: truncation
    # of |v| * 2^frac with the sign reapplied; 0 below 2^-frac
    L = olib()
    for v in (0.0, 1.0, -1.0, 3.8, -6.7, 99.99, 5204029348.6, 1e-9):
        for frac in (0, 5, 8, 11, 15):
            want = int(np.trunc(abs(v) * 2.0 ** frac)) * (1 if v >= 0 else -1)
            if abs(v) < 2.0 ** -frac:
                want = 0
            assert L.orc_buff_fixed(v, frac) == want, (v, frac)


def test_zigzag_known_answers():
    # compressor/sprintz.rs:222-231 and chunk_reader/sprintz.rs:508-517
    L = olib()
    i64max = (1 << 63) - 1
    for a, b in [(0, 0), (-1, 1), (2, 4), (-100000, 199999), (i64max // 2, i64max - 1)]:
        assert L.orc_zigzag(a) == b
        assert L.orc_unzigzag(b) == a


def test_elf_round_up_and_sp():
    # compression_common/elf.rs:51-58, 111-130
    L = olib()
    assert L.orc_elf_round_up(3.2999999999999998, 1) == 3.3
    assert L.orc_elf_round_up(-3.2999999999999998, 1) == -3.3
    assert [L.orc_elf_get_sp(v) for v in (1.0, 9.99, 10.0, 0.5, 0.1, 0.099, 1e-10, 123456789.0, 1e12)] == \
        [0, 0, 1, -1, -1, -2, -10, 8, 12]


def test_file_layout_formula():
    # encoder.rs:420-530: 1000 values, Uncompressed, RecordCount(100)
    vals = np.arange(1000, dtype=np.float64)
    data = bytes(E.encode(vals, E.CompressorConfig.uncompressed(capacity=1000), E.ChunkOption.record_count(100)))
    assert data[:4] == b"EBI1" and data[4] == 0 and data[5] == 0 and data[6] == 1
    assert struct.unpack_from("<Q", data, 8)[0] == 100
    assert struct.unpack_from("<3H", data, 24) == (0, 1, 14)
    fo = struct.unpack_from("<Q", data, 16)[0]
    n_records, n_chunks = struct.unpack_from("<QQ", data, fo)
    assert n_records == 1000 and n_chunks == 10
    phys, logi = struct.unpack_from("<QQ", data, fo + 16 + 5 * 16)
    assert data[phys:phys + 8] == struct.pack("<d", vals[logi])
    assert len(data) == fo + 16 + 16 * n_chunks + 1 + 8 + 164
    info = E.file_info(data)
    assert (info.number_of_records, info.number_of_chunks, info.scheme) == (1000, 10, 0)


def test_byte_size_chunking():
    # encoder.rs:347-362: ByteSizeBestEffort(n) reads ceil(n / 8 * 1.5) records per chunk
    data = E.encode(np.arange(2000, dtype=np.float64), E.CompressorConfig.gorilla(),
                    E.ChunkOption.byte_size_best_effort(5274))
    _, _, logi, nrec = E.file_chunks(data)
    assert nrec.tolist()[0] == 989 and logi.tolist() == [0, 989, 1978]


def brute_force_gorilla(body: bytes, n: int):
    """Bit-by-bit Gorilla interpreter written from the format description."""
    bits = "".join(format(b, "08b") for b in body)
    p = 0

    def rd(k):
        nonlocal p
        v = int(bits[p:p + k], 2) if k else 0
        p += k
        return v

    out = [rd(64)]
    L = T = 0
    while len(out) < n:
        if rd(1) == 0:
            out.append(out[-1])
            continue
        if rd(1) == 1:
            L = rd(5)
            T = 64 - L - (rd(6) + 1)
        out.append(out[-1] ^ (rd(64 - L - T) << T))
    return out


def test_gorilla_brute_force_agrees():
    rng = np.random.default_rng(5)
    vals = np.round(100 + np.cumsum(rng.normal(0, 1, 300)), 1)
    data = E.encode(vals, E.CompressorConfig.gorilla(), E.ChunkOption.record_count(1000))
    bf = brute_force_gorilla(chunk_body(data), len(vals))
    assert bf == vals.view(np.uint64).tolist()
    assert OracleDecoder(data).materialize().view(np.uint64).tolist() == bf
