"""The oracle's BUFF filters (index-list planes, 32-record first plane)
against a model read off the reference's plane loop, evaluated by a pure-Python
interpreter of the chunk bytes.

Reading buff_simd256_cmp_filter (decoder/chunk_reader/buff/filter_cmp.rs:
298-451) by hand: after the early exits (fixed(pred) < base; B = 0; B below
the width of dp = fixed(pred) - base), comparing the planes most significant
first and then the LSB tail is the numeric comparison of the record's delta
with dp.  buff_simd256_equal_filter (filter_eq.rs:73-201) compares each plane
with the truncated byte of dp and the tail with dp's low bits, so it selects
delta == dp mod 2^B, except with no bitmask and B < 8, where the tail is
compared with `dp as u8` unmasked."""
import struct

import numpy as np
import pytest

import ebi_amd as E
from oracle.oracle import OracleDecoder, lib as orc


def _chunk_deltas(body: bytes):
    base = struct.unpack_from("<q", body, 0)[0]
    n, B, frac = struct.unpack_from("<3I", body, 8)
    planes, r = (B // 8, B % 8) if B >= 8 else (0, B)
    d = [0] * n
    for pl in range(planes):
        off = 20 + pl * n
        d = [(x << 8) | (body[off + i] ^ 0x80) for i, x in enumerate(d)]
    if r:
        tail = int.from_bytes(body[20 + planes * n:], "little")
        d = [(x << r) | ((tail >> (i * r)) & ((1 << r) - 1)) for i, x in enumerate(d)]
    return base, B, frac, d


def _model_cmp(base, B, d, fp, greater, inclusive, mask):
    n = len(d)
    if fp < base:
        return [greater] * n
    if B == 0:
        v = (base > fp or (inclusive and base == fp)) if greater else (base < fp or (inclusive and base == fp))
        return [v] * n
    dp = fp - base
    if B < dp.bit_length():
        return [not greater] * n
    out = []
    for i, x in enumerate(d):
        if mask is not None and not mask[i]:
            out.append(False)
        elif greater:
            out.append(x > dp or (inclusive and x == dp))
        else:
            out.append(x < dp or (inclusive and x == dp))
    return out


def _model_eq(base, B, d, fp, mask):
    n = len(d)
    if fp < base:
        return [False] * n
    if B == 0:
        return [fp == base] * n
    dp = fp - base
    if mask is None and B < 8:
        return [x == (dp & 0xFF) for x in d]
    return [(mask is None or bool(mask[i])) and x == dp % (1 << B) for i, x in enumerate(d)]


def _model(body, pred, mask):
    base, B, frac, d = _chunk_deltas(body)
    fx = lambda v: orc().orc_buff_fixed(float(v), frac)  # noqa: E731
    if pred.kind in (0, 1):
        r = _model_eq(base, B, d, fx(pred.value), mask)
        if pred.kind == 1:
            r = [(mask is None or bool(mask[i])) and not x for i, x in enumerate(r)]
        return r
    left = right = None
    if pred.start.kind:
        left = _model_cmp(base, B, d, fx(pred.start.value), True, pred.start.kind == 1, mask)
    if pred.end.kind:
        right = _model_cmp(base, B, d, fx(pred.end.value), False, pred.end.kind == 1, mask)
    out = []
    for i in range(len(d)):
        if left is None and right is None:
            out.append(mask is None or bool(mask[i]))
        else:
            out.append((left is None or left[i]) and (right is None or right[i]))
    return out


# (scale, values): widths B = 3, 8, 9, 5, 17 and 26, record counts that leave
# a remainder after the 32-record steps
def _columns():
    rng = np.random.default_rng(11)
    yield 10, 1.0 + rng.integers(0, 8, 1000) / 32           # frac 5, B = 3
    yield 100, 1.0 + rng.integers(0, 256, 999) / 256        # frac 8, B = 8
    yield 100, 1.0 + rng.integers(0, 512, 1031) / 256       # B = 9
    yield 1000, rng.integers(0, 32, 777) / 2048             # frac 11, B = 5
    yield 100, rng.integers(0, 1 << 17, 513) / 256           # B = 17
    yield 10000, rng.integers(-(1 << 25), 1 << 25, 4097) / 32768  # frac 15, B = 26


def _preds(vals, rng):
    picks = [float(x) for x in rng.choice(vals, 4)] + [float(vals.min()) - 1, float(vals.max()) + 1, 1e9]
    I, X, N = E.RangeValue.inclusive, E.RangeValue.exclusive, E.RangeValue.none
    out = []
    for v in picks:
        out += [E.Predicate.eq(v), E.Predicate.ne(v), E.Predicate.range(I(v), N()), E.Predicate.range(X(v), N()),
                E.Predicate.range(N(), I(v)), E.Predicate.range(N(), X(v))]
    a, b = sorted(picks[:2])
    out += [E.Predicate.range(I(a), I(b)), E.Predicate.range(X(a), X(b)), E.Predicate.range(N(), N())]
    return out


@pytest.mark.parametrize("col", range(6))
@pytest.mark.parametrize("masked", [False, True])
def test_buff_filter_against_plane_model(col, masked):
    scale, vals = list(_columns())[col]
    data = E.encode(vals, E.CompressorConfig.buff(scale=scale), E.ChunkOption.record_count(1 << 20))
    phys, size, _, _ = E.file_chunks(data)
    body = bytes(data[int(phys[0]):int(phys[0]) + int(size[0])])
    rng = np.random.default_rng(col)
    mask = (rng.random(len(vals)) < 0.6) if masked else None
    od = OracleDecoder(data)
    ids = np.flatnonzero(mask) if masked else None
    for p in _preds(vals, rng):
        want = np.flatnonzero(np.array(_model(body, p, mask), dtype=bool))
        got = od.filter(p, bitmask=ids)
        assert np.array_equal(got, want), (col, masked, p)


def test_buff_widths_covered():
    bs = []
    for scale, vals in _columns():
        data = E.encode(vals, E.CompressorConfig.buff(scale=scale), E.ChunkOption.record_count(1 << 20))
        phys, _, _, _ = E.file_chunks(data)
        bs.append(struct.unpack_from("<I", bytes(data[int(phys[0]) + 12:int(phys[0]) + 16]))[0])
    assert bs == [3, 8, 9, 5, 17, 26], bs
