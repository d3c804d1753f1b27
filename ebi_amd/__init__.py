"""ebi_amd -- MI355X (gfx950) chunked decode and in-situ query path for EBI1 files.

Python mirror of the reference's public API (lemolatoon/ebi, ebi/src/api/
decoder.rs and api/encoder.rs) over the C ABI of ``libebi_hip.so``
(include/ebi_hip.h).  Every query runs in HIP kernels on the GPU; there is no
CPU fallback: if the library or a GPU is missing, the calls raise.

    cfg = CompressorConfig.buff(scale=100)
    file = encode(values, cfg, ChunkOption.record_count(131072))
    with Context(0) as ctx:
        dec = Decoder(ctx, file)
        ids = dec.filter(Predicate.range(RangeValue.inclusive(99.5),
                                         RangeValue.inclusive(100.5)))
        total = dec.sum(bitmask=ids)
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass
from pathlib import Path
from typing import Iterable, Optional

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libebi_hip.so"
_lib = None

OK, UNEXPECTED_END, CONVERSION, PRECONDITION, HIP, OOM, REJECTED, INVALID_ARGUMENT, UNSUPPORTED, PENDING = range(10)
OP_MATERIALIZE, OP_FILTER, OP_FILTER_MATERIALIZE, OP_SUM, OP_MIN, OP_MAX, OP_FILTER_SUM, OP_DECODE = range(8)
_STATUS = {
    1: "UnexpectedEndOfChunk",
    2: "ConversionError",
    3: "PreconditionsNotMet",
    4: "HipError",
    5: "OutOfMemory",
    6: "Rejected",
    7: "InvalidArgument",
    8: "Unsupported",
    9: "Pending",
}

SCHEMES = {
    "uncompressed": 0,
    "rle": 1,
    "buff": 2,
    "gorilla": 3,
    "chimp": 4,
    "chimp128": 5,
    "elf_on_chimp": 6,
    "elf": 7,
    "delta_sprintz": 8,
}
SCHEME_NAMES = {v: k for k, v in SCHEMES.items()}


class EbiError(RuntimeError):
    """A non-zero ebi_status; mirrors decoder::DecoderError (decoder/error.rs:6-21)."""

    def __init__(self, status: int, message: str = ""):
        self.status = status
        super().__init__(f"{_STATUS.get(status, status)}: {message}")


class _Pred(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("start_kind", C.c_int32),
        ("end_kind", C.c_int32),
        ("reserved", C.c_int32),
        ("value", C.c_double),
        ("start", C.c_double),
        ("end", C.c_double),
    ]


class _Query(C.Structure):
    _fields_ = [
        ("op", C.c_int32),
        ("reserved", C.c_int32),
        ("col", C.c_void_p),
        ("pred", _Pred),
        ("bitset", C.c_void_p),
        ("chunk_id", C.c_int64),
        ("out_values", C.c_void_p),
        ("out_bitset", C.c_void_p),
        ("out_scalar", C.c_void_p),
        ("n_out", C.c_void_p),
    ]


class _KnnResult(C.Structure):
    _fields_ = [("index", C.c_uint64), ("distance", C.c_double)]


class _ExprNode(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_uint32), ("literal", C.c_double)]


class _Aggregation(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("n_nodes", C.c_uint32),
        ("nodes", C.c_void_p),
        ("out_scalar", C.c_void_p),
        ("out_values", C.c_void_p),
        ("capacity", C.c_size_t),
        ("n_out", C.c_void_p),
    ]


class _ServeOpts(C.Structure):
    _fields_ = [
        ("device_budget", C.c_uint64),
        ("pinned_budget", C.c_uint64),
        ("max_inflight", C.c_uint32),
        ("n_streams", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class _ServeStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("submitted", "completed", "rejected", "batches", "max_inflight",
                                           "device_bytes_peak")]


class _BatchCfg(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_uint32), ("mixed", C.c_uint32), ("tiles_per_job", C.c_uint32),
                ("cached", C.c_uint32), ("tuned_ms", C.c_double)]


class _EncCfg(C.Structure):
    _fields_ = [
        ("scheme", C.c_int32),
        ("chunk_kind", C.c_int32),
        ("chunk_value", C.c_uint64),
        ("capacity", C.c_uint64),
        ("scale", C.c_uint64),
        ("n_threads", C.c_int32),
        ("reserved", C.c_int32),
    ]


class _FileInfo(C.Structure):
    _fields_ = [
        ("number_of_records", C.c_uint64),
        ("number_of_chunks", C.c_uint64),
        ("footer_offset", C.c_uint64),
        ("chunk_option_value", C.c_uint64),
        ("config_capacity", C.c_uint64),
        ("config_scale", C.c_uint64),
        ("scheme", C.c_int32),
        ("chunk_option_kind", C.c_int32),
        ("version", C.c_uint16 * 3),
        ("reserved", C.c_uint16),
    ]


_P = C.c_void_p
_SIGS = {
    "ebi_encode": (C.c_int, [_P, C.c_size_t, C.POINTER(_EncCfg), C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "ebi_buffer_free": (None, [_P]),
    "ebi_synth_column": (C.c_int, [C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_int, _P]),
    "ebi_file_parse": (C.c_int, [_P, C.c_size_t, C.POINTER(_FileInfo)]),
    "ebi_file_chunks": (C.c_int, [_P, C.c_size_t, _P, _P, _P, _P]),
    "ebi_hip_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "ebi_hip_ctx_destroy": (C.c_int, [_P]),
    "ebi_hip_last_error": (C.c_char_p, [_P]),
    "ebi_hip_last_device_ms": (C.c_double, [_P]),
    "ebi_hip_last_phase_ms": (C.c_double, [_P, C.c_int]),
    "ebi_hip_column_load": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(_P)]),
    "ebi_hip_column_free": (C.c_int, [_P, _P]),
    "ebi_hip_column_info": (C.c_int, [_P, C.POINTER(_FileInfo)]),
    "ebi_hip_column_compressed_bytes": (C.c_uint64, [_P]),
    "ebi_hip_column_device_bytes": (C.c_uint64, [_P]),
    "ebi_hip_column_set_precision": (C.c_int, [_P, C.c_int32]),
    "ebi_hip_decode_device": (C.c_int, [_P, _P, _P]),
    "ebi_hip_decode_per_chunk": (C.c_int, [_P, _P, _P]),
    "ebi_hip_decode_batch": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "ebi_hip_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "ebi_hip_device_free": (C.c_int, [_P, _P]),
    "ebi_hip_copy_to_host": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "ebi_hip_copy_to_device": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "ebi_hip_encode_device": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(_EncCfg), C.POINTER(_P),
                                        C.POINTER(C.c_size_t)]),
    "ebi_hip_column_encode": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(_EncCfg), C.POINTER(_P)]),
    "ebi_hip_device_zero": (C.c_int, [_P, _P, C.c_size_t]),
    "ebi_hip_materialize": (C.c_int, [_P, _P, _P, C.c_int64, _P, C.POINTER(C.c_size_t)]),
    "ebi_hip_filter": (C.c_int, [_P, _P, C.POINTER(_Pred), _P, C.c_int64, _P]),
    "ebi_hip_filter_materialize": (C.c_int, [_P, _P, C.POINTER(_Pred), _P, C.c_int64, _P, C.POINTER(C.c_size_t)]),
    "ebi_hip_sum": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(C.c_double)]),
    "ebi_hip_min": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(C.c_double)]),
    "ebi_hip_max": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(C.c_double)]),
    "ebi_hip_query": (C.c_int, [_P, C.POINTER(_Query)]),
    "ebi_hip_query_noalloc": (C.c_int, [_P, C.POINTER(_Query)]),
    "ebi_hip_batch_create": (C.c_int, [_P, _P, C.c_size_t, _P, C.c_int, C.POINTER(_P)]),
    "ebi_hip_batch_run": (C.c_int, [_P, _P, C.c_int]),
    "ebi_hip_batch_config_get": (C.c_int, [_P, C.POINTER(_BatchCfg)]),
    "ebi_hip_batch_config_set": (C.c_int, [_P, _P, C.POINTER(_BatchCfg)]),
    "ebi_hip_batch_destroy": (C.c_int, [_P, _P]),
    "ebi_hip_serve_create": (C.c_int, [_P, C.POINTER(_ServeOpts), C.POINTER(_P)]),
    "ebi_hip_serve_submit": (C.c_int, [_P, C.POINTER(_Query), C.c_size_t, C.POINTER(C.c_uint64)]),
    "ebi_hip_serve_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int)]),
    "ebi_hip_serve_wait": (C.c_int, [_P, C.c_uint64, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "ebi_hip_serve_stats": (C.c_int, [_P, C.POINTER(_ServeStats)]),
    "ebi_hip_serve_destroy": (C.c_int, [_P]),
    "ebi_hip_matmul": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, _P]),
    "ebi_hip_gemm_decoded": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_size_t, C.c_size_t, _P, _P]),
    "ebi_hip_dgemm_device": (C.c_int, [_P, _P, _P, _P, C.c_size_t, C.c_size_t, C.c_size_t]),
    "ebi_hip_knn1": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(_KnnResult)]),
    "ebi_hip_distance_squared": (C.c_int, [_P, _P, C.c_uint64, C.c_size_t, _P, C.c_size_t, C.POINTER(C.c_double)]),
    "ebi_hip_dot_product": (C.c_int, [_P, _P, C.c_uint64, C.c_size_t, _P, C.c_size_t, C.POINTER(C.c_double)]),
    "ebi_hip_matmul_dot": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, _P]),
    "ebi_hip_aggregate": (C.c_int, [_P, _P, C.c_size_t, _P, C.c_size_t, _P]),
}


def lib():
    """Load libebi_hip.so from this package directory (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise EbiError(HIP, f"{_LIB_PATH} is not built; run `python -m ebi_amd.build`")
        L = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, ctx: Optional["Context"] = None, what: str = ""):
    if rc != OK:
        msg = what
        if ctx is not None and ctx._h:
            m = lib().ebi_hip_last_error(ctx._h)
            if m:
                msg = f"{what}: {m.decode(errors='replace')}"
        raise EbiError(rc, msg)


# ----------------------------------------------------------------- config
@dataclass(frozen=True)
class ChunkOption:
    """encoder::ChunkOption (ebi/src/encoder.rs:26-64)."""

    kind: int
    value: int = 0

    @staticmethod
    def full() -> "ChunkOption":
        return ChunkOption(0, 0)

    @staticmethod
    def record_count(n: int) -> "ChunkOption":
        return ChunkOption(1, int(n))

    @staticmethod
    def byte_size_best_effort(n: int) -> "ChunkOption":
        return ChunkOption(2, int(n))


@dataclass(frozen=True)
class CompressorConfig:
    """compressor::CompressorConfig builders (ebi/src/compressor.rs:404-458)."""

    scheme: int
    scale: int = 1
    capacity: int = 8192

    @staticmethod
    def uncompressed(capacity: int = 8192):
        return CompressorConfig(0, 1, capacity)

    @staticmethod
    def rle(capacity: int = 8192):
        return CompressorConfig(1, 1, capacity)

    @staticmethod
    def buff(scale: int = 1):
        return CompressorConfig(2, scale, 0)

    @staticmethod
    def gorilla(capacity: int = 8192):
        return CompressorConfig(3, 1, capacity)

    @staticmethod
    def chimp(capacity: int = 8192):
        return CompressorConfig(4, 1, capacity)

    @staticmethod
    def chimp128(capacity: int = 8192):
        return CompressorConfig(5, 1, capacity)

    @staticmethod
    def elf(capacity: int = 8192):
        return CompressorConfig(7, 1, capacity)

    @staticmethod
    def elf_on_chimp(capacity: int = 8192):
        return CompressorConfig(6, 1, capacity)

    @staticmethod
    def delta_sprintz(scale: int = 1, capacity: int = 8192):
        return CompressorConfig(8, scale, capacity)

    @staticmethod
    def by_name(name: str, scale: int = 1) -> "CompressorConfig":
        return CompressorConfig(SCHEMES[name], scale, 8192)


def encode(values, config: CompressorConfig, chunk_option: ChunkOption, threads: int = 0) -> np.ndarray:
    """Encoder::encode (ebi/src/api/encoder.rs:140-178) on the host.

    Returns the EBI1 file as a uint8 array that owns the library's buffer
    (no copy); ``bytes(arr)`` gives a bytes object."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    cfg = _EncCfg(config.scheme, chunk_option.kind, chunk_option.value, config.capacity,
                  config.scale, threads or min(16, os.cpu_count() or 1), 0)
    out = _P()
    n = C.c_size_t()
    L = lib()
    _check(L.ebi_encode(v.ctypes.data if v.size else None, v.size, C.byref(cfg), C.byref(out), C.byref(n)),
           None, "encode")
    arr = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint8)), shape=(n.value,))
    weakref.finalize(arr, L.ebi_buffer_free, out.value)
    return arr


def synth(kind: int, n: int, seed: int, decimals: int, p0: float = 0.0, p1: float = 0.0, p2: float = 0.0,
          threads: int = 0) -> np.ndarray:
    """Seeded synthetic column (ebi_synth_column)."""
    out = np.empty(max(int(n), 1), dtype=np.float64)
    _check(lib().ebi_synth_column(int(kind), int(n), int(seed), int(decimals), float(p0), float(p1), float(p2),
                                  threads or min(16, os.cpu_count() or 1), out.ctypes.data), None, "synth")
    return out[: int(n)]


@dataclass
class FileInfo:
    number_of_records: int
    number_of_chunks: int
    scheme: int
    footer_offset: int
    chunk_option_kind: int
    chunk_option_value: int
    capacity: int
    scale: int
    version: tuple


def file_info(data: bytes) -> FileInfo:
    fi = _FileInfo()
    buf = np.frombuffer(data, dtype=np.uint8)
    _check(lib().ebi_file_parse(buf.ctypes.data, len(data), C.byref(fi)), None, "file_parse")
    return FileInfo(fi.number_of_records, fi.number_of_chunks, fi.scheme, fi.footer_offset,
                    fi.chunk_option_kind, fi.chunk_option_value, fi.config_capacity,
                    fi.config_scale, tuple(fi.version))


def file_chunks(data: bytes):
    """Per-chunk (physical_offset, size, logical_offset, number_of_records) arrays."""
    n = file_info(data).number_of_chunks
    arrs = [np.zeros(n, dtype=np.uint64) for _ in range(4)]
    buf = np.frombuffer(data, dtype=np.uint8)
    _check(lib().ebi_file_chunks(buf.ctypes.data, len(data), *[a.ctypes.data for a in arrs]), None, "file_chunks")
    return tuple(arrs)


# -------------------------------------------------------------- predicates
@dataclass(frozen=True)
class RangeValue:
    """decoder::query::RangeValue (decoder/query.rs:291-305)."""

    kind: int  # 0 None, 1 Inclusive, 2 Exclusive
    value: float = 0.0

    @staticmethod
    def inclusive(v: float) -> "RangeValue":
        return RangeValue(1, float(v))

    @staticmethod
    def exclusive(v: float) -> "RangeValue":
        return RangeValue(2, float(v))

    @staticmethod
    def none() -> "RangeValue":
        return RangeValue(0, 0.0)


@dataclass(frozen=True)
class Predicate:
    """decoder::query::Predicate (decoder/query.rs:396-446)."""

    kind: int
    value: float = 0.0
    start: RangeValue = RangeValue(0)
    end: RangeValue = RangeValue(0)

    @staticmethod
    def eq(v: float) -> "Predicate":
        return Predicate(0, float(v))

    @staticmethod
    def ne(v: float) -> "Predicate":
        return Predicate(1, float(v))

    @staticmethod
    def range(start: RangeValue, end: RangeValue) -> "Predicate":
        return Predicate(2, 0.0, start, end)

    def to_c(self) -> _Pred:
        return _Pred(self.kind, self.start.kind, self.end.kind, 0, self.value, self.start.value, self.end.value)


# ----------------------------------------------------------------- bitmasks
def ids_to_bitset(ids, n_records: int) -> np.ndarray:
    """Global record ids -> dense u64 words (bit r % 64 of word r // 64)."""
    words = np.zeros((n_records + 63) // 64 + 1, dtype=np.uint64)
    ids = np.asarray(list(ids) if not isinstance(ids, np.ndarray) else ids, dtype=np.int64)
    ids = ids[(ids >= 0) & (ids < n_records)]
    if ids.size:
        np.bitwise_or.at(words, ids >> 6, np.left_shift(np.uint64(1), (ids & 63).astype(np.uint64)))
    return words


def bitset_to_ids(words: np.ndarray, n_records: int) -> np.ndarray:
    """Dense u64 words -> ascending global record ids below n_records (only
    the nonzero words are unpacked)."""
    w = np.ascontiguousarray(np.asarray(words, dtype=np.uint64)[: (n_records + 63) // 64])
    nz = np.flatnonzero(w)
    if nz.size == 0:
        return np.empty(0, dtype=np.uint32)
    bits = np.unpackbits(w[nz].view(np.uint8), bitorder="little").reshape(nz.size, 64)
    r, c = np.nonzero(bits)
    ids = (nz[r].astype(np.int64) << 6) + c
    return ids[ids < n_records].astype(np.uint32)


# ------------------------------------------------------------------ device
class Context:
    """A device context: replaces cuda_binding::Context (cuda-binding/src/lib.rs:32-181)."""

    def __init__(self, device: int = 0):
        self._h = None
        h = _P()
        rc = lib().ebi_hip_ctx_create(int(device), C.byref(h))
        if rc != OK:
            raise EbiError(rc, f"ebi_hip_ctx_create(device={device})")
        self._h = h

    def close(self):
        if self._h:
            lib().ebi_hip_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def last_device_ms(self) -> float:
        return float(lib().ebi_hip_last_device_ms(self._h))

    def last_phase_ms(self, phase: int) -> float:
        """Device ms of kernel `phase` of the last decode (0: tile kernel,
        2: XOR segment kernel; 1 is 0)."""
        return float(lib().ebi_hip_last_phase_ms(self._h, int(phase)))

    def alloc(self, nbytes: int) -> "DeviceBuffer":
        p = _P()
        _check(lib().ebi_hip_device_alloc(self._h, int(nbytes), C.byref(p)), self, "device_alloc")
        return DeviceBuffer(self, p, int(nbytes))

    def upload(self, values) -> "DeviceBuffer":
        """f64 values copied into a new device buffer."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        buf = self.alloc(max(v.nbytes, 8))
        _check(lib().ebi_hip_copy_to_device(self._h, buf.ptr, v.ctypes.data, v.nbytes), self, "copy_to_device")
        return buf


def _enc_cfg(config: "CompressorConfig", chunk_option: "ChunkOption") -> "_EncCfg":
    return _EncCfg(config.scheme, chunk_option.kind, chunk_option.value, config.capacity, config.scale, 1, 0)


def encode_device(ctx: Context, values: "DeviceBuffer", n: int, config: "CompressorConfig",
                  chunk_option: "ChunkOption") -> np.ndarray:
    """ebi_hip_encode_device: n f64 values already in device memory encoded
    by GPU kernels (BUFF, DeltaSprintz); the EBI1 file as a uint8 array."""
    cfg = _enc_cfg(config, chunk_option)
    out, ln = _P(), C.c_size_t()
    L = lib()
    _check(L.ebi_hip_encode_device(ctx._h, values.ptr, int(n), C.byref(cfg), C.byref(out), C.byref(ln)), ctx,
           "encode_device")
    arr = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint8)), shape=(ln.value,))
    weakref.finalize(arr, L.ebi_buffer_free, out.value)
    return arr


class DeviceBuffer:
    def __init__(self, ctx: Context, ptr, nbytes: int):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def to_host(self, dtype=np.float64, count: Optional[int] = None, offset: int = 0) -> np.ndarray:
        """count elements from element `offset` copied to a new host array."""
        it = np.dtype(dtype).itemsize
        n = self.nbytes // it - offset if count is None else count
        if offset < 0 or n < 0 or (offset + n) * it > self.nbytes:
            raise EbiError(INVALID_ARGUMENT, "to_host: range outside the buffer")
        out = np.empty(n, dtype=dtype)
        src = C.c_void_p((self.ptr.value or 0) + offset * it)
        _check(lib().ebi_hip_copy_to_host(self.ctx._h, out.ctypes.data, src, n * it), self.ctx, "copy_to_host")
        return out

    def zero(self):
        _check(lib().ebi_hip_device_zero(self.ctx._h, self.ptr, self.nbytes & ~3), self.ctx, "device_zero")

    def free(self):
        if self.ptr:
            lib().ebi_hip_device_free(self.ctx._h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ctx._h:
                self.free()
        except Exception:
            pass


class Decoder:
    """api::decoder::Decoder (ebi/src/api/decoder.rs:121-753) over a resident column.

    Bitmask arguments are iterables of global record ids (the reference's
    RoaringBitmap) or dense u64 word arrays; filter results are sorted u32 id
    arrays.  ``chunk_id`` selects one chunk (ChunkId, api/decoder.rs:97-106).
    """

    def __init__(self, ctx: Context, data: bytes = None, _handle=None):
        self.ctx = ctx
        if _handle is not None:
            h = _handle
        else:
            self._data = np.frombuffer(data, dtype=np.uint8)
            h = _P()
            _check(lib().ebi_hip_column_load(ctx._h, self._data.ctypes.data, len(data), C.byref(h)), ctx,
                   "column_load")
        self._h = h
        fi = _FileInfo()
        _check(lib().ebi_hip_column_info(h, C.byref(fi)), ctx, "column_info")
        self.number_of_records = int(fi.number_of_records)
        self.number_of_chunks = int(fi.number_of_chunks)
        self.scheme = int(fi.scheme)

    @staticmethod
    def encode_on_device(ctx: Context, values: "DeviceBuffer", n: int, config: "CompressorConfig",
                         chunk_option: "ChunkOption") -> "Decoder":
        """ebi_hip_column_encode: device values encoded by GPU kernels and
        loaded as a resident column."""
        cfg = _enc_cfg(config, chunk_option)
        h = _P()
        _check(lib().ebi_hip_column_encode(ctx._h, values.ptr, int(n), C.byref(cfg), C.byref(h)), ctx,
               "column_encode")
        return Decoder(ctx, _handle=h)

    def close(self):
        if getattr(self, "_h", None) and self.ctx._h:
            lib().ebi_hip_column_free(self.ctx._h, self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def compressed_bytes(self) -> int:
        return int(lib().ebi_hip_column_compressed_bytes(self._h))

    @property
    def device_bytes(self) -> int:
        """Device bytes the resident column holds (bodies + load-time index)."""
        return int(lib().ebi_hip_column_device_bytes(self._h))

    def with_precision(self, precision: int) -> None:
        """Decoder::with_precision (api/decoder.rs:159-161): BUFF decodes keep
        `precision` decimal digits (controlled_precision.rs:5-75); -1 restores
        full precision."""
        _check(lib().ebi_hip_column_set_precision(self._h, int(precision)), self.ctx, "with_precision")

    def _mask(self, bitmask):
        if bitmask is None:
            return None, None
        if isinstance(bitmask, np.ndarray) and bitmask.dtype == np.uint64:
            w = np.ascontiguousarray(bitmask)
            if w.size < (self.number_of_records + 63) // 64:
                raise EbiError(INVALID_ARGUMENT, "bitset shorter than ceil(N/64) words")
        else:
            w = ids_to_bitset(bitmask, self.number_of_records)
        return w, w.ctypes.data

    @staticmethod
    def _cid(chunk_id):
        return -1 if chunk_id is None else int(chunk_id)

    def materialize(self, bitmask=None, chunk_id=None) -> np.ndarray:
        _, mp = self._mask(bitmask)
        out = np.empty(max(self.number_of_records, 1), dtype=np.float64)
        n = C.c_size_t()
        _check(lib().ebi_hip_materialize(self.ctx._h, self._h, mp, self._cid(chunk_id), out.ctypes.data, C.byref(n)),
               self.ctx, "materialize")
        return out[: n.value].copy()

    def filter_bitset(self, predicate: Predicate, bitmask=None, chunk_id=None) -> np.ndarray:
        _, mp = self._mask(bitmask)
        out = np.zeros((self.number_of_records + 63) // 64 + 1, dtype=np.uint64)
        p = predicate.to_c()
        _check(lib().ebi_hip_filter(self.ctx._h, self._h, C.byref(p), mp, self._cid(chunk_id), out.ctypes.data),
               self.ctx, "filter")
        return out

    def filter(self, predicate: Predicate, bitmask=None, chunk_id=None) -> np.ndarray:
        return bitset_to_ids(self.filter_bitset(predicate, bitmask, chunk_id), self.number_of_records)

    def filter_materialize(self, predicate: Predicate, bitmask=None, chunk_id=None) -> np.ndarray:
        _, mp = self._mask(bitmask)
        out = np.empty(max(self.number_of_records, 1), dtype=np.float64)
        n = C.c_size_t()
        p = predicate.to_c()
        _check(lib().ebi_hip_filter_materialize(self.ctx._h, self._h, C.byref(p), mp, self._cid(chunk_id),
                                                out.ctypes.data, C.byref(n)), self.ctx, "filter_materialize")
        return out[: n.value].copy()

    def _agg(self, fn, bitmask, chunk_id) -> float:
        _, mp = self._mask(bitmask)
        r = C.c_double()
        _check(fn(self.ctx._h, self._h, mp, self._cid(chunk_id), C.byref(r)), self.ctx, "aggregate")
        return float(r.value)

    def sum(self, bitmask=None, chunk_id=None) -> float:
        return self._agg(lib().ebi_hip_sum, bitmask, chunk_id)

    def min(self, bitmask=None, chunk_id=None) -> float:
        return self._agg(lib().ebi_hip_min, bitmask, chunk_id)

    def max(self, bitmask=None, chunk_id=None) -> float:
        return self._agg(lib().ebi_hip_max, bitmask, chunk_id)

    def decode_device(self, out: DeviceBuffer) -> None:
        """Whole-column decode into a device buffer of N doubles (stays in HBM)."""
        if out.nbytes < 8 * self.number_of_records:
            raise EbiError(INVALID_ARGUMENT, "device buffer too small")
        _check(lib().ebi_hip_decode_device(self.ctx._h, self._h, out.ptr), self.ctx, "decode_device")

    def decode_per_chunk(self, out: DeviceBuffer) -> None:
        """The plain per-chunk-launch comparator (one launch + host sync per chunk)."""
        if out.nbytes < 8 * self.number_of_records:
            raise EbiError(INVALID_ARGUMENT, "device buffer too small")
        _check(lib().ebi_hip_decode_per_chunk(self.ctx._h, self._h, out.ptr), self.ctx, "decode_per_chunk")


    def matmul_cuda(self, target, target_matrix_shape, data_matrix_shape) -> np.ndarray:
        """Decoder::matmul_cuda (api/decoder.rs:659-724): batch row-major
        products A_b * target of the column's data matrices, flattened."""
        k, n = (int(x) for x in target_matrix_shape)
        m, kd = (int(x) for x in data_matrix_shape)
        t = np.ascontiguousarray(target, dtype=np.float64).ravel()
        if t.size != k * n:
            raise EbiError(PRECONDITION, "target length does not match its shape")
        size = m * kd
        batch = self.number_of_records // size if size else 0
        out = np.empty(max(batch * m * n, 1), dtype=np.float64)
        _check(lib().ebi_hip_matmul(self.ctx._h, self._h, t.ctypes.data, k, n, m, kd, out.ctypes.data), self.ctx,
               "matmul")
        return out[: batch * m * n]


    # ---- ML queries (api/decoder.rs:469-656) ----
    def knn1(self, target):
        """Decoder::knn1: (index, squared distance) of the vector nearest to
        `target` (KnnResult; index 2**64 - 1 when none)."""
        t = np.ascontiguousarray(target, dtype=np.float64)
        r = _KnnResult()
        _check(lib().ebi_hip_knn1(self.ctx._h, self._h, t.ctypes.data if t.size else None, t.size, C.byref(r)),
               self.ctx, "knn1")
        return int(r.index), float(r.distance)

    def _fold(self, fn, chunk_id, offset, target):
        t = np.ascontiguousarray(target, dtype=np.float64)
        r = C.c_double()
        _check(fn(self.ctx._h, self._h, int(chunk_id), int(offset), t.ctypes.data if t.size else None, t.size,
                  C.byref(r)), self.ctx, "chunk fold")
        return float(r.value)

    def distance_squared(self, chunk_id, offset_in_chunk, target) -> float:
        """GeneralChunkReader::distance_squared of chunk `chunk_id`."""
        return self._fold(lib().ebi_hip_distance_squared, chunk_id, offset_in_chunk, target)

    def dot_product(self, chunk_id, offset_in_chunk, target) -> float:
        """GeneralChunkReader::dot_product of chunk `chunk_id`."""
        return self._fold(lib().ebi_hip_dot_product, chunk_id, offset_in_chunk, target)

    def matmul(self, target_matrix, target_matrix_shape, data_matrix_shape) -> np.ndarray:
        """Decoder::matmul (api/decoder.rs:606-656): target column-first,
        data matrices row-first; the flat row-first products."""
        tr, tc = (int(x) for x in target_matrix_shape)
        dr, dc = (int(x) for x in data_matrix_shape)
        t = np.ascontiguousarray(target_matrix, dtype=np.float64).ravel()
        if t.size < tr * tc:
            raise EbiError(INVALID_ARGUMENT, "target shorter than its shape")
        batch = self.number_of_records // (dr * dc) if dr * dc else 0
        out = np.empty(max(batch * dr * tc, 1), dtype=np.float64)
        _check(lib().ebi_hip_matmul_dot(self.ctx._h, self._h, t.ctypes.data, tr, tc, dr, dc, out.ctypes.data),
               self.ctx, "matmul")
        return out[: batch * dr * tc]


# --------------------------------------------------- aggregations (Expr)
EXPR_INDEX, EXPR_LITERAL, EXPR_ADD, EXPR_SUB, EXPR_MUL, EXPR_DIV = range(6)
AGG_NONE, AGG_MAX, AGG_MIN, AGG_SUM = range(4)


class Expr:
    """decoder::expr::Expr (decoder/expr.rs:31-40): Index(i), Literal(v) or
    Binary {op, lhs, rhs}; kept as a postfix program for the C ABI."""

    def __init__(self, nodes):
        self.nodes = list(nodes)

    @staticmethod
    def index(i: int) -> "Expr":
        return Expr([(EXPR_INDEX, int(i), 0.0)])

    @staticmethod
    def literal(v: float) -> "Expr":
        return Expr([(EXPR_LITERAL, 0, float(v))])

    @staticmethod
    def binary(op: str, lhs: "Expr", rhs: "Expr") -> "Expr":
        code = {"add": EXPR_ADD, "sub": EXPR_SUB, "mul": EXPR_MUL, "div": EXPR_DIV}[op]
        return Expr(lhs.nodes + rhs.nodes + [(code, 0, 0.0)])

    def __add__(self, o):
        return Expr.binary("add", self, _as_expr(o))

    def __sub__(self, o):
        return Expr.binary("sub", self, _as_expr(o))

    def __mul__(self, o):
        return Expr.binary("mul", self, _as_expr(o))

    def __truediv__(self, o):
        return Expr.binary("div", self, _as_expr(o))


def _as_expr(x):
    return x if isinstance(x, Expr) else Expr.literal(float(x))


class Aggregation:
    """api::decoder::Aggregation (api/decoder.rs:835-917)."""

    def __init__(self, kind: int, expr: Expr):
        self.kind, self.expr = int(kind), expr

    @staticmethod
    def sum(expr: Expr) -> "Aggregation":
        return Aggregation(AGG_SUM, expr)

    @staticmethod
    def min(expr: Expr) -> "Aggregation":
        return Aggregation(AGG_MIN, expr)

    @staticmethod
    def max(expr: Expr) -> "Aggregation":
        return Aggregation(AGG_MAX, expr)

    @staticmethod
    def none(expr: Expr) -> "Aggregation":
        return Aggregation(AGG_NONE, expr)

    def compute(self, decoders, bitmask=None):
        return Aggregation.compute_multiple([self], decoders, bitmask)[0]

    @staticmethod
    def compute_multiple(aggregations, decoders, bitmask=None) -> list:
        """Scalars for MAX / MIN / SUM, arrays for NONE."""
        decs = list(decoders)
        ctx = decs[0].ctx
        n_max = max(d.number_of_records for d in decs)
        mask = None
        if bitmask is not None:
            mask = np.ascontiguousarray(bitmask) if isinstance(bitmask, np.ndarray) and bitmask.dtype == np.uint64 \
                else ids_to_bitset(bitmask, n_max)
        keep, arr = [], (_Aggregation * len(aggregations))()
        for a, agg in zip(arr, aggregations):
            nodes = (_ExprNode * len(agg.expr.nodes))(*[_ExprNode(k, i, v) for k, i, v in agg.expr.nodes])
            scalar = C.c_double(0.0)
            n_out = C.c_size_t(0)
            vals = np.empty(max(n_max, 1), dtype=np.float64) if agg.kind == AGG_NONE else None
            a.kind, a.n_nodes, a.nodes = agg.kind, len(agg.expr.nodes), C.cast(nodes, C.c_void_p)
            a.out_scalar = C.addressof(scalar)
            a.out_values = vals.ctypes.data if vals is not None else None
            a.capacity = vals.size if vals is not None else 0
            a.n_out = C.addressof(n_out)
            keep.append((nodes, scalar, n_out, vals))
        cols = (_P * len(decs))(*[d._h for d in decs])
        _check(lib().ebi_hip_aggregate(ctx._h, cols, len(decs), arr, len(aggregations),
                                       mask.ctypes.data if mask is not None else None), ctx, "aggregate")
        out = []
        for agg, (_, scalar, n_out, vals) in zip(aggregations, keep):
            out.append(vals[: n_out.value].copy() if agg.kind == AGG_NONE else float(scalar.value))
        return out


class Query:
    """One query as data (ebi_query): op is one of OP_*; `decoder` names the
    resident column; predicate for the filters; bitmask / chunk_id as in the
    Decoder methods.  Holds its own output buffers until it completes."""

    def __init__(self, op: int, decoder: "Decoder", predicate: Optional[Predicate] = None, bitmask=None,
                 chunk_id=None, out: Optional["DeviceBuffer"] = None):
        self.op, self.decoder, self.predicate = int(op), decoder, predicate
        self._mask, mp = decoder._mask(bitmask)
        n = decoder.number_of_records
        self._values = self._bits = None
        self._scalar = C.c_double(0.0)
        self._n = C.c_size_t(0)
        q = _Query()
        q.op = self.op
        q.col = decoder._h
        if predicate is not None:
            q.pred = predicate.to_c()
        q.bitset = mp
        q.chunk_id = Decoder._cid(chunk_id)
        if self.op in (OP_MATERIALIZE, OP_FILTER_MATERIALIZE):
            self._values = np.empty(max(n, 1), dtype=np.float64)
            q.out_values = self._values.ctypes.data
            q.n_out = C.addressof(self._n)
        elif self.op == OP_FILTER:
            self._bits = np.zeros((n + 63) // 64 + 1, dtype=np.uint64)
            q.out_bitset = self._bits.ctypes.data
        elif self.op == OP_DECODE:
            if out is None or out.nbytes < 8 * n:
                raise EbiError(INVALID_ARGUMENT, "OP_DECODE needs a device buffer of N doubles")
            self._out = out
            q.out_values = out.ptr
        else:
            q.out_scalar = C.addressof(self._scalar)
        self.c = q

    def result(self):
        if self.op in (OP_MATERIALIZE, OP_FILTER_MATERIALIZE):
            return self._values[: self._n.value].copy()
        if self.op == OP_FILTER:
            return bitset_to_ids(self._bits, self.decoder.number_of_records)
        if self.op == OP_DECODE:
            return None
        return float(self._scalar.value)


def run_query(ctx: "Context", q: Query, noalloc: bool = False):
    """ebi_hip_query: one query synchronously; returns its result.  With
    noalloc, ebi_hip_query_noalloc (fixed per-GPU scratch, no allocation)."""
    fn = lib().ebi_hip_query_noalloc if noalloc else lib().ebi_hip_query
    _check(fn(ctx._h, C.byref(q.c)), ctx, "query")
    return q.result()


class Server:
    """ebi_hip_serve_*: continuous batching of concurrent queries."""

    def __init__(self, ctx: "Context", device_budget: int = 0, pinned_budget: int = 0, max_inflight: int = 0,
                 n_streams: int = 0, max_batch: int = 0):
        self.ctx = ctx
        o = _ServeOpts(device_budget, pinned_budget, max_inflight, n_streams, max_batch, 0)
        h = _P()
        _check(lib().ebi_hip_serve_create(ctx._h, C.byref(o), C.byref(h)), ctx, "serve_create")
        self._h = h
        self._live = {}

    def submit(self, queries) -> list:
        qs = list(queries)
        arr = (_Query * len(qs))(*[q.c for q in qs])
        t = (C.c_uint64 * len(qs))()
        _check(lib().ebi_hip_serve_submit(self._h, arr, len(qs), t), None, "serve_submit")
        ids = [int(x) for x in t]
        for i, q in zip(ids, qs):
            self._live[i] = q
        return ids

    def poll(self, ticket: int):
        st = C.c_int()
        rc = lib().ebi_hip_serve_poll(self._h, ticket, C.byref(st))
        if rc == PENDING:
            return None
        _check(rc, None, "serve_poll")
        return int(st.value)

    def wait(self, ticket: int, timeout_ms: int = -1):
        """(result, latency_ms); raises EbiError with the query's status."""
        st, lat = C.c_int(), C.c_double()
        rc = lib().ebi_hip_serve_wait(self._h, ticket, timeout_ms, C.byref(st), C.byref(lat))
        if rc == PENDING:
            raise TimeoutError(f"ticket {ticket} still pending")
        _check(rc, None, "serve_wait")
        q = self._live.pop(ticket)
        if st.value != OK:
            raise EbiError(st.value, f"query (ticket {ticket})")
        return q.result(), float(lat.value)

    def stats(self) -> dict:
        s = _ServeStats()
        _check(lib().ebi_hip_serve_stats(self._h, C.byref(s)), None, "serve_stats")
        return {f: int(getattr(s, f)) for f, _ in _ServeStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().ebi_hip_serve_destroy(self._h)
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gemm_decoded(ctx: Context, a: Decoder, b: Decoder, m: int, n: int, k: int,
                 out: Optional[DeviceBuffer] = None) -> Optional[np.ndarray]:
    """C = A * B of two decoded row-major columns (A: m x k, B: k x n).
    Returns C (m x n) on the host, or writes it into the device buffer `out`."""
    if out is not None:
        if out.nbytes < 8 * m * n:
            raise EbiError(INVALID_ARGUMENT, "device buffer too small")
        _check(lib().ebi_hip_gemm_decoded(ctx._h, a._h, b._h, m, n, k, None, out.ptr), ctx, "gemm_decoded")
        return None
    c = np.empty((m, n), dtype=np.float64)
    _check(lib().ebi_hip_gemm_decoded(ctx._h, a._h, b._h, m, n, k, c.ctypes.data, None), ctx, "gemm_decoded")
    return c


def dgemm_device(ctx: Context, a: DeviceBuffer, b: DeviceBuffer, c: DeviceBuffer, m: int, n: int, k: int) -> None:
    """The DGEMM kernel alone on device buffers (row-major C = A * B)."""
    if a.nbytes < 8 * m * k or b.nbytes < 8 * k * n or c.nbytes < 8 * m * n:
        raise EbiError(INVALID_ARGUMENT, "device buffer too small")
    _check(lib().ebi_hip_dgemm_device(ctx._h, a.ptr, b.ptr, c.ptr, m, n, k), ctx, "dgemm_device")


def decode_batch(ctx: Context, decoders, outs) -> None:
    """One persistent launch over every chunk of several resident columns."""
    n = len(decoders)
    cols = (_P * n)(*[d._h for d in decoders])
    ptrs = (_P * n)(*[o.ptr for o in outs])
    _check(lib().ebi_hip_decode_batch(ctx._h, cols, n, ptrs), ctx, "decode_batch")


class Batch:
    """ebi_hip_batch_*: a prepared (optionally autotuned) batched decode of
    several resident columns into device buffers, replayed as a hipGraph."""

    def __init__(self, ctx: Context, decoders, outs, autotune: bool = True):
        self.ctx = ctx
        self._keep = (list(decoders), list(outs))
        n = len(decoders)
        cols = (_P * n)(*[d._h for d in decoders])
        ptrs = (_P * n)(*[o.ptr for o in outs])
        h = _P()
        _check(lib().ebi_hip_batch_create(ctx._h, cols, n, ptrs, 1 if autotune else 0, C.byref(h)), ctx,
               "batch_create")
        self._h = h

    def run(self, eager: bool = False) -> float:
        """One replay; returns its device ms."""
        _check(lib().ebi_hip_batch_run(self.ctx._h, self._h, 1 if eager else 0), self.ctx, "batch_run")
        return self.ctx.last_device_ms

    @property
    def config(self) -> dict:
        c = _BatchCfg()
        _check(lib().ebi_hip_batch_config_get(self._h, C.byref(c)), None, "batch_config")
        return {"blocks_per_cu": c.blocks_per_cu, "mixed": bool(c.mixed), "tiles_per_job": c.tiles_per_job,
                "cached": bool(c.cached), "tuned_ms": c.tuned_ms}

    def set_config(self, blocks_per_cu: int, mixed: bool, tiles_per_job: int) -> None:
        """Forces a launch shape (ebi_hip_batch_config_set); re-captures the graph."""
        c = _BatchCfg(blocks_per_cu=blocks_per_cu, mixed=1 if mixed else 0, tiles_per_job=tiles_per_job)
        _check(lib().ebi_hip_batch_config_set(self.ctx._h, self._h, C.byref(c)), self.ctx, "batch_config_set")

    def close(self):
        if getattr(self, "_h", None) and self.ctx._h:
            lib().ebi_hip_batch_destroy(self.ctx._h, self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gpu_available() -> bool:
    try:
        h = _P()
        rc = lib().ebi_hip_ctx_create(0, C.byref(h))
        if rc == OK:
            lib().ebi_hip_ctx_destroy(h)
            return True
    except Exception:
        pass
    return False
