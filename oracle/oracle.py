"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, where it is the checker (and the timed CPU
reference), never the product path.  Mirrors the reference Decoder API
(lemolatoon/ebi ebi/src/api/decoder.rs:202-439) on an in-memory EBI1 file.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


class OracleError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"oracle status {status} in {what}")


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


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P = C.c_void_p
        for name in ("orc_sum", "orc_min", "orc_max", "orc_sum_compensated"):
            getattr(L, name).argtypes = [P, C.c_size_t, P, C.c_int64, C.POINTER(C.c_double)]
        L.orc_materialize.argtypes = [P, C.c_size_t, P, C.c_int64, P, C.c_size_t, C.POINTER(C.c_size_t)]
        L.orc_filter.argtypes = [P, C.c_size_t, C.POINTER(_Pred), P, C.c_int64, P]
        L.orc_filter_materialize.argtypes = [P, C.c_size_t, C.POINTER(_Pred), P, C.c_int64, P, C.c_size_t,
                                             C.POINTER(C.c_size_t)]
        L.orc_materialize_p.argtypes = [P, C.c_size_t, P, C.c_int64, C.c_int32, P, C.c_size_t,
                                        C.POINTER(C.c_size_t)]
        L.orc_filter_materialize_p.argtypes = [P, C.c_size_t, C.POINTER(_Pred), P, C.c_int64, C.c_int32, P,
                                               C.c_size_t, C.POINTER(C.c_size_t)]
        L.orc_chunk_fold.argtypes = [P, C.c_size_t, C.c_uint64, C.c_size_t, P, C.c_size_t, C.c_int,
                                     C.POINTER(C.c_double)]
        L.orc_knn1.argtypes = [P, C.c_size_t, P, C.c_size_t, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
        L.orc_matmul_dot.argtypes = [P, C.c_size_t, P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, P]
        L.orc_aggregate.argtypes = [P, P, C.c_size_t, P, C.c_size_t, C.c_int, P, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), P, C.c_size_t, C.POINTER(C.c_size_t)]
        L.orc_file_records.argtypes = [P, C.c_size_t, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_zigzag.restype = C.c_uint64
        L.orc_zigzag.argtypes = [C.c_int64]
        L.orc_unzigzag.restype = C.c_int64
        L.orc_unzigzag.argtypes = [C.c_uint64]
        L.orc_buff_fixed.restype = C.c_int64
        L.orc_buff_fixed.argtypes = [C.c_double, C.c_int32]
        L.orc_elf_round_up.restype = C.c_double
        L.orc_elf_round_up.argtypes = [C.c_double, C.c_uint32]
        L.orc_elf_get_sp.restype = C.c_int32
        L.orc_elf_get_sp.argtypes = [C.c_double]
        _lib = L
    return _lib


def _pred(p) -> _Pred:
    return _Pred(p.kind, p.start.kind, p.end.kind, 0, p.value, p.start.value, p.end.value)


def _ck(rc, what):
    if rc != 0:
        raise OracleError(rc, what)


class OracleDecoder:
    """CPU restatement of Decoder over one EBI1 file held in memory."""

    def __init__(self, data: bytes):
        self.precision = -1  # Decoder::with_precision (api/decoder.rs:159-161); -1: full
        self._data = data  # keep the caller's buffer alive; it is only read
        self._buf = np.frombuffer(data, dtype=np.uint8)
        self._p = self._buf.ctypes.data
        self._n = len(data)
        nr, nc = C.c_uint64(), C.c_uint64()
        _ck(lib().orc_file_records(self._p, self._n, C.byref(nr), C.byref(nc)), "file")
        self.number_of_records = int(nr.value)
        self.number_of_chunks = int(nc.value)

    def _mask(self, bitmask):
        if bitmask is None:
            return None, None
        if isinstance(bitmask, np.ndarray) and bitmask.dtype == np.uint64:
            w = np.ascontiguousarray(bitmask)
        else:
            w = np.zeros((self.number_of_records + 63) // 64 + 1, dtype=np.uint64)
            ids = np.asarray(list(bitmask) if not isinstance(bitmask, np.ndarray) else bitmask, dtype=np.int64)
            ids = ids[(ids >= 0) & (ids < self.number_of_records)]
            if ids.size:
                np.bitwise_or.at(w, ids >> 6, np.left_shift(np.uint64(1), (ids & 63).astype(np.uint64)))
        return w, w.ctypes.data

    @staticmethod
    def _cid(c):
        return -1 if c is None else int(c)

    def materialize(self, bitmask=None, chunk_id=None) -> np.ndarray:
        _, mp = self._mask(bitmask)
        cap = max(self.number_of_records * 2 + 64, 64)
        out = np.empty(cap, dtype=np.float64)
        n = C.c_size_t()
        _ck(lib().orc_materialize_p(self._p, self._n, mp, self._cid(chunk_id), self.precision, out.ctypes.data, cap,
                                    C.byref(n)), "materialize")
        return out[: n.value].copy()

    def with_precision(self, precision: int) -> None:
        self.precision = int(precision)

    def filter_bitset(self, predicate, bitmask=None, chunk_id=None) -> np.ndarray:
        _, mp = self._mask(bitmask)
        out = np.zeros((self.number_of_records + 63) // 64 + 1, dtype=np.uint64)
        p = _pred(predicate)
        _ck(lib().orc_filter(self._p, self._n, C.byref(p), mp, self._cid(chunk_id), out.ctypes.data), "filter")
        return out

    def filter(self, predicate, bitmask=None, chunk_id=None) -> np.ndarray:
        w = self.filter_bitset(predicate, bitmask, chunk_id)
        bits = np.unpackbits(w.view(np.uint8), bitorder="little")[: self.number_of_records]
        return np.flatnonzero(bits).astype(np.uint32)

    def filter_materialize(self, predicate, bitmask=None, chunk_id=None) -> np.ndarray:
        _, mp = self._mask(bitmask)
        cap = max(self.number_of_records * 2 + 64, 64)
        out = np.empty(cap, dtype=np.float64)
        n = C.c_size_t()
        p = _pred(predicate)
        _ck(lib().orc_filter_materialize_p(self._p, self._n, C.byref(p), mp, self._cid(chunk_id), self.precision,
                                           out.ctypes.data, cap, C.byref(n)), "filter_materialize")
        return out[: n.value].copy()

    def _agg(self, fn, bitmask, chunk_id, what):
        _, mp = self._mask(bitmask)
        r = C.c_double()
        _ck(fn(self._p, self._n, mp, self._cid(chunk_id), C.byref(r)), what)
        return float(r.value)

    def sum(self, bitmask=None, chunk_id=None):
        return self._agg(lib().orc_sum, bitmask, chunk_id, "sum")

    def min(self, bitmask=None, chunk_id=None):
        return self._agg(lib().orc_min, bitmask, chunk_id, "min")

    def max(self, bitmask=None, chunk_id=None):
        return self._agg(lib().orc_max, bitmask, chunk_id, "max")

    def sum_compensated(self, bitmask=None, chunk_id=None):
        return self._agg(lib().orc_sum_compensated, bitmask, chunk_id, "sum_compensated")


    # ---- ML queries (api/decoder.rs:469-656, decoder/query.rs:157-192) ----
    def knn1(self, target):
        t = np.ascontiguousarray(target, dtype=np.float64)
        idx, dist = C.c_uint64(), C.c_double()
        _ck(lib().orc_knn1(self._p, self._n, t.ctypes.data, t.size, C.byref(idx), C.byref(dist)), "knn1")
        return int(idx.value), float(dist.value)

    def _fold(self, chunk_id, offset, target, dot):
        t = np.ascontiguousarray(target, dtype=np.float64)
        r = C.c_double()
        _ck(lib().orc_chunk_fold(self._p, self._n, int(chunk_id), int(offset), t.ctypes.data, t.size, dot,
                                 C.byref(r)), "chunk fold")
        return float(r.value)

    def distance_squared(self, chunk_id, offset, target):
        return self._fold(chunk_id, offset, target, 0)

    def dot_product(self, chunk_id, offset, target):
        return self._fold(chunk_id, offset, target, 1)

    def matmul(self, target, target_shape, data_shape):
        """Decoder::matmul: target column-first (target_shape[1] columns of
        target_shape[0] values); returns the flat row-first products."""
        tr, tc = (int(x) for x in target_shape)
        dr, dc = (int(x) for x in data_shape)
        t = np.ascontiguousarray(target, dtype=np.float64)
        batch = self.number_of_records // (dr * dc) if dr * dc else 0
        out = np.zeros(max(batch * dr * tc, 1), dtype=np.float64)
        _ck(lib().orc_matmul_dot(self._p, self._n, t.ctypes.data, tr, tc, dr, dc, out.ctypes.data), "matmul")
        return out[: batch * dr * tc]


class _Node(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_uint32), ("literal", C.c_double)]


def aggregate(files, nodes, kind, bitmask=None):
    """Aggregation::compute of one aggregation over several EBI1 files:
    nodes = postfix [(kind, index, literal)], kind 0 none / 1 max / 2 min /
    3 sum.  Returns (result, compensated_sum) -- result an array for kind 0."""
    bufs = [np.frombuffer(f, dtype=np.uint8) for f in files]
    ptrs = (C.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
    lens = (C.c_size_t * len(bufs))(*[len(b) for b in bufs])
    arr = (_Node * len(nodes))(*[_Node(int(k), int(i), float(v)) for k, i, v in nodes])
    n_max = max(OracleDecoder(f).number_of_records for f in files)
    words = None
    if bitmask is not None:
        words = np.zeros((n_max + 63) // 64 + 1, dtype=np.uint64)
        ids = np.asarray(list(bitmask), dtype=np.int64)
        if ids.size:
            np.bitwise_or.at(words, ids >> 6, np.left_shift(np.uint64(1), (ids & 63).astype(np.uint64)))
    out = np.zeros(max(n_max, 1), dtype=np.float64)
    sc, comp, n = C.c_double(), C.c_double(), C.c_size_t()
    _ck(lib().orc_aggregate(ptrs, lens, len(bufs), arr, len(nodes), int(kind),
                            None if words is None else words.ctypes.data, C.byref(sc), C.byref(comp),
                            out.ctypes.data, out.size, C.byref(n)), "aggregate")
    if kind == 0:
        return out[: n.value].copy(), float(comp.value)
    return float(sc.value), float(comp.value)
