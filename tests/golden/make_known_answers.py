"""Writes tests/golden/known_answers.json: the known-answer vectors of the
reference's own tests, transcribed as data (inputs and expected outputs).

Sources (lemolatoon/ebi):
  filter          ebi/tests/query.rs:285-760
  materialize     ebi/tests/query.rs:1116-1199
  sum / max / min ebi/tests/aggregation.rs:195-251, 303-316, 351-364
  matmul_cuda     ebi/tests/ml_query.rs:455-576 (run on uncompressed, rle,
                  gorilla, chimp, chimp128, elf: ml_query.rs:605-617)
  knn1            ebi/tests/ml_query.rs:266-290 (same codecs, and ElfOnChimp)
  matmul (CPU)    ebi/tests/ml_query.rs:292-450 (target column-first)
  expr            ebi/tests/aggregation.rs:472-750 (uncompressed, 3 records
                  per chunk: aggregation.rs:172-188)
The reference runs the filter / materialize vectors on every codec, with BUFF
at scale 100 (filters) or 10 (materialize) and DeltaSprintz at scale 10
(query.rs:1280-1367); aggregations use scale 100 for the scaled codecs.
Run: python tests/golden/make_known_answers.py
"""
import json
from pathlib import Path

R = lambda s, e: {"kind": 2, "start": s, "end": e}  # noqa: E731
INC = lambda v: [1, v]  # noqa: E731
EXC = lambda v: [2, v]  # noqa: E731
NONE = [0, 0.0]
EQ = lambda v: {"kind": 0, "value": v}  # noqa: E731
NE = lambda v: {"kind": 1, "value": v}  # noqa: E731
RC = lambda n: [1, n]  # noqa: E731
BS = lambda n: [2, n]  # noqa: E731

SIMD = [1.0] * 5 + [2.0] * 15 + [10.5, 10.6, 10.7, 10.9, 10.9] + [100.0] * 5 + [10.0] * 5
V7 = [3.3, 5.6, 3.8, 6.7, 2.1, 9.9, 10.1]
V6 = [14.22, 5.57, 6.61, 9.25, 2.72, 8.23, 13.2, -0.68, 12.95, 13.85, 12.21, 10.36, 2.33,
      8.5, 18.13, 19.54, 16.42, 19.12, 11.68, 1.78, 7.73, -2.19, 17.43, 21.88, 3.01,
      10.76, 25.08, 4.01, 8.84, 23.35, -4.32, 12.5, 5.72, -7.0, 14.02, -1.23, 19.57,
      6.73, 8.4, 11.27, 15.61, -3.01, -15.35, 14.38, -3.09, 19.94, 23.16, 10.28, 1.46,
      9.2, 20.33, 8.93, 2.18, 1.54, -0.48, -4.13, 16.71, 3.71, 9.09, -10.67, 13.75,
      14.24, 10.93, 8.0, 23.34, 8.92, 1.81, 5.89, 13.97, -7.69, 18.02, 4.85, 4.21, 17.07,
      -0.65, 11.39, 10.36, 17.98, 4.27, 2.95, 18.71, 0.72, 1.85, -5.6, 6.43, 12.12,
      -2.33, -0.91, 14.51, 5.36, 5.49, 5.11, 10.65, 0.39, -1.55, 10.05, -2.89, 21.09,
      -3.97, 7.83, 22.17, 20.66, 7.24, 15.89, 1.96, 13.53, 10.87, 9.67, 12.27, 12.62,
      3.52, 24.96, 13.25, 25.06]
E6 = [0, 1, 2, 3, 4, 5, 6, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22, 23, 24,
      25, 26, 27, 28, 29, 31, 32, 34, 36, 37, 38, 39, 40, 43, 45, 46, 47, 48, 49, 50, 51,
      52, 53, 56, 57, 58, 60, 61, 62, 63, 64, 65, 66, 67, 68, 70, 71, 72, 73, 75, 76, 77,
      78, 79, 80, 81, 82, 84, 85, 88, 89, 90, 91, 92, 93, 95, 97, 99, 100, 101, 102, 103,
      104, 105, 106, 107, 108, 109, 110, 111, 112, 113]


def f(name, values, chunk, pred, expected, bitmask=None, chunk_id=None):
    return {"name": name, "values": values, "chunk": chunk, "pred": pred, "bitmask": bitmask,
            "chunk_id": chunk_id, "expected": expected}


FILTER = [
    f("range incl-none negative", [-300.3, -5.6, -3.8, -699.7, 2.1, 9.9, 100.1], RC(5), R(INC(-3.8), NONE), [2, 4, 5, 6]),
    f("range incl-incl", [-300.3, -5.6, -3.8, -699.7, 2.1, 9.9, 100.1], RC(5), R(INC(-3.8), INC(9.9)), [2, 4, 5]),
    f("range excl-incl", V7, RC(5), R(EXC(3.8), INC(9.9)), [1, 3, 5]),
    f("range incl-excl", V7, RC(5), R(INC(3.8), EXC(9.9)), [1, 2, 3]),
    f("range excl-excl", V7, RC(5), R(EXC(3.8), EXC(9.9)), [1, 3]),
    f("range excl-none", [3.3, 52.6, 3.8, 6.7, 2.1, 9.9, 100.1], RC(5), R(EXC(3.8), NONE), [1, 3, 5, 6]),
    f("range incl-none", V7, RC(5), R(INC(3.8), NONE), [1, 2, 3, 5, 6]),
    f("range none-excl", V7, RC(5), R(NONE, EXC(9.9)), [0, 1, 2, 3, 4]),
    f("range none-none with bitmask", V7, RC(5), R(NONE, NONE), [0, 1, 3, 4, 6], bitmask=[0, 1, 3, 4, 6]),
    f("range large", [9.9, 10.1], RC(5), R(INC(-1000.0), INC(10000000.0)), [0, 1]),
    f("range simd incl-none", SIMD, RC(35), R(INC(10.7), NONE), list(range(22, 30))),
    f("range simd none-excl", SIMD, RC(35), R(NONE, EXC(10.7)), list(range(0, 22)) + list(range(30, 35))),
    f("eq", [322.3, 5204029348.6, 3.8, -6.7, 2.1, 99.9, 1000220.1, 3.8], RC(5), EQ(3.8), [2, 7]),
    f("ne", [3.7, 522.6, 3.8, 6099.7, -22390.1, 908.9, -10.1, 3.8], RC(5), NE(3.8), [0, 1, 3, 4, 5, 6]),
    f("eq simd", SIMD, RC(35), EQ(10.5), [20]),
    f("ne simd", SIMD, RC(35), NE(10.5), list(range(0, 20)) + list(range(21, 35))),
    f("ne simd random failure", [-7072.1] * 32 + [-7071.2], RC(35), NE(-7071.2), list(range(32))),
    f("range low selectivity 2 chunks", ([41.4] * 15 + [41.5] * 15) * 2, RC(33), R(INC(41.4), NONE), list(range(60))),
    f("range random failure bytesize",
      [5714.14, 5710.09, 5718.45, 5712.45, 5714.98, 5715.13, 5717.12, 5714.03, 5712.14, 5714.21, 5716.51],
      BS(5274), R(INC(5715.88), NONE), [10], bitmask=[0, 3, 5, 8, 9, 10]),
    f("range random failure 2", V6, RC(50), R(INC(0.37), NONE), E6),
    f("with bitmask", [3.3, 5.6, 3.8, 6.7, 2.1, 9.9, 10.1, 3.8], RC(5), R(EXC(3.8), NONE), [3, 6],
      bitmask=[0, 2, 3, 4, 6, 7]),
    f("with bitmask ne", [3.3, 5.6, 3.8, 3.3, 2.1, 9.9, 10.1, 3.8], RC(5), NE(3.3), [2, 4, 6, 7],
      bitmask=[0, 2, 3, 4, 6, 7]),
    f("with chunk_id", [3.3, 5.6, 3.8, 6.7, 2.1, 9.9, 10.1, 3.8], RC(3), R(INC(3.8), EXC(9.9)), [7], chunk_id=2),
    f("with bitmask and chunk_id", [3.3, 5.6, 3.8, 6.7, 2.1, 9.9, 10.1, 3.8], RC(3), R(EXC(3.8), INC(9.9)), [3, 5],
      bitmask=[0, 2, 3, 5, 7], chunk_id=1),
]

MATERIALIZE = [
    {"name": "no bitmask", "values": [33.3, 5.6, 3.8, 6.7, 2.1, 9.9, 10.1], "chunk": RC(5), "bitmask": None,
     "chunk_id": None, "expected": [33.3, 5.6, 3.8, 6.7, 2.1, 9.9, 10.1]},
    {"name": "almost the same values", "values": [0.4, 0.4, 0.5], "chunk": RC(5), "bitmask": None,
     "chunk_id": None, "expected": [0.4, 0.4, 0.5]},
    {"name": "with bitmask", "values": [33.3, 5.6, 3.8, 6.7, 2.1, 9.9, 10.1], "chunk": RC(5),
     "bitmask": [0, 2, 3, 5], "chunk_id": None, "expected": [33.3, 3.8, 6.7, 9.9]},
    {"name": "with chunk_id", "values": [33.3, 5.6, 3.8, 6.7, 2.1, 9.9, 10.1, 200.3], "chunk": RC(3),
     "bitmask": None, "chunk_id": 2, "expected": [10.1, 200.3]},
    {"name": "with bitmask and chunk_id", "values": [33.3, 5.6, 3.8, 67.7, -2.1, 9999.9, 10.1], "chunk": RC(3),
     "bitmask": [0, 2, 3, 5], "chunk_id": 1, "expected": [67.7, 9999.9]},
]

AGGREGATE = [
    {"name": "sum base0", "op": "sum", "values": [0.0, 2.0, 3.0, 4.0, 5.0], "chunk": RC(8), "bitmask": None,
     "expected": 14.0},
    {"name": "sum bitmask", "op": "sum", "values": [11.0, 12.0], "chunk": RC(8), "bitmask": [0], "expected": 11.0},
    {"name": "sum multi-subcolumn", "op": "sum", "values": [0.0, 2.0, 3.0, 4.0, 500.0], "chunk": RC(8),
     "bitmask": None, "expected": 509.0},
    {"name": "sum base -1", "op": "sum", "values": [-1.0, 2.0, 3.0, 4.0, 5.0], "chunk": RC(8), "bitmask": None,
     "expected": 13.0},
    {"name": "max", "op": "max", "values": [-1.0, 2.0, 3.0, 4.0, 5.0], "chunk": RC(8), "bitmask": None,
     "expected": 5.0},
    {"name": "min", "op": "min", "values": [-1.0, 2.0, 3.0, 4.0, 5.0], "chunk": RC(8), "bitmask": None,
     "expected": -1.0},
]


def mm(name, data, chunk, target, expected):
    """data: batch x m x k, target: k x n, expected: batch x m x n."""
    return {"name": name, "data": data, "chunk": RC(chunk), "target": target, "expected": expected}


MATMUL = [
    mm("Test 1: matmul m=2 k=3 n=4",
       [[[-1.0, -2.0, 3.0], [-3.0, 3.0, 1.0]], [[-3.0, -1.0, 0.0], [2.0, 0.0, -3.0]]], 12,
       [[1.0, -2.0, -1.0, -1.0], [2.0, -2.0, 1.0, 0.0], [0.0, 1.0, 1.0, -2.0]],
       [[[-5.0, 9.0, 2.0, -5.0], [3.0, 1.0, 7.0, 1.0]], [[-5.0, 8.0, 2.0, 3.0], [2.0, -7.0, -5.0, 4.0]]]),
    mm("Test 2: matmul m=3 k=2 n=2", [[[1.0, 1.0], [-3.0, -1.0], [0.0, -1.0]]], 6,
       [[0.0, 2.0], [-2.0, 1.0]], [[[-2.0, 3.0], [2.0, -7.0], [2.0, -1.0]]]),
    mm("Test 3: matmul m=1 k=2 n=3", [[[-1.0, -3.0]], [[3.0, -1.0]], [[2.0, 0.0]]], 6,
       [[-2.0, 0.0, 1.0], [0.0, -2.0, -1.0]], [[[2.0, 6.0, 2.0]], [[-6.0, 2.0, 4.0]], [[-4.0, 0.0, 2.0]]]),
    mm("Test 4: matmul m=2 k=2 n=1", [[[1.0, 3.0], [-1.0, -1.0]], [[2.0, -1.0], [-2.0, -2.0]]], 8,
       [[0.0], [-1.0]], [[[-3.0], [1.0]], [[1.0], [2.0]]]),
    mm("Test 5: matmul m=1 k=3 n=2", [[[1.0, -2.0, 0.0]], [[3.0, 1.0, -3.0]], [[-2.0, 3.0, 1.0]], [[-3.0, 1.0, 3.0]]],
       12, [[-2.0, -1.0], [2.0, 0.0], [1.0, -2.0]], [[[-6.0, -1.0]], [[-7.0, 3.0]], [[11.0, 0.0]], [[11.0, -3.0]]]),
    mm("Test 6: matmul m=2 k=2 n=5",
       [[[1.0, -2.0], [1.0, 0.0]], [[1.0, -1.0], [-1.0, -1.0]], [[0.0, 2.0], [0.0, 2.0]]], 12,
       [[1.0, 2.0, 0.0, 2.0, 0.0], [2.0, 0.0, 1.0, 1.0, 0.0]],
       [[[-3.0, 2.0, -2.0, 0.0, 0.0], [1.0, 2.0, 0.0, 2.0, 0.0]],
        [[-1.0, 2.0, -1.0, 1.0, 0.0], [-3.0, -2.0, -1.0, -3.0, 0.0]],
        [[4.0, 0.0, 2.0, 2.0, 0.0], [4.0, 0.0, 2.0, 2.0, 0.0]]]),
]
MATMUL_CODECS = ["uncompressed", "rle", "gorilla", "chimp", "chimp128", "elf"]

# knn1: data, chunk records, target, expected index (ml_query.rs:266-290)
KNN_DATA = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
KNN1 = [{"name": f"Test {t}: {i}", "data": KNN_DATA, "chunk": RC(c), "target": KNN_DATA[2 * i:2 * i + 2],
         "expected": i} for t, c in ((1, 2), (2, 3)) for i in range(3)]


def mmc(name, data, chunk, target, expected):
    """data: batch x H x W; target: columns (target_w lists of target_h
    values, column-first); expected: batch x H x target_w."""
    return {"name": name, "data": data, "chunk": RC(chunk), "target": target, "expected": expected}


MATMUL_CPU = [
    mmc("Test 1: Simple matmul", [[[1.0, 2.0], [3.0, 4.0], [0.0, 1.0]], [[3.0, 2.0], [-1.0, 8.0], [0.0, 1.0]]], 6,
        [[1.0, 1.0], [5.0, 4.0], [6.0, 3.0], [2.0, 1.0]],
        [[[3.0, 13.0, 12.0, 4.0], [7.0, 31.0, 30.0, 10.0], [1.0, 4.0, 3.0, 1.0]],
         [[5.0, 23.0, 24.0, 8.0], [7.0, 27.0, 18.0, 6.0], [1.0, 4.0, 3.0, 1.0]]]),
    mmc("Test 2: Updated matmul with four matrices",
        [[[1.0, 2.0], [3.0, 4.0], [0.0, 1.0]], [[3.0, 2.0], [-1.0, 8.0], [0.0, 1.0]],
         [[2.0, 1.0], [0.0, 3.0], [1.0, 0.0]], [[4.0, -1.0], [2.0, 2.0], [0.0, 0.5]]], 6,
        [[1.0, 1.0], [5.0, 4.0], [6.0, 3.0], [2.0, 1.0]],
        [[[3.0, 13.0, 12.0, 4.0], [7.0, 31.0, 30.0, 10.0], [1.0, 4.0, 3.0, 1.0]],
         [[5.0, 23.0, 24.0, 8.0], [7.0, 27.0, 18.0, 6.0], [1.0, 4.0, 3.0, 1.0]],
         [[3.0, 14.0, 15.0, 5.0], [3.0, 12.0, 9.0, 3.0], [1.0, 5.0, 6.0, 2.0]],
         [[3.0, 16.0, 21.0, 7.0], [4.0, 18.0, 18.0, 6.0], [0.5, 2.0, 1.5, 0.5]]]),
    mmc("Test 3: Updated matmul with 5x3 matrices and target 3x9",
        [[[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 9.0], [0.0, 1.0, 2.0], [3.0, 4.0, 5.0]],
         [[2.0, 1.0, 0.0], [3.0, 2.0, 1.0], [4.0, 3.0, 2.0], [5.0, 4.0, 3.0], [6.0, 5.0, 4.0]],
         [[0.0, 1.0, 2.0], [1.0, 0.0, 3.0], [2.0, 1.0, 4.0], [3.0, 2.0, 5.0], [4.0, 3.0, 6.0]],
         [[3.0, 2.0, 1.0], [4.0, 3.0, 2.0], [5.0, 4.0, 3.0], [6.0, 5.0, 4.0], [7.0, 6.0, 5.0]]], 30,
        [[float(k * j) for j in (1, 2, 3)] for k in range(1, 10)],
        [[[14.0 * m for m in range(1, 10)], [32.0 * m for m in range(1, 10)], [50.0 * m for m in range(1, 10)],
          [8.0 * m for m in range(1, 10)], [26.0 * m for m in range(1, 10)]],
         [[4.0 * m for m in range(1, 10)], [10.0 * m for m in range(1, 10)], [16.0 * m for m in range(1, 10)],
          [22.0 * m for m in range(1, 10)], [28.0 * m for m in range(1, 10)]],
         [[8.0 * m for m in range(1, 10)], [10.0 * m for m in range(1, 10)], [16.0 * m for m in range(1, 10)],
          [22.0 * m for m in range(1, 10)], [28.0 * m for m in range(1, 10)]],
         [[10.0 * m for m in range(1, 10)], [16.0 * m for m in range(1, 10)], [22.0 * m for m in range(1, 10)],
          [28.0 * m for m in range(1, 10)], [34.0 * m for m in range(1, 10)]]]),
]

# expressions in postfix form: ["i", k] column k, ["l", v] literal, ["+"|"-"|"*"|"/"] operators
def ex(name, columns, aggs, expected, bitmask=None):
    return {"name": name, "columns": columns, "aggs": aggs, "expected": expected, "bitmask": bitmask}


A01 = [["i", 0], ["i", 1], ["*"], ["i", 1], ["+"]]
EXPR = [
    ex("test_expr1", [[0.0, 2.0, 3.0, 4.0, 5.0]], [["sum", [["i", 0]]]], [14.0]),
    ex("test_expr2", [[0.0, 2.0, -3.0, 4.0, 7.0], [1.0, 2.0, 3.0, 4.0, 5.0]], [["sum", A01]], [61.0]),
    ex("test_expr3", [[0.0, 2.0, 3.0, 4.0, 5.0]], [["sum", [["i", 0]]]], [14.0]),
    ex("test_expr4", [[0.0, 2.0, -3.0, 4.0, 7.0], [1.0, 2.0, 3.0, 4.0, 5.0]], [["sum", A01]], [61.0]),
    ex("test_expr5", [[1.0, 5.0, 2.0, 4.0]], [["max", [["i", 0]]]], [5.0]),
    ex("test_expr6", [[1.0, 2.0, 3.0], [2.0, 1.0, 4.0]], [["max", [["i", 0], ["i", 1], ["*"]]]], [12.0]),
    ex("test_expr7", [[1.0, 5.0, 2.0, 4.0]], [["min", [["i", 0]]]], [1.0]),
    ex("test_expr8", [[3.0, 1.0, 4.0], [2.0, 2.0, 1.0]], [["min", [["i", 0], ["i", 1], ["-"]]]], [-1.0]),
    ex("test_expr9", [[0.0, 0.8, 0.3], [1.0, 1.0, 20.0]], [["sum", [["i", 0]]]], [0.8], bitmask=[0, 1]),
    ex("test_expr10", [[1.0, 2.2, 5.0, 4.0]], [["sum", [["i", 0]]]], [6.0], bitmask=[0, 2]),
    ex("test_multiple_sum_index_and_binary", [[1.0, 2.0, 3.0], [10.0, 20.0, 30.0]],
       [["sum", [["i", 0]]], ["sum", [["i", 0], ["i", 1], ["+"]]]], [6.0, 66.0]),
    ex("test_multiple_max_index_and_binary", [[3.0, 8.0, 5.0], [2.0, 1.0, 7.0]],
       [["max", [["i", 1]]], ["max", [["i", 0], ["i", 1], ["*"]]]], [7.0, 35.0]),
    ex("test_multiple_min_index_and_binary", [[6.0, 2.0, 9.0], [1.0, 2.0, 3.0]],
       [["min", [["i", 0]]], ["min", [["i", 0], ["i", 1], ["-"]]]], [2.0, 0.0]),
    ex("test_multiple_count_index_and_binary", [[1.0, 0.0, 3.0, 4.0], [5.0, 6.0, 7.0, 8.0]],
       [["sum", [["i", 0]]], ["max", [["i", 0], ["i", 1], ["+"]]]], [4.0, 10.0], bitmask=[0, 2]),
]
EXPR_CHUNK = RC(3)
ML_CODECS = ["uncompressed", "rle", "gorilla", "chimp", "chimp128", "elf_on_chimp", "elf"]

# scales the reference tests use per codec
SCALES = {"filter": {"buff": 100, "delta_sprintz": 10}, "materialize": {"buff": 10, "delta_sprintz": 10},
          "aggregate": {"buff": 100, "delta_sprintz": 100}}

if __name__ == "__main__":
    out = Path(__file__).resolve().parent / "known_answers.json"
    out.write_text(json.dumps({"filter": FILTER, "materialize": MATERIALIZE, "aggregate": AGGREGATE,
                               "matmul": MATMUL, "matmul_codecs": MATMUL_CODECS, "scales": SCALES,
                               "knn1": KNN1, "matmul_cpu": MATMUL_CPU, "expr": EXPR, "expr_chunk": EXPR_CHUNK,
                               "ml_codecs": ML_CODECS}, indent=1))
    print(out)
