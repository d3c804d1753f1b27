#!/usr/bin/env python3
"""Benchmark of the MI355X chunked-decode path (one JSON line on stdout).

Workload (one "step"): one batched persistent launch that decodes every
resident column of the plain-decode workload -- 2^28 f64 values per column, four generators (2-decimal random
walk, 3-decimal uniform[-1000, 1000], 2-decimal step series with runs of
length ~Geom(0.01), 4-decimal sinusoid + N(0, 0.01) noise), each encoded with
the eight codecs (Uncompressed, RLE, Gorilla, Chimp, Chimp128, Elf, BUFF,
DeltaSprintz) at 65,536 records per chunk -- into HBM.  Columns are encoded
on the host and loaded once before the timed region; the timed region
decodes all 32 columns K times.

value = algorithmic bytes (compressed chunk bytes + 8 bytes per decoded
record, summed over all columns) / wall time of the step, in GB/s.

Also reported: the roofline of the batched step's dominant kernel
(algorithmic bytes of the columns it decodes / its mean launch time from HIP
events on the library's stream, around each launch of the timed steps),
per-codec GB/s from one prepared batch per codec, the plain per-chunk-launch
comparator on Gorilla and Chimp, and the CPU oracle timed on a bounded sample
of the same workload on one host core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# kernels of each codec in a batch: (reads the compressed stream, writes the
# decoded records); the XOR codecs: k_decode_seg; the rest: k_decode
KERNELS_OF = {"chimp128": ("k_decode_seg", "k_decode_seg"), "gorilla": ("k_decode_seg", "k_decode_seg"),
              "chimp": ("k_decode_seg", "k_decode_seg"), "elf": ("k_decode_seg", "k_decode_seg")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2n", type=int, default=28, help="log2 of values per column")
    ap.add_argument("--chunk", type=int, default=65536)
    ap.add_argument("--codecs", default="all")
    ap.add_argument("--gens", default="all")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-query", action="store_true", help="skip the in-situ query latency leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (Poisson stream) leg")
    ap.add_argument("--no-matmul", action="store_true", help="skip the decompress-then-DGEMM leg")
    ap.add_argument("--serve-rates", default="2000,5000,10000,20000,50000", help="Poisson arrival rates, queries/s")
    ap.add_argument("--serve-queries", type=int, default=1000, help="queries per arrival rate")
    ap.add_argument("--serve-log2max", type=int, default=26, help="largest serving column, log2 values")
    ap.add_argument("--matmul-n", type=int, default=8192)
    ap.add_argument("--no-tune", action="store_true",
                    help="use the context's launch shape (EBI_HIP_* overrides) instead of autotuning the batch, "
                         "so a profiler sees only the timed launches")
    ap.add_argument("--no-graph", action="store_true", help="skip the hipGraph replays after the timed region")
    ap.add_argument("--serve-budget-gib", type=int, default=16, help="server device arena (query scratch), GiB")
    ap.add_argument("--no-breakdown", action="store_true",
                    help="skip the per-column breakdown and the comparator (after the timed region), so a "
                         "profiler sees only the batched launches")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import ebi_amd as E
    from ebi_amd import datagen as G

    n = 1 << args.log2n
    codecs = G.CODECS if args.codecs == "all" else args.codecs.split(",")
    gens = list(G.GENERATORS) if args.gens == "all" else args.gens.split(",")
    threads = min(16, os.cpu_count() or 1)

    ctx = E.Context(0)
    cols = []  # (gen, codec, decoder, compressed_bytes)
    cpu = {"seconds": 0.0, "bytes": 0, "values": 0, "chunks": 0,
           "par_seconds": 0.0, "par_bytes": 0, "par_chunks": 0, "par_threads": cpu_threads(), "core": None}
    sample_q = {}  # generator -> every 64th value, for predicate quantiles
    cold = {}  # codec -> the random walk column's cold load and first decode
    checks = []  # (column, first record, oracle's records): sampled chunks of every column
    rng_check = np.random.default_rng(5)
    t_setup = time.time()
    for g in gens:
        fn, dec = G.GENERATORS[g]
        vals = fn(n)
        sample_q[g] = np.sort(vals[::64])
        for codec in codecs:
            cfg = E.CompressorConfig.by_name(codec, scale=10 ** dec)
            data = E.encode(vals, cfg, E.ChunkOption.record_count(args.chunk), threads=threads)
            t = time.perf_counter()
            d = E.Decoder(ctx, data)
            load_ms = (time.perf_counter() - t) * 1e3
            phases_ms = [ctx.last_phase_ms(i) for i in range(3)]
            cols.append((g, codec, d, d.compressed_bytes))
            if g == "walk2":
                cold[codec] = cold_entry(E, ctx, d, n, load_ms, phases_ms)
            checks += sampled_chunks(data, len(cols) - 1, n, args.chunk, rng_check)
            log(f"[setup] {g:6s} {codec:14s} {len(data) / n:6.3f} B/value  loaded")
            if not args.no_cpu:
                cpu_sample(data, n, args.chunk, cpu, args.cpu_seconds / (len(gens) * len(codecs)))
            del data
        del vals
    log(f"[setup] {len(cols)} columns resident in {time.time() - t_setup:.1f}s")

    total_bytes = sum(cb + 8 * n for _, _, _, cb in cols)
    outs = [ctx.alloc(8 * n) for _ in cols]
    decs = [d for _, _, d, _ in cols]

    # prepared batch: arguments uploaded once, launch shape autotuned on this
    # batch (before the timed region), replayed every step
    t_tune = time.perf_counter()
    batch = E.Batch(ctx, decs, outs, autotune=not args.no_tune)
    tune_s = time.perf_counter() - t_tune
    log(f"[setup] batch tuned in {tune_s:.1f}s: {batch.config}")

    def step():
        # one batched decode over every chunk of every column (mixed codecs):
        # k_decode_mix over the tile jobs of all columns and the XOR
        # checkpoint segments (or, when the tuner finds it faster, k_decode
        # then k_decode_seg), launched eagerly so each kernel is bracketed by
        # HIP events
        batch.run(eager=True)
        return ctx.last_device_ms, ctx.last_phase_ms(0), ctx.last_phase_ms(1), ctx.last_phase_ms(2)

    for _ in range(args.warmup):
        step()
    batch_dev_ms, phase_ms = [], [[], [], []]
    t0 = time.perf_counter()
    for s in range(args.steps):
        tot, p0, p1, p2 = step()
        batch_dev_ms.append(tot)
        phase_ms[0].append(p0)
        phase_ms[1].append(p1)
        phase_ms[2].append(p2)
        log(f"[timed] step {s + 1}/{args.steps} done")
    t1 = time.perf_counter()
    output_check = check_outputs(outs, checks)
    log(f"[check] {output_check}")
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = total_bytes / (ms_per_step / 1000.0) / 1e9
    graph_ms = []
    for _ in range(0 if args.no_graph else 2):  # the same batch as one hipGraph launch (after the timed region)
        t = time.perf_counter()
        batch.run()
        graph_ms.append((time.perf_counter() - t) * 1e3)
    batch_cfg = batch.config
    batch.close()

    # roofline of the batched step's kernels: algorithmic bytes of the columns
    # each kernel decodes / its mean launch time (HIP events on the library's
    # stream, around each launch of the timed steps)
    kernels_of = {c: KERNELS_OF.get(c, ("k_decode", "k_decode")) for c in codecs}
    kernel_of = {c: "+".join(dict.fromkeys(k)) for c, k in kernels_of.items()}  # per-codec grouping
    phases = ("k_decode", "", "k_decode_seg")  # phase 1 is unused
    if batch_cfg.get("mixed"):  # both job lists in one k_decode_mix launch (phase 0)
        kernels_of = {c: ("k_decode_mix", "k_decode_mix") for c in codecs}
        phases = ("k_decode_mix", "", "")
    batch_kernels = {}
    for k, name in enumerate(phases):
        # algorithmic bytes: the compressed stream to its reader, 8 B a record to its writer
        b = sum((cb if kernels_of[codec][0] == name else 0) + (8 * n if kernels_of[codec][1] == name else 0)
                for _, codec, _, cb in cols)
        if b:
            ms = float(np.mean(phase_ms[k]))
            batch_kernels[name] = {"bytes_per_launch": b, "mean_launch_ms": round(ms, 4),
                                   "gbs": round(b / (ms / 1000.0) / 1e9, 2)}
    dominant = max(batch_kernels, key=lambda k: batch_kernels[k]["mean_launch_ms"])
    achieved = batch_kernels[dominant]["gbs"]

    per_codec, per_kernel, comparator = {}, {}, {}
    if not args.no_breakdown:
        per_codec, per_kernel = breakdown(E, ctx, cols, decs, outs, n, kernel_of)
        comparator = run_comparator(E, ctx, cols, n)
    for o in outs:
        o.free()
    outs = []
    query = serving = matmul = c1 = None
    cold_obj = {"workload": f"2^{args.log2n} f64 2-decimal random walk per codec, {args.chunk} records/chunk, "
                            "file bytes in host memory", "codecs": cold} if cold else None
    if not args.no_query:
        query = query_leg(E, ctx, cols, sample_q, n)
        c1 = c1_leg(E, G, ctx)
    if not args.no_serve:
        serving = serve_leg(E, G, ctx, args, threads)
    if not args.no_matmul:
        matmul = matmul_leg(E, ctx, args.matmul_n, threads)

    workload = (f"plain decode: 2^{args.log2n} f64 x {len(gens)} generators x {len(codecs)} codecs, "
                f"{args.chunk} records/chunk, all columns per step")
    res = {
        "metric": "decode GB/s (% HBM peak) + in-situ query p50/p99 vs per-chunk-launch form",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 generators, encoded on the host)",
        "config": {
            "workload": workload,
            "columns": len(cols),
            "values_per_column": n,
            "records_per_chunk": args.chunk,
            "algorithmic_bytes_per_step": total_bytes,
            "hbm_frac_of_value": round(value / HBM_PEAK_GBS, 4),
            "batch_device_ms": round(float(np.mean(batch_dev_ms)), 3),
            "batch_launch_shape": batch_cfg,
            "batch_tuning_s": round(tune_s, 2),
            "graph_replay_ms": round(float(np.min(graph_ms)), 3) if graph_ms else None,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_of(workload, dominant),
            "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE passes of this workload, calibrated on a copy "
                              "(profiles/hbm_traffic.json); null when the workload differs",
        },
        "batch_kernels": batch_kernels,
        "per_codec_gbs": {c: round(v["bytes"] / (v["ms"] / 1000.0) / 1e9, 2) for c, v in per_codec.items()},
        "per_codec_tuned_shape": {c: v["shape"] for c, v in per_codec.items()},
        "per_kernel": {k: {"gbs": round(v["bytes"] / (v["ms"] / 1000.0) / 1e9, 2),
                           "ms_per_step": round(v["ms"], 3), "launches": v["launches"]}
                       for k, v in per_kernel.items()},
        "per_chunk_launch_comparator": comparator,
        "output_check": output_check,
        "query": query,
        "cold": cold_obj,
        "cpu_plain_case": c1,
        "serving": serving,
        "matmul": matmul,
    }
    if not args.no_cpu and cpu["seconds"] > 0:
        res["cpu_baseline"] = {
            "value": round(cpu["bytes"] / cpu["seconds"] / 1e9, 4),
            "unit": "GB/s",
            "cores": 1,
            "kind": "port",
            "sample": f"oracle decode of the first {cpu['chunks']} chunks across the {len(cols)} columns "
                      f"({cpu['values']} values, {cpu['seconds']:.1f}s on one core)",
            "pinned_core": cpu["core"],
            "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count(),
        }
        if cpu["par_seconds"] > 0:  # secondary: chunk-parallel over the host cores this process may use
            res["cpu_baseline"]["chunk_parallel"] = {
                "value": round(cpu["par_bytes"] / cpu["par_seconds"] / 1e9, 4), "unit": "GB/s",
                "cores": cpu["par_threads"],
                "sample": f"oracle decode of {cpu['par_chunks']} chunks across the {len(cols)} columns, "
                          f"one chunk per task on {cpu['par_threads']} threads ({cpu['par_seconds']:.1f}s wall)"}
        res["gpu_vs_cpu"] = round(value / res["cpu_baseline"]["value"], 1)
    for _, _, d, _ in cols:
        d.close()
    ctx.close()
    print(json.dumps(res), flush=True)


def pct(xs, q):
    return round(float(np.percentile(np.asarray(xs), q)), 4) if len(xs) else None


def sampled_chunks(data, col, n, chunk, rng, per_column=2):
    """Two random chunks of a column decoded by the oracle (the checker)."""
    from oracle.oracle import OracleDecoder

    od = OracleDecoder(data)
    nch = (n + chunk - 1) // chunk
    out = []
    for k in rng.choice(nch, min(per_column, nch), replace=False):
        out.append((col, int(k) * chunk, np.asarray(od.materialize(chunk_id=int(k)))))
    return out


def check_outputs(outs, checks):
    """The timed step's decoded columns against the oracle on the sampled
    chunks: f64 bit patterns compared."""
    mism, recs = 0, 0
    for ci, r0, want in checks:
        got = outs[ci].to_host(count=len(want), offset=r0)
        recs += len(want)
        mism += int(np.count_nonzero(got.view(np.uint64) != want.view(np.uint64)))
    return {"chunks": len(checks), "records": recs, "mismatched_records": mism, "bit_exact": mism == 0,
            "checker": "oracle (C restatement of the reference decoders), 2 random chunks per column"}


def cold_entry(E, ctx, d, n, load_ms, phases_ms):
    """The cold path of one column, as the reference's Decoder::new +
    materialize meets it: the file's bytes in host memory are parsed,
    repacked and uploaded, the load-time index is built (tile prefixes;
    XOR checkpoints by a lane per chunk; the lane-interleaved stream copy;
    Chimp128's ring checkpoints by a wavefront over its jobs), then the first
    decode of the whole column into HBM.  index_bytes_per_record: the
    resident bytes beside the chunk bodies (checkpoints, interleaved copy,
    ring, tile prefixes, descriptors, job lists) per record."""
    out = ctx.alloc(8 * n)
    E.run_query(ctx, E.Query(E.OP_DECODE, d, out=out))
    dec_ms = ctx.last_device_ms
    out.free()
    cb = d.compressed_bytes
    dev_ms = phases_ms[1] + phases_ms[2]
    return {"load_ms": round(load_ms, 2), "upload_ms": round(phases_ms[0], 2), "index_ms": round(phases_ms[1], 2),
            "copy_ring_ms": round(phases_ms[2], 2), "first_decode_ms": round(dec_ms, 3),
            "index_over_decode": round(dev_ms / dec_ms, 2) if dec_ms > 0 else None,
            "index_bytes_per_record": round((d.device_bytes - cb) / n, 3),
            "cold_gbs_device": round((cb + 8 * n) / ((dev_ms + dec_ms) / 1e3) / 1e9, 1)}


QUERY_CODECS = ("buff", "delta_sprintz", "rle", "uncompressed", "gorilla", "chimp", "elf", "chimp128")


def query_leg(E, ctx, cols, sample_q, n, reps=20):
    """In-situ query latency on the resident 2^log2n random-walk columns:
    one-pass range-filter + sum (greater-than at the (1 - s) quantile, s =
    0.1, 1, 10, 50 %) and max, one synchronous query at a time.  BUFF and
    DeltaSprintz run on their codes (BUFF chunks whose range excludes the
    predicate are skipped by the reference's own early exits), RLE per run
    (on the random walk, nearly one record per run, and on the step series),
    Uncompressed in place; Gorilla, Chimp and Elf decode and query in one
    kernel (records stay in registers), Chimp128 decodes into HBM first.
    p50 is host wall time of the call; device_p50 the library's HIP-event
    time of the query's launches (block upload through the result copy).
    GB/s count the column's compressed bytes + 8 (no bitset leaves the
    device in a filter + sum), at wall and at device time; for BUFF's
    filters the bytes of skipped chunks are counted too, so their GB/s are
    effective rates (max reads every chunk)."""
    out = {}
    if sample_q.get("walk2") is None:
        return None

    def run(q):
        E.run_query(ctx, q)
        lat, dev = [], []
        for _ in range(reps):
            t = time.perf_counter()
            E.run_query(ctx, q)
            lat.append((time.perf_counter() - t) * 1e3)
            dev.append(ctx.last_device_ms)
        return lat, dev

    for g, codec, d, cb in cols:
        # every codec on the random walk; RLE also on the step series, whose
        # runs (~Geom(0.01), at most 255) are what RLE is for
        if not ((g == "walk2" and codec in QUERY_CODECS) or (g == "step2" and codec == "rle")):
            continue
        sq = sample_q[g]
        res = {}
        byt = cb + 8
        for s in (0.001, 0.01, 0.1, 0.5):
            thr = float(sq[int((1.0 - s) * (len(sq) - 1))])
            p = E.Predicate.range(E.RangeValue.exclusive(thr), E.RangeValue.none())
            lat, dev = run(E.Query(E.OP_FILTER_SUM, d, predicate=p))
            res[f"filter_sum_sel{s * 100:g}%"] = {
                "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "device_p50_ms": pct(dev, 50),
                "gbs_at_p50": round(byt / (pct(lat, 50) / 1e3) / 1e9, 1),
                "gbs_device": round(byt / (pct(dev, 50) / 1e3) / 1e9, 1)}
        lat, dev = run(E.Query(E.OP_MAX, d))
        res["max"] = {"p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "device_p50_ms": pct(dev, 50),
                      "gbs_at_p50": round(byt / (pct(lat, 50) / 1e3) / 1e9, 1),
                      "gbs_device": round(byt / (pct(dev, 50) / 1e3) / 1e9, 1),
                      "device_hbm_frac": round(byt / (pct(dev, 50) / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
        res["compressed_bytes"] = int(cb)
        out[codec if g == "walk2" else f"{codec}_{g}"] = res
        log(f"[query] {codec} done")
    return {"workload": f"2^{int(np.log2(n))} f64 random walk (2 decimals) per column, one query at a time",
            "codecs": out}


def c1_leg(E, G, ctx, reps=5):
    """The reference's CPU-plain case: 1,048,576-value 2-decimal random walk,
    Gorilla and BUFF (scale 100) at 131,072 records per chunk; full decode,
    range filter [99.5, 100.5], sum over the filtered rows.  Median of 5 on
    the CPU oracle (one core) and on the GPU (synchronous calls, results to
    the host)."""
    from oracle.oracle import OracleDecoder

    vals = G.random_walk(1 << 20, 100.0, 0.1, 2, seed=1)
    pred = E.Predicate.range(E.RangeValue.inclusive(99.5), E.RangeValue.inclusive(100.5))
    out = {}
    for name, cfg in (("gorilla", E.CompressorConfig.gorilla()), ("buff", E.CompressorConfig.buff(scale=100))):
        f = E.encode(vals, cfg, E.ChunkOption.record_count(131072))
        res = {}
        for side, dec in (("cpu", OracleDecoder(f)), ("gpu", E.Decoder(ctx, f))):
            t_dec, t_fil, t_sum = [], [], []
            for _ in range(reps):
                t = time.perf_counter()
                dec.materialize()
                t_dec.append(time.perf_counter() - t)
                t = time.perf_counter()
                ids = dec.filter(pred)
                t_fil.append(time.perf_counter() - t)
                t = time.perf_counter()
                s_ = dec.sum(ids)
                t_sum.append(time.perf_counter() - t)
            res[side] = {"decode_ms": round(float(np.median(t_dec)) * 1e3, 3),
                         "filter_ms": round(float(np.median(t_fil)) * 1e3, 3),
                         "sum_ms": round(float(np.median(t_sum)) * 1e3, 3), "selected": int(ids.size), "sum": s_}
            if side == "gpu":
                dec.close()
        res["cores_cpu"] = 1
        out[name] = res
    return {"workload": "2^20 f64 random walk (2 decimals), 131072 records/chunk", "codecs": out}


def serve_leg(E, G, ctx, args, threads, n_cols=64):
    """Serving: Poisson arrivals over 64 resident 4-decimal random walks
    (sizes log-uniform on [2^20, 2^log2max], codecs uniform over the eight,
    chunk sizes log-uniform on [8192, 262144]); mix 50 % range-filter + sum,
    25 % min / max, 20 % range-filter + materialize, 5 % full-column decode
    into HBM.  Each rate runs through the server and, on the same arrival
    sequence, through the per-query-launch form (one synchronous query at a
    time).  Latency = completion - arrival."""
    rng = np.random.default_rng(7)
    t0 = time.time()
    cols = []
    for i in range(n_cols):
        nv = int(round(2 ** rng.uniform(20, args.serve_log2max)))
        chunk = int(round(2 ** rng.uniform(13, 18)))
        codec = G.CODECS[i % len(G.CODECS)]
        vals = G.random_walk(nv, 100.0, 0.1, 4, seed=100 + i)
        f = E.encode(vals, E.CompressorConfig.by_name(codec, scale=10000), E.ChunkOption.record_count(chunk),
                     threads=threads)
        cols.append((E.Decoder(ctx, f), np.sort(vals[::256]), nv))
        del f, vals
    log(f"[serve] {n_cols} columns resident in {time.time() - t0:.1f}s")
    c3 = batched_queries(E, ctx, cols)
    dec_bufs = [ctx.alloc(8 * max(nv for _, _, nv in cols)) for _ in range(8)]

    def make(k):
        d, sq, nv = cols[int(rng.integers(len(cols)))]
        u = rng.random()
        if u < 0.5 or 0.75 <= u < 0.95:  # range filter + sum, range filter + materialize
            s = float(rng.choice([0.001, 0.01, 0.1]))
            a = float(rng.uniform(0, 1 - s))
            lo, hi = float(sq[int(a * (len(sq) - 1))]), float(sq[int((a + s) * (len(sq) - 1))])
            p = E.Predicate.range(E.RangeValue.inclusive(lo), E.RangeValue.inclusive(hi))
            return E.Query(E.OP_FILTER_SUM if u < 0.5 else E.OP_FILTER_MATERIALIZE, d, predicate=p)
        if u < 0.75:  # min / max
            return E.Query(E.OP_MIN if u < 0.625 else E.OP_MAX, d)
        return E.Query(E.OP_DECODE, d, out=dec_bufs[k % len(dec_bufs)])  # full-column decode into HBM

    rates = [float(x) for x in args.serve_rates.split(",") if x]
    out = {}
    for lam in rates:
        nq = args.serve_queries
        qs_served = [make(k) for k in range(nq)]
        arrivals = np.cumsum(rng.exponential(1.0 / lam, nq))
        # served: submit each query at its arrival time (in batches of those due)
        lat = np.zeros(nq)
        with E.Server(ctx, device_budget=args.serve_budget_gib << 30, max_batch=64, n_streams=4) as srv:
            tickets = [0] * nq
            lag = np.zeros(nq)
            start = time.perf_counter()
            i = 0
            while i < nq:
                now = time.perf_counter() - start
                if arrivals[i] > now:
                    time.sleep(min(arrivals[i] - now, 0.001))
                    continue
                j = i
                while j < nq and arrivals[j] <= now:
                    j += 1
                ts = srv.submit(qs_served[i:j])
                sub = time.perf_counter() - start
                tickets[i:j] = ts
                lag[i:j] = sub - arrivals[i:j]
                i = j
            for k in range(nq):
                _, ms = srv.wait(tickets[k])
                lat[k] = lag[k] * 1e3 + ms
            wall = time.perf_counter() - start
            st = srv.stats()
        # per-query-launch form on the same arrival sequence
        qs_plain = [E.Query(q.op, q.decoder, q.predicate, out=getattr(q, "_out", None)) for q in qs_served]
        lat_p = np.zeros(nq)
        start = time.perf_counter()
        for k, q in enumerate(qs_plain):
            now = time.perf_counter() - start
            if arrivals[k] > now:
                time.sleep(arrivals[k] - now)
            E.run_query(ctx, q)
            lat_p[k] = (time.perf_counter() - start - arrivals[k]) * 1e3
        wall_p = time.perf_counter() - start
        same = all(qs_served[k].op not in (E.OP_FILTER_SUM, E.OP_MIN, E.OP_MAX) or
                   np.float64(qs_served[k].result()).view(np.uint64) == np.float64(qs_plain[k].result()).view(np.uint64)
                   for k in range(nq))
        out[f"{int(lam)}qps"] = {
            "served": {"p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "throughput_qps": round(nq / wall, 1),
                       "batches": st["batches"], "max_inflight": st["max_inflight"]},
            "per_query_launch": {"p50_ms": pct(lat_p, 50), "p99_ms": pct(lat_p, 99),
                                 "throughput_qps": round(nq / wall_p, 1)},
            "scalar_results_equal": bool(same),
        }
        log(f"[serve] {int(lam)} q/s: served p50 {pct(lat, 50)} p99 {pct(lat, 99)} ms; "
            f"per-query p50 {pct(lat_p, 50)} p99 {pct(lat_p, 99)} ms")
        del qs_served, qs_plain
    for b in dec_bufs:
        b.free()
    for d, _, _ in cols:
        d.close()
    return {"workload": f"{n_cols} columns, 4-decimal random walks, 2^20..2^{args.serve_log2max} values, "
                        f"mixed codecs, chunks of 2^13..2^18 records, {args.serve_queries} Poisson arrivals per rate",
            "rates": out, "batched": c3}


def batched_queries(E, ctx, cols):
    """The batched workload over the same 64 resident columns: (1) every column
    decoded by one tuned prepared batch replayed as a hipGraph; (2) per
    selectivity s (0.1, 1, 10, 50 %), on every column: a greater-than filter
    + sum at the (1 - s) quantile, a range filter [q(a), q(a + s)] (the
    experimenter's two-sided predicate), an eq filter on a sampled value,
    max and min, and for s <= 1 % the range filter + materialize (results to
    the host), all submitted together to the server (in-situ on BUFF /
    DeltaSprintz codes, per run on RLE).  GB/s counts the compressed bytes
    each query scans (+ 8 per decoded record for the decode)."""
    decs = [d for d, _, _ in cols]
    nrec = sum(nv for _, _, nv in cols)
    comp = sum(d.compressed_bytes for d in decs)
    outs = [ctx.alloc(8 * nv) for _, _, nv in cols]
    b = E.Batch(ctx, decs, outs, autotune=True)
    b.run()
    ms = []
    for _ in range(5):
        t = time.perf_counter()
        b.run()
        ms.append((time.perf_counter() - t) * 1e3)
    cfg = b.config
    b.close()
    for o in outs:
        o.free()
    res = {"decode": {"graph_ms": round(float(np.median(ms)), 3),
                      "gbs": round((comp + 8 * nrec) / (float(np.median(ms)) / 1e3) / 1e9, 1),
                      "launch_shape": cfg}}
    with E.Server(ctx, device_budget=16 << 30, max_batch=64, n_streams=4) as srv:
        rng = np.random.default_rng(11)
        for sel in (0.001, 0.01, 0.1, 0.5):
            qs, scanned = [], 0
            for d, sq, nv in cols:
                thr = float(sq[int((1.0 - sel) * (len(sq) - 1))])
                a = float(rng.uniform(0, 1 - sel))
                lo, hi = float(sq[int(a * (len(sq) - 1))]), float(sq[int((a + sel) * (len(sq) - 1))])
                gt = E.Predicate.range(E.RangeValue.exclusive(thr), E.RangeValue.none())
                rg = E.Predicate.range(E.RangeValue.inclusive(lo), E.RangeValue.inclusive(hi))
                eq = E.Predicate.eq(float(sq[int(rng.integers(len(sq)))]))
                qs += [E.Query(E.OP_FILTER_SUM, d, predicate=gt), E.Query(E.OP_FILTER_SUM, d, predicate=rg),
                       E.Query(E.OP_FILTER, d, predicate=eq), E.Query(E.OP_MAX, d), E.Query(E.OP_MIN, d)]
                if sel <= 0.01:
                    qs.append(E.Query(E.OP_FILTER_MATERIALIZE, d, predicate=rg))
                scanned += (6 if sel <= 0.01 else 5) * d.compressed_bytes
            for x in srv.submit(qs[:2]):  # warm
                srv.wait(x)
            t = time.perf_counter()
            ts = srv.submit(qs)
            lat = [srv.wait(x)[1] for x in ts]
            wall = time.perf_counter() - t
            ops = "filter_sum(gt,range)+filter(eq)+max+min" + ("+filter_materialize(range)" if sel <= 0.01 else "")
            res[f"mix_sel{sel * 100:g}%"] = {
                "ops": ops, "queries": len(qs), "wall_ms": round(wall * 1e3, 3), "p50_ms": pct(lat, 50),
                "p99_ms": pct(lat, 99), "gbs": round(scanned / wall / 1e9, 1)}
    return res


def matmul_leg(E, ctx, nmat, threads):
    """Decompress-then-DGEMM: two nmat x nmat matrices, uniform[-10, 10] at 3
    decimals, stored one row per chunk; decode both into HBM, then the FP64
    MFMA DGEMM.  Decoded rows are checked bit-exact against the oracle on 16
    sampled rows of A and of B; C is checked on those rows of A against a
    float64 numpy product over the decoded data (relative Frobenius error)."""
    from oracle.oracle import OracleDecoder

    a = E.synth(1, nmat * nmat, 21, 3, -10.0, 10.0, threads=threads)
    b = E.synth(1, nmat * nmat, 22, 3, -10.0, 10.0, threads=threads)
    out = {}
    cbuf = ctx.alloc(8 * nmat * nmat)
    rows = np.random.default_rng(3).choice(nmat, 16, replace=False)
    for codec, scale in (("chimp128", 1), ("elf", 1), ("buff", 1000)):
        cfg = E.CompressorConfig.by_name(codec, scale=scale)
        fa = E.encode(a, cfg, E.ChunkOption.record_count(nmat), threads=threads)
        fb = E.encode(b, cfg, E.ChunkOption.record_count(nmat), threads=threads)
        da, db = E.Decoder(ctx, fa), E.Decoder(ctx, fb)
        E.gemm_decoded(ctx, da, db, nmat, nmat, nmat, out=cbuf)
        ph = []
        for _ in range(3):
            t = time.perf_counter()
            E.gemm_decoded(ctx, da, db, nmat, nmat, nmat, out=cbuf)
            wall = (time.perf_counter() - t) * 1e3
            ph.append((ctx.last_phase_ms(0), ctx.last_phase_ms(1), ctx.last_phase_ms(2), wall))
        ph = np.median(np.array(ph), axis=0)
        c = cbuf.to_host(count=nmat * nmat).reshape(nmat, nmat)
        oa, ob = OracleDecoder(fa), OracleDecoder(fb)
        arows = np.stack([oa.materialize(chunk_id=int(r)) for r in rows])
        bdec = db.materialize().reshape(nmat, nmat)
        exact = all(np.array_equal(da.materialize(chunk_id=int(r)).view(np.uint64), arows[i].view(np.uint64)) and
                    np.array_equal(bdec[r].view(np.uint64), ob.materialize(chunk_id=int(r)).view(np.uint64))
                    for i, r in enumerate(rows))
        ref = arows @ bdec
        err = float(np.linalg.norm(c[rows] - ref) / np.linalg.norm(ref))
        flops = 2.0 * nmat ** 3
        out[codec] = {"decode_a_ms": round(ph[0], 3), "decode_b_ms": round(ph[1], 3), "dgemm_ms": round(ph[2], 3),
                      "dgemm_tflops": round(flops / (ph[2] / 1e3) / 1e12, 2),
                      "dgemm_frac_fp64_peak": round(flops / (ph[2] / 1e3) / 78.6e12, 4),
                      "end_to_end_ms": round(ph[3], 3),
                      "decode_gbs": round((da.compressed_bytes + db.compressed_bytes + 16 * nmat * nmat)
                                          / ((ph[0] + ph[1]) / 1e3) / 1e9, 1),
                      "sampled_rows_decoded_bit_exact": bool(exact),
                      "sampled_rel_fro_err": err, "within_1e-12": err <= 1e-12}
        log(f"[matmul] {codec}: dgemm {ph[2]:.2f} ms, decode {ph[0] + ph[1]:.2f} ms, err {err:.2e}")
        da.close()
        db.close()
        del fa, fb, oa, ob, bdec
    cbuf.free()
    return {"workload": f"{nmat}x{nmat} f64 x2, uniform[-10,10] 3 decimals, one row per chunk",
            "codecs": out}


def traffic_of(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload`, from the committed
    rocprofv3 counter passes (None if they cover another workload)."""
    try:
        t = json.loads((ROOT / "profiles" / "hbm_traffic.json").read_text())
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload:
        return None
    return (t.get("kernels", {}).get(kernel) or {}).get("traffic")


def breakdown(E, ctx, cols, decs, outs, n, kernel_of):
    """Per-codec GB/s: one tuned prepared batch per codec over its columns
    (every generator), median of three eager replays (HIP event time);
    per-kernel totals from the same runs.  Outside the timed region."""
    per_codec, per_kernel = {}, {}
    for codec in dict.fromkeys(c for _, c, _, _ in cols):
        idx = [i for i, (_, c, _, _) in enumerate(cols) if c == codec]
        b = E.Batch(ctx, [decs[i] for i in idx], [outs[i] for i in idx], autotune=True)
        ms = []
        for _ in range(3):
            b.run(eager=True)
            ms.append(ctx.last_device_ms)
        shape = b.config
        b.close()
        t = float(np.median(ms))
        byt = sum(cols[i][3] + 8 * n for i in idx)
        per_codec[codec] = {"bytes": byt, "ms": t,
                            "shape": {k: shape[k] for k in ("blocks_per_cu", "mixed", "tiles_per_job", "cached")}}
        pk = per_kernel.setdefault(kernel_of[codec], {"bytes": 0, "ms": 0.0, "launches": 0})
        pk["bytes"] += byt
        pk["ms"] += t
        pk["launches"] += 1
    return per_codec, per_kernel


def run_comparator(E, ctx, cols, n):
    """Batched decode vs the plain form (the same decode kernel launched once
    per chunk over that chunk's jobs, a host sync after each) on the full
    Gorilla and Chimp columns of every generator (C2: 4,096 chunks of 65,536
    records per column at 2^28)."""
    per, tot = {}, {"chunks": 0, "batched_ms": 0.0, "per_chunk_launch_ms": 0.0}
    buf = ctx.alloc(8 * n)
    for g, codec, d, _ in cols:
        if codec not in ("gorilla", "chimp"):
            continue
        d.decode_device(buf)
        t = time.perf_counter()
        for _ in range(3):
            d.decode_device(buf)
        batched = (time.perf_counter() - t) / 3
        t = time.perf_counter()
        d.decode_per_chunk(buf)
        plain = time.perf_counter() - t
        per[f"{codec}_{g}"] = {"chunks": d.number_of_chunks, "batched_ms": round(batched * 1e3, 3),
                               "per_chunk_launch_ms": round(plain * 1e3, 3), "speedup": round(plain / batched, 1)}
        tot["chunks"] += d.number_of_chunks
        tot["batched_ms"] += batched * 1e3
        tot["per_chunk_launch_ms"] += plain * 1e3
    buf.free()
    if tot["chunks"]:
        tot["speedup"] = round(tot["per_chunk_launch_ms"] / tot["batched_ms"], 1)
        tot["batched_ms"] = round(tot["batched_ms"], 3)
        tot["per_chunk_launch_ms"] = round(tot["per_chunk_launch_ms"], 3)
    return {"columns": per, "total": tot}


def cpu_threads():
    """Host cores this process may use (at most 16, the box's CPU share)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_sample(data, n, chunk, acc, budget_s):
    """Time the CPU oracle on the first chunks of this column: one core (the
    calling thread pinned to the last core it may use) until roughly
    `budget_s` seconds have been spent, then the chunk-parallel form (one
    chunk per task on acc["par_threads"] threads; the oracle's ctypes calls
    release the GIL) over as many chunks as the single core would decode in
    about half that time on every thread."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle.oracle import OracleDecoder
    import ebi_amd as E

    phys, size, logi, nrec = E.file_chunks(data)
    od = OracleDecoder(data)
    mask = os.sched_getaffinity(0)
    acc["core"] = max(mask)
    os.sched_setaffinity(0, {acc["core"]})  # this thread only
    try:
        t_end = time.perf_counter() + budget_s
        k = 0
        while k < len(size) and time.perf_counter() < t_end:
            t = time.perf_counter()
            v = od.materialize(chunk_id=k)
            acc["seconds"] += time.perf_counter() - t
            acc["bytes"] += int(size[k]) + 8 * v.size
            acc["values"] += v.size
            acc["chunks"] += 1
            k += 1
    finally:
        os.sched_setaffinity(0, mask)
    T = acc["par_threads"]
    if T > 1 and k:
        m = min(len(size), max(T, k * T // 2))
        t = time.perf_counter()
        with ThreadPoolExecutor(T) as pool:
            got = list(pool.map(lambda c: od.materialize(chunk_id=c).size, range(m)))
        acc["par_seconds"] += time.perf_counter() - t
        acc["par_bytes"] += int(np.sum(size[:m])) + 8 * sum(got)
        acc["par_chunks"] += m


if __name__ == "__main__":
    main()
