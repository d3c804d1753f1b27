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
per-codec GB/s from one launch per column, the plain per-chunk-launch
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
    ap.add_argument("--comparator-chunks", type=int, default=128,
                    help="chunks of the Gorilla/Chimp column used for the per-chunk-launch comparator")
    ap.add_argument("--no-cpu", action="store_true")
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
    cpu = {"seconds": 0.0, "bytes": 0, "values": 0, "chunks": 0}
    comparator_files = {}
    t_setup = time.time()
    for g in gens:
        fn, dec = G.GENERATORS[g]
        vals = fn(n)
        for codec in codecs:
            cfg = E.CompressorConfig.by_name(codec, scale=10 ** dec)
            data = E.encode(vals, cfg, E.ChunkOption.record_count(args.chunk), threads=threads)
            d = E.Decoder(ctx, data)
            cols.append((g, codec, d, d.compressed_bytes))
            log(f"[setup] {g:6s} {codec:14s} {len(data) / n:6.3f} B/value  loaded")
            if g == "walk2" and codec in ("gorilla", "chimp"):
                comparator_files[codec] = E.encode(vals[: args.chunk * args.comparator_chunks], cfg,
                                                   E.ChunkOption.record_count(args.chunk), threads=threads)
            if not args.no_cpu:
                cpu_sample(data, n, args.chunk, cpu, args.cpu_seconds / (len(gens) * len(codecs)))
            del data
        del vals
    log(f"[setup] {len(cols)} columns resident in {time.time() - t_setup:.1f}s")

    total_bytes = sum(cb + 8 * n for _, _, _, cb in cols)
    outs = [ctx.alloc(8 * n) for _ in cols]
    decs = [d for _, _, d, _ in cols]

    def step():
        # one batched decode over every chunk of every column (mixed codecs):
        # k_decode (tile + segment jobs of all columns), then k_chimp128_par
        E.decode_batch(ctx, decs, outs)
        return ctx.last_device_ms, ctx.last_phase_ms(0), ctx.last_phase_ms(1)

    for _ in range(args.warmup):
        step()
    batch_dev_ms, phase_ms = [], [[], []]
    t0 = time.perf_counter()
    for s in range(args.steps):
        tot, p0, p1 = step()
        batch_dev_ms.append(tot)
        phase_ms[0].append(p0)
        phase_ms[1].append(p1)
        log(f"[timed] step {s + 1}/{args.steps} done")
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = total_bytes / (ms_per_step / 1000.0) / 1e9

    # roofline of the batched step's kernels: algorithmic bytes of the columns
    # each kernel decodes / its mean launch time (HIP events on the library's
    # stream, around each launch of the timed steps)
    kernel_of = {c: ("k_chimp128_par" if c == "chimp128" else "k_decode") for c in codecs}
    batch_kernels = {}
    for k, name in enumerate(("k_decode", "k_chimp128_par")):
        b = sum(cb + 8 * n for _, codec, _, cb in cols if kernel_of[codec] == name)
        if b:
            ms = float(np.mean(phase_ms[k]))
            batch_kernels[name] = {"bytes_per_launch": b, "mean_launch_ms": round(ms, 4),
                                   "gbs": round(b / (ms / 1000.0) / 1e9, 2)}
    dominant = max(batch_kernels, key=lambda k: batch_kernels[k]["mean_launch_ms"])
    achieved = batch_kernels[dominant]["gbs"]

    per_codec, per_kernel, comparator = {}, {}, {}
    if not args.no_breakdown:
        per_codec, per_kernel = breakdown(ctx, cols, decs, outs, n, kernel_of)
        comparator = run_comparator(E, ctx, comparator_files)

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
        "per_kernel": {k: {"gbs": round(v["bytes"] / (v["ms"] / 1000.0) / 1e9, 2),
                           "ms_per_step": round(v["ms"], 3), "launches": v["launches"]}
                       for k, v in per_kernel.items()},
        "per_chunk_launch_comparator": comparator,
    }
    if not args.no_cpu and cpu["seconds"] > 0:
        res["cpu_baseline"] = {
            "value": round(cpu["bytes"] / cpu["seconds"] / 1e9, 4),
            "unit": "GB/s",
            "cores": 1,
            "kind": "port",
            "sample": f"oracle decode of the first {cpu['chunks']} chunks across the {len(cols)} columns "
                      f"({cpu['values']} values, {cpu['seconds']:.1f}s on one core)",
        }
        res["gpu_vs_cpu"] = round(value / res["cpu_baseline"]["value"], 1)
    for o in outs:
        o.free()
    for _, _, d, _ in cols:
        d.close()
    ctx.close()
    print(json.dumps(res), flush=True)


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


def breakdown(ctx, cols, decs, outs, n, kernel_of):
    """Per-codec and per-kernel GB/s from one decode_device per column (two
    passes, HIP event time), outside the timed region."""
    per_col_ms = {i: [] for i in range(len(cols))}
    for _ in range(2):
        for i, d in enumerate(decs):
            d.decode_device(outs[i])
            per_col_ms[i].append(ctx.last_device_ms)
    per_codec, per_kernel = {}, {}
    for i, (g, codec, d, cb) in enumerate(cols):
        ms = float(np.mean(per_col_ms[i]))
        b = cb + 8 * n
        pc = per_codec.setdefault(codec, {"bytes": 0, "ms": 0.0})
        pc["bytes"] += b
        pc["ms"] += ms
        pk = per_kernel.setdefault(kernel_of[codec], {"bytes": 0, "ms": 0.0, "launches": 0})
        pk["bytes"] += b
        pk["ms"] += ms
        pk["launches"] += 1
    return per_codec, per_kernel


def run_comparator(E, ctx, comparator_files):
    """Batched decode vs the plain form (one launch + host sync per chunk) on
    the Gorilla / Chimp random-walk columns."""
    comparator = {}
    for codec, data in comparator_files.items():
        d = E.Decoder(ctx, data)
        m = d.number_of_records
        buf = ctx.alloc(8 * m)
        d.decode_device(buf)
        t = time.perf_counter()
        for _ in range(3):
            d.decode_device(buf)
        batched = (time.perf_counter() - t) / 3
        t = time.perf_counter()
        d.decode_per_chunk(buf)
        plain = time.perf_counter() - t
        comparator[codec] = {"chunks": d.number_of_chunks, "batched_ms": batched * 1e3,
                             "per_chunk_launch_ms": plain * 1e3, "speedup": plain / batched}
        buf.free()
        d.close()

    return comparator


def cpu_sample(data, n, chunk, acc, budget_s):
    """Time the CPU oracle (one core) on the first chunks of this column until
    roughly `budget_s` seconds have been spent."""
    from oracle.oracle import OracleDecoder
    import ebi_amd as E

    phys, size, logi, nrec = E.file_chunks(data)
    od = OracleDecoder(data)
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


if __name__ == "__main__":
    main()
