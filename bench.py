#!/usr/bin/env python3
"""Packets/s of the Lucy transport pass on one MI355X.

A step is one Lucy iteration of the hot path over one batch of packets: the
transport pass (launch, cell walk, scattering, absorption/re-emission and the
absorbed-energy tallies) followed by the device temperature update.  Inputs,
tallies and temperatures stay resident in HBM for the whole timed region.

The default workload is the 200 x 199 x 200-cell (~8M cells) fractal disk +
envelope + star model on the FP64 plain path; warm-up steps are Lucy
iterations that bring the dust temperatures close to equilibrium before the
timed steps.  The CPU baseline is the FP64 restatement of the same path
(oracle/, C library elementary functions), one thread per host core on
disjoint packet ranges, over a bounded sample of the same workload.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(model, seconds: float, threads: int, seg_base: int):
    """Oracle packets/s over ~`seconds` of wall time on `threads` threads."""
    from oracle.pyoracle import Oracle

    orc = Oracle(model, "libm")
    tallies = [orc.new_tallies() for _ in range(threads)]
    done = [0] * threads
    stop = time.perf_counter() + seconds
    chunk = 50

    def work(i):
        p = seg_base + i * 10_000_000
        while time.perf_counter() < stop:
            orc.run_into(tallies[i], 1_000_000, 0, chunk, seg0=p)
            p += chunk
            done[i] += chunk

    t0 = time.perf_counter()
    ts = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    wall = time.perf_counter() - t0
    return sum(done) / wall, sum(done), wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="config2", choices=["config1", "config2"])
    ap.add_argument("--np", type=int, default=2_000_000, help="packets per step")
    ap.add_argument("--warmup-np", type=int, default=1_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs the transport path on one GPU; "
                          "multi-GPU sharding (independent packet streams per GPU, one "
                          "all-reduce of the per-cell tallies per pass) is not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from hochunk3d_amd import build_model, config1, config2
    from hochunk3d_amd import roofline as rf
    from hochunk3d_amd._lib import Transport

    t0 = time.time()
    cfg = config2() if args.config == "config2" else config1()
    model = build_model(cfg)
    log(f"model {args.config}: {model.nrg - 1}x{model.ntg - 1}x{model.npg - 1} cells, "
        f"built in {time.time() - t0:.1f}s")
    tr = Transport(model)
    seg = 0
    for w in range(args.warmup):
        s, ms = tr.run_scalars(args.warmup_np, seg0=seg)
        seg += args.warmup_np
        tr.temperature(args.warmup_np, apply=True)
        log(f"warmup {w}: {args.warmup_np} packets, kernel {ms:.0f} ms, "
            f"{args.warmup_np / (ms / 1e3):.3e} packets/s")

    kernel_ms = 0.0
    tot = {}
    t_start = time.perf_counter()
    for k in range(args.steps):
        s, ms = tr.run_scalars(args.np, seg0=seg)
        seg += args.np
        tr.temperature(args.np, apply=True)
        kernel_ms += ms
        for key, v in s.items():
            tot[key] = tot.get(key, 0.0) + v
        log(f"step {k}: kernel {ms:.0f} ms ({args.np / (ms / 1e3):.3e} packets/s) "
            f"errors wall={s['n_err_nowall']:.0f} dust={s['n_err_dust']:.0f}")
    wall = time.perf_counter() - t_start
    npk = args.np * args.steps
    value = npk / wall
    flops = rf.pass_flops(tot, npk)
    achieved = flops / (kernel_ms / 1e3) / 1e12
    roof = {
        "bound": "mfma", "achieved": round(achieved, 4), "peak": rf.PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s", "frac": round(achieved / rf.PEAK_FP64_TFLOPS, 5),
        "traffic": None,
        "note": "VALU FP64 (no MFMA instructions; MI355X's FP64 vector and matrix peaks are "
                "both 78.6 TFLOP/s); algorithmic FLOPs from the work counters of the pass "
                "times the per-unit figures of hochunk3d_amd/roofline.py",
        "kernel_ms_per_step": round(kernel_ms / args.steps, 3),
    }

    cpu = None
    if args.cpu_seconds > 0:
        # temperatures of the GPU run so the CPU sample sees the same workload
        t2, _ = tr.temperature(args.np)
        model.tdust[...] = t2
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {threads} threads x {args.cpu_seconds:.0f}s")
        rate, n, w = cpu_baseline(model, args.cpu_seconds, threads, seg_base=seg + 10**9)
        cpu = {"value": round(rate, 2), "unit": "packets/s", "cores": threads, "kind": "port",
               "sample": f"{n} packets of the same model and temperatures in {w:.1f}s wall, "
                         f"FP64 CPU restatement (oracle/, -O2), one thread per core"}
    tr.close()

    line = {
        "metric": "photon packets/s on 1 MI355X, FP32-batched vs FP64 plain, at <=1% SED error",
        "value": round(value, 2),
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.config}: {model.nrg - 1}x{model.ntg - 1}x{model.npg - 1}-cell "
                         f"disk+envelope+star" + (" with fractal clumps" if cfg.fractal else "")
                         + f", Lucy pass of {args.np} packets + temperature update per step, "
                           f"FP64 plain path"),
            "packets_per_step": args.np,
            "cells": int(model.ncells),
            "path": "fp64-plain",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
