#!/usr/bin/env python3
"""Packets/s of the Lucy transport pass on one MI355X.

A step is one Lucy iteration of the hot path: the transport pass over one
batch of packets (launch, cell walk, scattering, absorption and re-emission,
the absorbed-energy tallies) followed by the device temperature update.
Grid, tables, tallies and temperatures stay resident in HBM for the whole
timed region; the library's calls are synchronous on their own stream, so
each step ends with the device idle.

Workload (default): the 200 x 199 x 200-cell (~8M cells) disk + Ulrich
envelope + star model with a fractal clumpy envelope, 3 x N_cells = 23.9M
packets per Lucy iteration as the reference sizes its temperature
iterations (newtts.F90:168-177).  Warm-up Lucy iterations (untimed, fewer
packets) bring the dust temperatures towards equilibrium first: the event
count per packet grows from ~2 at the initial 0.1 K to ~100 near
equilibrium, so timing the first iterations would measure an almost
transparent model.

`value` is the plain FP64 path.  The FP32 path (FP32 packet state and event
arithmetic; FP64 positions, wall roots and accumulators) is then timed over
the same steps from the same temperatures, and the SED of its first step is
compared with the FP64 SED of the same packets.

The CPU baseline is the FP64 restatement of the same path (oracle/, C library
elementary functions), one thread per host core on disjoint packet ranges,
over a bounded sample of the same workload and temperatures.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(model, seconds: float, threads: int, seg_base: int, np_total: int):
    """Oracle packets/s over ~`seconds` of wall time on `threads` threads."""
    from oracle.pyoracle import Oracle

    orc = Oracle(model, "libm")
    tallies = [orc.new_tallies() for _ in range(threads)]
    done = [0] * threads
    stop = time.perf_counter() + seconds
    chunk = 20

    def work(i):
        p = seg_base + i * 100_000_000
        while time.perf_counter() < stop:
            orc.run_into(tallies[i], np_total, 0, chunk, seg0=p)
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


def cpu_model_name() -> str:
    """The host CPU's model name (lscpu's 'Model name')."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sed_errors(s64: np.ndarray, s32: np.ndarray) -> dict:
    """SED differences of two passes over the same packets: per inclination
    bin (frequency-integrated, all photons, largest aperture) and the
    luminosity-weighted spectral difference."""
    a = s64[0, :, :, :, -1, 0]   # I, (nfreq, nmu, nph), largest aperture, all photons
    b = s32[0, :, :, :, -1, 0]
    inc64, inc32 = a.sum(axis=(0, 2)), b.sum(axis=(0, 2))
    per_inc = np.abs(inc32 - inc64) / np.maximum(inc64, 1e-300)
    spec64, spec32 = a.sum(axis=(1, 2)), b.sum(axis=(1, 2))
    return {"per_inclination_max": float(per_inc.max()),
            "luminosity_weighted": float(np.abs(spec32 - spec64).sum() / spec64.sum())}


def temperature_errors(model, t64: np.ndarray, t32: np.ndarray) -> dict:
    """Dust-temperature differences of the two paths after the same Lucy
    steps from the same temperatures, over the (cell, species) with dust:
    mass-weighted mean, median and 99th percentile of |T32/T64 - 1|."""
    dusty = model.densarr > 0.0
    rel = np.abs(t32[dusty] / t64[dusty] - 1.0)
    w = (model.densarr * model.vcell[..., None])[dusty]
    return {"mass_weighted_mean": float(np.sum(w * rel) / np.sum(w)),
            "median": float(np.median(rel)), "p99": float(np.percentile(rel, 99))}


def timed_steps(tr, steps: int, np_: int, seg: int, label: str):
    """`steps` Lucy iterations; returns wall s, kernel ms, summed counters and
    the SED summed over the steps."""
    kernel_ms = 0.0
    tot = {}
    sed_sum = None
    t_start = time.perf_counter()
    for k in range(steps):
        s, ms, sed = tr.run_sed(np_, seg0=seg + k * np_)
        sed_sum = sed if sed_sum is None else sed_sum + sed
        tr.temperature(np_, apply=True)
        kernel_ms += ms
        for key, v in s.items():
            tot[key] = tot.get(key, 0.0) + v
        # the pass cannot end before its longest packet's serial walk
        tot["longest_packet_work"] = max(tot.get("longest_packet_work", 0.0),
                                         s["max_packet_work"])
        log(f"{label} step {k}: kernel {ms:.0f} ms ({np_ / (ms / 1e3):.3e} packets/s), "
            f"{(s['nscat_ev'] + s['nabs_ev']) / np_:.1f} events and {s['ncross'] / np_:.0f} "
            f"crossings per packet, longest packet {s['max_packet_work']:.0f} events + "
            f"crossings, errors wall={s['n_err_nowall']:.0f} "
            f"dust={s['n_err_dust']:.0f} warn={s['n_warn']:.0f}")
    wall = time.perf_counter() - t_start
    return wall, kernel_ms, tot, sed_sum


def batched_grid(args) -> dict:
    """Parameter-grid models, the same models, packets and iterations (from
    each model's initial temperatures) run four ways:
      - fp32_batched: the FP32 path, every model's passes in one launch
        (hc_run_batch, each iteration of the grid one replayed graph);
      - fp64_batched: the plain FP64 path in the same single launch, which
        separates the precision's gain from the batching's;
      - fp32_streams: the FP32 path with `batch_streams` models on their own
        contexts and streams, one graph-replayed Lucy loop each;
      - fp64_plain: the plain FP64 path, one model at a time."""
    from hochunk3d_amd._lib import HC_FP32, HC_FP64
    from hochunk3d_amd.batch import build_models, parameter_grid, run_batch

    models = parameter_grid(args.batch_models, packet_scale=args.batch_scale)
    t0 = time.perf_counter()
    built = build_models(models, workers=8)
    cells = sum(m.cfg.nr * m.cfg.ntheta * m.cfg.nphi for m in models)
    log(f"batch: {len(models)} grid models, {cells} cells, "
        f"{sum(m.packets for m in models)} packets per iteration, built in "
        f"{time.perf_counter() - t0:.1f}s")
    runs = {}
    runs["fp32_batched"] = run_batch(models, args.batch_iters, precision=HC_FP32, built=built,
                                     sed_last=True, fused=True)
    runs["fp64_batched"] = run_batch(models, args.batch_iters, precision=HC_FP64, built=built,
                                     sed_last=True, fused=True)
    runs["fp32_streams"] = run_batch(models, args.batch_iters, streams=args.batch_streams,
                                     precision=HC_FP32, built=built, sed_last=True, graph=True)
    runs["fp64_plain"] = run_batch(models, args.batch_iters, streams=1, precision=HC_FP64,
                                   built=built, sed_last=True)
    for k, r in runs.items():
        log(f"batch {k}: {r.rate:.3e} packets/s ({r.wall_s:.2f}s)")
    b64 = runs["fp64_plain"]
    errs = [sed_errors(a["sed"], b["sed"])
            for a, b in zip(b64.per_model, runs["fp32_batched"].per_model)]
    lw = np.array([e["luminosity_weighted"] for e in errs])
    inc = np.array([e["per_inclination_max"] for e in errs])
    out = {
        "workload": (f"{len(models)} parameter-grid models (nr 50-250, ntheta 60-200, nphi "
                     f"1-128, {cells} cells in all), {args.batch_scale:g} x their 1e6-2e7 "
                     f"packets = {sum(m.packets for m in models)} packets per Lucy iteration, "
                     f"{args.batch_iters} iterations from the initial temperatures; grids "
                     f"resident in HBM before the timed region"),
    }
    how = {"fp32_batched": "one launch for all models per iteration (hc_run_batch), the "
                           "iterations replayed from one captured graph",
           "fp64_batched": "as fp32_batched on the plain FP64 path",
           "fp32_streams": f"{args.batch_streams} models at a time on their own streams, each "
                           f"model's Lucy loop replayed from its own graph",
           "fp64_plain": "one model at a time, one pass per call"}
    for k, r in runs.items():
        out[k] = {"value": round(r.rate, 2), "unit": "packets/s", "wall_s": round(r.wall_s, 3),
                  "how": how[k]}
    out["speedup_fp32_batched_vs_fp64_plain"] = round(b64.wall_s / runs["fp32_batched"].wall_s, 3)
    out["speedup_fp64_batched_vs_fp64_plain"] = round(b64.wall_s / runs["fp64_batched"].wall_s, 3)
    out["speedup_fp32_vs_fp64_both_batched"] = round(runs["fp64_batched"].wall_s
                                                     / runs["fp32_batched"].wall_s, 3)
    out["sed_error_last_iteration"] = {
        "luminosity_weighted_median": round(float(np.median(lw)), 6),
        "luminosity_weighted_max": round(float(lw.max()), 6),
        "per_inclination_max_median": round(float(np.median(inc)), 6),
        "note": "fp32_batched vs fp64_plain SED of each model's last iteration: same packets, "
                "temperatures iterated separately; includes the Monte Carlo noise of histories "
                "that diverge"}
    return out


def config5_section(args) -> dict:
    """The large hydro-like model (400 x 300 x 256 cells, log-normal
    densities, 20 observers x 500^2 pixels x 200 frequencies): a Lucy pass
    on the FP32 path against the plain FP64 path on the same packets and
    temperatures, then a peeled pass with the FP32 image cube in HBM."""
    from hochunk3d_amd import build_model, config5
    from hochunk3d_amd._lib import HC_FP32, HC_FP64, Transport
    from hochunk3d_amd.testing import synthetic_temperature

    t0 = time.perf_counter()
    m = build_model(config5(), device=0)
    m.tdust[...] = synthetic_temperature(m)
    build_s = time.perf_counter() - t0
    n = int(args.config5_np)
    log(f"config5: {m.ncells} cells built in {build_s:.1f}s (device densities, log-normal "
        f"multiplier); Lucy passes of {n} packets")
    out = {"workload": f"config5: {m.nrg - 1}x{m.ntg - 1}x{m.npg - 1} cells, disk+envelope x "
                       f"log-normal (sigma 1.5, k^-11/3, seed 11), synthetic temperatures; "
                       f"one Lucy pass of {n} packets per path, then a peeled pass of "
                       f"{int(args.config5_peel_np)} packets with 20 observers x 500^2 pixels "
                       f"x 200 frequencies (FP32 IQUV cube, 44 GB, in HBM)",
           "build_s": round(build_s, 2)}
    for prec, name in ((HC_FP32, "fp32"), (HC_FP64, "fp64")):
        tr = Transport(m, precision=prec, cube_fp32=True)
        try:
            t1 = time.perf_counter()
            sc, ms = tr.run_scalars(n, seg0=0)
            wall = time.perf_counter() - t1
            out[name] = {"value": round(n / wall, 2), "unit": "packets/s",
                         "kernel_ms": round(ms, 1),
                         "longest_packet": int(sc["max_packet_work"]),
                         "crossings_per_packet": round(sc["ncross"] / n, 1),
                         "events_per_packet": round((sc["nscat_ev"] + sc["nabs_ev"]) / n, 1),
                         "dropped": int(sc["n_warn"]), "lost": int(sc["n_err_nowall"])}
            log(f"config5 {name}: {n / wall:.3e} packets/s, longest packet "
                f"{sc['max_packet_work']:.0f}, dropped {sc['n_warn']:.0f}")
            if prec == HC_FP32 and args.config5_peel_np > 0:
                pn = int(args.config5_peel_np)
                t1 = time.perf_counter()
                sp, pms = tr.run_scalars(pn, ipeel=1, seg0=n)
                pw = time.perf_counter() - t1
                out["fp32_peeled"] = {"value": round(pn / pw, 2), "unit": "packets/s",
                                      "kernel_ms": round(pms, 1),
                                      "peel_rays_per_packet": round(sp["npeelray"] / pn, 1)}
                log(f"config5 fp32 peeled: {pn / pw:.3e} packets/s")
        finally:
            tr.close()
    out["fp32_speedup_vs_fp64"] = round(out["fp32"]["value"] / out["fp64"]["value"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=5, help="untimed Lucy iterations")
    ap.add_argument("--config", default="config2", choices=["config1", "config2"])
    ap.add_argument("--np", type=float, default=0,
                    help="packets per step (default 3 x cells, as the reference)")
    ap.add_argument("--warmup-np", type=float, default=0,
                    help="packets per warm-up iteration (default a quarter of --np)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 path")
    ap.add_argument("--peel-np", type=float, default=2_000_000,
                    help="packets of the final peeled pass timed after the Lucy steps (0: skip)")
    ap.add_argument("--autotune-np", type=float, default=-1,
                    help="packets per autotuning probe pass (-1: the step's packet count; "
                         "0: default launch shape)")
    ap.add_argument("--batch-models", type=int, default=64,
                    help="parameter-grid models of the batched section (0: skip)")
    ap.add_argument("--batch-scale", type=float, default=0.02,
                    help="fraction of each grid model's packet count run per Lucy iteration")
    ap.add_argument("--batch-iters", type=int, default=3)
    ap.add_argument("--batch-streams", type=int, default=4)
    ap.add_argument("--config5-np", type=float, default=4_000_000,
                    help="packets of the config-5 Lucy passes (0: skip the section)")
    ap.add_argument("--config5-peel-np", type=float, default=20_000)
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", type=float, default=0,
                    help="only copy this many bytes device-to-device once (a plain copy kernel "
                         "of known size that scales the HBM byte counters) and exit")
    args = ap.parse_args()

    if args.copy_calibration > 0:
        import torch
        n = int(args.copy_calibration)
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {n} bytes")
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs the transport path on one GPU; "
                          "multi-GPU sharding (independent packet streams per GPU, one "
                          "all-reduce of the per-cell tallies per pass) is not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from hochunk3d_amd import build_model, config1, config2
    from hochunk3d_amd import roofline as rf
    from hochunk3d_amd._lib import HC_FP32, HC_FP64, Transport

    t0 = time.time()
    cfg = config2() if args.config == "config2" else config1()
    model = build_model(cfg)
    ncells = int(model.ncells)
    np_ = int(args.np) if args.np > 0 else 3 * ncells
    wnp = int(args.warmup_np) if args.warmup_np > 0 else max(np_ // 4, 1)
    log(f"model {args.config}: {model.nrg - 1}x{model.ntg - 1}x{model.npg - 1} cells, "
        f"built in {time.time() - t0:.1f}s; {np_} packets per step")

    tr = Transport(model, precision=HC_FP64)
    seg = 0
    for w in range(args.warmup):
        s, ms = tr.run_scalars(wnp, seg0=seg)
        seg += wnp
        _, nch = tr.temperature(wnp, apply=True)
        log(f"warmup {w}: {wnp} packets, kernel {ms:.0f} ms, {wnp / (ms / 1e3):.3e} packets/s, "
            f"{(s['nscat_ev'] + s['nabs_ev']) / wnp:.1f} events per packet, "
            f"T changed in {nch} cell-species")
    t_start = tr.get_temperature()  # where the timed steps start
    tuned = {}
    # probes as large as a step: a shorter pass is mostly its longest packet,
    # whose walk the launch shape hardly changes
    anp = np_ if args.autotune_np < 0 else int(args.autotune_np)
    if anp > 0:
        from hochunk3d_amd.autotune import tune_transport
        r = tune_transport(tr, anp)
        tuned["fp64"] = {"options": r.options, "gain_vs_default": round(r.gain, 3),
                         "probes": len(r.trials), "probe_packets": anp}
        log(f"autotune fp64: {r.options}, x{r.gain:.3f} vs default on {len(r.trials)} probes "
            f"of {anp} packets")

    wall64, kms64, tot64, sed64 = timed_steps(tr, args.steps, np_, seg, "fp64")
    t64_end = tr.get_temperature()
    npk = np_ * args.steps
    value = npk / wall64

    # the final pass of the run (newtts.F90:347-349): peel-off towards every
    # observer from every emission and scattering, images on
    peel = None
    if args.peel_np > 0:
        pnp = int(args.peel_np)
        if anp > 0:
            r = tune_transport(tr, max(pnp // 20, 1000), ipeel=1)
            tuned["fp64_peeled"] = {"options": r.options, "gain_vs_default": round(r.gain, 3),
                                    "probes": len(r.trials)}
            log(f"autotune peeled pass: {r.options}, x{r.gain:.3f} vs default")
        # a quarter-size pass first: the difference of the two passes' times
        # is the rate of the bulk, which a 1e8-packet final pass runs at (a
        # pass also carries its longest packet's serial walk, ~1 s)
        q = max(pnp // 4, 1)
        t0 = time.perf_counter()
        tr.run_scalars(q, ipeel=1, seg0=seg + npk)
        wallq = time.perf_counter() - t0
        t0 = time.perf_counter()
        s, ms = tr.run_scalars(pnp, ipeel=1, seg0=seg + npk + q)
        wallp = time.perf_counter() - t0
        marginal = (pnp - q) / (wallp - wallq) if wallp > wallq else None
        peel = {"packets": pnp, "value": round(pnp / wallp, 2), "unit": "packets/s",
                "kernel_ms": round(ms, 3),
                "bulk_packets_per_s": round(marginal, 2) if marginal else None,
                "final_pass_1e8_packets_s": round(1e8 / marginal, 1) if marginal else None,
                "peel_rays_per_packet": round(s["npeelray"] / pnp, 1),
                "peel_crossings_per_packet": round(s["npeelcross"] / pnp, 1),
                "observers": cfg.npeel, "image_pixels": f"{cfg.nximg}x{cfg.nximg}",
                "note": f"bulk rate from passes of {q} and {pnp} packets "
                        f"({wallq:.2f} s and {wallp:.2f} s)"}
        log(f"peeled pass: {pnp} packets in {wallp * 1e3:.0f} ms ({pnp / wallp:.3e} packets/s), "
            f"{s['npeelray'] / pnp:.0f} peel rays and {s['npeelcross'] / pnp:.0f} peel-ray "
            f"crossings per packet; bulk {marginal or 0:.3e} packets/s, a 1e8-packet final "
            f"pass ~{1e8 / marginal if marginal else 0:.0f} s")
    tr.close()

    fp32 = None
    if not args.no_fp32:
        model.tdust[...] = t_start
        tr32 = Transport(model, precision=HC_FP32)
        if anp > 0:
            r = tune_transport(tr32, anp)
            tuned["fp32"] = {"options": r.options, "gain_vs_default": round(r.gain, 3),
                             "probes": len(r.trials)}
            log(f"autotune fp32: {r.options}, x{r.gain:.3f} vs default")
        wall32, kms32, tot32, sed32 = timed_steps(tr32, args.steps, np_, seg, "fp32")
        t32_end = tr32.get_temperature()
        tr32.close()
        err = sed_errors(sed64, sed32)
        terr = temperature_errors(model, t64_end, t32_end)
        fp32 = {"value": round(npk / wall32, 2), "unit": "packets/s",
                "speedup_vs_fp64": round(wall64 / wall32, 3),
                "kernel_ms_per_step": round(kms32 / args.steps, 3),
                "sed_error_vs_fp64_all_steps": {k: round(v, 6) for k, v in err.items()},
                "dropped_packets": {"fp32": int(tot32["n_warn"]), "fp64": int(tot64["n_warn"]),
                                    "note": "packets dropped at the per-packet work bound over "
                                            "the timed steps (their energy is missing from "
                                            "the tallies)"},
                "longest_packet": int(tot32["longest_packet_work"]),
                "temperature_error_vs_fp64_after_steps": {k: round(v, 6) for k, v in terr.items()},
                "note": "same packets (L'Ecuyer-Cote segments) and starting temperatures as "
                        "the FP64 steps; SED error includes the Monte Carlo noise of histories "
                        "that diverge"}
        log(f"fp32: {fp32['value']:.3e} packets/s, x{fp32['speedup_vs_fp64']} vs fp64, "
            f"SED error {err}, temperature error {terr}")

    batch = None
    if args.batch_models > 0:
        batch = batched_grid(args)
    c5 = config5_section(args) if args.config5_np > 0 else None

    # the binding roof of the Lucy pass: HBM (a crossing's density gather and
    # the per-cell tally read-modify-writes) against FP64 VALU (its FLOPs)
    kernel_s = kms64 / 1e3
    nbytes = rf.pass_bytes(tot64)
    flops = rf.pass_flops(tot64, npk)
    hbm_gbs = nbytes / kernel_s / 1e9
    valu_tfs = flops / kernel_s / 1e12
    hbm_frac = hbm_gbs / rf.PEAK_HBM_GBS
    valu_frac = valu_tfs / rf.PEAK_FP64_TFLOPS
    profiled = args.config == "config2" and np_ == rf.CONFIG2_STEP_PACKETS
    roof = {
        "bound": "hbm", "achieved": round(hbm_gbs, 2),
        "peak": rf.PEAK_HBM_GBS, "unit": "GB/s", "frac": round(hbm_frac, 5),
        "traffic": (rf.CONFIG2_STEP_HBM_READ + rf.CONFIG2_STEP_HBM_WRITE if profiled else None),
        "traffic_source": (rf.CONFIG2_TRAFFIC_PROFILE if profiled else None),
        "kernel": "hc::transport<double, 4, false, false>",
        "algorithmic_bytes_per_step": round(nbytes / args.steps),
        "per_crossing_bytes": round(nbytes / tot64["ncross"], 2),
        "valu": {"achieved": round(valu_tfs, 4), "peak": rf.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                 "frac": round(valu_frac, 5)},
        "note": "HBM binds: a crossing moves the densities of its cell (4 x 8 B) and the "
                "tally words its cell flush read-modify-writes (16 B per dtauabs/jmean/nabs/"
                "dbig atomic, counted by the kernel), against 149 FP64 FLOPs "
                "(hochunk3d_amd/roofline.py): 8 TB/s / ~70 B = 1.1e11 crossings/s against "
                "78.6 TFLOP/s / 149 = 5.3e11.  achieved = those algorithmic bytes over the "
                "kernel time measured with HIP events on the library's stream; traffic = HBM "
                "bytes per step from the rocprofv3 FETCH_SIZE/WRITE_SIZE profile named in "
                "traffic_source (a separate profiled run of this workload, scaled by a 256 MiB "
                "copy's readings), null when the workload differs",
        "kernel_ms_per_step": round(kms64 / args.steps, 3),
    }

    cpu = None
    if args.cpu_seconds > 0:
        model.tdust[...] = t_start
        avail = len(os.sched_getaffinity(0))
        # the host cores this process may use, at most the 16-core share of
        # one GPU on the GPU box
        threads = args.cpu_threads or min(16, avail)
        cpu_model = cpu_model_name()
        log(f"cpu baseline: {threads} threads x {args.cpu_seconds:.0f}s on {cpu_model} "
            f"({avail} cores available)")
        rate, n, w = cpu_baseline(model, args.cpu_seconds, threads, seg_base=seg + 10**10,
                                  np_total=np_)
        cpu = {"value": round(rate, 2), "unit": "packets/s", "cores": threads, "kind": "port",
               "cpu": cpu_model, "cores_available": avail, "os_cpu_count": os.cpu_count(),
               "sample": f"{n} packets of the same model and starting temperatures in {w:.1f}s "
                         f"wall, FP64 CPU restatement (oracle/, gcc -O3, C library elementary "
                         f"functions), one independent seeded instance (thread with its own "
                         f"packet segments and tallies) per core"}

    line = {
        "metric": "photon packets/s on 1 MI355X, FP32-batched vs FP64 plain, at <=1% SED error",
        "value": round(value, 2),
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall64 / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.config}: {model.nrg - 1}x{model.ntg - 1}x{model.npg - 1}-cell "
                         f"disk+envelope+star" + (" with fractal clumps" if cfg.fractal else "")
                         + f"; per step one Lucy iteration: transport pass of {np_} packets + "
                           f"temperature update, after {args.warmup} warm-up iterations of "
                           f"{wnp} packets; FP64 plain path"),
            "packets_per_step": np_,
            "cells": ncells,
            "path": "fp64-plain",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "longest_packet_fp64": {"events_plus_crossings": int(tot64["longest_packet_work"]),
                                "mean_per_packet": round((tot64["nscat_ev"] + tot64["nabs_ev"]
                                                          + tot64["ncross"]) / npk, 1),
                                "note": "the longest packet of the timed steps: a pass "
                                        "cannot end before its serial walk does"},
        "fp32_path": fp32,
        "peeled_pass_fp64": peel,
        "batched_grid": batch,
        "config5": c5,
        "autotune": tuned or None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
