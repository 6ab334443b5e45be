"""Batched parameter-grid runs: many models' Lucy iterations in flight at once.

A pass cannot end before its longest packet, a random walk of 1e5-1e6 events
in an optically thick disk, and while that packet finishes most of the GPU is
idle.  Independent models (a parameter-grid fit) fill that time: each model
gets its own context, and with it its own HIP stream, and a pool of host
threads keeps `streams` models' passes queued, so the blocks of one model's
pass run beside another model's tail.  Every model's results are those of
running it alone (its packets draw from their own generator segments); only
the order of the floating-point tally sums may differ.

With fused=True all models' passes run in one launch instead
(hc_run_batch): workgroups are dealt to the models in proportion to their
packets, and every iteration of the whole grid is one replayed graph, so the
models' longest packets overlap each other rather than each ending its own
pass.

The model parameters follow the parameter-grid configuration of the build's
benchmark set: radial cells uniform in 50-250, theta cells 60-200, phi cells
1-128, disk mass log-uniform 1e-4-0.1 Msun, infall rate log-uniform
1e-7-1e-4 Msun/yr, cavity opening 5-40 deg, stellar temperature 3000-10000 K,
packets per model log-uniform 1e6-2e7.
"""
from __future__ import annotations

import concurrent.futures as cf
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .model import ModelConfig, build_model


@dataclass
class GridModel:
    cfg: ModelConfig
    packets: int


def parameter_grid(n: int, seed: int = 4, packet_scale: float = 1.0,
                   max_cells: Optional[int] = None) -> List[GridModel]:
    """n model configurations drawn from the parameter-grid ranges (fixed
    seed).  packet_scale multiplies the drawn packet counts (a bounded
    sample of the grid); max_cells redraws models with more cells."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        nr = int(rng.integers(50, 251))
        nth = int(rng.integers(60, 201))
        nph = int(rng.integers(1, 129))
        if max_cells is not None and nr * nth * nph > max_cells:
            continue
        cfg = ModelConfig(
            nr=nr, ntheta=nth, nphi=nph,
            massd=float(10.0 ** rng.uniform(-4.0, -1.0)),
            rate=float(10.0 ** rng.uniform(-7.0, -4.0)),
            cavity_deg=float(rng.uniform(5.0, 40.0)),
            tstar=float(rng.uniform(3000.0, 10000.0)),
            npeel=1, nximg=1, imcube=False, nfreq=50, nmu=10)
        npk = int(10.0 ** rng.uniform(6.0, np.log10(2.0e7)) * packet_scale)
        out.append(GridModel(cfg, max(npk, 1)))
    return out


@dataclass
class BatchResult:
    packets: int
    wall_s: float          # the Lucy iterations of the whole batch
    setup_s: float         # building the models and uploading them
    kernel_ms: float       # summed over models (passes overlap)
    per_model: list = field(default_factory=list)

    @property
    def rate(self) -> float:
        return self.packets / self.wall_s


def build_models(models: List[GridModel], workers: int = 4) -> list:
    """The grids, densities and dust of every model (host arrays)."""
    with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        return list(pool.map(lambda gm: build_model(gm.cfg), models))


def _iterate(tr, packets: int, iterations: int, sed_last: bool, graph: bool):
    kms, seg, scal, sed = 0.0, 0, [], None
    if graph:
        # all iterations but an SED-downloading last one from one captured
        # graph; only the last one's scalars come back
        ng = iterations - 1 if sed_last else iterations
        if ng > 0:
            s, ms, _ = tr.run_lucy(ng, packets, seg0=0)
            kms += ms
            seg += ng * packets
            scal.append(s)
        if sed_last:
            s, ms, sed = tr.run_sed(packets, seg0=seg)
            tr.temperature(packets, apply=True)
            kms += ms
            scal.append(s)
        return kms, scal, sed
    for i in range(iterations):
        if sed_last and i == iterations - 1:
            s, ms, sed = tr.run_sed(packets, seg0=seg)
        else:
            s, ms = tr.run_scalars(packets, seg0=seg)
        seg += packets
        tr.temperature(packets, apply=True)
        kms += ms
        scal.append(s)
    return kms, scal, sed


def run_batch(models: List[GridModel], iterations: int = 3, streams: int = 4,
              precision: int = 0, options: Optional[dict] = None,
              built: Optional[list] = None, sed_last: bool = False,
              graph: bool = False, fused: bool = False,
              oversubscribe: int = 0) -> BatchResult:
    """Lucy iterations of every model from its initial temperatures.

    Every model is built (unless `built` holds build_models' result) and
    uploaded first, and all of them stay resident in HBM; then a pool of
    `streams` host threads runs the models' iterations, largest packet count
    first, each model on its own context and stream.  `wall_s` covers the
    iterations only.  sed_last downloads the SED of every model's last
    iteration (per_model[i]["sed"]).  graph runs the iterations through
    hc_run_lucy (one captured HIP graph per model, replayed); the scalars
    then come from the last replay only.  fused runs every model's passes in
    one launch (hc_run_batch: one captured graph for all models, replayed;
    `streams` is then unused)."""
    from ._lib import Transport, run_batch_lucy
    t0 = time.perf_counter()
    if built is None:
        built = build_models(models)
    trs = [Transport(m, precision=precision) for m in built]
    for tr in trs:
        for k, v in (options or {}).items():
            tr.set_option(k, v)
    setup = time.perf_counter() - t0
    order = sorted(range(len(models)), key=lambda i: -models[i].packets)
    try:
        if fused:
            t1 = time.perf_counter()
            res, ms = run_batch_lucy(trs, iterations, [m.packets for m in models],
                                     oversubscribe=oversubscribe, sed=sed_last)
            wall = time.perf_counter() - t1
            per = [{"packets": models[i].packets * iterations, "kernel_ms": ms,
                    "scalars": [res[i][0]], "sed": res[i][2],
                    "tdust": trs[i].get_temperature()} for i in range(len(models))]
            return BatchResult(sum(p["packets"] for p in per), wall, setup, ms, per)
        t1 = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=max(1, streams)) as pool:
            futs = {i: pool.submit(_iterate, trs[i], models[i].packets, iterations, sed_last,
                                   graph)
                    for i in order}
            res = {i: f.result() for i, f in futs.items()}
        wall = time.perf_counter() - t1
        per = [{"packets": models[i].packets * iterations, "kernel_ms": res[i][0],
                "scalars": res[i][1], "sed": res[i][2], "tdust": trs[i].get_temperature()}
               for i in range(len(models))]
    finally:
        for tr in trs:
            tr.close()
    return BatchResult(sum(p["packets"] for p in per), wall, setup,
                       sum(p["kernel_ms"] for p in per), per)


def measure_grid(mode: str, nmodels: int = 256, iterations: int = 4, first: int = 0,
                 log=print, oversubscribe: int = 0) -> dict:
    """The parameter grid at its drawn packet counts, iteration by iteration
    from the initial temperatures, on the GPU: mode "fp32fused"/"fp64fused"
    runs every model's passes in one launch per iteration (hc_run_batch),
    "fp64plain" one model after the other on the plain FP64 path (a pass and
    its temperature update each).  Models are built with device densities.
    oversubscribe: the fused launch's workgroups per resident workgroup (0:
    the library's default).  Returns the packets/s of each iteration and of
    all of them."""
    from ._lib import HC_FP32, HC_FP64, Transport, run_batch_lucy
    grid = parameter_grid(nmodels + first)[first:]
    pk = [g.packets for g in grid]
    ncell = sum(g.cfg.nr * g.cfg.ntheta * g.cfg.nphi for g in grid)
    log(f"{mode}: {len(grid)} models, {ncell / 1e6:.1f}M cells, {sum(pk) / 1e9:.3f}e9 packets "
        f"per iteration")
    t0 = time.perf_counter()
    prec = HC_FP32 if mode.startswith("fp32") else HC_FP64
    trs = []
    try:
        for i, g in enumerate(grid):
            trs.append(Transport(build_model(g.cfg, device=0), precision=prec))
            if i % 64 == 63:
                log(f"built and uploaded {i + 1} models in {time.perf_counter() - t0:.0f} s")
        rates, total = [], 0.0
        for it in range(iterations):
            t1 = time.perf_counter()
            if mode.endswith("fused"):
                run_batch_lucy(trs, 1, pk, seg0=[it * p for p in pk], oversubscribe=oversubscribe)
            else:
                for i, tr in enumerate(trs):
                    tr.run_scalars(pk[i], seg0=it * pk[i])
                    tr.temperature(pk[i], apply=True)
                    if i % 32 == 31:
                        log(f"  iteration {it}: {i + 1} models in "
                            f"{time.perf_counter() - t1:.0f} s")
            w = time.perf_counter() - t1
            total += w
            rates.append(sum(pk) / w)
            log(f"iteration {it}: {sum(pk) / w:.4e} packets/s ({w:.1f} s)")
    finally:
        for tr in trs:
            tr.close()
    out = {"mode": mode, "models": len(grid), "cells": ncell, "packets_per_iteration": sum(pk),
           "oversubscribe": oversubscribe,
           "iteration_packets_per_s": [round(r, 1) for r in rates],
           "packets_per_s": round(iterations * sum(pk) / total, 1)}
    log(f"{mode}: {out['packets_per_s']:.4e} packets/s over {iterations} iterations")
    return out


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="the parameter grid at its drawn packet counts")
    ap.add_argument("mode", choices=["fp32fused", "fp64fused", "fp64plain"])
    ap.add_argument("--models", type=int, default=256)
    ap.add_argument("--iterations", type=int, default=4)
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--oversubscribe", type=int, default=0)
    a = ap.parse_args()
    print(json.dumps(measure_grid(a.mode, a.models, a.iterations, a.first,
                                  log=lambda m: print(m, flush=True),
                                  oversubscribe=a.oversubscribe)))
