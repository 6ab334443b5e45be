"""Launch-shape autotuner of the transport kernel.

The knobs (include/hc.h, hc_set_option) change how a wavefront schedules
its lanes' packets, not what the packets do:
- regen_lanes: waiting lanes of one event kind that end a walk phase;
- regen_run_lanes: waiters that make an event kind run once the walk has
  stopped;
- blocks_per_cu: resident workgroups per compute unit;
- peel_queue: depth of the LDS peel-off queues (peeled passes only).

Every candidate is timed on the same packets (the same generator segments)
from the same temperatures, so the probes do the same work and only the
schedule differs.  The search is coordinate-wise: each knob in turn takes
the value that gives the shortest kernel time with the others fixed.  The
tuning passes leave the temperatures alone.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

# the library's defaults (include/hc.h): regen 8/8 on the FP64 path, 12/8 on
# the FP32 path, the fastest of 8/8, 12/8 and 16/16 on full passes of the
# 200x199x200-cell model
DEFAULTS = {"regen_lanes": 8, "regen_run_lanes": 8, "blocks_per_cu": 0, "peel_queue": 0}
DEFAULTS_FP32 = dict(DEFAULTS, regen_lanes=12)

# regen_lanes 24 and 32 can win on the short probes, whose time is mostly
# the longest packet, and then run a full 2.4e7-packet pass of the
# 200x199x200-cell model ~40 % slower: they are left out.
LUCY_SPACE: Dict[str, Sequence[int]] = {
    "regen_lanes": (8, 12, 16),
    "regen_run_lanes": (4, 8, 12, 16),
    "blocks_per_cu": (1, 0),
}
PEEL_SPACE: Dict[str, Sequence[int]] = {
    "peel_queue": (65, 80, 96, 128),
    "regen_lanes": (8, 16, 32),
}


@dataclass
class TuneResult:
    options: Dict[str, int]
    best_ms: float
    default_ms: float
    trials: List[tuple] = field(default_factory=list)   # (options, ms)

    @property
    def gain(self) -> float:
        return self.default_ms / self.best_ms


def coordinate_search(time_fn: Callable[[Dict[str, int]], float],
                      space: Dict[str, Sequence[int]],
                      start: Optional[Dict[str, int]] = None, rounds: int = 1) -> TuneResult:
    """Coordinate descent over `space` from `start`; time_fn(options) -> ms.
    Candidates already timed are not timed again.  A candidate with
    regen_run_lanes above regen_lanes is timed with it lowered to
    regen_lanes, the value the kernel would use."""
    cur = dict(DEFAULTS if start is None else start)
    seen: Dict[tuple, float] = {}
    trials = []

    def timed(opts):
        key = tuple(sorted(opts.items()))
        if key not in seen:
            seen[key] = float(time_fn(dict(opts)))
            trials.append((dict(opts), seen[key]))
        return seen[key]

    default_ms = timed(cur)
    best_ms = default_ms
    for _ in range(max(1, rounds)):
        changed = False
        for name, values in space.items():
            for v in values:
                cand = dict(cur, **{name: v})
                if "regen_run_lanes" in cand and "regen_lanes" in cand:
                    cand["regen_run_lanes"] = min(cand["regen_run_lanes"], cand["regen_lanes"])
                ms = timed(cand)
                if ms < best_ms:
                    best_ms, cur, changed = ms, cand, True
        if not changed:
            break
    return TuneResult(cur, best_ms, default_ms, trials)


def apply(tr, options: Dict[str, int]) -> None:
    for k, v in options.items():
        tr.set_option(k, v)


def tune_transport(tr, packets: int, seg0: int = 1 << 40, ipeel: int = 0,
                   space: Optional[Dict[str, Sequence[int]]] = None,
                   rounds: int = 1, log: Optional[Callable[[str], None]] = None) -> TuneResult:
    """Tune the launch shape of a Transport on passes of `packets` packets
    drawn from segments seg0.. (apart from the production passes' segments),
    then leave the best options set on it."""
    space = space or (PEEL_SPACE if ipeel else LUCY_SPACE)

    from ._lib import HcError

    def time_fn(opts):
        apply(tr, opts)
        try:
            _, ms = tr.run_scalars(packets, ipeel=ipeel, seg0=seg0)
        except HcError:         # a launch shape this grid's LDS cannot hold
            ms = float("inf")
        if log:
            log(f"autotune {opts}: {ms:.1f} ms")
        return ms

    start = DEFAULTS_FP32 if getattr(tr, "precision", 0) == 1 else DEFAULTS
    res = coordinate_search(time_fn, space, start=start, rounds=rounds)
    apply(tr, res.options)
    return res
