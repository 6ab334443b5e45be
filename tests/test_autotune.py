"""The autotuner's search on synthetic timing surfaces (no GPU)."""
from hochunk3d_amd.autotune import DEFAULTS, LUCY_SPACE, coordinate_search


def test_finds_the_minimum_of_a_separable_surface():
    best = {"regen_lanes": 12, "regen_run_lanes": 8, "blocks_per_cu": 1}

    def ms(o):
        return (100.0 + 3 * abs(o["regen_lanes"] - best["regen_lanes"])
                + 5 * abs(o["regen_run_lanes"] - best["regen_run_lanes"])
                + 40 * (o["blocks_per_cu"] != best["blocks_per_cu"]))

    r = coordinate_search(ms, LUCY_SPACE)
    for k, v in best.items():
        assert r.options[k] == v
    assert r.best_ms == 100.0
    assert r.default_ms == ms(DEFAULTS)
    assert r.gain > 1.0


def test_never_times_a_candidate_twice_and_respects_the_lane_constraint():
    calls = []

    def ms(o):
        calls.append(tuple(sorted(o.items())))
        assert o["regen_run_lanes"] <= o["regen_lanes"]
        return 50.0 + o["regen_lanes"] - 0.5 * o["regen_run_lanes"]

    r = coordinate_search(ms, LUCY_SPACE, rounds=3)
    assert len(calls) == len(set(calls))
    assert r.options["regen_lanes"] == 8
    assert r.options["regen_run_lanes"] == 8


def test_keeps_the_defaults_when_nothing_is_faster():
    r = coordinate_search(lambda o: 10.0, LUCY_SPACE)
    assert r.options == DEFAULTS
    assert r.gain == 1.0
