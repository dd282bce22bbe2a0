"""The boat-test runner's host logic: level selection, the test plan (one
base per level the variants use, built without boats) and the step count a
run to the end takes."""
import math

import numpy as np
import pytest

from water_flow_simulation_amd import boat_tests as B
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.geometry import Section
from water_flow_simulation_amd.hydraulics1d import q_ref

SEC = Section(B=6.0, Hb=2.5, z=1.0, Bf=0.0, n=0.03)
SLOPE = 2.2e-4


def small_test():
    return dict(key="T", title="small", objective="uneven cases", variants=[
        dict(label="tug x1 low", boat="tug", rows=1, per_row=1, level="low"),
        dict(label="long-tail x2 mid", boat="longtail", rows=1, per_row=2, level="mid"),
        dict(label="jet x1 mid", boat="jet_hp", rows=1, per_row=1, level="mid", over=dict(spread_deg=30)),
    ])


def small_plan(cfg=None):
    cfg = cfg or R.ReachConfig(length=24.0, dx=0.25, run_flows=0.25)
    return B.plan_test(small_test(), SEC, [2.4, 1.5, 2.0], 2.0, cfg, lambda h: (float(q_ref(h, 1.0, SLOPE, SEC)), SLOPE),
                       x_first=6.0), cfg


@pytest.mark.parametrize("levels,which,Hb,want", [
    ([3.0, 1.0, 2.0, 4.0], "low", 2.5, 1.0),
    ([3.0, 1.0, 2.0, 4.0], "high", 2.5, 4.0),
    ([3.0, 1.0, 2.0, 4.0], "mid", 2.5, 2.0),      # in bank: 1, 2 -> the second
    ([3.0, 1.0, 2.0, 4.0], "mid", 3.5, 2.0),      # in bank: 1, 2, 3 -> the middle one
    ([5.0, 6.0], "mid", 2.0, 6.0),                # none in bank: all levels count
    ([1.0, 2.0], 1.7, 2.0, 1.7),                  # a number passes through
])
def test_pick_level(levels, which, Hb, want):
    assert B.pick_level(levels, which, Hb) == want


def test_plan_has_one_boatless_base_per_level():
    plan, _ = small_plan()
    assert len(plan.cases) == 5                    # 2 levels -> 2 bases, 3 variants
    bases = [c for c in plan.cases if c[3] is None]
    assert sorted(c[1] for c in bases) == [1.5, 2.0]
    for label, h, reach, kb in plan.cases:
        if kb is None:
            assert not reach.props and not reach.hulls and not np.any(reach.Fx)
        else:
            base = plan.cases[kb]
            assert base[3] is None and base[1] == h and reach.props
            assert reach.solid.shape == base[2].solid.shape
    assert [c[0] for _, c in plan.variants()] == [v["label"] for v in small_test()["variants"]]
    # each variant's flow is the 1-D discharge at its level; the base shares it
    for _, (label, h, reach, kb) in plan.variants():
        assert reach.Q == plan.cases[kb][2].Q == pytest.approx(float(q_ref(h, 1.0, SLOPE, SEC)), rel=0, abs=0)


def _run_stub(reach, t=0.0):
    run = R.Run3D.__new__(R.Run3D)
    run.r, run.t = reach, t
    run.t_total = reach.cfg.run_flows * reach.cfg.length / max(reach.V, 0.05)
    return run


def test_steps_to_end_counts_whole_advance_calls():
    """steps_to_end is the count of run_to_end's loop: whole chunks until the
    float64 time, accumulated 20 steps at a time, reaches t_total."""
    plan, _ = small_plan()
    counts = []
    for _, _, reach, _ in plan.cases:
        run = _run_stub(reach)
        n = run.steps_to_end(chunk=400)
        assert n % 400 == 0 and n > 0
        t = 0.0
        for _ in range(n // 20):
            t += 20 * reach.dt
        assert t >= run.t_total
        t_short = 0.0
        for _ in range((n - 400) // 20):
            t_short += 20 * reach.dt
        assert not t_short >= run.t_total
        assert math.isclose(n, math.ceil(run.t_total / (400 * reach.dt)) * 400, rel_tol=0, abs_tol=400)
        counts.append(n)
        assert _run_stub(reach, t=run.t_total).steps_to_end() == 0
    assert len(set(counts)) > 1                    # the cases end at different step counts
