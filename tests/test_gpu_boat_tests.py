"""A whole boat test (every variant and one base per level, each run to its
own end time) on the batched path against the same plan run one case after
another: every reported figure is identical."""
import numpy as np
import pytest

from tests.test_boat_tests import SEC, SLOPE, small_test
from water_flow_simulation_amd import boat_tests as B
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.hydraulics1d import q_ref

pytestmark = pytest.mark.gpu


def _run(batched, **kw):
    cfg = R.ReachConfig(length=24.0, dx=0.25, run_flows=0.25)
    seen = []
    res = B.run_test(small_test(), None, SEC, [2.4, 1.5, 2.0], 2.0, cfg, None,
                     lambda h: (float(q_ref(h, 1.0, SLOPE, SEC)), SLOPE), x_first=6.0, batched=batched,
                     on_result=lambda c, run: seen.append((c["label"], run.lbm.nstep)), **kw)
    return res, seen


def _same(a, b):
    if isinstance(a, dict):
        assert a.keys() == b.keys()
        for k in a:
            _same(a[k], b[k])
    elif isinstance(a, np.ndarray):
        assert np.array_equal(a, b, equal_nan=True)
    elif isinstance(a, float):
        assert a == b or (np.isnan(a) and np.isnan(b))
    else:
        assert a == b


def test_run_test_batched_equals_sequential():
    print("case: boat test with 3 variants at 2 levels (5 runs, 4200-8600 steps), batched vs sequential", flush=True)
    steps = []
    res_b, seen_b = _run(True, progress=lambda f, label: steps.append(f))
    res_s, seen_s = _run(False)
    assert [c["label"] for c in res_b] == [v["label"] for v in small_test()["variants"]]
    assert seen_b == seen_s and len({n for _, n in seen_b}) > 1      # uneven step counts
    for a, b in zip(res_b, res_s):
        _same(a, b)
    assert steps and steps[-1] == 1.0 and all(0 < f <= 1 for f in steps)
    keys = {"d_eta_cm", "jet_reach", "us_gain", "tau_area", "dQ_net_pct", "F_net_kN", "profiles", "base_profiles"}
    assert keys <= set(res_b[0])
    assert all(np.isfinite(c["profiles"]["eta"]).all() for c in res_b)


def test_run_test_cancel_stops_all_runs():
    print("case: boat test cancelled after the first segment", flush=True)
    calls = []

    def cancel():
        calls.append(1)
        return len(calls) > 2      # the first check is before planning

    res, seen = _run(True, cancel=cancel)
    assert len(res) == 3
    assert all(n == 800 for _, n in seen)          # two 400-step segments, then the stop


def test_run_test_carries_fp64_errors():
    """With fp64_errors every result carries |value - FP64 value| for each
    reported quantity; the FP64 values are those of the same plan run with
    FP64 storage on its own."""
    print("case: boat test, FP16 storage with its FP64 twin", flush=True)
    from water_flow_simulation_amd.boat_tests import FP64_ERROR_KEYS
    res, _ = _run(True, storage="f16dev", fp64_errors=True)
    ref, _ = _run(True, storage="f64")
    for c, c64 in zip(res, ref):
        assert set(FP64_ERROR_KEYS) <= set(c["err_fp64"])
        for k in FP64_ERROR_KEYS:
            assert c["fp64"][k] == c64[k] or (np.isnan(c["fp64"][k]) and np.isnan(c64[k])), k
            assert c["err_fp64"][k] == abs(float(c[k]) - float(c64[k])), k
        for k in ("eta_x_max_mm", "us_x_max", "q_x_max_rel", "tau_bed_max"):
            assert np.isfinite(c["err_fp64"][k]) and c["err_fp64"][k] >= 0, k
