"""The two largest workloads at their real sizes.

- The boat-test sweep ensemble: 32 seeded reaches of 100-200 m x 20-60 m at
  dx 0.2 m (111 M cells in all) run as one batch must equal the same reaches
  run one after another, bit for bit (populations, fields and time-mean sums
  over every fluid cell).
- The 1 km meander with eight pusher tugs: 1.56 G-cell box, 0.70 G cells
  stored, FP16 storage streamed in place with the learned closure, must stay
  finite over 220 steps with its FP64 budgets (mass, inlet discharge) where
  the physics puts them.
"""
import numpy as np
import pytest

from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.batch import BatchRun
from water_flow_simulation_amd.sweep import ensemble

pytestmark = pytest.mark.gpu


def test_full_ensemble_batched_equals_sequential_bitwise():
    print("case: 32-reach boat-test ensemble (seed 0, full size, FP32), 60 + 40 steps, batched vs sequential",
          flush=True)
    reaches, labels = ensemble(32, seed=0)
    cells = sum(r.cells for r in reaches)
    assert cells > 100e6
    pattern = (60, 40)
    seq = []
    for r in reaches:
        run = R.Run3D(r)
        for m in pattern:
            run.advance(m)
        seq.append(run)
    b = BatchRun(reaches)
    for m in pattern:
        b.advance(m, use_graph=True)
    for k, (rs, rb) in enumerate(zip(seq, b.runs)):
        fl = ~reaches[k].solid.ravel()
        assert rs.t == rb.t and rb.lbm.nstep == sum(pattern), labels[k]
        (s_sums, s_n), (b_sums, b_n) = rs.lbm.sums(), rb.lbm.sums()
        assert s_n == b_n, labels[k]
        for name in s_sums:
            assert np.array_equal(s_sums[name][fl], b_sums[name][fl]), (labels[k], name)
        for a, c in zip(rs.lbm.macro(), rb.lbm.macro()):
            assert np.array_equal(a[fl], c[fl]), labels[k]
        m19 = np.tile(fl, 19)
        assert np.array_equal(rs.lbm.populations()[m19], rb.lbm.populations()[m19]), labels[k]
        rs.close()
    print(f"{len(reaches)} cases, {cells / 1e6:.1f} M cells: batched == sequential", flush=True)
    b.close()


def test_full_meander_fp16_in_place_learned():
    print("case: 1 km meander (scale 1.0), FP16 in place, learned closure, 20 + 200 steps", flush=True)
    from water_flow_simulation_amd import meander
    logs = []
    res = meander.run(scale=1.0, steps=200, warmup=20, storage="f16dev", learned=True, streaming="inplace",
                      report_every=100, log=lambda m: (logs.append(m), print(m, flush=True)))
    assert res["box_cells"] > 1.5e9 and res["stored_cells"] > 0.65e9
    assert res["closure"] == "learned" and res["streaming"] == "inplace"
    # the run checks the whole state for non-finite values every 100 steps and at the end
    mass = res["mass_fp64"] / res["fluid_cells"]
    assert abs(mass - 1.0) < 3e-3, mass
    q_in, q_out = res["discharge_in_out"]
    assert abs(q_in - res["Q"]) < 0.02 * res["Q"], (q_in, res["Q"])
    # the start-up state is the inlet profile copied along x, which the bends shift sideways: the outlet
    # discharge stays away from Q until the flow has crossed the reach (tens of thousands of steps)
    assert np.isfinite(q_out) and abs(q_out) < 2.0 * res["Q"], (q_out, res["Q"])
    # the thrust ramps in over 20 s of simulated time: 0.11 s in, it is a small fraction of the design
    for name, f in res["forces_N"].items():
        assert np.all(np.isfinite(f)), name
    assert 0.0 < abs(res["forces_N"]["thrust"][0]) < 0.01 * res["thrust_design_N"]
    # 38 B of FP16 populations, 29 B of per-cell fields and 2 B of closure geometry per stored cell
    assert res["device_gb"] < 55.0, res["device_gb"]
    print("kernel ms", res["kernel_ms"], "device GB", res["device_gb"], flush=True)
