"""Per-step column velocity for the wall model (wfs_set_column_u).

With the mode on, every step refreshes the wall model's column speed from the
previous step's velocity inside the step kernel. That is the reference's
Run3D cadence (floodsim/reach3d.py:427-440, uref refreshed once per 20-step
chunk) run with a chunk of one step, so the FP32 oracle driven with cadence 1
must agree bit for bit; the tile-sparse layout must agree with the dense one;
and the departure from the 20-step cadence is measured."""
import numpy as np
import pytest
from dataclasses import replace

from oracle.oracle import OracleLBM, drive
from tests.helpers import channel128, rel_l2, solver_inputs
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.lbm3d import LBM3D
from water_flow_simulation_amd.meander import MeanderConfig, build_meander

pytestmark = pytest.mark.gpu


def _gpu(reach, storage="f32"):
    a = solver_inputs(reach)
    g = LBM3D(None, a["solid"], a["Fx"], a["Fy"], a["Fz"], a["uin"], a["tau0"], a["cs"], a["cf_wall"],
              storage=storage)
    g.set_bedwall(reach.bedwall)
    g.update_uref()
    return g, a


def test_per_step_column_u_matches_oracle_with_cadence_one():
    print("case: 128x32x16 channel, f32, 60 steps, per-step column speed vs oracle cadence 1", flush=True)
    reach = channel128()
    g, a = _gpu(reach)
    g.set_column_u(True)
    g.step(25)
    g.step(35)
    orc = OracleLBM(a["solid"], a["Fx"], a["Fy"], a["Fz"], a["uin"], a["tau0"], a["cs"], a["cf_wall"],
                    precision="f32")
    drive(orc, reach.bedwall, 60, cadence=1)
    rho, ux, uy, uz = g.macro()
    fl = ~reach.solid.ravel()
    for u, v in ((rho, orc.rho), (ux, orc.ux), (uy, orc.uy), (uz, orc.uz)):
        assert np.array_equal(u[fl], v[fl])
    # and it is not the 20-step cadence
    orc20 = OracleLBM(a["solid"], a["Fx"], a["Fy"], a["Fz"], a["uin"], a["tau0"], a["cs"], a["cf_wall"],
                      precision="f32")
    drive(orc20, reach.bedwall, 60)
    assert not np.array_equal(ux[fl], orc20.ux[fl])
    # switching back restores the cadence (uref from update_uref reaches the step again)
    g.set_column_u(False)
    g.update_uref()
    g.step(20)
    drive(orc, reach.bedwall, 20)
    rho, ux, uy, uz = g.macro()
    assert np.array_equal(ux[fl], orc.ux[fl])
    g.close()


@pytest.mark.parametrize("storage", ["f32", "f16dev"])
def test_per_step_column_u_tiles_match_dense(storage):
    print(f"case: small meander with a hull, {storage}, per-step column speed, dense vs tiles", flush=True)
    mc = MeanderConfig(length=30.0, amplitude=2.0, wavelength=30.0, bed=6.0, bank=2.0, depth=1.6,
                       rows=(10.0,), per_row=1)
    r = build_meander(mc)
    runs = [R.Run3D(replace(r, layout=lay), storage=storage, column_u="per_step") for lay in ("dense", "tiles")]
    for run in runs:
        run.advance(60)
    fl = ~r.solid.ravel()
    (a, b) = (run.lbm.macro() for run in runs)
    for u, v in zip(a, b):
        assert np.array_equal(u[fl], v[fl])
    for run in runs:
        run.close()


def test_per_step_column_u_departure_from_cadence():
    """The measured deviation: after 400 steps of the 128x32x16 channel the
    per-step refresh moves the velocity field by a small fraction of the
    difference the wall model makes at all."""
    print("case: 128x32x16 channel, f32, 400 steps, per-step vs 20-step column speed", flush=True)
    reach = channel128()
    fl = ~reach.solid.ravel()
    out = {}
    for mode in ("cadence", "per_step", "none"):
        g, a = _gpu(reach)
        if mode == "per_step":
            g.set_column_u(True)
        if mode == "none":
            g.uref = np.zeros(g.N, np.float32)   # wall model without the column speed
        for _ in range(20):
            if mode == "cadence":
                g.update_uref()
            g.step(20)
        out[mode] = np.concatenate([u[fl] for u in g.macro()[1:]])
        g.close()
    d_step = rel_l2(out["per_step"], out["cadence"])
    d_wall = rel_l2(out["none"], out["cadence"])
    print(f"velocity rel. L2, per-step vs 20-step column speed: {d_step:.3e}; "
          f"vs no column speed: {d_wall:.3e}", flush=True)
    assert np.isfinite(out["per_step"]).all()
    assert 0 < d_step < 0.5 * d_wall
