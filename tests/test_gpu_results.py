"""A web-viewer record and base profiles from a real run of the reference's
cached base case (B50 trapezoid with floodplain, h = 3 m, 600 x 250 x 12):
same sizes and solid blocks as the reference's record, and the decoded
fields agree with the run's own time means."""
import hashlib
import json
from dataclasses import replace

import numpy as np
import pytest

from tests.test_results import FIX, _reach
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd import results as RS

pytestmark = pytest.mark.gpu


def test_record_from_a_run(tmp_path):
    print("case: reference base case 600x250x12, 200 steps averaged, web record + base profiles", flush=True)
    r = _reach(False)
    r = replace(r, cfg=replace(r.cfg, avg_frac=1.0))   # average from the first chunk
    run = R.Run3D(r)
    run.advance(200)
    assert run.navg == 10
    res = R.profiles(run)
    rec = json.loads(json.dumps(RS.export_case(run, "base", None, res)))
    ref = FIX["records"]["base"]
    for k in ("nx", "ny", "nz", "zb", "hulls", "props"):
        assert rec[k] == ref[k], k
    vel, fluid = RS.decode_velocity(rec)
    solid = ~fluid
    assert hashlib.sha256(np.packbits(solid.ravel()).tobytes()).hexdigest() == ref["solid_sha256"]
    u = np.stack([run.field(k, True) for k in ("ux", "uy", "uz")], -1)
    mean_u, _ = RS.pool(u, (~r.solid).astype(np.float32))
    assert np.isfinite(vel).all() and rec["scale"] > 0
    # half a quantum, plus the float32 rounding of the block mean the record quantizes
    f32 = 4 * np.finfo(np.float32).eps * np.abs(mean_u).max()
    assert np.abs(vel[fluid] - mean_u[fluid]).max() <= 0.5 * rec["scale"] + f32
    assert np.isfinite(RS.decode_tau(rec)).all() and rec["tau_top"] >= 1.0
    p = str(tmp_path / f"base_profiles_{FIX['tag']}.npz")
    RS.save_base_profiles(p, res)
    back = RS.load_base_profiles(p)
    assert {k: [list(v.shape), str(v.dtype)] for k, v in back.items()} == FIX["base_profiles"]
    run.close()
