"""Device-memory layout of the solver: body forces in a side table over
64-cell groups, the time-mean sums allocated by the first accumulation, the
bed/bank mask as a flag bit and the learned closure's C^2 field optional.
Each is a storage choice with no effect on results: forces given as dense
arrays or as a cell list step alike, a run that never averages holds no
sums and its state loads either way, and the learned step gives the same
fields with or without its C^2 field."""
import numpy as np
import pytest

from tests.helpers import channel128, solver_inputs
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.lbm3d import LBM3D

pytestmark = pytest.mark.gpu


def _lbm(reach, storage="f32", forces=True, **kw):
    a = solver_inputs(reach)
    F = (a["Fx"], a["Fy"], a["Fz"]) if forces else (None, None, None)
    g = LBM3D(None, a["solid"], *F, a["uin"], a["tau0"], a["cs"], a["cf_wall"], storage=storage, **kw)
    g.set_bedwall(reach.bedwall)
    g.update_uref()
    return g


def _same(g, h, fluid):
    for u, v in zip(g.macro(np.float64), h.macro(np.float64)):
        assert np.array_equal(u[fluid], v[fluid])


@pytest.mark.parametrize("storage", ["f32", "f16dev", "f64"])
def test_forces_as_dense_arrays_or_cell_list(storage):
    print(f"case: 128x32x16 channel with its boat, {storage}, dense forces vs a cell list", flush=True)
    reach = channel128()
    a = solver_inputs(reach)
    idx = np.flatnonzero((a["Fx"] != 0) | (a["Fy"] != 0) | (a["Fz"] != 0))
    assert idx.size > 0
    dense, listed = _lbm(reach, storage), _lbm(reach, storage, forces=False)
    # a first entry of the same cell with other values: the later entry wins
    cells = np.concatenate([idx[:5], idx])
    vals = [np.concatenate([np.full(5, 7.0, np.float32), np.asarray(a[k]).ravel()[idx]]) for k in ("Fx", "Fy", "Fz")]
    listed.set_force_cells(cells, *vals)
    for g in (dense, listed):
        g.step(40)
    _same(dense, listed, ~reach.solid.ravel())
    # the force ramp scales the table
    for g in (dense, listed):
        g.set_force_scale(0.5)
        g.step(20)
    _same(dense, listed, ~reach.solid.ravel())
    dense.close()
    listed.close()


def test_time_mean_sums_allocated_by_the_first_accumulation(tmp_path):
    print("case: 128x32x16 channel, f32, memory before and after the first accumulation; states across", flush=True)
    reach = channel128()
    g = _lbm(reach)
    cells = g.stored_cells
    m0 = g.memory_mb()
    g.step(20)
    g.reset_sums()
    assert g.memory_mb() == m0                         # clearing sums that do not exist allocates nothing
    st = g.state()                                     # a state without sums: zeros in their place
    h = _lbm(channel128())
    h.load_state(st)
    assert h.memory_mb() == m0                         # and loading it allocates none
    for s in (g, h):
        s.accumulate()
    grown = g.memory_mb() - m0
    assert abs(grown - 6 * 4 * cells / 1e6) < 6 * 256 / 1e6 + 1e-9, grown
    for s in (g, h):
        s.step(20)
        s.accumulate()
    sg, ng = g.sums()
    sh, nh = h.sums()
    fluid = ~reach.solid.ravel()
    assert ng == nh == 2
    for k in sg:
        assert np.array_equal(sg[k][fluid], sh[k][fluid]), k
    # a state with sums loads into a solver that has none yet
    k2 = _lbm(channel128())
    k2.load_state(g.state())
    assert k2.memory_mb() == g.memory_mb()
    for s in (g, k2):
        s.step(20)
        s.accumulate()
    s1, _ = g.sums()
    s2, _ = k2.sums()
    for k in s1:
        assert np.array_equal(s1[k][fluid], s2[k][fluid]), k
    for s in (g, h, k2):
        s.close()


def test_sums_of_a_run_that_never_averaged_read_as_zero():
    print("case: 128x32x16 channel, f32, sums read before any accumulation", flush=True)
    g = _lbm(channel128())
    g.step(10)
    sums, n = g.sums()
    assert n == 0
    for k, v in sums.items():
        assert not np.any(v), k
    g.close()


def test_learned_closure_without_its_field():
    print("case: 128x32x16 channel, learned closure with and without the C^2 field", flush=True)
    from water_flow_simulation_amd import closure as CL
    w = CL.init_params(5)
    reach = channel128()
    on = R.Run3D(reach, closure=w)
    off = R.Run3D(channel128(), closure=w, closure_field=False)
    assert abs((on.lbm.memory_mb() - off.lbm.memory_mb()) * 1e6 - 4 * on.lbm.stored_cells) <= 256
    for run in (on, off):
        run.advance(40)
    fluid = ~reach.solid.ravel()
    _same(on.lbm, off.lbm, fluid)
    with pytest.raises(ValueError, match="field is off"):
        off.lbm.closure_field()
    off.lbm.keep_closure_field(True)
    assert not np.any(off.lbm.closure_field())     # zeros until the next step
    for run in (on, off):
        run.lbm.step(1)
    assert np.array_equal(on.lbm.closure_field()[fluid], off.lbm.closure_field()[fluid])
    on.close()
    off.close()
