"""The tile-sparse layout (only the 64-cell x-runs of a row that hold water
are stored) against the dense layout on the same geometry: the step, the
inlet/outlet rows, the run cadence (uref, ramp, time-mean sums) and the FP64
readouts must agree bit for bit on every fluid cell."""
from dataclasses import replace

import numpy as np
import pytest

from tests.helpers import channel128
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.meander import meander_case

pytestmark = pytest.mark.gpu


def _state(run):
    sums, navg = run.lbm.sums()
    return sums, navg, run.lbm.macro(), run.lbm.populations()


def _compare(reach, storage, pattern):
    dense = R.Run3D(replace(reach, layout="dense"), storage=storage)
    tiles = R.Run3D(replace(reach, layout="tiles"), storage=storage)
    assert tiles.lbm.stored_cells < dense.lbm.stored_cells
    for m in pattern:
        dense.advance(m)
        tiles.advance(m)
    fl = ~reach.solid.ravel()
    s_d, n_d, m_d, p_d = _state(dense)
    s_t, n_t, m_t, p_t = _state(tiles)
    assert n_d == n_t
    for k in s_d:
        assert np.array_equal(s_d[k][fl], s_t[k][fl]), k
    for a, b in zip(m_d, m_t):
        assert np.array_equal(a[fl], b[fl])
    m19 = np.tile(fl, 19)
    assert np.array_equal(p_d[m19], p_t[m19])
    for field in ("rho", "ux", "sum_ux"):
        assert np.array_equal(dense.lbm.profile_x(field), tiles.lbm.profile_x(field)), field
    assert dense.lbm.total("rho") == tiles.lbm.total("rho")
    tiles.lbm.check_finite()
    frac = tiles.lbm.stored_cells / dense.lbm.stored_cells
    dense.close()
    tiles.close()
    return frac


@pytest.mark.parametrize("storage", ["f32", "f16dev"])
def test_tiles_match_dense_on_the_parity_channel(storage):
    print(f"case: channel128 dense vs tiles, {storage}", flush=True)
    frac = _compare(channel128(), storage, (100, 35, 40))
    print("stored fraction", frac, flush=True)


@pytest.mark.parametrize("storage", ["f32", "f16dev"])
def test_tiles_match_dense_on_a_meander(storage):
    print(f"case: meander at scale 0.05 (625x317x50 box) dense vs tiles, {storage}", flush=True)
    r = meander_case(scale=0.05)
    assert r.props and r.force_cells[0].size > 0
    frac = _compare(r, storage, (60, 40))
    print("stored fraction", frac, flush=True)
    assert frac < 0.75


def test_tiles_learned_closure_and_force_budget():
    """The learned closure's step and the FP64 momentum budget on the tiled
    layout against the dense one (the budget's sums run in another order)."""
    print("case: meander at scale 0.05, learned closure and forces, dense vs tiles", flush=True)
    from water_flow_simulation_amd import closure as CL
    r = meander_case(scale=0.05)
    w = CL.init_params(3)
    runs = [R.Run3D(replace(r, layout=lay), storage="f32", closure=w) for lay in ("dense", "tiles")]
    for run in runs:
        run.advance(40)
    fl = ~r.solid.ravel()
    (a, b) = (run.lbm.macro() for run in runs)
    for u, v in zip(a, b):
        assert np.array_equal(u[fl], v[fl])
    c2a, c2b = (run.lbm.closure_field() for run in runs)
    assert np.array_equal(c2a[fl], c2b[fl])
    fa, fb = (run.forces() for run in runs)
    for k in fa:
        np.testing.assert_allclose(fb[k], fa[k], rtol=1e-12, atol=1e-9 * max(1.0, np.abs(fa[k]).max()))
    assert np.abs(fa["body"]).max() > 0 and np.abs(fa["thrust"]).max() > 0
    for run in runs:
        run.close()


def test_tiles_force_split_with_hull_longer_than_a_tile():
    """A 12 m hull is 150 cells long, so some 64-cell runs inside it are all
    solid and the tile layout stores no tile for them. The hull's share of the
    momentum budget must still match the dense layout (the body mask is looked
    up by dense cell, not by stored cell)."""
    print("case: small meander, one 12 m hull, f32, forces dense vs tiles", flush=True)
    from water_flow_simulation_amd.meander import MeanderConfig, build_meander
    mc = MeanderConfig(length=40.0, amplitude=3.0, wavelength=40.0, bed=8.0, bank=2.0, depth=2.0,
                       rows=(14.0,), per_row=1)
    r = build_meander(mc)
    hull = r.hull_cells
    nz, ny, nx = r.solid.shape
    ntx = nx // 64
    runs_solid = r.solid[:, :, :ntx * 64].reshape(nz, ny, ntx, 64).all(-1)
    runs_hull = hull[:, :, :ntx * 64].reshape(nz, ny, ntx, 64).any(-1)
    assert (runs_solid & runs_hull).any(), "the case must hold an all-solid tile with hull cells"
    runs = [R.Run3D(replace(r, layout=lay), storage="f32") for lay in ("dense", "tiles")]
    for run in runs:
        run.advance(20)
    fa, fb = (run.forces() for run in runs)
    for k in fa:
        np.testing.assert_allclose(fb[k], fa[k], rtol=1e-12, atol=1e-9 * max(1.0, np.abs(fa[k]).max()))
    assert np.abs(fa["body"][2]) > 0
    for run in runs:
        run.close()

