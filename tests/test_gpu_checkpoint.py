"""Checkpoints (wfs_state_save / wfs_state_load, Run3D.checkpoint/restore):
a run saved mid-way and restored into a fresh solver built from the same
reach continues bit for bit: macroscopic fields, time-mean sums, populations
and the FP64 force budget, in every storage format, on both layouts, with
in-place streaming (also after an odd step count) and with the per-step
column speed."""
from dataclasses import replace

import numpy as np
import pytest

from tests.helpers import channel128
from water_flow_simulation_amd import reach3d as R

pytestmark = pytest.mark.gpu


def _snapshot(run):
    sums, navg = run.lbm.sums()
    return dict(macro=run.lbm.macro(np.float64), sums=sums, navg=navg, nstep=run.lbm.nstep, t=run.t)


def _same(a, b, fluid):
    assert a["navg"] == b["navg"] and a["nstep"] == b["nstep"] and a["t"] == b["t"]
    for u, v in zip(a["macro"], b["macro"]):
        assert np.array_equal(u[fluid], v[fluid])
    for k in a["sums"]:
        assert np.array_equal(a["sums"][k][fluid], b["sums"][k][fluid]), k


@pytest.mark.parametrize("storage,layout,streaming,column_u", [
    ("f32", "dense", "ab", "cadence"),
    ("f16dev", "dense", "ab", "cadence"),
    ("bf16dev", "tiles", "ab", "cadence"),
    ("f64", "dense", "ab", "cadence"),
    ("f16dev", "tiles", "inplace", "cadence"),
    ("f32", "dense", "inplace", "per_step"),
])
def test_checkpoint_restore_continues_bitwise(tmp_path, storage, layout, streaming, column_u):
    print(f"case: 128x32x16 channel, {storage}, {layout}, {streaming}, column speed {column_u}", flush=True)
    reach = replace(channel128(), layout=layout)
    kw = dict(storage=storage, streaming=streaming, column_u=column_u)
    a = R.Run3D(reach, **kw)
    a.t_avg = 0.0                     # average from the start, so the sums take part
    a.advance(60)
    a.lbm.step(3)                     # an odd count: in place, mid-pair
    path = tmp_path / "run.npz"
    a.checkpoint(path)
    b = R.Run3D(replace(channel128(), layout=layout), **kw)
    b.t_avg = 0.0
    b.restore(path)
    for run in (a, b):
        run.lbm.step(1)               # back to an even count
        run.advance(40)
    fluid = ~reach.solid.ravel()
    _same(_snapshot(a), _snapshot(b), fluid)
    pa, pb = a.lbm.populations(), b.lbm.populations()
    m19 = np.tile(fluid, 19)
    assert np.array_equal(pa[m19], pb[m19])
    fa, fb = a.forces(), b.forces()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    a.close()
    b.close()


def test_checkpoint_refuses_another_solver(tmp_path):
    print("case: 128x32x16 channel, state loaded into solvers built differently", flush=True)
    reach = channel128()
    a = R.Run3D(reach)
    a.advance(20)
    st = a.lbm.state()
    for kw, msg in ((dict(storage="f16dev"), "storage format, layout or size"),
                    (dict(streaming="inplace"), "streaming mode"),
                    (dict(column_u="per_step"), "column speed")):
        b = R.Run3D(channel128(), **kw)
        with pytest.raises(ValueError, match=msg):
            b.lbm.load_state(st)
        b.close()
    with pytest.raises(ValueError, match="not a solver state"):
        a.lbm.load_state(np.zeros(st.size, np.uint8))
    with pytest.raises(ValueError, match="size"):
        a.lbm.load_state(st[:-8])
    a.close()


def test_checkpoint_refuses_another_geometry_of_the_same_size():
    """Two reaches with the same stored-cell count but a different lattice
    shape or solid mask: the state of one is refused by the other."""
    print("case: 128x32x16 channel vs 128x16x32 box and a shifted hull", flush=True)
    from water_flow_simulation_amd.lbm3d import LBM3D
    def solver(solid):
        nz, ny, nx = solid.shape
        z = np.zeros(solid.shape, np.float32)
        return LBM3D(None, solid, z, z, z, np.full((nz, ny), 0.05, np.float32), 0.5006)
    s1 = np.zeros((16, 32, 128), bool)
    s1[:, 0] = s1[:, -1] = True
    s2 = np.zeros((32, 16, 128), bool)
    s2[:, 0] = s2[:, -1] = True
    s3 = s1.copy()
    s3[8:, 14:18, 40:60] = True
    a = solver(s1)
    a.step(4)
    st = a.state()
    for other in (s2, s3):
        b = solver(other)
        with pytest.raises(ValueError, match="another geometry"):
            b.load_state(st)
        b.close()
    c = solver(s1.copy())
    c.load_state(st)
    c.step(2)
    a.step(2)
    assert np.array_equal(a.ux, c.ux)
    for L in (a, c):
        L.close()
