"""Population stagger (wfs_set_pop_stagger): a gap between the 19 population
direction arrays changes where they lie in memory and nothing else. A
staggered solver must give the packed solver's macros, populations, force
budget and state bit for bit, in every storage format, with two buffers and
in place, on both layouts; a state saved at one stagger loads at another;
the autotuner times the candidates and caches its pick."""
from dataclasses import replace

import numpy as np
import pytest

from tests.helpers import channel128
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd import _native as nat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage,layout,streaming,stagger", [
    ("f32", "dense", "ab", 1024),
    ("f16dev", "dense", "ab", 1920),
    ("bf16dev", "dense", "inplace", 128),
    ("f64", "dense", "ab", 3584),
    ("f16dev", "tiles", "inplace", 512),
    ("f32fast", "tiles", "ab", 2048),
])
def test_stagger_leaves_results_unchanged(storage, layout, streaming, stagger):
    print(f"case: 128x32x16 channel, {storage}, {layout}, {streaming}, stagger {stagger} bytes", flush=True)
    reach = replace(channel128(), layout=layout)
    runs = [R.Run3D(reach, storage=storage, streaming=streaming, pop_stagger=s) for s in (0, stagger)]
    packed, gapped = runs
    elem = {"f32": 4, "f32fast": 4, "f16dev": 2, "bf16dev": 2, "f64": 8}[storage]
    np_cells = nat.lib().wfs_stored_cells(packed.lbm.handle)
    assert nat.lib().wfs_pop_stride(packed.lbm.handle) == np_cells
    assert nat.lib().wfs_pop_stride(gapped.lbm.handle) == np_cells + stagger // elem
    assert gapped.lbm.pop_stagger == stagger and packed.lbm.pop_stagger == 0
    fluid = ~reach.solid.ravel()
    for run in runs:
        run.t_avg = 0.0
        run.advance(60)
        run.lbm.step(2)
    for u, v in zip(packed.lbm.macro(np.float64), gapped.lbm.macro(np.float64)):
        assert np.array_equal(u[fluid], v[fluid])
    m19 = np.tile(fluid, 19)
    assert np.array_equal(packed.lbm.populations()[m19], gapped.lbm.populations()[m19])
    fa, fb = packed.forces(), gapped.forces()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    sa, _ = packed.lbm.sums()
    sb, _ = gapped.lbm.sums()
    for k in sa:
        assert np.array_equal(sa[k][fluid], sb[k][fluid]), k
    # the saved state does not depend on the stagger
    assert np.array_equal(packed.lbm.state(), gapped.lbm.state())
    for run in runs:
        run.close()


@pytest.mark.parametrize("layout,streaming", [("dense", "inplace"), ("tiles", "ab")])
def test_stagger_with_the_learned_closure(layout, streaming):
    print(f"case: 128x32x16 channel, f16dev, learned closure, {layout}, {streaming}, stagger 0 vs 1536 bytes",
          flush=True)
    from water_flow_simulation_amd import closure as CL
    reach = replace(channel128(), layout=layout)
    w = CL.init_params(4)
    runs = [R.Run3D(reach, storage="f16dev", streaming=streaming, closure=w, pop_stagger=s) for s in (0, 1536)]
    for run in runs:
        run.advance(40)
    fluid = ~reach.solid.ravel()
    for u, v in zip(runs[0].lbm.macro(np.float64), runs[1].lbm.macro(np.float64)):
        assert np.array_equal(u[fluid], v[fluid])
    assert np.array_equal(runs[0].lbm.closure_field()[fluid], runs[1].lbm.closure_field()[fluid])
    for run in runs:
        run.close()


def test_state_loads_across_staggers(tmp_path):
    print("case: 128x32x16 channel, f16dev, state saved packed, loaded at a 1 KiB stagger", flush=True)
    reach = channel128()
    a = R.Run3D(reach, storage="f16dev")
    a.advance(40)
    a.checkpoint(tmp_path / "run.npz")
    b = R.Run3D(channel128(), storage="f16dev", pop_stagger=1024)
    b.restore(tmp_path / "run.npz")
    for run in (a, b):
        run.advance(40)
    fluid = ~reach.solid.ravel()
    for u, v in zip(a.lbm.macro(np.float64), b.lbm.macro(np.float64)):
        assert np.array_equal(u[fluid], v[fluid])
    a.close()
    b.close()


def test_stagger_rejections():
    print("case: 128x32x16 channel, invalid staggers", flush=True)
    for bad in (-128, 100, (1 << 20) + 128):
        with pytest.raises(ValueError, match="wfs_set_pop_stagger"):
            R.Run3D(channel128(), pop_stagger=bad)
    run = R.Run3D(channel128())
    with pytest.raises(ValueError, match="before the upload"):
        nat.check(nat.lib().wfs_set_pop_stagger(run.lbm.handle, 1024), "wfs_set_pop_stagger")
    run.close()


def test_stagger_autotuner(tmp_path, monkeypatch):
    print("case: channel128, stagger autotune over 0 and 512 bytes", flush=True)
    monkeypatch.setenv("WFS_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    from water_flow_simulation_amd import autotune
    monkeypatch.setattr(autotune, "_cache", None)
    monkeypatch.setattr(autotune, "STAGGERS", (0, 512))
    r = channel128()
    t = autotune.time_staggers(r, "f16dev", steps=10, warmup=2, staggers=(0, 512), rounds=1)
    assert sorted(t) == [0, 512] and all(0 < v < float("inf") for v in t.values())
    monkeypatch.setattr(autotune, "time_staggers",
                        lambda reach, storage: {0: 2.0, 512: 1.0})
    s = autotune.pop_stagger_for(r, "f16dev")
    assert s == 512 and (tmp_path / "tune.json").exists()
    run = R.Run3D(r, storage="f16dev", pop_stagger="auto")
    assert run.lbm.pop_stagger == 512
    run.close()
    tiles = R.Run3D(replace(channel128(), layout="tiles"), pop_stagger="auto")
    assert tiles.lbm.pop_stagger == 0
    tiles.close()


def test_released_buffers_return_device_memory():
    """Dropping a buffer (a new stagger re-allocates the populations,
    in-place streaming switched on after the upload frees the second buffer,
    the column buffers switched off) gives its memory back to the device, and
    the solver's count follows."""
    import ctypes
    from water_flow_simulation_amd.lbm3d import LBM3D
    hip = ctypes.CDLL("libamdhip64.so.7")    # the runtime instance the solver library uses

    def free_bytes():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value
    print("case: 500x128x64 box, f32, stagger 1 KiB, in place after the upload, column buffers on and off", flush=True)
    nz, ny, nx = 64, 128, 500
    solid = np.zeros((nz, ny, nx), bool)
    solid[0] = True
    z3 = np.zeros(solid.shape, np.float32)
    uin = np.full((nz, ny), 0.02, np.float32)
    nat.device_count()                       # the device is initialised
    free0 = free_bytes()
    g = LBM3D(None, solid, z3, z3, z3, uin, 0.51, pop_stagger=1024)
    g.step(2)
    g.set_inplace(True)
    bw = np.zeros(solid.shape, np.uint8)
    bw[1] = 1                                # the bed layer carries the wall model
    g.set_bedwall(bw)
    g.set_column_u(True)
    g.step(2)
    g.set_column_u(False)
    g.step(2)
    held = g.memory_mb() * 1e6
    used = free0 - free_bytes()
    pops = 19 * 4 * g.stored_cells
    assert used < held + 0.25 * pops, (used, held, pops)   # a dropped buffer would add one or more pops
    g.close()
    assert free0 - free_bytes() < 0.05 * pops
