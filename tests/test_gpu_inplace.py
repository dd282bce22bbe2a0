"""In-place streaming (wfs_set_inplace) against the two-buffer step.

The reference streams from one population array into another
(floodsim/lbm3d.py:50-60). With in-place streaming the solver keeps one
buffer and alternates an odd step (pull from the neighbours' exchanged slots,
push to where the next step reads) and an even step (own slots only). The
two-buffer path is pinned to the oracle bit for bit elsewhere; here the
in-place path must give the same macros after every step count, the same
populations and force budget after every even count, on the parity channel
(hull, propellers, wall model), uneven builder rows, raw reference fixtures,
the tile-sparse meander and with the learned closure, in every storage
format."""
import numpy as np
import pytest

from tests.helpers import channel128, solver_inputs, step_fixture
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.lbm3d import LBM3D, CX

pytestmark = pytest.mark.gpu


def _pair(reach, storage="f32"):
    a = solver_inputs(reach)
    out = []
    for streaming in ("ab", "inplace"):
        g = LBM3D(None, a["solid"], a["Fx"], a["Fy"], a["Fz"], a["uin"], a["tau0"], a["cs"], a["cf_wall"],
                  storage=storage, streaming=streaming)
        if reach.bedwall is not None:
            g.set_bedwall(reach.bedwall)
            g.update_uref()
        out.append(g)
    return out


def _kept(shape):
    """Population entries both schemes keep: all but those leaving the box
    through an x-end row (cx = -1 at x = 0, cx = +1 at x = nx - 1), which no
    step reads."""
    nz, ny, nx = shape
    m = np.ones((19, nz, ny, nx), bool)
    for i in range(19):
        if CX[i] == -1:
            m[i, :, :, 0] = False
        elif CX[i] == 1:
            m[i, :, :, nx - 1] = False
    return m.ravel()


def _same_macros(g, h, fluid):
    for u, v in zip(g.macro(), h.macro()):
        assert np.array_equal(u[fluid], v[fluid])


@pytest.mark.parametrize("storage", ["f32", "f32fast", "f16dev", "bf16dev", "f64"])
def test_inplace_matches_two_buffers_on_channel(storage):
    print(f"case: 128x32x16 channel, {storage}, in-place vs two buffers", flush=True)
    reach = channel128()
    ab, aa = _pair(reach, storage)
    fluid = ~reach.solid.ravel()
    kept = _kept(reach.solid.shape) & np.tile(fluid, 19)
    # half the population memory (the second buffer is never allocated)
    esz = {"f32": 4, "f32fast": 4, "f16dev": 2, "bf16dev": 2, "f64": 8}[storage]
    assert ab.memory_mb() - aa.memory_mb() > 0.99 * 19 * reach.solid.size * esz / 1e6
    assert aa.inplace and not ab.inplace
    for n in (1, 6, 13, 20):          # odd and even running counts: 1, 7, 20, 40
        ab.step(n)
        aa.step(n)
        _same_macros(ab, aa, fluid)
    fa, fb = ab.populations(np.float64), aa.populations(np.float64)
    assert np.array_equal(fa[kept], fb[kept])
    # and again after the readout converted the buffer there and back
    ab.step(10)
    aa.step(10)
    _same_macros(ab, aa, fluid)
    fa, fb = ab.populations(np.float64), aa.populations(np.float64)
    assert np.array_equal(fa[kept], fb[kept])
    ab.close()
    aa.close()


def test_inplace_force_budget_and_population_upload():
    print("case: 128x32x16 channel, f32, forces, set_populations, switching off", flush=True)
    reach = channel128()
    ab, aa = _pair(reach)
    fluid = ~reach.solid.ravel()
    hull = np.zeros(reach.solid.shape, bool)
    hull[reach.solid] = True
    for g in (ab, aa):
        g.set_body_mask(hull)
        g.step(30)
    fa, fb = ab.forces(), aa.forces()
    for k in fa:
        assert np.array_equal(np.asarray(fa[k]), np.asarray(fb[k])), k
    # an upload of the two-buffer state continues bit for bit
    aa.set_populations(ab.populations())
    for g in (ab, aa):
        g.step(11)
    _same_macros(ab, aa, fluid)
    aa.step(1)
    ab.step(1)
    # back to two buffers: the same path onwards
    aa.set_inplace(False)
    assert not aa.inplace
    for g in (ab, aa):
        g.step(9)
    _same_macros(ab, aa, fluid)
    kept = _kept(reach.solid.shape) & np.tile(fluid, 19)
    assert np.array_equal(ab.populations()[kept], aa.populations()[kept])
    ab.close()
    aa.close()


def test_inplace_readouts_need_an_even_count():
    print("case: 128x32x16 channel, readouts after an odd number of in-place steps", flush=True)
    reach = channel128()
    _, aa = _pair(reach)
    aa.step(3)
    for call in (aa.populations, aa.forces, lambda: aa.set_inplace(False)):
        with pytest.raises(ValueError, match="even number"):
            call()
    aa.macro()            # macros are written by every step
    aa.step(1)
    aa.populations()
    aa.close()


@pytest.mark.parametrize("name", ["trap_jets", "trap_tug", "nowall_longtail"])
def test_inplace_matches_two_buffers_on_builder_cases(name):
    """Padded row pitch, trapezoid banks, floodplain, under-hull jets."""
    print(f"case: builder {name}, f32, 60 steps in-place vs two buffers", flush=True)
    from tests.test_builder import CASES
    reach = R.build_reach(*CASES[name]())
    ab, aa = _pair(reach)
    fluid = ~reach.solid.ravel()
    for n in (17, 43):
        ab.step(n)
        aa.step(n)
        _same_macros(ab, aa, fluid)
    kept = _kept(reach.solid.shape) & np.tile(fluid, 19)
    assert np.array_equal(ab.populations()[kept], aa.populations()[kept])
    ab.close()
    aa.close()


@pytest.mark.parametrize("case", ["a", "b"])
def test_inplace_on_raw_fixture_state(case):
    """The reference's raw step fixtures (non-equilibrium populations on
    every cell, x-end rows included): uploaded, stepped both ways."""
    print(f"case: fixture step_bc_{case}, in-place vs two buffers", flush=True)
    from tests.test_gpu_parity import _lbm_from_fixture
    inp, _ = step_fixture(case)
    ab, aa = _lbm_from_fixture(inp), _lbm_from_fixture(inp)
    aa.set_inplace(True)
    nx, ny, nz = (int(v) for v in inp["dims"])
    fluid = inp["solid"] == 0
    kept = _kept((nz, ny, nx)) & np.tile(fluid, 19)
    for n in (1, 1, 2):
        ab.step(n)
        aa.step(n)
        _same_macros(ab, aa, fluid)
    assert np.array_equal(ab.populations()[kept], aa.populations()[kept])
    ab.close()
    aa.close()


@pytest.mark.parametrize("nx,ny,nz,storage", [(3, 2, 2, "f32"), (4, 3, 2, "f16dev"), (5, 2, 3, "f64"), (67, 3, 3, "f32")])
def test_inplace_on_tiny_grids(nx, ny, nz, storage):
    """The x-end rows side by side (nx = 3), a lid over one layer of fluid,
    a solid cell, random non-equilibrium populations everywhere; nx = 67
    puts the outlet row in the second wavefront."""
    print(f"case: {nx}x{ny}x{nz} random state, {storage}, in-place vs two buffers", flush=True)
    rng = np.random.default_rng(nx * 100 + ny * 10 + nz)
    solid = np.zeros((nz, ny, nx), bool)
    solid[0, 0, nx // 2] = True
    sh = solid.shape
    F = [(1e-4 * rng.standard_normal(sh)).astype(np.float32) for _ in range(3)]
    cf = np.where(solid, 0, 2e-3 * rng.random(sh)).astype(np.float32)
    uin = np.full((nz, ny), 0.02, np.float32)
    f0 = (np.tile(np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, np.float32)[:, None], (1, solid.size))
          * (1 + 0.05 * rng.standard_normal((19, solid.size)))).astype(np.float32)
    runs = []
    for streaming in ("ab", "inplace"):
        g = LBM3D(None, solid, *F, uin, 0.52, cf_wall=cf, storage=storage, streaming=streaming)
        g.set_populations(f0)
        runs.append(g)
    fluid = ~solid.ravel()
    kept = _kept(sh) & np.tile(fluid, 19)
    for n in (1, 3, 4):
        for g in runs:
            g.step(n)
        _same_macros(runs[0], runs[1], fluid)
    assert np.array_equal(runs[0].populations(np.float64)[kept], runs[1].populations(np.float64)[kept])
    for g in runs:
        g.close()


def test_inplace_with_per_step_column_u():
    print("case: 128x32x16 channel, f16dev, per-step column speed, in-place vs two buffers", flush=True)
    reach = channel128()
    ab, aa = _pair(reach, "f16dev")
    for g in (ab, aa):
        g.step(2)
        g.set_column_u(True)
        g.step(25)
    _same_macros(ab, aa, ~reach.solid.ravel())
    ab.close()
    aa.close()


def test_run3d_inplace_matches_two_buffers():
    print("case: Run3D on the channel, f32, streaming inplace vs ab, 60 steps", flush=True)
    runs = [R.Run3D(channel128(), streaming=s) for s in ("ab", "inplace")]
    for r in runs:
        r.advance(60)
    assert runs[1].lbm.inplace and runs[0].navg == runs[1].navg
    fl = ~runs[0].r.solid
    for k in ("ux", "uy", "uz", "rho"):
        assert np.array_equal(runs[0].field(k)[fl], runs[1].field(k)[fl]), k
    fa, fb = (r.forces() for r in runs)
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    for r in runs:
        r.close()


def test_inplace_beyond_32_bit_offsets():
    """A population buffer over 4 GiB (FP64, 28.4 M cells) takes the kernels
    with 64-bit addresses; smaller ones take 32-bit buffer offsets."""
    nz, ny, nx = 300, 296, 320
    print(f"case: {nx}x{ny}x{nz} open channel, f64 (buffer > 4 GiB), in-place vs two buffers, 4 steps", flush=True)
    assert 19 * nx * ny * nz * 8 >= 2 ** 32
    solid = np.zeros((nz, ny, nx), bool)
    solid[0] = True
    solid[:, :8, 100:120] = True                   # a groyne from the left bank
    z3 = np.zeros(solid.shape, np.float32)
    z3[150, 140:150, 60] = 1e-5                    # a point force
    uin = np.where(solid[:, :, 0], 0.0, 0.04).astype(np.float32)
    out = []
    for inplace in (False, True):
        g = LBM3D(None, solid, z3, z3, z3, uin, 0.51, storage="f64")
        if inplace:               # switched on after the upload: the state is converted, a buffer freed
            g.set_inplace(True)
        g.step(4)
        out.append(g.macro(np.float64))
        g.close()
    fluid = ~solid.ravel()
    for u, v in zip(*out):
        assert np.array_equal(u[fluid], v[fluid])


def _runs(reach, **kw):
    return [R.Run3D(reach, streaming=s, **kw) for s in ("ab", "inplace")]


def _same_runs(runs, fluid, closure=False):
    (a, b) = runs
    sa, na = a.lbm.sums()
    sb, nb = b.lbm.sums()
    assert na == nb
    for k in sa:
        assert np.array_equal(sa[k][fluid], sb[k][fluid]), k
    _same_macros(a.lbm, b.lbm, fluid)
    if closure:
        assert np.array_equal(a.lbm.closure_field()[fluid], b.lbm.closure_field()[fluid])
    fa, fb = a.forces(), b.forces()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k


@pytest.mark.parametrize("storage", ["f32", "f16dev"])
def test_inplace_on_tiles_matches_two_buffers(storage):
    print(f"case: meander at scale 0.05 (625x317x50 box), tiles, {storage}, in-place vs two buffers", flush=True)
    from dataclasses import replace
    from water_flow_simulation_amd.meander import meander_case
    r = replace(meander_case(scale=0.05), layout="tiles")
    runs = _runs(r, storage=storage)
    assert runs[1].lbm.memory_mb() < runs[0].lbm.memory_mb()
    fluid = ~r.solid.ravel()
    for m in (40, 20):
        for run in runs:
            run.advance(m)
    _same_runs(runs, fluid)
    kept = _kept(r.solid.shape) & np.tile(fluid, 19)
    assert np.array_equal(runs[0].lbm.populations()[kept], runs[1].lbm.populations()[kept])
    for run in runs:                  # an odd count: macros still agree
        run.lbm.step(7)
    _same_macros(runs[0].lbm, runs[1].lbm, fluid)
    for run in runs:
        run.close()


def test_inplace_on_tiles_beyond_32_bit_offsets():
    """A tile-sparse buffer over 4 GiB (62 M stored cells, FP32): one buffer
    resource per access instead of one for the whole buffer."""
    print("case: meander at scale 0.25 (3125x777x50 box), tiles, f32 (buffer > 4 GiB), in-place vs two buffers",
          flush=True)
    from dataclasses import replace
    from water_flow_simulation_amd.meander import meander_case
    r = replace(meander_case(scale=0.25), layout="tiles")
    runs = _runs(r, storage="f32")
    assert 19 * 4 * runs[1].lbm.stored_cells >= 2 ** 32
    for run in runs:
        run.advance(20)
    _same_runs(runs, ~r.solid.ravel())
    for run in runs:
        run.close()


@pytest.mark.parametrize("layout,storage", [("dense", "f32"), ("dense", "f16dev"), ("tiles", "f16dev")])
def test_inplace_with_learned_closure(layout, storage):
    print(f"case: meander at scale 0.05, learned closure, {layout}, {storage}, in-place vs two buffers", flush=True)
    from dataclasses import replace
    from water_flow_simulation_amd import closure as CL
    from water_flow_simulation_amd.meander import meander_case
    r = replace(meander_case(scale=0.05), layout=layout)
    runs = _runs(r, storage=storage, closure=CL.init_params(3))
    fluid = ~r.solid.ravel()
    for run in runs:
        run.advance(40)
    _same_runs(runs, fluid, closure=True)
    for run in runs:
        run.lbm.step(3)
    _same_macros(runs[0].lbm, runs[1].lbm, fluid)
    for run in runs:
        run.close()


def test_inplace_refused_for_a_batch_case():
    """A batch holds pointers to both population buffers of its cases, so a
    case cannot switch to one buffer while the batch exists."""
    print("case: two small seeded reaches in a batch, set_inplace on a case", flush=True)
    from water_flow_simulation_amd.batch import BatchRun
    from water_flow_simulation_amd.sweep import ensemble
    reaches, _ = ensemble(2, seed=3, scale=0.2)
    b = BatchRun(reaches)
    b.advance(20)
    b.synchronize()
    with pytest.raises(ValueError, match="case of a batch"):
        b.runs[0].lbm.set_inplace(True)
    b.set_group(1)                    # new batches, same refusal
    with pytest.raises(ValueError, match="case of a batch"):
        b.runs[1].lbm.set_inplace(True)
    b.close()


def test_inplace_rejections():
    print("case: in-place streaming with the outlet row opening a wavefront", flush=True)
    nz, ny = 6, 8
    solid = np.zeros((nz, ny, 129), bool)
    solid[0] = True
    z3 = np.zeros(solid.shape, np.float32)
    with pytest.raises(ValueError, match="multiple of 64"):
        LBM3D(None, solid, z3, z3, z3, np.full((nz, ny), 0.02, np.float32), 0.51, streaming="inplace")
    g = LBM3D(None, solid, z3, z3, z3, np.full((nz, ny), 0.02, np.float32), 0.51)
    with pytest.raises(ValueError, match="multiple of 64"):
        g.set_inplace(True)
    g.step(2)                         # the two-buffer path is unaffected
    g.close()
