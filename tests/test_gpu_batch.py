"""The batched sweep (one launch per step over all cases, device cadence,
HIP-graph replay) against the same cases run one after another."""
import numpy as np
import pytest

from tests.helpers import channel128
from tests.test_gpu_inplace import _kept
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.batch import BatchRun
from water_flow_simulation_amd.sweep import ensemble

pytestmark = pytest.mark.gpu


def _cases():
    small, _ = ensemble(3, seed=1, scale=0.2)
    from tests.test_builder import CASES
    nowall = R.build_reach(*CASES["nowall_longtail"]())        # no wall model, no uref
    return small + [channel128(), nowall]


def _state(run):
    sums, navg = run.lbm.sums()
    return sums, navg, run.lbm.macro(), run.lbm.populations()


@pytest.mark.parametrize("use_graph,group", [(True, 0), (False, 0), (True, 2)])
def test_batch_matches_sequential_bitwise(use_graph, group):
    print(f"case: batch of 5 reaches, graph={use_graph}, cases per launch={group or 'all'}", flush=True)
    reaches = _cases()
    pattern = (100, 35, 40)              # graph chunks, an odd remainder (parity flip), a re-captured graph
    seq = []
    for r in reaches:
        run = R.Run3D(r)
        for m in pattern:
            run.advance(m)
        seq.append(run)
    b = BatchRun(reaches, group=group)
    assert len(b._hs) == (1 if group == 0 else 3)
    for m in pattern:
        b.advance(m, use_graph=use_graph)
        # eager readout between graph replays
        _ = b.runs[0].lbm.macro()
    for k, (rs, rb) in enumerate(zip(seq, b.runs)):
        s_sums, s_navg, s_mac, s_pop = _state(rs)
        b_sums, b_navg, b_mac, b_pop = _state(rb)
        assert s_navg == b_navg and rs.t == rb.t, k
        fl = ~reaches[k].solid.ravel()
        for name in s_sums:
            assert np.array_equal(s_sums[name][fl], b_sums[name][fl]), (k, name)
        for a, c in zip(s_mac, b_mac):
            assert np.array_equal(a[fl], c[fl]), k
        m19 = np.tile(fl, 19)
        assert np.array_equal(s_pop[m19], b_pop[m19]), k
        assert rb.lbm.nstep == sum(pattern)
    for run in seq:
        run.close()
    b.close()


def test_batch_of_staggered_cases_matches_sequential_bitwise():
    """A case's population stride comes with its step arguments, so a batch
    of cases with their direction arrays 1.5 KiB apart steps exactly as the
    packed cases run one after another."""
    print("case: batch of 3 reaches, population stagger 1536 bytes vs packed sequential runs", flush=True)
    reaches = _cases()[:3]
    seq = []
    for r in reaches:
        run = R.Run3D(r)
        run.advance(60)
        seq.append(run)
    b = BatchRun(reaches, pop_stagger=1536)
    assert all(run.lbm.pop_stagger == 1536 for run in b.runs)
    b.advance(60)
    for k, (rs, rb) in enumerate(zip(seq, b.runs)):
        fl = np.tile(~reaches[k].solid.ravel(), 19)
        assert np.array_equal(rs.lbm.populations()[fl], rb.lbm.populations()[fl]), k
    for run in seq:
        run.close()
    b.close()


def test_block_width_switch_mid_run_and_autotuner(tmp_path, monkeypatch):
    """Changing the work-group width between steps leaves the result bit for
    bit unchanged, and the autotuner times every width and caches its pick."""
    print("case: channel128, width switch and autotune", flush=True)
    monkeypatch.setenv("WFS_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    from water_flow_simulation_amd import autotune
    monkeypatch.setattr(autotune, "_cache", None)
    r = channel128()
    a, b = R.Run3D(r, block_x=128), R.Run3D(r, block_x=128)
    a.advance(60)
    b.advance(20)
    b.lbm.block_x = 64
    b.advance(20)
    b.lbm.block_x = 256
    b.advance(20)
    assert b.lbm.block_x == 256
    fl = np.tile(~r.solid.ravel(), 19)
    assert np.array_equal(a.lbm.populations()[fl], b.lbm.populations()[fl])
    with pytest.raises(ValueError):
        b.lbm.block_x = 96
    a.close()
    b.close()
    t = autotune.time_widths(r, "f32", steps=10, warmup=2)
    assert sorted(t) == [64, 128, 192, 256] and all(v > 0 for v in t.values())
    bx = autotune.block_x_for(r, "f32")
    assert bx in (64, 128, 192, 256)
    assert (tmp_path / "tune.json").exists()
    run = R.Run3D(r, block_x="auto")
    assert run.lbm.block_x == bx
    run.close()
    reaches, _ = ensemble(3, seed=1, scale=0.2)
    batch = BatchRun(reaches, group="auto")
    assert batch.group in (1, 2, 3) and len(batch.tuning) >= 2
    batch.close()


def test_batched_case_refuses_direct_stepping_and_state_changes():
    """While a batch holds a solver, stepping it, loading a state or
    replacing its populations, macroscopic fields or forces directly is
    refused (the batch's next run would overwrite it, or it holds the old
    force table); once the batch is gone it works."""
    print("case: batch of 2 reaches, direct calls on a held case", flush=True)
    reaches, _ = ensemble(2, seed=3, scale=0.2)
    b = BatchRun(reaches)
    b.advance(20)
    L = b.runs[0].lbm
    st = L.state()
    rho, ux, uy, uz = L.macro()
    one = np.zeros(1, np.float32)
    for call in (lambda: L.step(1), lambda: L.load_state(st), lambda: L.set_populations(L.populations()),
                 lambda: L.set_macro(rho, ux, uy, uz), lambda: L.set_force_cells(np.array([0]), one, one, one)):
        with pytest.raises(ValueError, match="case of a batch"):
            call()
    assert L.nstep == 20
    b._destroy()
    L.step(1)
    assert L.nstep == 21
    b.close()


def _nx65_reach():
    """A reach whose outlet row opens a wavefront ((nx - 1) % 64 == 0), so the
    batch runs its outlet row as a separate pass."""
    from water_flow_simulation_amd.geometry import Section
    sec = Section(B=6.0, Hb=2.5, z=1.0, Bf=0.0, n=0.03)
    fleet = R.Fleet(R.boat_type("longtail"), rows=1, per_row=1, x_first=4.0)
    r = R.build_reach(sec, 2.0, 9.0, 2.2e-4, fleet, R.ReachConfig(length=16.25, dx=0.25))
    assert (r.nx - 1) % 64 == 0
    return r


def test_batch_budgets_match_sequential_bitwise():
    """Cases with uneven step budgets (some ending inside a 20-step chunk,
    one with none) share launches; each ends exactly like a sequential run
    of its own length, and after the budgets are lifted all continue."""
    print("case: batch of 6 reaches with budgets 60/35/100/0/47/81, then 40 more steps", flush=True)
    reaches = _cases() + [_nx65_reach()]
    budgets = [60, 35, 100, 0, 47, 81]
    seq = []
    for r, n in zip(reaches, budgets):
        run = R.Run3D(r)
        run.t_avg = 0.0
        if n:
            run.advance(n)
        seq.append(run)
    b = BatchRun(reaches)
    for run in b.runs:
        run.t_avg = 0.0
    b.set_group(len(reaches))          # pushes the schedules again
    b.set_budgets(budgets)
    b.advance(100)
    assert b.budgets_left() == [0] * len(reaches)
    b.set_budgets([None] * len(reaches))
    b.advance(40)
    for run in seq:
        run.advance(40)
    for k, (rs, rb) in enumerate(zip(seq, b.runs)):
        s_sums, s_navg, s_mac, s_pop = _state(rs)
        b_sums, b_navg, b_mac, b_pop = _state(rb)
        assert rb.lbm.nstep == rs.lbm.nstep == budgets[k] + 40, k
        assert s_navg == b_navg and rs.t == rb.t, k
        fl = ~reaches[k].solid.ravel()
        for name in s_sums:
            assert np.array_equal(s_sums[name][fl], b_sums[name][fl]), (k, name)
        for a, c in zip(s_mac, b_mac):
            assert np.array_equal(a[fl], c[fl]), k
        m19 = np.tile(fl, 19)
        assert np.array_equal(s_pop[m19], b_pop[m19]), k
    for run in seq:
        run.close()
    b.close()


@pytest.mark.parametrize("storage,use_graph", [("f32", True), ("f16dev", True), ("f32", False)])
def test_inplace_batch_matches_two_buffer_sequential_bitwise(storage, use_graph):
    """Cases streamed in place share the batched launches: with uneven
    budgets (odd ones leave the cases on different streaming phases for the
    next chunk) every case ends exactly like a sequential two-buffer run of
    its own length, with one population buffer per case."""
    print(f"case: in-place batch of 5 reaches ({storage}, graph={use_graph}), budgets 61/35/100/47/81 "
          "then 39/45/0/53/19", flush=True)
    reaches = _cases()
    rounds = [([61, 35, 100, 47, 81], 100), ([39, 45, 0, 53, 19], 60)]
    seq = []
    for r in reaches:
        run = R.Run3D(r, storage=storage)
        run.t_avg = 0.0
        seq.append(run)
    b = BatchRun(reaches, storage=storage, streaming="inplace")
    assert all(run.lbm.inplace for run in b.runs)
    for run in b.runs:
        run.t_avg = 0.0
    b.set_group(len(reaches))          # pushes the schedules again
    for budgets, m in rounds:
        for run, n in zip(seq, budgets):
            if n:
                run.advance(n)
        b.set_budgets(budgets)
        b.advance(m, use_graph=use_graph)
        assert b.budgets_left() == [0] * len(reaches)
        for k, (rs, rb) in enumerate(zip(seq, b.runs)):   # macroscopic fields are readable at any phase
            fl = ~reaches[k].solid.ravel()
            assert rb.lbm.nstep == rs.lbm.nstep, k
            for a, c in zip(rs.lbm.macro(), rb.lbm.macro()):
                assert np.array_equal(a[fl], c[fl]), k
    b.set_budgets([None] * len(reaches))
    size = 4 if storage == "f32" else 2
    for rs, rb, r in zip(seq, b.runs, reaches):   # both hold their time-mean sums now: one population buffer fewer
        assert rs.lbm.memory_mb() - rb.lbm.memory_mb() >= 19 * r.nx * r.ny * r.nz * size / 1e6
    for k, (rs, rb) in enumerate(zip(seq, b.runs)):
        s_sums, s_navg, s_mac, s_pop = _state(rs)
        b_sums, b_navg, b_mac, b_pop = _state(rb)      # an even number of steps: populations readable
        assert s_navg == b_navg and rs.t == rb.t, k
        fl = ~reaches[k].solid.ravel()
        for name in s_sums:
            assert np.array_equal(s_sums[name][fl], b_sums[name][fl]), (k, name)
        kept = _kept(reaches[k].solid.shape) & np.tile(fl, 19)   # entries leaving through an x-end: never read
        assert np.array_equal(s_pop[kept], b_pop[kept]), k
    for run in seq:
        run.close()
    b.close()


def test_batch_refuses_mixed_streaming_modes():
    import ctypes
    from water_flow_simulation_amd import _native as nat
    print("case: batch of one two-buffer and one in-place reach", flush=True)
    reaches, _ = ensemble(2, seed=3, scale=0.2)
    runs = [R.Run3D(reaches[0]), R.Run3D(reaches[1], streaming="inplace")]
    hs = (ctypes.c_void_p * 2)(*[run.lbm.handle.value for run in runs])
    h = ctypes.c_void_p()
    L = nat.lib()
    assert L.wfs_batch_create(hs, 2, ctypes.byref(h)) != 0
    assert b"streaming mode" in L.wfs_last_error()
    for run in runs:
        run.close()


def test_inplace_batch_with_a_buffer_beyond_4_gib():
    """A batch holding a case whose one population buffer exceeds 4 GiB
    (FP64, 28.4 M cells) takes the 64-bit-address kernel for all its cases;
    each case equals its own in-place solver stepped alone."""
    import ctypes
    from water_flow_simulation_amd import _native as nat
    from water_flow_simulation_amd.lbm3d import LBM3D
    nz, ny, nx = 300, 296, 320
    print(f"case: batch of a {nx}x{ny}x{nz} channel (f64, buffer > 4 GiB) and the 128x32x16 channel, in place, "
          "6 steps", flush=True)
    assert 19 * nx * ny * nz * 8 >= 2 ** 32
    big = np.zeros((nz, ny, nx), bool)
    big[0] = True
    big[:, :8, 100:120] = True
    fb = np.zeros(big.shape, np.float32)
    fb[150, 140:150, 60] = 1e-5
    uin_b = np.where(big[:, :, 0], 0.0, 0.04).astype(np.float32)
    from tests.helpers import solver_inputs
    a = solver_inputs(channel128())

    def make():
        return [LBM3D(None, big, fb, fb, fb, uin_b, 0.51, storage="f64", streaming="inplace"),
                LBM3D(None, a["solid"], a["Fx"], a["Fy"], a["Fz"], a["uin"], a["tau0"], a["cs"], a["cf_wall"],
                      storage="f64", streaming="inplace")]

    want = []
    for g in make():
        g.step(6)
        want.append(g.macro(np.float64))
        g.close()
    gs = make()
    hs = (ctypes.c_void_p * 2)(*[g.handle.value for g in gs])
    h = ctypes.c_void_p()
    L = nat.lib()
    nat.check(L.wfs_batch_create(hs, 2, ctypes.byref(h)), "wfs_batch_create")
    nat.check(L.wfs_batch_run(h, 6, 0), "wfs_batch_run")
    nat.check(L.wfs_batch_synchronize(h), "wfs_batch_synchronize")
    nat.check(L.wfs_batch_destroy(h), "wfs_batch_destroy")
    for g, w, solid in zip(gs, want, (big, a["solid"])):
        g.nstep = int(L.wfs_nstep(g.handle))
        assert g.nstep == 6
        fl = ~np.asarray(solid, bool).ravel()
        for u, v in zip(w, g.macro(np.float64)):
            assert np.array_equal(u[fl], v[fl])
        g.close()
