"""Writes the golden fixtures in this folder from the reference implementation.

Run in the development container only (it reads the reference checkout; the
fixtures it writes are data: inputs and the reference's outputs):

    python tests/golden/make_golden.py /root/reference

What it runs, all on tiny inputs, and why (each is a function whose exact
output no closed form or invariant fixes on its own):

* floodsim/lbm3d.py ``make_numba_kernels`` ``step``/``bc`` (lbm3d.py:186-328):
  the per-cell update with bounce-back, lid mirror, wall function, forcing,
  Smagorinsky and the inlet/outlet rows. ``numba`` is absent here, so a stub
  module makes ``njit`` the identity and ``prange`` plain ``range``; the
  kernels then run as ordinary Python/numpy code (numpy 2 scalar promotion,
  i.e. mostly float32 arithmetic with some float64 terms).
* floodsim/reach3d.py ``build_reach`` (reach3d.py:201-297) incl. ``wall_cf``
  and ``_add_prop``: the solid mask, inlet profile (EDT), forces (disk cell
  selection), friction coefficients and time step for small cases and for the
  128x32x16 parity channel.
* ``Run3D`` + ``profiles`` + ``compare`` (reach3d.py:381-618) on a 16x18x4 toy
  reach (100 steps at lattice speed ~0.03, a 1 kW actuator disk whose hull is
  too thin to take a cell, ramp and averaging both active), to pin the 20-step
  cadence, ramp, averaging and readouts. A case whose FP32 and FP64 runs
  agree closely is chosen so that the fixture pins the algorithm, not the
  rounding of one chaotic trajectory.
"""
from __future__ import annotations

import os
import sys
import types
from dataclasses import replace

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))


def _stub_numba():
    m = types.ModuleType("numba")

    def njit(*a, **k):
        if a and callable(a[0]) and not k:
            return a[0]
        return lambda fn: fn

    m.njit = njit
    m.prange = range
    sys.modules["numba"] = m


def _step_cases(lbm3d, rng):
    """Tiny grids exercising every branch of step/bc."""
    step, bc = lbm3d.make_numba_kernels()
    W = lbm3d.W
    cases = {}
    for name, (nx, ny, nz) in (("a", (8, 6, 5)), ("b", (7, 5, 4))):
        N = nx * ny * nz
        solid = np.zeros((nz, ny, nx), bool)
        solid[0, :, :] = True               # bed
        solid[:, 0, :] = True               # one bank
        if name == "a":
            solid[3:, 2:4, 3:5] = True      # a hull block touching the lid
            solid[1, 4, 6] = True           # a lone obstacle
        else:
            solid[1:3, 2, 2:4] = True
            solid[2, 3, 0] = True           # solid at the inlet row
            solid[1, 3, nx - 2] = True      # solid next to the outlet row
        f = (W[:, None] * (1.0 + 0.08 * rng.standard_normal((19, N)))).astype(np.float32)
        f += (W[:, None] * 3 * lbm3d.CX[:, None] * 0.04).astype(np.float32)
        Fx = np.zeros(N, np.float32)
        Fy = np.zeros(N, np.float32)
        Fz = np.zeros(N, np.float32)
        pick = rng.choice(np.flatnonzero(~solid.ravel()), 6, replace=False)
        Fx[pick] = rng.uniform(-2e-3, 2e-3, 6).astype(np.float32)
        Fy[pick] = rng.uniform(-2e-3, 2e-3, 6).astype(np.float32)
        Fz[pick] = rng.uniform(-2e-3, 2e-3, 6).astype(np.float32)
        cfw = np.where(rng.random(N) < 0.7, rng.uniform(1e-3, 2e-2, N), 0.0).astype(np.float32)
        cfw[solid.ravel()] = 0
        uref = np.where(rng.random(N) < 0.5, rng.uniform(0.0, 0.08, N), 0.0).astype(np.float32)
        rho_o = (1.0 + 0.01 * rng.standard_normal(N)).astype(np.float32)
        ux_o = (0.05 + 0.02 * rng.standard_normal(N)).astype(np.float32)
        uy_o = (0.02 * rng.standard_normal(N)).astype(np.float32)
        uz_o = (0.02 * rng.standard_normal(N)).astype(np.float32)
        uin = np.where(solid[:, :, 0], 0.0, rng.uniform(0.03, 0.06, (nz, ny))).astype(np.float32).ravel()
        tau0 = np.float32(0.5006)
        cs2 = np.float32(0.17 * 0.17)
        inp = dict(f=f.ravel().copy(), solid=solid.ravel().astype(np.uint8), Fx=Fx, Fy=Fy, Fz=Fz, cfw=cfw,
                   uref=uref, rho=rho_o, ux=ux_o, uy=uy_o, uz=uz_o, uin=uin,
                   dims=np.array([nx, ny, nz]), tau0=np.array(tau0), csmag2=np.array(cs2))
        fs = f.ravel().copy()
        fd = np.zeros_like(fs)
        rho, ux, uy, uz = rho_o.copy(), ux_o.copy(), uy_o.copy(), uz_o.copy()
        outs = {}
        for k in range(3):
            step(fs, fd, inp["solid"], Fx, Fy, Fz, rho, ux, uy, uz, cfw, uref, nx, ny, nz, tau0, cs2)
            if k == 0:
                outs.update(fd_step1_prebc=fd.copy(), rho1=rho.copy(), ux1=ux.copy(), uy1=uy.copy(), uz1=uz.copy())
            bc(fd, inp["solid"], uin, nx, ny, nz, np.float32(1.0))
            if k == 0:
                outs["fd_step1"] = fd.copy()
            fs, fd = fd, fs
        outs.update(f_step3=fs.copy(), rho3=rho.copy(), ux3=ux.copy(), uy3=uy.copy(), uz3=uz.copy())
        cases[name] = (inp, outs)
    return cases


def _reach_arrays(r):
    d = dict(solid=np.packbits(r.solid.ravel()), shape=np.array(r.solid.shape), uin_phys=r.uin_phys,
             Fx=r.Fx, Fy=r.Fy, Fz=r.Fz, dt=np.array(r.dt), dx=np.array(r.dx), A=np.array(r.A), V=np.array(r.V),
             W=np.array(r.W), y=r.y, zc=r.zc)
    if r.cf is not None:
        d["cf"] = r.cf
        d["bedwall"] = np.packbits(r.bedwall.ravel())
    d["props"] = np.array([[p["x"], p["y"], p["z"], p["D"], p["Deff"], p["tilt"], p["yaw"], p["T"], p["Vjet"],
                            p["cells"]] for p in r.props]).reshape(-1, 10)
    d["hulls"] = np.array(r.hulls).reshape(-1, 6)
    return d


def main(ref_root, parts=("step", "reach", "run")):
    _stub_numba()
    sys.path.insert(0, ref_root)
    from floodsim import lbm3d, reach3d  # noqa: E402
    from floodsim.backend import Backend  # noqa: E402
    from floodsim.geometry import Section  # noqa: E402

    rng = np.random.default_rng(20261203)

    # ---- step / bc on tiny grids
    for name, (inp, outs) in (_step_cases(lbm3d, rng).items() if "step" in parts else ()):
        np.savez_compressed(os.path.join(OUT, f"step_bc_{name}.npz"), **{f"in_{k}": v for k, v in inp.items()},
                            **{f"out_{k}": v for k, v in outs.items()})
        print("step/bc case", name, "written")

    # ---- build_reach cases (reference Δt rules)
    cases = {}
    cfg = reach3d.ReachConfig(length=32.0, dx=0.25)
    sec = Section(B=7.5, Hb=4.0, z=0.0, Bf=0.0, n=0.025)
    boat = replace(reach3d.boat_type("custom"), hull_L=8.0, hull_B=2.5, draft=1.0, n_props=1, prop_D=0.30,
                   prop_depth=0.35, tilt_deg=12.0, prop_x=2.5, prop_spacing=0.0, power_kW=75.0, merit=0.6)
    cases["channel128"] = (sec, 4.0, 150.0, 2.2222e-4, reach3d.Fleet(boat, rows=1, per_row=1, x_first=8.0), cfg)
    sec2 = Section(B=6.0, Hb=1.5, z=2.0, Bf=3.0, n=0.03, nf=0.06)
    cfg2 = reach3d.ReachConfig(length=20.0, dx=0.25)
    cases["trap_jets"] = (sec2, 2.0, 9.0, 3e-4,
                          reach3d.Fleet(replace(reach3d.boat_type("jet_hp"), hull_L=2.0, hull_B=1.5, draft=0.4,
                                                spread_deg=30.0), rows=2, per_row=2, row_spacing=6.0, x_first=4.0),
                          cfg2)
    cases["trap_tug"] = (Section(B=8.0, Hb=2.5, z=2.0, Bf=0.0, n=0.03), 2.5, 20.0, 2e-4,
                         reach3d.Fleet(replace(reach3d.boat_type("tug"), hull_L=5.0, hull_B=2.0, draft=1.0),
                                       rows=1, per_row=1, x_first=5.0), reach3d.ReachConfig(length=24.0, dx=0.2))
    cases["nowall_longtail"] = (Section(B=5.0, Hb=2.0, z=1.0, Bf=0.0, n=0.025), 1.5, 6.0, 2e-4,
                                reach3d.Fleet(reach3d.boat_type("longtail"), rows=1, per_row=2, x_first=3.0),
                                reach3d.ReachConfig(length=16.0, dx=0.25, wall_model=False))
    cases["auto_dx"] = (Section(B=30.0, Hb=3.0, z=2.0, Bf=50.0, n=0.03), 3.5, 100.0, 2e-4,
                        reach3d.Fleet(reach3d.boat_type("pontoon"), rows=1, per_row=3, x_first=20.0),
                        reach3d.ReachConfig(length=60.0, max_cells=200_000))
    for name, (sec_, h, Q, S, fleet, cfg_) in (cases.items() if "reach" in parts else ()):
        r = reach3d.build_reach(sec_, h, Q, S, fleet, cfg_)
        np.savez_compressed(os.path.join(OUT, f"reach_{name}.npz"), **_reach_arrays(r),
                            notes=np.array(len(r.notes)))
        print("reach", name, r.nx, r.ny, r.nz, "dt", r.dt, "props", len(r.props))

    # ---- a toy Run3D with a boat and its no-boat base, then profiles/compare
    be = Backend()
    sec3 = Section(B=2.0, Hb=1.5, z=1.0, Bf=0.0, n=0.03)
    cfg3 = reach3d.ReachConfig(length=4.0, dx=0.25, run_flows=1.0, avg_frac=0.99)
    tb = replace(reach3d.boat_type("custom"), hull_L=1.0, hull_B=0.05, draft=0.05, n_props=1, prop_D=0.1,
                 prop_depth=0.45, tilt_deg=10.0, prop_x=0.2, power_kW=1.0)
    fl = reach3d.Fleet(tb, rows=1, per_row=1, x_first=1.0)
    out = {}
    if "run" not in parts:
        return
    for tag, fleet in (("boat", fl), ("base", replace(fl, enabled=False))):
        r = reach3d.build_reach(sec3, 1.0, 1.2, 2e-4, fleet, cfg3)
        run = reach3d.Run3D(r, be)
        for m in (40, 30, 30):             # chunks of 20 plus a short one of 10
            run.advance(m)
        prof = reach3d.profiles(run)
        out[tag] = (r, run, prof)
        np.savez_compressed(os.path.join(OUT, f"run_toy_{tag}.npz"), **_reach_arrays(r),
                            nsteps=np.array(run.lbm.nstep), navg=np.array(run.navg), t=np.array(run.t),
                            t_total=np.array(run.t_total), t_avg=np.array(run.t_avg), t_ramp=np.array(run.t_ramp),
                            tau0=np.array(run.lbm.tau0), **{f"sum_{k}": v for k, v in run.sum.items()},
                            rho=run.lbm.rho, ux=run.lbm.ux, uy=run.lbm.uy, uz=run.lbm.uz,
                            **{f"prof_{k}": np.asarray(v) for k, v in prof.items()})
        print("run", tag, "steps", run.lbm.nstep, "navg", run.navg, "t", run.t, "t_total", run.t_total)
    r, run, prof = out["boat"]
    c = reach3d.compare(prof, out["base"][2], r, cfg3)
    keys = sorted(k for k, v in c.items() if isinstance(v, (int, float, bool, np.floating)))
    np.savez_compressed(os.path.join(OUT, "compare_toy.npz"), keys=np.array(keys),
                        vals=np.array([float(c[k]) for k in keys]), label=np.array(c["label"]))
    print("compare", {k: c[k] for k in keys})


if __name__ == "__main__":
    # usage: make_golden.py [reference root] [parts: step,reach,run]
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference",
         tuple(sys.argv[2].split(",")) if len(sys.argv) > 2 else ("step", "reach", "run"))
