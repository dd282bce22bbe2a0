"""Meandering reach on the tile-sparse layout.

A trapezoidal channel (bed width B, banks 1:bank, depth h, Manning n)
follows a sine centreline y_c(x) = A sin(2 pi x / L_w) through a bounding box
wide enough for the meander; a cell is solid when it lies below the bed or
beyond the top width, measured along the local normal to the centreline,
n = (y - y_c(x)) cos(theta), tan(theta) = dy_c/dx. Only the 64-cell x-runs of
each (y, z) row that hold water are stored on the GPU (wfs_create_tiled), so
the channel costs memory and bandwidth, not its bounding box.

Boats sit in cross-channel rows at given x, on the crests and troughs of the
centreline where it runs along x; each is a hull box with actuator disks at
the stern (the hull and disk rules of floodsim/reach3d.py:234-276, 346-374).
Forces travel as a cell list. The inlet is the reference's 1/7-power profile
of the section at x = 0, scaled to the Manning discharge; wall friction uses
the Manning coefficient on cells next to the bed and banks (contact in the
(y, z) plane of each x) and the log law next to hulls.

The full configuration (1 km, A = 60 m, L_w = 400 m, 80 m top width, 4 m deep,
dx = 0.08 m, 8 pushers in two rows 200 m apart) is 12500 x 2502 x 50 cells in
its box (1.56 x 10^9), of which 0.70 x 10^9 are stored and 0.67 x 10^9 fluid. ``meander_case(scale=...)``
shrinks lengths and widths for tests.

    python -m water_flow_simulation_amd.meander [--scale 1.0] [--steps 200] [--storage f16dev] [--learned]
        [--report-every 2000]
"""
from __future__ import annotations

import argparse
import json
import math
import time
from dataclasses import dataclass, replace

import numpy as np

from . import reach3d as R
from .geometry import Section
from .hydraulics1d import q_ref
from .workloads import E3_SLOPE


@dataclass
class MeanderConfig:
    length: float = 1000.0       # along x (m)
    amplitude: float = 60.0      # centreline amplitude (m)
    wavelength: float = 400.0    # centreline wavelength (m)
    bed: float = 64.0            # bed width (m); top width = bed + 2 bank depth
    bank: float = 2.0            # bank slope 1:bank
    depth: float = 4.0           # water depth (m)
    n: float = 0.035             # Manning n
    dx: float = 0.08             # cell size (m)
    rows: tuple = (300.0, 500.0)  # x of the boat rows (m): a trough and a crest of the centreline
    per_row: int = 4             # boats across each row
    boat: str = "tug"            # preset
    hull: tuple = (12.0, 4.0, 1.5)   # hull length, beam, draft (m)
    power_kW: float = 300.0      # per propeller
    slope: float = E3_SLOPE
    run_flows: float = 1.0
    avg_frac: float = 0.4


def meander_case(scale=1.0, **kw):
    """The config-5 reach, with lengths and widths times ``scale`` (depth and
    cell size kept)."""
    mc = MeanderConfig(**kw)
    if scale != 1.0:
        mc = replace(mc, length=mc.length * scale, amplitude=mc.amplitude * scale,
                     wavelength=mc.wavelength * scale, bed=max(mc.bed * scale, 4 * mc.dx),
                     rows=tuple(r * scale for r in mc.rows),
                     hull=(mc.hull[0] * max(scale, 0.25), mc.hull[1] * max(scale, 0.25), mc.hull[2]))
    return build_meander(mc)


def _centreline(mc, X):
    k = 2 * math.pi / mc.wavelength
    return mc.amplitude * np.sin(k * X), mc.amplitude * k * np.cos(k * X)


def build_meander(mc: MeanderConfig) -> R.Reach:
    nz = max(int(round(mc.depth / mc.dx)), 4)
    dx = mc.depth / nz
    h = mc.depth
    sec = Section(B=mc.bed, Hb=h + 0.5, z=mc.bank, Bf=0.0, n=mc.n)
    w_top = mc.bed + 2 * mc.bank * h
    half = mc.amplitude + w_top / 2
    ny = int(math.ceil(2 * half / dx)) + 2
    nx = max(int(round(mc.length / dx)), 16)
    if (nx - 1) % 64 == 0:
        nx += 1                       # the tile-sparse step fuses the outlet row into its wavefront
    y = (np.arange(ny) - ny / 2 + 0.5) * dx
    zc = (np.arange(nz) + 0.5) * dx
    Xc = (np.arange(nx) + 0.5) * dx
    yc, dyc = _centreline(mc, Xc)
    cos_t = 1.0 / np.sqrt(1.0 + dyc * dyc)

    solid = np.empty((nz, ny, nx), bool)
    bedwall = np.zeros((nz, ny, nx), bool)
    hcol = np.zeros((ny, nx), np.float32)
    step = 512
    for x0 in range(0, nx, step):
        sl = slice(x0, min(nx, x0 + step))
        nrm = (y[:, None] - yc[None, sl]) * cos_t[None, sl]          # (ny, k) distance from the centreline
        bed = R.bed_level(nrm, sec)
        s = (zc[:, None, None] < bed[None]) | (np.abs(nrm)[None] > w_top / 2)
        solid[:, :, sl] = s
        # Manning contact: a solid D3Q19 neighbour in the (y, z) plane, or the floor / y edges
        pad = np.pad(s, ((1, 1), (1, 1), (0, 0)), constant_values=False)
        pad[0] = True
        pad[:, 0] = True
        pad[:, -1] = True
        touch = np.zeros_like(s)
        for dz, dy in ((-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (-1, 1), (1, -1), (1, 1)):
            touch |= pad[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny]
        bedwall[:, :, sl] = touch & ~s
        hcol[:, sl] = np.maximum((~s).sum(0) * dx, dx)

    # inlet: the reference's 1/7-power profile of the section at x = 0, Manning discharge
    fluid2 = ~solid[:, :, 0]
    prof = R._inlet_profile(fluid2, dx, True)
    n_fluid = int(fluid2.sum())
    A = n_fluid * dx * dx
    Q = float(q_ref(h, 1.0, mc.slope, sec))
    V = Q / A
    uin = prof * (V * n_fluid / prof.sum())

    # boats: hull boxes and actuator disks, on local boxes (forces as a cell list)
    preset = replace(R.boat_type(mc.boat), hull_L=mc.hull[0], hull_B=mc.hull[1], draft=mc.hull[2],
                     power_kW=mc.power_kW)
    D = preset.outlet_D()
    T = preset.thrust()
    hulls, props, cells, fvals = [], [], [], []
    usable = w_top * 0.7
    hull_mask = np.zeros((nz, ny, nx), bool) if mc.per_row > 0 and mc.rows else None
    for xr in mc.rows:
        yr, _ = _centreline(mc, np.array([xr + preset.hull_L / 2]))
        for k in range(mc.per_row):
            yb = float(yr[0]) - usable / 2 + (k + 0.5) * usable / mc.per_row
            x0, x1 = xr, xr + preset.hull_L
            draft = min(preset.draft, 0.6 * h)
            hulls.append((x0, x1, yb - preset.hull_B / 2, yb + preset.hull_B / 2, h - draft, h))
            ix = np.flatnonzero((Xc >= x0 - 1) & (Xc <= x1 + 1))
            sub = solid[:, :, ix[0]:ix[-1] + 1]
            hsub = sub.copy()
            R._carve_hull(sub, Xc[ix[0]:ix[-1] + 1], y, zc, x0, preset.hull_L, yb, preset.hull_B, h - draft)
            hull_mask[:, :, ix[0]:ix[-1] + 1] |= sub & ~hsub
            for p in range(preset.n_props):
                yp = yb + (p - (preset.n_props - 1) / 2) * preset.prop_spacing
                zp = min(max(preset.prop_depth, draft + 0.5 * D), h - 0.5 * D)
                xp = x1 + preset.prop_x
                reach_ = max(D, 3 * dx) / 2 + 3 * dx
                jx = np.flatnonzero(np.abs(Xc - xp) <= reach_)
                jy = np.flatnonzero(np.abs(y - yp) <= reach_)
                Fx = np.zeros((nz, len(jy), len(jx)), np.float32)
                Fy, Fz = np.zeros_like(Fx), np.zeros_like(Fx)
                loc_solid = solid[:, jy[0]:jy[-1] + 1, jx[0]:jx[-1] + 1]
                pr = R._add_prop(Fx, Fy, Fz, loc_solid, Xc[jx], y[jy], zc, dx, xp, yp, h - zp, D, preset.tilt_deg, T)
                props.append(pr)
                nzmask = (Fx != 0) | (Fy != 0) | (Fz != 0)
                zz, yy, xx = np.nonzero(nzmask)
                cells.append(jx[xx] + nx * (jy[yy] + ny * zz).astype(np.int64))
                fvals.append(np.stack([Fx[nzmask], Fy[nzmask], Fz[nzmask]]))
    bedwall &= ~solid

    cfg = R.ReachConfig(length=mc.length, dx=dx, run_flows=mc.run_flows, avg_frac=mc.avg_frac)
    fmax = max((float(np.sqrt((f ** 2).sum(0)).max()) for f in fvals), default=0.0)
    u_peak = float(uin.max())
    u_jet = 0.8 * max([p["Vjet"] for p in props], default=0.0) + V
    dt = cfg.u_lb * dx / max(1.6 * u_peak, u_jet, 0.2)
    t_win = cfg.run_flows * cfg.length / max(V, 0.05) * cfg.avg_frac
    dt = float(min(dt, t_win * dx / (R.N_SLOSH * 2 * math.sqrt(3) * nx * dx)))
    if fmax > 0 and fmax * dt * dt / (R.RHO * dx) > R.F_LB_MAX:
        dt = math.sqrt(R.F_LB_MAX * R.RHO * dx / fmax)
    kf = dt * dt / (R.RHO * dx)
    idx = np.concatenate(cells) if cells else np.zeros(0, np.int64)
    fv = (np.concatenate(fvals, axis=1) * kf).astype(np.float32) if fvals else np.zeros((3, 0), np.float32)

    # wall friction: Manning next to bed and banks, log law next to hulls; none next to the end rows
    cf = np.zeros((nz, ny, nx), np.float32)
    if hull_mask is not None and hull_mask.any():
        hz = np.pad(hull_mask, 1)
        near = np.zeros_like(hull_mask)
        for dz, dy, dxx in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
            near |= hz[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dxx:1 + dxx + nx]
        cf[near & ~solid] = R.log_cf(0.5 * dx, cfg.hull_ks)
    cf_col = (R.G * mc.n * mc.n / hcol ** (1 / 3)).astype(np.float32)
    np.copyto(cf, np.broadcast_to(cf_col[None], cf.shape), where=bedwall)
    for arr in (cf, bedwall):
        arr[:, :, :2] = 0
        arr[:, :, -2:] = 0
    fleet = R.Fleet(preset, rows=len(mc.rows), per_row=mc.per_row, x_first=mc.rows[0] if mc.rows else 0.0,
                    row_spacing=(mc.rows[1] - mc.rows[0]) if len(mc.rows) > 1 else 0.0,
                    enabled=mc.per_row > 0 and bool(mc.rows))
    return R.Reach(sec=sec, cfg=cfg, fleet=fleet, h=h, Q=Q, S=mc.slope, dx=dx, dt=dt, nx=nx, ny=ny, nz=nz, y=y,
                   zc=zc, solid=solid, uin_phys=uin, Fx=None, Fy=None, Fz=None, cf=cf, bedwall=bedwall,
                   hulls=hulls, props=props, notes=[], W=w_top, A=A, V=V,
                   force_cells=(idx, fv[0], fv[1], fv[2]), layout="tiles", hull_cells=hull_mask)


def budgets(run_, reach):
    """FP64 budgets of the current state: mass (lattice units), discharge
    through the second and the third-last cross-section (m^3/s) and the
    thrust, hull and bed/bank forces (N)."""
    L = run_.lbm
    q = L.profile_x("ux")
    return {"mass_fp64": L.total("rho"),
            "discharge_in_out": [float(q[2] * reach.vel * reach.dx ** 2), float(q[-3] * reach.vel * reach.dx ** 2)],
            "forces_N": {k: [round(float(v), 1) for v in f] for k, f in run_.forces().items()}}


def run(scale=1.0, steps=200, warmup=20, storage="f16dev", learned=False, log=print, report_every=0,
        streaming="ab"):
    """Builds the case, steps it inside the run cadence and reports MLUPS over
    stored and fluid cells, device memory, and the FP64 budgets: mass,
    discharge at the inlet and outlet, thrust and the forces on the hulls and
    on bed and banks. learned: the committed learned closure instead of
    constant-Cs Smagorinsky. report_every > 0: every that many steps, check
    the state is finite and log one JSON line with the step, simulated time
    and the budgets. streaming: "ab" (two population buffers) or "inplace"
    (one)."""
    t0 = time.time()
    reach = meander_case(scale)
    t_build = time.time() - t0
    fluid = int((~reach.solid).sum())
    if log:
        log(f"meander {reach.nx}x{reach.ny}x{reach.nz} box ({reach.cells / 1e9:.3f} G cells), {fluid / 1e9:.3f} G "
            f"fluid, {len(reach.props)} disks; built in {t_build:.1f} s")
    t1 = time.time()
    closure = None
    if learned:
        from .closure import load_params
        closure = load_params()[0]
    run_ = R.Run3D(reach, storage=storage, closure=closure, streaming=streaming, closure_field=False)
    t_up = time.time() - t1
    L = run_.lbm
    stored = L.stored_cells
    run_.advance(warmup)
    L.timing(True)
    t2 = time.time()
    if report_every > 0:
        done = 0
        while done < steps:
            n = min(report_every, steps - done)
            run_.advance(n)
            done += n
            L.check_finite()
            if log:
                log(json.dumps({"step": warmup + done, "t_s": round(run_.t, 3), "wall_s": round(time.time() - t2, 1),
                                **budgets(run_, reach)}))
    else:
        run_.advance(steps)
    wall = time.time() - t2
    ms, n = L.timing_read()
    L.timing(False)
    L.check_finite()
    res = {"box": [reach.nx, reach.ny, reach.nz], "box_cells": reach.cells, "stored_cells": stored,
           "fluid_cells": fluid, "storage": storage, "steps": steps, "build_s": round(t_build, 1),
           "upload_s": round(t_up, 1), "wall_s": round(wall, 3), "kernel_ms": round(ms / max(n, 1), 4),
           "mlups_fluid": round(steps * fluid / wall / 1e6, 1), "mlups_stored": round(steps * stored / wall / 1e6, 1),
           "device_gb": round(L.memory_mb() / 1e3, 2), **budgets(run_, reach), "t_s": round(run_.t, 3),
           "dt_s": reach.dt, "Q": reach.Q, "closure": "learned" if learned else "smagorinsky", "streaming": streaming,
           "thrust_design_N": float(sum(p["T"] for p in reach.props))}
    run_.close()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--storage", default="f16dev")
    ap.add_argument("--learned", action="store_true", help="the learned sub-grid closure")
    ap.add_argument("--report-every", type=int, default=0, help="steps between budget reports (0: none)")
    ap.add_argument("--streaming", default="ab", choices=("ab", "inplace"),
                    help="two population buffers or one streamed in place")
    a = ap.parse_args()
    print(json.dumps(run(a.scale, a.steps, storage=a.storage, learned=a.learned,
                         log=lambda m: print(m, flush=True), report_every=a.report_every,
                         streaming=a.streaming)), flush=True)


if __name__ == "__main__":
    main()
