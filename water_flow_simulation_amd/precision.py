"""Error of each population storage format against an FP64 run of the same
case, on the quantities the boat tests report.

Every case is run twice per format through the full run driver (ramp,
20-step cadence, time averaging over the final part of the run): with the
boats and without them (the base). The readouts are the time-mean
profiles of ``profiles`` and the comparison of ``compare``:

    level profile eta_x (m), surface velocity us_x (m/s), discharge q_x
    (m^3/s), bed shear tau_bed (N/m^2), and the boat-induced water-level
    change d_eta_cm.

``study()`` returns, per format, the largest absolute difference from the
FP64 run of each readout, plus the run time. Usage:

    python -m water_flow_simulation_amd.precision [--length 60] [--dx 0.2]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import workloads
from .reach3d import Run3D, compare, profiles


def run_case(reach, storage):
    t0 = time.perf_counter()
    run = Run3D(reach, storage=storage)
    run.run_to_end(chunk=2000)
    prof = profiles(run)
    steps = run.lbm.nstep
    run.close()
    return prof, steps, time.perf_counter() - t0


def study(length=60.0, dx=0.2, x_first=20.0, storages=("f32", "f32fast", "f16dev", "bf16dev"), run_flows=1.0,
          avg_frac=0.4):
    """run_flows / avg_frac: simulated flow-through times and the averaged
    fraction (the reference defaults are 1.0 / 0.4; longer windows shrink the
    sampling noise of the LES time means that every format carries)."""
    # the time step of the reference-default run, kept for longer windows (the slosh rule
    # would otherwise raise it with the window and change the lattice speed)
    dt = workloads.trapezoid_tug(length=length, dx=dx, x_first=x_first).dt
    boat = workloads.trapezoid_tug(length=length, dx=dx, x_first=x_first, run_flows=run_flows, avg_frac=avg_frac,
                                   dt=dt)
    base = workloads.trapezoid_tug(length=length, dx=dx, x_first=x_first, with_boat=False, run_flows=run_flows,
                                   avg_frac=avg_frac, dt=dt)
    out = {"case": f"trapezoidal reach {length:g} m at dx {dx:g} m ({boat.nx}x{boat.ny}x{boat.nz} cells), "
                   f"one pusher tug at x = {x_first:g} m, {run_flows:g} flow-throughs, last {avg_frac:g} averaged",
           "formats": {}}
    ref, ref_base = {}, {}
    for st in ("f64",) + tuple(storages):
        pb, steps, secs = run_case(boat, st)
        p0, _, secs0 = run_case(base, st)
        c = compare(pb, p0, boat, boat.cfg)
        if st == "f64":
            ref, ref_base, ref_c = pb, p0, c
            out["steps"] = steps
            out["f64_seconds"] = round(secs + secs0, 2)
            out["f64_d_eta_cm"] = c["d_eta_cm"]
            out["Q"] = boat.Q
            continue
        e = {
            "eta_x_max_abs_mm": float(np.abs(pb["eta"] - ref["eta"]).max() * 1e3),
            "eta_x_base_max_abs_mm": float(np.abs(p0["eta"] - ref_base["eta"]).max() * 1e3),
            "d_eta_cm_abs_mm": float(abs(c["d_eta_cm"] - ref_c["d_eta_cm"]) * 10),
            "us_x_max_abs_m_s": float(np.abs(pb["us"] - ref["us"]).max()),
            "q_x_max_rel": float(np.abs(pb["q"] - ref["q"]).max() / boat.Q),
            "tau_bed_max_abs_Pa": float(np.abs(pb["tau_bed"] - ref["tau_bed"]).max()),
            "tau_bed_rel_l2": float(np.linalg.norm(pb["tau_bed"] - ref["tau_bed"]) /
                                    max(np.linalg.norm(ref["tau_bed"]), 1e-30)),
            "seconds": round(secs + secs0, 2),
        }
        out["formats"][st] = e
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--length", type=float, default=60.0)
    ap.add_argument("--dx", type=float, default=0.2)
    ap.add_argument("--x-first", type=float, default=20.0)
    ap.add_argument("--formats", default="f32,f32fast,f16dev,bf16dev")
    ap.add_argument("--run-flows", type=float, default=1.0)
    ap.add_argument("--avg-frac", type=float, default=0.4)
    a = ap.parse_args()
    print(json.dumps(study(a.length, a.dx, a.x_first, tuple(a.formats.split(",")), a.run_flows, a.avg_frac),
                     indent=1))


if __name__ == "__main__":
    main()
