"""Named river-reach cases used by the parity tests and the benchmark.

channel128      128 x 32 x 16 rectangular channel, one boat, inlet at lattice
                speed 0.05 (parity case, CPU-runnable).
trapezoid_tug   1500 x 522 x 30 trapezoidal reach (40 m bed, 1:2 banks, 3 m
                deep, 150 m at dx = 0.1 m, Manning 0.03) with one pusher tug,
                12 x 4 x 1.5 m hull, two 1.2 m propellers under the stern at
                300 kW each (single-GPU throughput case, 23.5 M cells).
"""
from __future__ import annotations

from dataclasses import replace

from . import reach3d as R
from .geometry import Section
from .hydraulics1d import q_ref

E3_SLOPE = 2.2222222222222817e-4   # mean slope of the reference's default river (E3)


def channel128(dt=2.5e-3, with_boat=True):
    """32 m x 7.5 m x 4 m at dx = 0.25 m, Manning n = 0.025, V = 5 m/s
    (Q = 150 m^3/s); hull 8 x 2.5 x 1 m at x = 8..16 m with a 0.30 m, 75 kW
    actuator disk tilted 12 deg at 0.35 m depth. dt = 2.5 ms puts the inlet
    at lattice speed 0.05."""
    sec = Section(B=7.5, Hb=4.0, z=0.0, Bf=0.0, n=0.025)
    boat = replace(R.boat_type("custom"), hull_L=8.0, hull_B=2.5, draft=1.0, n_props=1, prop_D=0.30,
                   prop_depth=0.35, tilt_deg=12.0, prop_x=2.5, prop_spacing=0.0, power_kW=75.0, merit=0.6)
    fleet = R.Fleet(boat, rows=1, per_row=1, x_first=8.0, enabled=with_boat)
    cfg = R.ReachConfig(length=32.0, dx=0.25)
    return R.build_reach(sec, 4.0, 150.0, 2.2222e-4, fleet, cfg, dt=dt)


def trapezoid_tug(with_boat=True, length=150.0, dx=0.1, x_first=40.0, run_flows=1.0, avg_frac=0.4, dt=None):
    """Trapezoidal reach with one pusher tug (bow at x_first); Q from the
    Manning rule at h = 3 m on the E3 slope; reference time-step rules."""
    sec = Section(B=40.0, Hb=3.0, z=2.0, Bf=0.0, n=0.03)
    h = 3.0
    Q = float(q_ref(h, 1.0, E3_SLOPE, sec))
    tug = replace(R.boat_type("tug"), hull_L=12.0, hull_B=4.0, draft=1.5)
    fleet = R.Fleet(tug, rows=1, per_row=1, x_first=x_first, enabled=with_boat)
    cfg = R.ReachConfig(length=length, dx=dx, run_flows=run_flows, avg_frac=avg_frac)
    return R.build_reach(sec, h, Q, E3_SLOPE, fleet, cfg, dt=dt)
