"""Host case builder, thrust formulas and discharge against the reference.

* build_reach outputs vs the reference's own build_reach on five cases
  (tests/golden/reach_*.npz): the parity channel, a trapezoid with floodplain
  and spread water jets, a tug with under-hull propellers, a run without the
  wall model, and the automatic cell size. Exact equality: these are float64
  selections and float32 roundings that must not move by one cell.
* BoatType thrust against the thrust columns the reference's CSV results
  hold (results/wallmodel/test*/boat_results.csv).
* q_ref against the discharge stored with a reference case.
* Closed forms: uniform inlet of the rectangular channel, brute-force
  Euclidean distance and D3Q19 contact on tiny masks, disk force sum.
"""
from dataclasses import replace

import math

import numpy as np
import pytest
from scipy.ndimage import distance_transform_edt

from tests.conftest import golden, unpack_bool
from tests.helpers import channel128
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.geometry import Section
from water_flow_simulation_amd.hydraulics1d import q_ref

CASES = {
    "channel128": lambda: (Section(B=7.5, Hb=4.0, z=0.0, Bf=0.0, n=0.025), 4.0, 150.0, 2.2222e-4,
                           R.Fleet(replace(R.boat_type("custom"), hull_L=8.0, hull_B=2.5, draft=1.0, n_props=1,
                                           prop_D=0.30, prop_depth=0.35, tilt_deg=12.0, prop_x=2.5,
                                           prop_spacing=0.0, power_kW=75.0, merit=0.6),
                                   rows=1, per_row=1, x_first=8.0),
                           R.ReachConfig(length=32.0, dx=0.25)),
    "trap_jets": lambda: (Section(B=6.0, Hb=1.5, z=2.0, Bf=3.0, n=0.03, nf=0.06), 2.0, 9.0, 3e-4,
                          R.Fleet(replace(R.boat_type("jet_hp"), hull_L=2.0, hull_B=1.5, draft=0.4,
                                          spread_deg=30.0), rows=2, per_row=2, row_spacing=6.0, x_first=4.0),
                          R.ReachConfig(length=20.0, dx=0.25)),
    "trap_tug": lambda: (Section(B=8.0, Hb=2.5, z=2.0, Bf=0.0, n=0.03), 2.5, 20.0, 2e-4,
                         R.Fleet(replace(R.boat_type("tug"), hull_L=5.0, hull_B=2.0, draft=1.0),
                                 rows=1, per_row=1, x_first=5.0), R.ReachConfig(length=24.0, dx=0.2)),
    "nowall_longtail": lambda: (Section(B=5.0, Hb=2.0, z=1.0, Bf=0.0, n=0.025), 1.5, 6.0, 2e-4,
                                R.Fleet(R.boat_type("longtail"), rows=1, per_row=2, x_first=3.0),
                                R.ReachConfig(length=16.0, dx=0.25, wall_model=False)),
    "auto_dx": lambda: (Section(B=30.0, Hb=3.0, z=2.0, Bf=50.0, n=0.03), 3.5, 100.0, 2e-4,
                        R.Fleet(R.boat_type("pontoon"), rows=1, per_row=3, x_first=20.0),
                        R.ReachConfig(length=60.0, max_cells=200_000)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_build_reach_matches_reference(name):
    g = golden(f"reach_{name}.npz")
    r = R.build_reach(*CASES[name]())
    shape = tuple(int(s) for s in g["shape"])
    assert r.solid.shape == shape
    assert np.array_equal(r.solid, unpack_bool(g["solid"], shape))
    assert r.dt == float(g["dt"]) and r.dx == float(g["dx"])
    assert r.A == float(g["A"]) and r.V == float(g["V"]) and r.W == float(g["W"])
    for k in ("uin_phys", "Fx", "Fy", "Fz", "y", "zc"):
        assert np.array_equal(getattr(r, k), g[k]), k
    if "cf" in g.files:
        assert np.array_equal(r.cf, g["cf"])
        assert np.array_equal(r.bedwall, unpack_bool(g["bedwall"], shape))
    else:
        assert r.cf is None and r.bedwall is None
    props = np.array([[p["x"], p["y"], p["z"], p["D"], p["Deff"], p["tilt"], p["yaw"], p["T"], p["Vjet"],
                       p["cells"]] for p in r.props]).reshape(-1, 10)
    assert np.array_equal(props, g["props"])
    assert np.array_equal(np.array(r.hulls).reshape(-1, 6), g["hulls"])


# thrust per unit from results/wallmodel/test*/boat_results.csv (T_kN / units)
@pytest.mark.parametrize("key,over,units,t_kn", [
    ("longtail", {}, 3, 19.772),
    ("outboard", {}, 3, 28.822),
    ("tug", {}, 6, 251.09),
    ("pontoon", {}, 3, 32.766),
])
def test_thrust_matches_reference_csv(key, over, units, t_kn):
    b = replace(R.boat_type(key), **over)
    assert b.thrust() * units / 1000 == pytest.approx(t_kn, rel=5e-5)


def test_jet_thrust_closed_form():
    b = R.boat_type("jet_hp")
    v = math.sqrt(2 * 100e5 / 1000)
    assert b.jet_speed() == pytest.approx(v)
    assert b.thrust() == pytest.approx(2 * 0.7 * 25e3 / v)
    assert b.outlet_D() == pytest.approx(math.sqrt(4 * b.thrust() / (math.pi * 1000 * v * v)))


def test_q_ref_matches_reference_case():
    # results/web3d/cache/base_e3_B24_Hb4.5_z2_Bf0_n0.03_S0.0184_x100_h4_*.json: straight canal, h = 4 m
    sec = Section(B=24, Hb=4.5, z=2, Bf=0, n=0.03)
    assert q_ref(4.0, 1.0, 1.84e-5, sec) == pytest.approx(38.539620226879194, rel=1e-12)


def test_grid_and_area_match_reference_case():
    # results/web3d/cache/base_e3_B50_Hb4_z2_Bf300_n0.03_h3_*.json: Q, V -> A = 168 m^2, 600 x 250 x 12
    sec = Section(B=50, Hb=4, z=2, Bf=300, n=0.03, nf=0.06)
    Q = 128.89374645934603
    r = R.build_reach(sec, 3.0, Q, 0.00022222222222222817, R.Fleet(R.boat_type("tug"), enabled=False),
                      R.ReachConfig(length=150.0, max_cells=2_000_000))
    assert (r.nx, r.ny, r.nz) == (600, 250, 12)
    assert r.V == pytest.approx(0.7672246813056312, rel=1e-12)
    assert r.y[0] - 0.5 * r.dx == -31.25   # the cached case stores the domain edge y0


def test_rectangular_channel_inlet_is_uniform():
    r = channel128()
    fl = ~r.solid[:, :, 0]
    assert np.allclose(r.uin_phys[fl], 5.0, rtol=1e-14)
    assert r.V == 5.0 and r.A == 30.0
    assert r.vel * 0.05 == pytest.approx(5.0)


def test_disk_cells_and_force_sum():
    r = channel128()
    (p,) = r.props
    assert p["cells"] == 12
    k = r.dt ** 2 / (R.RHO * r.dx)
    tilt = math.radians(12.0)
    fx = r.Fx.astype(np.float64).sum() / k * r.dx ** 3
    fz = r.Fz.astype(np.float64).sum() / k * r.dx ** 3
    assert fx == pytest.approx(p["T"] * math.cos(tilt), rel=1e-6)
    assert fz == pytest.approx(-p["T"] * math.sin(tilt), rel=1e-6)
    assert float(np.abs(r.Fx).max()) <= R.F_LB_MAX


def test_edt_brute_force():
    rng = np.random.default_rng(1)
    m = rng.random((6, 7)) < 0.75
    d = distance_transform_edt(m)
    zeros = np.argwhere(~m)
    for (i, j), v in np.ndenumerate(d):
        if m[i, j]:
            assert v == pytest.approx(np.sqrt(((zeros - [i, j]) ** 2).sum(1)).min())


def test_wall_contact_brute_force():
    rng = np.random.default_rng(2)
    nz, ny, nx = 4, 5, 6
    solid = rng.random((nz, ny, nx)) < 0.2
    got = R._touching(solid, False, solid)
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)
            if (a, b, c) != (0, 0, 0) and abs(a) + abs(b) + abs(c) < 3]
    want = np.zeros_like(solid)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                if solid[z, y, x]:
                    continue
                for a, b, c in offs:
                    zz, yy, xx = z + a, y + b, min(max(x + c, 0), nx - 1)
                    if 0 <= zz < nz and 0 <= yy < ny and solid[zz, yy, xx]:
                        want[z, y, x] = True
    assert np.array_equal(got, want)


def test_run3d_time_scalars_match_reference():
    g = golden("run_toy_boat.npz")
    from water_flow_simulation_amd.reach3d import ReachConfig
    sec = Section(B=2.0, Hb=1.5, z=1.0, Bf=0.0, n=0.03)
    cfg = ReachConfig(length=4.0, dx=0.25, run_flows=1.0, avg_frac=0.99)
    tb = replace(R.boat_type("custom"), hull_L=1.0, hull_B=0.05, draft=0.05, n_props=1, prop_D=0.1,
                 prop_depth=0.45, tilt_deg=10.0, prop_x=0.2, power_kW=1.0)
    r = R.build_reach(sec, 1.0, 1.2, 2e-4, R.Fleet(tb, rows=1, per_row=1, x_first=1.0), cfg)
    assert r.dt == float(g["dt"])
    t_total = cfg.run_flows * cfg.length / max(r.V, 0.05)
    assert t_total == float(g["t_total"])
    assert t_total * (1 - cfg.avg_frac) == float(g["t_avg"])
    assert min(0.15 * t_total, 20.0) == float(g["t_ramp"])
    assert 0.5 + 3 * max(1e-6 * r.dt / r.dx ** 2, 2e-4) == float(g["tau0"])
