"""River-reach cases for the boat tests: geometry, boats and actuator disks,
units and time step, the run driver and the readouts.

Public names follow the reference module floodsim/reach3d.py so that callers
switch by changing the import: ``BoatType``, ``BOAT_TYPES``, ``boat_type``,
``Fleet``, ``ReachConfig``, ``Reach``, ``bed_level``, ``auto_dx``,
``build_reach``, ``log_cf``, ``manning_ks``, ``wall_cf``, ``Run3D``,
``profiles``, ``compare``.

The case builder is host code (float64 numpy/scipy). Cell selections
(hull voxels, disk cells, grid sizes) depend on float64 comparisons, so every
formula keeps the reference's float64 operations; they are written out again
here, not imported.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass, field, replace

import numpy as np
from scipy.ndimage import binary_dilation, distance_transform_edt

from .geometry import Section
from .lbm3d import LBM3D

G = 9.81
RHO = 1000.0
N_SLOSH = 10          # pressure-wave round trips required inside the averaging window
F_LB_MAX = 1.5e-3     # largest lattice body force per cell accepted before dt is reduced
GPU_DUTY = float(os.environ.get("FLOODSIM_GPU_DUTY", "1"))


# ================================================================ boats
@dataclass
class BoatType:
    """A boat (or pump pontoon) type (floodsim/reach3d.py:24-72).

    Lengths in m, power per propeller or nozzle in kW. ``prop_x`` is the
    propeller position behind the stern (negative = under the hull);
    ``jet_bar`` > 0 makes it a water jet with nozzle speed sqrt(2 dp / rho);
    ``spread_deg`` splits a jet into three yaw directions -a, 0, +a.
    """
    key: str
    name: str
    hull_L: float
    hull_B: float
    draft: float
    n_props: int
    prop_D: float
    prop_depth: float
    tilt_deg: float
    prop_x: float
    prop_spacing: float
    power_kW: float
    merit: float = 0.6
    thrust_N: float = 0.0
    note: str = ""
    jet_bar: float = 0.0
    spread_deg: float = 0.0

    @property
    def is_jet(self) -> bool:
        return self.jet_bar > 0

    def thrust(self) -> float:
        """Thrust per propeller or nozzle (N): actuator disk T = (2 rho A (eta P)^2)^(1/3);
        jet T = 2 eta P / v."""
        if self.thrust_N > 0:
            return self.thrust_N
        shaft_w = self.merit * self.power_kW * 1000
        if self.is_jet:
            return 2 * self.merit * self.power_kW * 1000 / self.jet_speed()
        disk = math.pi * self.prop_D ** 2 / 4
        return (2 * RHO * disk * shaft_w ** 2) ** (1 / 3)

    def jet_speed(self) -> float:
        """Jet speed (m/s): nozzle sqrt(2 dp / rho); propeller slipstream 2 sqrt(T / (2 rho A))."""
        if self.is_jet:
            return math.sqrt(2 * self.jet_bar * 1e5 / RHO)
        disk = max(math.pi * self.prop_D ** 2 / 4, 1e-6)
        return 2 * math.sqrt(self.thrust() / (2 * RHO * disk))

    def outlet_D(self) -> float:
        """Propeller diameter, or the nozzle diameter giving the thrust at the jet speed (T = rho A v^2)."""
        if not self.is_jet:
            return self.prop_D
        v = self.jet_speed()
        return math.sqrt(4 * self.thrust() / (math.pi * RHO * v * v))


# The seven boat presets plus "custom" (floodsim/reach3d.py:75-93): the same
# table of dimensions, powers and nozzle pressures, with English names.
BOAT_TYPES = [
    BoatType("longtail", "long-tail boat", 9.0, 1.6, 0.35, 1, 0.30, 0.35, 12, 2.5, 0, 75,
             note="car engine ~100 hp, long shaft, small fast propeller, shallow and tilted down"),
    BoatType("outboard", "outboard boat", 7.0, 2.4, 0.5, 1, 0.36, 0.6, 0, 0.3, 0, 110,
             note="~150 hp outboard, horizontal propeller"),
    BoatType("tug", "tug / pusher boat", 16.0, 5.0, 1.8, 2, 1.2, 1.6, 0, -1.0, 2.4, 300,
             note="two large propellers under the stern, ~400 hp each"),
    BoatType("pontoon", "pusher on a pontoon", 6.0, 4.0, 0.5, 1, 1.6, 1.3, 0, 0.5, 0, 30,
             note="large slow propeller: low power, high thrust per kW"),
    BoatType("jet_hp", "high-pressure water jets (100 bar)", 6.0, 4.0, 0.5, 4, 0.0, 1.2, 0, 0.3, 0.8, 25,
             merit=0.7, jet_bar=100, note="4 nozzles, ~141 m/s jets, low thrust per kW (T = 2 eta P / v)"),
    BoatType("jet_mp", "medium-pressure water jets (10 bar)", 6.0, 4.0, 0.5, 4, 0.0, 1.2, 0, 0.3, 0.8, 25,
             merit=0.7, jet_bar=10, note="4 centrifugal-pump nozzles, ~45 m/s jets"),
    BoatType("curtain", "low-speed water curtain", 6.0, 6.0, 0.5, 4, 0.0, 1.3, 0, 0.3, 1.4, 25,
             merit=0.7, jet_bar=0.045, note="4 axial-flow pipes ~1.3 m across, ~3 m/s: highest thrust per kW"),
    BoatType("custom", "custom", 10.0, 3.0, 0.8, 1, 0.8, 1.0, 0, 0.5, 0, 100),
]


def boat_type(key: str) -> BoatType:
    """A fresh copy of the preset with this key."""
    for b in BOAT_TYPES:
        if b.key == key:
            return replace(b)
    raise KeyError(key)


@dataclass
class Fleet:
    """Boats moored in ``rows`` rows along the river, ``per_row`` across it."""
    boat: BoatType
    rows: int = 1
    per_row: int = 3
    row_spacing: float = 30.0
    x_first: float = 40.0      # bow of the first row, from the inlet (m)
    enabled: bool = True

    @property
    def n_boats(self):
        return self.rows * self.per_row if self.enabled else 0

    def total_thrust(self):
        return self.n_boats * self.boat.n_props * self.boat.thrust()

    def total_power(self):
        return self.n_boats * self.boat.n_props * self.boat.power_kW

    def label(self):
        if not self.enabled or self.n_boats == 0:
            return "no boats"
        return f"{self.boat.name} {self.rows}x{self.per_row}"


@dataclass
class ReachConfig:
    """Run settings with the reference defaults (floodsim/reach3d.py:125-138)."""
    length: float = 150.0      # reach length (m)
    dx: float = 0.0            # cell size (m); 0 = from max_cells
    max_cells: int = 2_000_000
    fp_max: float = 40.0       # widest floodplain strip kept in the domain (m per side)
    u_lb: float = 0.14         # lattice speed of the fastest flow (Mach control)
    cs: float = 0.17           # Smagorinsky constant
    run_flows: float = 1.0     # simulated time = run_flows * length / mean speed
    avg_frac: float = 0.4      # final fraction of the run that is time-averaged
    spacing_km: float = 1.0    # spacing of boat groups along the river (for the discharge estimate)
    tau_crit: float = 2.0      # critical bed shear stress (N/m^2)
    wall_model: bool = True    # Manning bed/banks and log-law hulls; False = plain no-slip walls
    hull_ks: float = 0.002     # hull roughness (m)


# ================================================================ geometry
def bed_level(yy, sec: Section):
    """Bed elevation across the section: flat bed |y| <= B/2, banks rising at
    slope 1:z (vertical steps when z = 0) up to Hb, floodplain at Hb."""
    ay = np.abs(yy)
    if sec.z > 0:
        bank = np.minimum((ay - sec.B / 2) / sec.z, sec.Hb)
    else:
        bank = np.where(ay > sec.B / 2, sec.Hb, 0.0)
    inner = ay <= sec.B / 2
    in_channel = ay <= sec.Tb / 2
    return np.where(inner, 0.0, np.where(in_channel, bank, sec.Hb))


@dataclass
class Reach:
    sec: Section
    cfg: ReachConfig
    fleet: Fleet
    h: float
    Q: float
    S: float
    dx: float
    dt: float
    nx: int
    ny: int
    nz: int
    y: np.ndarray
    zc: np.ndarray
    solid: np.ndarray
    uin_phys: np.ndarray
    Fx: np.ndarray
    Fy: np.ndarray
    Fz: np.ndarray
    cf: np.ndarray = None
    bedwall: np.ndarray = None
    hulls: list = field(default_factory=list)   # (x0, x1, y0, y1, z0, z1)
    props: list = field(default_factory=list)   # dict(x, y, z, D, Deff, tilt, yaw, T, Vjet, cells)
    notes: list = field(default_factory=list)
    W: float = 0.0
    A: float = 0.0
    V: float = 0.0
    force_cells: tuple = None   # (dense cell index, Fx, Fy, Fz) in lattice units instead of Fx/Fy/Fz arrays
    hull_cells: np.ndarray = None   # hull mask when the bed/bank section varies along x
    layout: str = "dense"       # "tiles": only the 64-cell x-runs that hold fluid are stored

    @property
    def vel(self):
        """Lattice-to-physical velocity factor dx/dt (m/s)."""
        return self.dx / self.dt

    @property
    def x(self):
        return (np.arange(self.nx) + 0.5) * self.dx

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    def eta_from_rho(self, rho):
        """Rigid-lid pressure to water level: eta = (rho - 1)/3 * (dx/dt)^2 / g (m), in float64."""
        return (np.asarray(rho, np.float64) - 1.0) / 3.0 * self.vel ** 2 / G


def auto_dx(L, Wd, h, cfg: ReachConfig):
    """Cell size from the cell budget, kept within [h/40, max(h/8, budget)]."""
    budget = (L * Wd * h / cfg.max_cells) ** (1 / 3)
    return float(min(max(budget, h / 40), max(h / 8, budget)))


def _grid(sec, h, cfg):
    fp = min(sec.Bf, cfg.fp_max) if h > sec.Hb else 0.0
    w_top = (sec.B + 2 * sec.z * min(h, sec.Hb)) + 2 * fp
    dx = cfg.dx if cfg.dx > 0 else auto_dx(cfg.length, w_top, h, cfg)
    nz = max(int(round(h / dx)), 4)
    dx = h / nz                       # the lid falls exactly on a cell face
    ny = int(math.ceil(w_top / dx)) + 2
    nx = max(int(round(cfg.length / dx)), 16)
    return fp, w_top, dx, nx, ny, nz


def _inlet_profile(fluid2, dx, wall_model):
    """1/7-power profile of the distance to bed and banks (the water surface
    mirrored, not a wall); with the wall model each column is rescaled to a
    mean proportional to its depth^(2/3) (Manning)."""
    nz = fluid2.shape[0]
    mirrored = np.concatenate([fluid2, fluid2[::-1]], axis=0)
    dist = distance_transform_edt(mirrored)[:nz] * dx
    prof = np.where(fluid2, np.maximum(dist, 0.5 * dx) ** (1 / 7), 0.0)
    if wall_model:
        ncol = fluid2.sum(0)
        cmean = np.where(ncol > 0, prof.sum(0) / np.maximum(ncol, 1), 1.0)
        prof = prof / cmean[None, :] * (np.maximum(ncol, 1) * dx) ** (2 / 3)
        prof = np.where(fluid2, prof, 0.0)
    return prof


def _carve_hull(solid, Xc, y, zc, x0, hull_L, yb, hull_B, z_keel):
    """Hull voxels: a box below the lid whose bow narrows linearly from 35 %
    of the beam to the full beam over the first quarter of its length."""
    along = (Xc >= x0) & (Xc <= x0 + hull_L)
    taper = np.clip((Xc[along] - x0) / (0.25 * hull_L), 0.35, 1.0)
    below_lid = zc >= z_keel
    for jx, t in zip(np.flatnonzero(along), taper):
        across = np.abs(y - yb) <= hull_B / 2 * t
        solid[np.ix_(below_lid, across, [jx])] = True


def build_reach(sec: Section, h: float, Q: float, S: float, fleet: Fleet, cfg: ReachConfig, *,
                dt: float | None = None) -> Reach:
    """Build one case (floodsim/reach3d.py:201-297): grid, solid mask, inlet
    profile, hulls and actuator disks, time step, lattice forces and wall
    friction. ``dt`` (s), if given, replaces the automatic time-step rules
    (used to set the lattice inlet speed directly)."""
    notes = []
    fp, w_top, dx, nx, ny, nz = _grid(sec, h, cfg)
    y = (np.arange(ny) - ny / 2 + 0.5) * dx
    zc = (np.arange(nz) + 0.5) * dx
    zb = bed_level(y, sec)
    sol2 = (zc[:, None] < zb[None, :]) | (np.abs(y)[None, :] > w_top / 2)
    if fp > 0 and sec.Bf > fp:
        notes.append(f"floodplain truncated to {fp:g} m per side (the domain edge is a wall)")
    solid = np.broadcast_to(sol2[:, :, None], (nz, ny, nx)).copy()

    fluid2 = ~sol2
    prof = _inlet_profile(fluid2, dx, cfg.wall_model)
    n_fluid = fluid2.sum()
    A = n_fluid * dx * dx
    V = Q / A
    uin = prof * (V * n_fluid / prof.sum())

    Fx = np.zeros(solid.shape, np.float32)
    Fy = np.zeros(solid.shape, np.float32)
    Fz = np.zeros(solid.shape, np.float32)
    hulls, props = [], []
    b = fleet.boat
    Xc = (np.arange(nx) + 0.5) * dx
    if fleet.enabled and fleet.n_boats > 0:
        usable = (sec.B + 2 * sec.z * min(h, sec.Hb)) * 0.9
        if b.draft >= h * 0.9:
            notes.append(f"draft {b.draft:g} m is close to the depth {h:g} m; reduced to {0.6 * h:.2f} m")
        draft = min(b.draft, 0.6 * h)
        D = b.outlet_D()
        T = b.thrust()
        yaws = (-b.spread_deg, 0.0, b.spread_deg) if b.spread_deg > 0 else (0.0,)
        for row in range(fleet.rows):
            xb = fleet.x_first + row * fleet.row_spacing
            for k in range(fleet.per_row):
                yb = -usable / 2 + (k + 0.5) * usable / fleet.per_row
                x0, x1 = xb, xb + b.hull_L
                hulls.append((x0, x1, yb - b.hull_B / 2, yb + b.hull_B / 2, h - draft, h))
                _carve_hull(solid, Xc, y, zc, x0, b.hull_L, yb, b.hull_B, h - draft)
                for p in range(b.n_props):
                    yp = yb + (p - (b.n_props - 1) / 2) * b.prop_spacing
                    depth = b.prop_depth
                    if depth > h - 0.6 * D:
                        depth = max(h - 0.6 * D, 0.5 * D)
                        notes.append(f"propeller/nozzle deeper than the water; raised to {depth:.2f} m")
                    if b.prop_x < 0 and depth < draft + 0.5 * D:
                        depth = min(draft + 0.5 * D, h - 0.5 * D)
                    xp = x1 + b.prop_x
                    for yaw in yaws:
                        pr = _add_prop(Fx, Fy, Fz, solid, Xc, y, zc, dx, xp, yp, h - depth, D, b.tilt_deg,
                                       T / len(yaws), yaw)
                        if b.is_jet:   # nozzle jets move at the pump's speed (limits dt for low Mach)
                            pr["Vjet"] = min(b.jet_speed(), pr["Vjet"] * 4)
                        props.append(pr)
        notes = list(dict.fromkeys(notes))

    if dt is None:
        dt = _time_step(uin, V, props, Fx, Fy, Fz, dx, nx, cfg)
    else:
        dt = float(dt)
        notes.append(f"time step set by the caller: dt = {dt:g} s")
    k = dt * dt / (RHO * dx)
    cf, bedwall = wall_cf(solid, sol2, dx, sec.n, cfg) if cfg.wall_model else (None, None)
    return Reach(sec=sec, cfg=cfg, fleet=fleet, h=h, Q=Q, S=S, dx=dx, dt=dt, nx=nx, ny=ny, nz=nz, y=y, zc=zc,
                 solid=solid, uin_phys=uin, Fx=(Fx * k).astype(np.float32), Fy=(Fy * k).astype(np.float32),
                 Fz=(Fz * k).astype(np.float32), cf=cf, bedwall=bedwall, hulls=hulls, props=props, notes=notes,
                 W=w_top, A=A, V=V)


def _time_step(uin, V, props, Fx, Fy, Fz, dx, nx, cfg):
    """Time step rules (floodsim/reach3d.py:278-292): lattice speed u_lb for
    the fastest flow (inlet peak x1.6, or 80 % of the fastest jet on top of
    the mean flow); at least N_SLOSH pressure-wave round trips inside the
    averaging window; lattice force per cell at most F_LB_MAX."""
    u_peak = float(uin.max())
    u_jet = 0.8 * max([p["Vjet"] for p in props], default=0.0) + V
    u_max = max(1.6 * u_peak, u_jet, 0.2)
    dt = cfg.u_lb * dx / u_max
    t_win = cfg.run_flows * cfg.length / max(V, 0.05) * cfg.avg_frac
    dt = float(min(dt, t_win * dx / (N_SLOSH * 2 * math.sqrt(3) * nx * dx)))
    fmax = float(np.sqrt(Fx ** 2 + Fy ** 2 + Fz ** 2).max()) if props else 0.0
    if fmax > 0 and fmax * dt * dt / (RHO * dx) > F_LB_MAX:
        dt = math.sqrt(F_LB_MAX * RHO * dx / fmax)
    return dt


# D3Q19 neighbourhood: the 18 face and edge neighbours, no corners.
_D3Q19 = np.ones((3, 3, 3), bool)
_D3Q19[::2, ::2, ::2] = False


def log_cf(z, ks):
    """Rough-wall friction coefficient tau = rho cf u^2 at distance z (log law, z0 = ks/30)."""
    return (0.41 / math.log(max(30.0 * z / max(ks, 1e-6), 2.0))) ** 2


def manning_ks(n):
    """Equivalent sand roughness (m) of a Manning n (Strickler)."""
    return (21.1 * n) ** 6


def _touching(mask, floor, solid):
    """Fluid cells with a D3Q19 neighbour in ``mask``. Outside the domain the
    bed side (z = -1) and the two y edges count as ``floor``; above the lid
    counts as open; x is extended by edge copies."""
    pad = np.pad(mask, ((1, 1), (1, 1), (0, 0)), constant_values=False)
    pad[0] = floor
    pad[:, 0] = floor
    pad[:, -1] = floor
    pad = np.pad(pad, ((0, 0), (0, 0), (1, 1)), mode="edge")
    return binary_dilation(pad, _D3Q19)[1:-1, 1:-1, 1:-1] & ~solid


def wall_cf(solid, bed2, dx, n, cfg: ReachConfig):
    """Per-cell wall friction (floodsim/reach3d.py:315-343). Bed/bank cells:
    Manning cf = g n^2 / h_col^(1/3) used with the depth-averaged speed; hull
    cells: log-law cf from hull_ks used with the local tangential speed; bed
    wins over hull; the two columns at each end carry none (plain
    bounce-back next to the inlet/outlet rows). Returns (cf, bedwall)."""
    bed = np.broadcast_to(bed2[:, :, None], solid.shape)
    hull = solid & ~bed
    cf = np.zeros(solid.shape, np.float32)
    cf[_touching(hull, False, solid)] = log_cf(0.5 * dx, cfg.hull_ks)
    bedwall = _touching(bed, True, solid)
    hcol = np.maximum((~bed2).sum(0) * dx, dx)
    cf_col = (G * n * n / hcol ** (1 / 3)).astype(np.float32)
    cf = np.where(bedwall, cf_col[None, :, None], cf)
    for arr in (cf, bedwall):
        arr[:, :, :2] = 0
        arr[:, :, -2:] = 0
    return cf.astype(np.float32), bedwall


def _add_prop(Fx, Fy, Fz, solid, Xc, y, zc, dx, xp, yp, zp, D, tilt, T, yaw=0.0):
    """Spread a thrust T (N) over a disk of cells (floodsim/reach3d.py:346-374):
    diameter Deff = max(D, 3 dx), thickness 1.5 dx along the axis (tilted down
    by ``tilt``, turned by ``yaw``), fluid cells only, uniform force density
    T / (n dx^3) along the axis. Returns the disk record."""
    Deff = max(D, 3 * dx)
    t, w = math.radians(tilt), math.radians(yaw)
    ax, ay, az = math.cos(t) * math.cos(w), math.cos(t) * math.sin(w), -math.sin(t)
    reach = Deff / 2 + 2 * dx
    ix = np.flatnonzero(np.abs(Xc - xp) <= reach)
    iy = np.flatnonzero(np.abs(y - yp) <= reach)
    iz = np.flatnonzero(np.abs(zc - zp) <= reach)
    X, Yy, Z = np.meshgrid(Xc[ix] - xp, y[iy] - yp, zc[iz] - zp, indexing="ij")
    axial = X * ax + Yy * ay + Z * az
    rx, ry, rz = X - axial * ax, Yy - axial * ay, Z - axial * az
    rad = np.sqrt(rx * rx + ry * ry + rz * rz)
    disk = (np.abs(axial) <= dx * 0.75) & (rad <= Deff / 2)
    disk &= ~solid[np.ix_(iz, iy, ix)].transpose(2, 1, 0)
    ncell = int(disk.sum())
    if ncell == 0:
        return dict(x=xp, y=yp, z=zp, D=D, Deff=Deff, tilt=tilt, yaw=yaw, T=0.0, Vjet=0.0, cells=0)
    density = T / (ncell * dx ** 3)   # N/m^3
    sub = np.zeros(disk.shape, np.float32)
    sub[disk] = density
    sub = sub.transpose(2, 1, 0)       # (z, y, x)
    box = np.ix_(iz, iy, ix)
    Fx[box] += sub * ax
    Fy[box] += sub * ay
    Fz[box] += sub * az
    area = math.pi * Deff ** 2 / 4
    return dict(x=xp, y=yp, z=zp, D=D, Deff=Deff, tilt=tilt, yaw=yaw, T=T, cells=ncell,
                Vjet=2 * math.sqrt(T / (2 * RHO * area)))


# ================================================================ run driver
class Run3D:
    """Runs one case (floodsim/reach3d.py:378-470): thrust ramp, wall-model
    reference speed and time averages on a 20-step cadence, wall-clock MLUPS.

    The ramp, the uref column reduction and the running sums all execute on
    the GPU (libwfs.so); only scalars cross to the host between chunks.
    """

    cadence = 20

    def __init__(self, reach: Reach, be=None, *, storage="f32", block_x=0, device=0, closure=None,
                 column_u="cadence", streaming="ab", pop_stagger=0, closure_field=True):
        """closure: None for the constant-Cs Smagorinsky model, or the weights
        of the learned closure (closure.py), evaluated inside the step.
        block_x: work-group width, 0 for the default, "auto" for the tuned one.
        column_u: "cadence" refreshes the wall model's column speed once per
        20-step chunk, as the reference does (reach3d.py:427-440); "per_step"
        refreshes it every step inside the step kernel.
        streaming: "ab" streams between two population buffers, as the
        reference does; "inplace" keeps one (half the population memory, the
        same results bit for bit).
        pop_stagger: bytes between the population direction arrays (a layout
        choice, results unchanged), "auto" for the tuned one.
        closure_field: False keeps no per-cell C^2 field of the learned
        closure (lbm.closure_field() then refuses; 4 bytes per cell less)."""
        if column_u not in ("cadence", "per_step"):
            raise ValueError("column_u: 'cadence' or 'per_step'")
        if streaming not in ("ab", "inplace"):
            raise ValueError("streaming: 'ab' or 'inplace'")
        self.column_u = column_u
        self.r, self.be = reach, be
        if block_x == "auto" and reach.layout == "dense":
            from .autotune import block_x_for
            block_x = block_x_for(reach, storage, closure)
        elif block_x == "auto":
            block_x = 0
        if pop_stagger == "auto" and reach.layout == "dense":
            from .autotune import pop_stagger_for
            pop_stagger = pop_stagger_for(reach, storage)
        elif pop_stagger == "auto":
            pop_stagger = 0
        nu_lb = max(1e-6 * reach.dt / reach.dx ** 2, 2e-4)
        uin_lb = (reach.uin_phys / reach.vel).astype(np.float32)
        self.lbm = LBM3D(be, reach.solid, reach.Fx, reach.Fy, reach.Fz, uin_lb, tau0=0.5 + 3 * nu_lb,
                         cs=reach.cfg.cs, cf_wall=reach.cf, storage=storage, block_x=block_x, device=device,
                         layout=reach.layout, streaming=streaming, pop_stagger=pop_stagger)
        if reach.force_cells is not None:
            self.lbm.set_force_cells(*reach.force_cells)
        if closure is not None:
            from .closure import geometry
            geo, hcol = geometry(reach.solid)
            if not closure_field:
                self.lbm.keep_closure_field(False)
            self.lbm.set_closure("mlp", closure, geo, hcol, float(uin_lb[uin_lb > 0].mean()))
        if reach.bedwall is not None:
            self.lbm.set_bedwall(reach.bedwall)
            self.lbm.update_uref()
            if column_u == "per_step":
                self.lbm.set_column_u(True)
        self.navg = 0
        self.t = 0.0
        self.t_total = reach.cfg.run_flows * reach.cfg.length / max(reach.V, 0.05)
        self.t_avg = self.t_total * (1 - reach.cfg.avg_frac)
        self.t_ramp = min(0.15 * self.t_total, 20.0) if reach.props else 0.0
        self._ramp_done = self.t_ramp <= 0
        if not self._ramp_done:
            self.lbm.set_force_scale(0.0)
        self.wall0 = time.time()
        self.steps_wall = 0.0

    @property
    def done(self):
        return self.t >= self.t_total

    def _update_uref(self):
        self.lbm.update_uref()

    def advance(self, nsteps: int):
        t0 = time.time()
        L = self.lbm
        k = 0
        while k < nsteps:
            m = min(self.cadence, nsteps - k)
            if not self._ramp_done:
                s = min(self.t / self.t_ramp, 1.0)
                s = np.float32(s * s * (3 - 2 * s))
                L.set_force_scale(s)
                self._ramp_done = s >= 1.0
            if self.r.bedwall is not None and self.column_u == "cadence":
                L.update_uref()
            L.step(m)
            k += m
            self.t += m * self.r.dt
            if self.t >= self.t_avg:
                L.accumulate()
                self.navg += 1
        L.synchronize()
        self.steps_wall += time.time() - t0

    @staticmethod
    def run_chunk(chunk=400):
        """Steps per advance() call of run_to_end: ``chunk``, or 100 when the
        GPU duty cycle is limited (reach3d.py:456-458)."""
        return 100 if min(max(GPU_DUTY, 0.05), 1.0) < 1 else chunk

    def steps_to_end(self, chunk=400):
        """How many steps run_to_end(chunk=chunk) takes from the current time:
        whole advance() calls until t >= t_total, with t accumulated chunk by
        chunk in the same float64 operations as advance()."""
        chunk = self.run_chunk(chunk)
        t, n = self.t, 0
        while not t >= self.t_total:
            k = 0
            while k < chunk:
                m = min(self.cadence, chunk - k)
                t += m * self.r.dt
                k += m
            n += chunk
        return n

    def run_to_end(self, progress=None, cancel=None, chunk=400):
        duty = min(max(GPU_DUTY, 0.05), 1.0)
        chunk = self.run_chunk(chunk)
        while not self.done:
            t0 = time.time()
            self.advance(chunk)
            if duty < 1:
                time.sleep((time.time() - t0) * (1 / duty - 1))
            if progress:
                progress(self.t / self.t_total)
            if cancel and cancel():
                break

    def tick(self, budget_s=0.08, chunk=20):
        """One frame of an interactive run (the reference's GUI loop,
        floodsim/gui.py:823-842): whole chunks of ``chunk`` steps until
        ``budget_s`` seconds of wall time are spent (at least one), tracer
        particles moved by the simulated time covered, and the figures a
        status line shows."""
        t0 = time.time()
        n = 0
        while True:
            self.advance(chunk)
            n += chunk
            if time.time() - t0 >= budget_s:
                break
        respawned = self.move_particles(n * self.r.dt)
        return {"steps": n, "t": self.t, "t_total": self.t_total, "nstep": self.lbm.nstep,
                "mlups": self.mlups(), "navg": self.navg, "respawned": respawned}

    def mlups(self):
        """Million lattice updates per second, all cells counted (reach3d.py:469-470)."""
        return self.lbm.nstep * self.r.cells / max(self.steps_wall, 1e-9) / 1e6

    def _raw(self, name, mean):
        wide = np.float64 if self.lbm.storage == "f64" else np.float32
        if mean and self.navg > 0:
            # FP64 readout: the sum of rho comes back as n + sum(rho - 1) without a float rounding
            sums, n = self.lbm.sums(np.float64)
            return sums[name] / n
        rho, ux, uy, uz = self.lbm.macro(wide)
        if name == "sp":
            return np.sqrt(ux * ux + uy * uy + uz * uz).astype(np.float64)
        if name == "uxx":
            return (ux * ux).astype(np.float64)
        return dict(rho=rho, ux=ux, uy=uy, uz=uz)[name].astype(np.float64)

    def field(self, name: str, mean: bool = False) -> np.ndarray:
        """Field in physical units as float64 (nz, ny, nx): rho -> water level
        (m), uxx -> m^2/s^2, velocities -> m/s."""
        r = self.r
        a = self._raw(name, mean).reshape(r.nz, r.ny, r.nx)
        if name == "rho":
            return r.eta_from_rho(a)
        if name == "uxx":
            return a * r.vel ** 2
        return a * r.vel

    def slice_np(self, name, axis, idx, mean=False):
        """One plane of a field (axis 0 = z, 1 = y, 2 = x) in physical units."""
        a = self.field(name, mean)
        return a[idx] if axis == 0 else (a[:, idx, :] if axis == 1 else a[:, :, idx])

    # ---- tracer particles (floodsim/reach3d.py:500-541), advected on the GPU
    def init_particles(self, n, seed=0):
        """n particles in random fluid cells (the reference's seeded draw, so the
        same start positions); ``seed`` feeds the device's restart draws."""
        r = self.r
        rng = np.random.default_rng(0)
        fluid = np.argwhere(~r.solid)
        pick = fluid[rng.integers(0, len(fluid), n)]
        P = (pick[:, ::-1] + rng.random((n, 3))) * r.dx   # (x, y index, z) in metres from the domain corner
        inlet = np.argwhere(~r.solid[:, :, 1])             # (z, y) of the cells particles restart in
        self.lbm.particles_set(P, inlet[:, 1] + r.ny * inlet[:, 0], seed)
        self.particles = True

    def move_particles(self, dt_phys):
        """One tracer step of dt_phys seconds through the current velocity field;
        returns the number of particles restarted at the inlet."""
        if not getattr(self, "particles", False):
            return 0
        return self.lbm.particles_move(dt_phys, self.r.vel, self.r.dx)

    @property
    def P(self):
        """Particle positions (n x 3, metres from the domain corner)."""
        return self.lbm.particles_get()

    def particles_xyz(self):
        """x from the inlet, y from the channel centre line, z from the bottom."""
        r = self.r
        P = self.P
        return P[:, 0], P[:, 1] + r.y[0] - 0.5 * r.dx, P[:, 2]

    def hull_mask(self):
        """Solid cells that are not part of the bed/bank cross-section (the hulls)."""
        if self.r.hull_cells is not None:
            return self.r.hull_cells
        s = self.r.solid
        return s & ~np.broadcast_to(s[:, :, :1], s.shape)

    def forces(self):
        """FP64 momentum budget of the current state in newtons: applied
        thrust, force of the water on the hulls, and on the bed and banks
        (x, y, z each; lattice momentum per step times rho dx^4 / dt^2)."""
        if not getattr(self, "_body_set", False):
            self.lbm.set_body_mask(self.hull_mask())
            self._body_set = True
        f = self.lbm.forces()
        k = RHO * self.r.dx ** 4 / self.r.dt ** 2
        return {name: v * k for name, v in f.items()}

    def checkpoint(self, path):
        """Save the run (solver state, simulated time, ramp and averaging
        counters) to an .npz file; restore() into a Run3D of the same reach
        and options continues it bit for bit. Particles are not saved."""
        np.savez(path, state=self.lbm.state(), t=np.float64(self.t), ramp_done=np.bool_(self._ramp_done),
                 navg=np.int64(self.navg), steps_wall=np.float64(self.steps_wall))

    def restore(self, path):
        with np.load(path) as d:
            self.lbm.load_state(d["state"])
            self.t = float(d["t"])
            self._ramp_done = bool(d["ramp_done"])
            self.navg = int(d["navg"])
            self.steps_wall = float(d["steps_wall"])

    def close(self):
        self.lbm.close()


# ================================================================ readouts
class CrossSections:
    """Reductions of (nz, ny, nx) time means to one value per cross-section x
    of the wet section (~solid), with cell faces dx by dx."""

    def __init__(self, solid: np.ndarray, dx: float):
        self.wet = ~solid
        self.dx = dx
        self.lid = self.wet[-1]                         # (ny, nx): fluid cells right under the lid
        self.n_lid = np.maximum(self.lid.sum(0), 1)
        dry_below = np.ones_like(self.wet)
        dry_below[1:] = ~self.wet[:-1]
        self.bed = self.wet & dry_below                  # first fluid cell above solid in each column

    def under_lid(self, f: np.ndarray) -> np.ndarray:
        """Mean of f over the top fluid layer."""
        return (f[-1] * self.lid).sum(0) / self.n_lid

    def integral(self, f: np.ndarray) -> np.ndarray:
        """Integral of f over the wet section."""
        return (f * self.wet).sum(axis=(0, 1)) * self.dx * self.dx

    def peak(self, f: np.ndarray) -> np.ndarray:
        return np.where(self.wet, f, 0).max(axis=(0, 1))

    def at_bed(self, f: np.ndarray) -> np.ndarray:
        """(ny, nx) map of f at the first fluid cell above the bed (the
        largest over the column where it meets solid more than once)."""
        return np.where(self.bed, f, 0.0).max(axis=0)


def profiles(run: Run3D) -> dict:
    """Along-river profiles of the time-mean fields (floodsim/reach3d.py:545-568):
    level under the lid relative to the outlet, surface velocity, peak speed,
    discharge, momentum + pressure flux (rho u^2 + rho g eta) and bed shear
    rho cf |u|^2 with the log-law cf of the Manning roughness at dx/2."""
    r = run.r
    mean = {k: run.field(k, mean=True) for k in ("ux", "sp", "rho", "uxx")}
    xs = CrossSections(r.solid, r.dx)
    level = xs.under_lid(mean["rho"])                 # field("rho") is the level eta
    cf_bed = log_cf(0.5 * r.dx, manning_ks(r.sec.n))
    tau_bed = xs.at_bed(RHO * cf_bed * mean["sp"] ** 2)
    return dict(x=r.x, eta=level - level[-1], us=xs.under_lid(mean["ux"]), umax=xs.peak(mean["sp"]),
                q=xs.integral(mean["ux"]), tau_bed=tau_bed, tau_max_x=tau_bed.max(0),
                mom=xs.integral(RHO * mean["uxx"] + RHO * G * mean["rho"]))


def _upstream_gauge(x: np.ndarray, hulls) -> int:
    """Index of the upstream level gauge: 8 % down the reach, but at least
    10 m ahead of the first bow, and never the inlet column."""
    i = int(0.08 * len(x))
    if hulls:
        i = min(i, int(np.searchsorted(x, min(h[0] for h in hulls) - 10.0)))
    return max(i, 1)


def _jet_reach(x: np.ndarray, excess: np.ndarray, x_from: float, threshold: float) -> tuple[float, bool]:
    """How far past x_from the peak speed still exceeds the base's by more
    than threshold, and whether that reaches the last three columns."""
    hit = np.flatnonzero((x > x_from) & (excess > threshold))
    if not len(hit):
        return 0.0, False
    return float(x[hit[-1]] - x_from), bool(hit[-1] >= len(x) - 3)


def _shear_window(x: np.ndarray, hulls, dx: float) -> np.ndarray:
    """Columns where bed shear is judged: from 5 m ahead of the first bow
    (of the point 15 % down the reach without hulls), but not within the
    first 10 % of the reach, to three cells before the outlet."""
    first_bow = min([h[0] for h in hulls], default=0.15 * x[-1])
    return (x >= max(first_bow - 5.0, 0.1 * x[-1])) & (x <= x[-1] - 3 * dx)


def compare(res: dict, base: dict, reach: Reach, cfg: ReachConfig) -> dict:
    """A boat case against the no-boat base at the same level
    (floodsim/reach3d.py:571-618): level change at the upstream gauge,
    jet reach, surface-speed gain downstream of the last propeller, bed area
    above the critical shear, the net force from the change of momentum +
    pressure flux, and the discharge gain if such a group stood every
    cfg.spacing_km (momentum balance against the bed slope)."""
    x = res["x"]
    fleet = reach.fleet
    thrust, power = fleet.total_thrust(), fleet.total_power()
    x_last = max((p["x"] for p in reach.props), default=0.0)
    i_up, i_out = _upstream_gauge(x, reach.hulls), len(x) - 4

    def rise(p):   # level at the gauge above the outlet
        return p["eta"][i_up] - p["eta"][-1]

    d_eta = rise(res) - rise(base)
    jet_m, jet_capped = _jet_reach(x, res["umax"] - base["umax"], x_last, max(0.2, 0.15 * reach.V))
    downstream = (x > x_last) & (x < x[-1] - 2 * reach.dx)
    us_gain = float((res["us"] - base["us"])[downstream].mean()) if downstream.any() else 0.0
    win = _shear_window(x, reach.hulls, reach.dx)
    tau, tau_b = res["tau_bed"][:, win], base["tau_bed"][:, win]
    cell = reach.dx * reach.dx
    d_mom = res["mom"] - base["mom"]
    f_net = float(d_mom[i_out] - d_mom[i_up])
    weight_along = RHO * G * reach.A * (cfg.spacing_km * 1000) * max(reach.S, 1e-7)

    def q_gain(force):
        return math.sqrt(max(1 + force / weight_along, 0.0)) - 1

    dq_t, dq_n = q_gain(thrust), q_gain(f_net)
    per_100kw = power / 100
    return dict(label=fleet.label(), n=fleet.n_boats, T_kN=thrust / 1000, P_kW=power, h=reach.h, Q=reach.Q,
                V=reach.V, d_eta_cm=100 * d_eta, jet_reach=jet_m, reach_capped=jet_capped,
                us_gain=us_gain, us_gain_pct=100 * us_gain / max(reach.V, 1e-6),
                umax=float(res["umax"].max()), tau_max=float(tau.max()), tau_base=float(tau_b.max()),
                tau_area=float((tau > cfg.tau_crit).sum() * cell),
                tau_area_base=float((tau_b > cfg.tau_crit).sum() * cell),
                dQ_pct=100 * dq_t, dQ_net_pct=100 * dq_n, F_net_kN=f_net / 1000,
                dQ_per_100kW=100 * dq_n / max(per_100kw, 1e-9) if power > 0 else 0.0,
                drop_cm_per_100kW=-100 * d_eta / per_100kw if power > 0 else 0.0,
                kW_per_cm=power / (-100 * d_eta) if (power > 0 and d_eta < -1e-5) else float("nan"),
                T_per_kW=thrust / max(power, 1e-9) if power > 0 else 0.0)
