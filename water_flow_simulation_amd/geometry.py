"""Channel cross-section and generated river centre lines.

``Section`` has the fields and meaning of the reference's dataclass
(floodsim/geometry.py:26-38): a trapezoidal main channel with bed width ``B``,
bank height ``Hb`` and bank slope ``z`` (horizontal:vertical), plus a
floodplain of width ``Bf`` on each side, Manning coefficients ``n`` (channel)
and ``nf`` (floodplain).

``generate_river`` builds the synthetic meandering river the reference's
experiments start from (floodsim/geometry.py:131-177): its sinuosity and
mean bed slope feed the 1-D discharge of a level (hydraulics1d.q_ref), which
is the inlet discharge of a 3-D reach.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Section:
    B: float = 40.0
    Hb: float = 3.0
    z: float = 2.0
    Bf: float = 200.0
    n: float = 0.032
    nf: float = 0.06

    @property
    def Tb(self) -> float:
        """Top width of the main channel at bank height."""
        return self.B + 2 * self.z * self.Hb


@dataclass
class RiverParams:
    """Valley length (km, straight line), requested sinuosity, valley slope
    (m/km), bend wavelength (multiples of the bed width), irregularity of the
    bends (0..1), random seed and downstream bed level (m), as the
    reference's RiverParams (floodsim/geometry.py:15-22)."""
    D_km: float = 10.0
    sigma: float = 1.6
    Sv_m_per_km: float = 0.5
    wavelength_B: float = 12.0
    irregularity: float = 0.25
    seed: int = 7
    z_down: float = 2.0


# The rivers of the reference's five experiments (floodsim/experiments.py),
# by key, with their sections.
EXPERIMENT_RIVERS = {
    "E1": (RiverParams(D_km=4, sigma=1.0, Sv_m_per_km=0.5, wavelength_B=12, irregularity=0.0, seed=1),
           Section(B=20, Hb=4, z=1.5, Bf=60, n=0.025, nf=0.05)),
    "E2": (RiverParams(D_km=3, sigma=1.15, Sv_m_per_km=15, wavelength_B=10, irregularity=0.4, seed=3),
           Section(B=12, Hb=1.5, z=1.0, Bf=25, n=0.045, nf=0.08)),
    "E3": (RiverParams(D_km=8, sigma=1.8, Sv_m_per_km=0.4, wavelength_B=12, irregularity=0.25, seed=7),
           Section(B=50, Hb=4, z=2, Bf=300, n=0.03, nf=0.06)),
}


@dataclass
class River:
    """A generated centre line: plan coordinates x, y (m, downstream along
    +x), arc length s, bed level zb along it, its length L, the straight
    distance D between its ends and the sinuosity L / D it reached."""
    x: np.ndarray
    y: np.ndarray
    s: np.ndarray
    zb: np.ndarray
    L: float
    D: float
    sigma: float

    @property
    def mean_slope(self) -> float:
        return max((self.zb[0] - self.zb[-1]) / self.L, 1e-6)


def _sine_generated(L, ds_target, lam_arc, amp, amp_mult, half_mult):
    """A sine-generated curve of arc length L: the heading is
    amp * amp_mult[k] * sin(phase) over successive half waves k of arc length
    lam_arc * half_mult[k] / 2; positions integrate the heading at the
    midpoints of n uniform steps. Returns s, x, y."""
    n = max(int(round(L / ds_target)), 2) + 1
    s = np.linspace(0, L, n)
    step = s[1] - s[0]
    half = 0.5 * lam_arc * half_mult
    starts = np.concatenate([[0.0], np.cumsum(half)])
    k = np.clip(np.searchsorted(starts, s, side="right") - 1, 0, len(half) - 1)
    heading = amp * amp_mult[k] * np.sin(np.pi * k + np.pi * (s - starts[k]) / half[k])
    mid = 0.5 * (heading[1:] + heading[:-1])
    x = np.concatenate([[0.0], np.cumsum(step * np.cos(mid))])
    y = np.concatenate([[0.0], np.cumsum(step * np.sin(mid))])
    return s, x, y


def generate_river(rp: RiverParams, sec: Section) -> River:
    """The reference's synthetic river (floodsim/geometry.py:131-177): a
    sine-generated centre line of length D * sigma whose bend amplitude is
    found by bisection so that its ends lie D apart, with half-wave lengths
    and amplitudes perturbed by ``irregularity`` (seeded); the bed falls
    linearly along the arc by the valley slope times the end-to-end distance."""
    D = rp.D_km * 1000.0
    sigma = max(1.0, rp.sigma)
    L = D * sigma
    lam_arc = max(rp.wavelength_B, 2.0) * sec.B * sigma
    ds = max(min(lam_arc / 40.0, L / 800.0), L / 20000.0)
    rng = np.random.default_rng(rp.seed)
    nh = int(2 * L / lam_arc) + 20
    irr = float(np.clip(rp.irregularity, 0, 1))
    amp_mult = 1 + irr * 0.45 * (2 * rng.random(nh) - 1)
    half_mult = 1 + irr * 0.5 * (2 * rng.random(nh) - 1)
    amp = 0.0
    if sigma > 1.0005:
        lo, hi = 0.0, 2.25
        for _ in range(45):
            mid = 0.5 * (lo + hi)
            _, x, y = _sine_generated(L, ds, lam_arc, mid, amp_mult, half_mult)
            if np.hypot(x[-1], y[-1]) > D:
                lo = mid
            else:
                hi = mid
        amp = 0.5 * (lo + hi)
    s, x, y = _sine_generated(L, ds, lam_arc, amp, amp_mult, half_mult)
    # turn the line so that it runs from the origin along +x
    a = np.arctan2(y[-1], x[-1])
    c, sn = np.cos(-a), np.sin(-a)
    x, y = x * c - y * sn, x * sn + y * c
    chord = np.hypot(x[-1], y[-1])
    drop = rp.Sv_m_per_km / 1000.0 * chord
    zb = rp.z_down + drop * (1 - s / s[-1])
    return River(x=x, y=y, s=s, zb=zb, L=float(s[-1]), D=float(chord), sigma=float(s[-1] / chord))

