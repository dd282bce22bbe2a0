"""Discharge for a water level from the divided-channel Manning formula.

This is the step that feeds the solver's inlet discharge Q for a chosen level
h (the reference's ``hydraulics1d.q_ref``, floodsim/hydraulics1d.py:102-105,
built on ``xs_props``/``conveyance``/``meander_mult``, lines 18-53). Main
channel and floodplains carry separate conveyances K = A R^(2/3) / n; the
channel roughness is raised by Cowan's meander factor for the river's
sinuosity sigma.
"""
from __future__ import annotations

import numpy as np

from .geometry import Section


def meander_mult(sig):
    """Cowan (1956) meander factor m5: 1.0 up to sigma 1.1, rising linearly to
    1.15 at 1.5 and to at most 1.30 beyond (floodsim/hydraulics1d.py:18-23)."""
    s = np.asarray(sig, float)
    low = 1.0 + 0.15 * (s - 1.1) / 0.4
    high = np.minimum(1.30, 1.15 + 0.15 * (s - 1.5) / 0.5)
    return np.where(s <= 1.1, 1.0, np.where(s <= 1.5, low, high))


def xs_props(h, sec: Section) -> dict:
    """Area, wetted perimeter and hydraulic radius of the main channel and the
    floodplains at depth h above the bed (floodsim/hydraulics1d.py:26-48)."""
    h = np.maximum(np.asarray(h, float), 0.0)
    B, Hb, z, Bf = sec.B, sec.Hb, sec.z, sec.Bf
    slope_len = np.sqrt(1 + z * z)
    in_bank = h <= Hb
    above = h - Hb
    Am = np.where(in_bank, (B + z * h) * h, (B + z * Hb) * Hb + sec.Tb * above)
    Pm = np.where(in_bank, B + 2 * h * slope_len,
                  B + 2 * Hb * slope_len + (0.0 if Bf > 0 else 2 * np.maximum(above, 0)))
    Tm = np.where(in_bank, B + 2 * z * h, sec.Tb)
    if Bf > 0:
        Af = np.where(in_bank, 0.0, 2 * Bf * np.maximum(above, 0))
        Pf = np.where(in_bank, 0.0, 2 * Bf + 2 * np.maximum(above, 0))
        T = np.where(in_bank, Tm, sec.Tb + 2 * Bf)
    else:
        Af = np.zeros_like(h)
        Pf = np.zeros_like(h)
        T = Tm
    Rm = Am / np.maximum(Pm, 1e-9)
    Rf = np.where(Pf > 0, Af / np.maximum(Pf, 1e-9), 0.0)
    return dict(A=Am + Af, Am=Am, Af=Af, Pm=Pm, Pf=Pf, Rm=Rm, Rf=Rf, T=T, Tm=Tm,
                Kmu=Am * Rm ** (2 / 3), Kfu=Af * Rf ** (2 / 3))


def conveyance(h, sec: Section, nm):
    p = xs_props(h, sec)
    return p["Kmu"] / nm + p["Kfu"] / sec.nf


def q_ref(h0, sigma: float, slope: float, sec: Section):
    """Discharge whose normal depth is h0 for a river of sinuosity ``sigma``
    and mean slope ``slope``. The reference takes both from a River object
    (river.sigma, river.mean_slope); they are passed directly here."""
    n_eff = sec.n * float(meander_mult(sigma))
    return conveyance(h0, sec, n_eff) * np.sqrt(slope)
