"""Channel cross-section type consumed by the case builder.

``Section`` has the fields and meaning of the reference's dataclass
(floodsim/geometry.py:26-38): a trapezoidal main channel with bed width ``B``,
bank height ``Hb`` and bank slope ``z`` (horizontal:vertical), plus a
floodplain of width ``Bf`` on each side, Manning coefficients ``n`` (channel)
and ``nf`` (floodplain).
"""
from __future__ import annotations

from dataclasses import dataclass


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
