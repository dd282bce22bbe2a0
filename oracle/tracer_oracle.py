"""CPU oracle for the particle tracer.

TEST INFRASTRUCTURE ONLY: the checker the device tracer (k_particles in
libwfs) is compared against. Only tests/ imports it.

Restates the deterministic part of the reference's Run3D.move_particles
(floodsim/reach3d.py:509-531) in numpy: the velocity of the cell holding a
particle (indices truncated toward zero and clamped, reach3d.py:513-517),
converted to m/s and multiplied by the time step in float32 as NumPy 2
evaluates `float32 array * python float` (reach3d.py:518-519), added to the
float64 position; then the test for leaving the box (x past nx - 1, y, z
past the grid) or landing in a solid cell (reach3d.py:520-527). Where the
reference draws restart positions from an unseeded generator
(reach3d.py:528-531), this oracle only reports which particles restart.

Pinned by tests/test_tracer_oracle.py: a uniform field moves every particle
by the closed-form step; cell lookup and the box/solid classification are
checked against brute-force loops.
"""
from __future__ import annotations

import numpy as np


def move(P, ux, uy, uz, solid, dx, vel, dt):
    """P: (n, 3) float64 metres from the domain corner; ux, uy, uz: (nz, ny,
    nx) float32 lattice velocity; solid: (nz, ny, nx) bool. Returns the moved
    positions and the mask of particles that must restart at the inlet."""
    nz, ny, nx = solid.shape
    P = np.array(P, dtype=np.float64, copy=True)
    ix = np.clip((P[:, 0] / dx).astype(np.int64), 0, nx - 1)
    iy = np.clip((P[:, 1] / dx).astype(np.int64), 0, ny - 1)
    iz = np.clip((P[:, 2] / dx).astype(np.int64), 0, nz - 1)
    u = np.stack([np.asarray(a, np.float32)[iz, iy, ix] for a in (ux, uy, uz)], 1)
    u = u * np.float32(vel)                                  # float32
    u = np.nan_to_num(u, nan=0.0, posinf=0.0, neginf=0.0)
    P += u * np.float32(dt)                                  # float32 product, float64 sum
    fin = np.isfinite(P).all(1)
    Q = np.where(fin[:, None], P, -1.0)
    jx, jy, jz = ((Q[:, k] / dx).astype(np.int64) for k in range(3))
    out = ~fin | (jx < 0) | (jx >= nx - 1) | (jy < 0) | (jy >= ny) | (jz < 0) | (jz >= nz)
    bad = out.copy()
    ok = ~out
    bad[ok] = solid[jz[ok], jy[ok], jx[ok]]
    return P, bad
