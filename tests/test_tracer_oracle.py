"""Pins of the tracer oracle (oracle/tracer_oracle.py): closed forms and
brute force."""
import numpy as np

from oracle import tracer_oracle as T


def _box(nz=5, ny=6, nx=9):
    solid = np.zeros((nz, ny, nx), bool)
    solid[0] = True
    solid[2:4, 2, 4] = True
    return solid


def test_uniform_field_translates_by_u_dt():
    solid = _box()
    rng = np.random.default_rng(1)
    P = np.column_stack([rng.uniform(2, 3, 50), rng.uniform(0.5, 1.5, 50), rng.uniform(0.6, 0.9, 50)])   # fluid cells
    f = [np.full(solid.shape, v, np.float32) for v in (0.031, -0.0042, 0.0017)]
    vel, dt, dx = 12.5, 0.37, 0.5
    Q, bad = T.move(P, *f, solid, dx, vel, dt)
    step = np.array([float(np.float32(np.float32(v) * np.float32(vel)) * np.float32(dt))
                     for v in (0.031, -0.0042, 0.0017)])
    assert np.array_equal(Q, P + step)
    assert not bad.any()


def test_lookup_and_restart_brute_force():
    solid = _box()
    nz, ny, nx = solid.shape
    rng = np.random.default_rng(2)
    dx = 0.25
    P = rng.uniform(-0.4, 1.1, (400, 3)) * np.array([nx, ny, nz]) * dx
    ux = rng.normal(0, 0.05, solid.shape).astype(np.float32)
    uy = rng.normal(0, 0.05, solid.shape).astype(np.float32)
    uz = rng.normal(0, 0.05, solid.shape).astype(np.float32)
    ux[1, 1, 1] = np.nan
    vel, dt = 4.0, 0.05
    Q, bad = T.move(P, ux, uy, uz, solid, dx, vel, dt)
    for k in range(len(P)):
        i = [min(max(int(P[k, a] / dx), 0), n - 1) for a, n in enumerate((nx, ny, nz))]
        q = P[k].copy()
        for a, f in enumerate((ux, uy, uz)):
            v = np.float32(f[i[2], i[1], i[0]]) * np.float32(vel)
            v = np.float32(0) if not np.isfinite(v) else v
            q[a] += float(np.float32(v * np.float32(dt)))
        assert np.array_equal(Q[k], q)
        j = [int(q[a] / dx) for a in range(3)]
        outside = not (0 <= j[0] < nx - 1 and 0 <= j[1] < ny and 0 <= j[2] < nz)
        assert bad[k] == (outside or bool(solid[j[2], j[1], j[0]]))
