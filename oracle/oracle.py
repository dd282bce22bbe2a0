"""CPU oracle driver for the river-reach D3Q19 solver.

TEST INFRASTRUCTURE ONLY: this module is the checker. Only tests/, the smoke
check in __graft_entry__.py and the cpu_baseline leg of bench.py may import it;
the product package never does.

It wraps oracle/_build/liblbm_oracle.so (a C++ restatement of floodsim/lbm3d.py
lbm_step / lbm_bc, see lbm_oracle.cpp) behind an object with the surface of the
reference's ``LBM3D`` (floodsim/lbm3d.py:333-418), and restates the run cadence
of ``Run3D.advance`` (floodsim/reach3d.py:424-453) in numpy so that the product's
own driver is checked against an independent copy.

precision: "f64" (FP64 oracle), "f32" (IEEE single, reference CUDA evaluation
order), "mixed" (float storage, double arithmetic, like the Numba path).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblbm_oracle.so")

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])
MIRZ = np.array([0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16])

_lib = None


def build():
    """Compile the oracle library with its Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, i64, f32, f64 = ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_double
        for name, rt in (("f64", f64), ("f32", f32), ("mixed", f32)):
            fn = getattr(L, f"orc_step_{name}")
            fn.argtypes = [P] * 12 + [i32, i32, i32, rt, rt]
            fn.restype = i32
            fn = getattr(L, f"orc_bc_{name}")
            fn.argtypes = [P, P, P, i32, i32, i32, rt]
            fn.restype = i32
        for name in ("f64", "f32"):
            fn = getattr(L, f"orc_init_{name}")
            fn.argtypes = [P, P, i64]
            fn.restype = i32
            fn = getattr(L, f"orc_uref_{name}")
            fn.argtypes = [P] * 5 + [i32, i32, i32]
            fn.restype = i32
        L.orc_num_threads.restype = i32
        _lib = L
    return _lib


def num_threads() -> int:
    return int(lib().orc_num_threads())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleLBM:
    """CPU restatement of ``LBM3D`` (floodsim/lbm3d.py:342-412).

    Same constructor arguments as the reference (the ``be`` backend argument is
    dropped) plus ``precision``. Arrays are flat numpy arrays of N cells; the
    caller may replace ``Fx/Fy/Fz/uref`` like the reference's callers do.
    """

    def __init__(self, solid, Fx, Fy, Fz, uin, tau0, cs=0.17, cf_wall=None, precision="f32"):
        if precision not in ("f64", "f32", "mixed"):
            raise ValueError(precision)
        self.precision = precision
        self.dt = np.float64 if precision == "f64" else np.float32
        self.nz, self.ny, self.nx = solid.shape
        self.N = solid.size
        self.tau0 = float(tau0)
        self.csmag2 = float(cs * cs)
        A = lambda a: np.array(np.asarray(a).ravel(), dtype=self.dt, copy=True)  # owns its data  # noqa: E731
        self.solid = np.ascontiguousarray(solid.ravel(), dtype=np.uint8)
        self.Fx, self.Fy, self.Fz = A(Fx), A(Fy), A(Fz)
        self.uin = A(uin)
        self.rho = np.ones(self.N, self.dt)
        self.ux = np.zeros(self.N, self.dt)
        self.uy = np.zeros(self.N, self.dt)
        self.uz = np.zeros(self.N, self.dt)
        self.cfw = A(cf_wall if cf_wall is not None else np.zeros(solid.shape))
        self.uref = np.zeros(self.N, self.dt)
        # lbm3d.py:362-364: inlet profile broadcast along x, zero in solid cells.
        u0 = np.broadcast_to(np.asarray(uin)[:, :, None], solid.shape).astype(self.dt)
        u0 = np.where(solid, 0, u0).astype(self.dt).ravel()
        self.f = np.empty(19 * self.N, self.dt)
        if self.dt is np.float64:
            lib().orc_init_f64(_ptr(self.f), _ptr(u0), self.N)
        else:
            lib().orc_init_f32(_ptr(self.f), _ptr(u0), self.N)
        self.f2 = np.zeros_like(self.f)
        self.nstep = 0

    def step(self, n=1):
        L = lib()
        st = getattr(L, f"orc_step_{self.precision}")
        bc = getattr(L, f"orc_bc_{self.precision}")
        for a, size in ((self.Fx, self.N), (self.Fy, self.N), (self.Fz, self.N), (self.uin, self.ny * self.nz),
                        (self.uref, self.N)):
            if a.dtype != self.dt or a.size != size or not a.flags.c_contiguous:
                raise TypeError("OracleLBM: force/inlet arrays must match the storage dtype and size")
        for _ in range(n):
            st(_ptr(self.f), _ptr(self.f2), _ptr(self.solid), _ptr(self.Fx), _ptr(self.Fy), _ptr(self.Fz),
               _ptr(self.rho), _ptr(self.ux), _ptr(self.uy), _ptr(self.uz), _ptr(self.cfw), _ptr(self.uref),
               self.nx, self.ny, self.nz, self.tau0, self.csmag2)
            bc(_ptr(self.f2), _ptr(self.solid), _ptr(self.uin), self.nx, self.ny, self.nz, 1.0)
            self.f, self.f2 = self.f2, self.f
        self.nstep += n

    def update_uref(self, bedwall):
        """Run3D._update_uref (reach3d.py:412-418)."""
        bw = np.ascontiguousarray(np.asarray(bedwall).ravel(), dtype=np.uint8)
        fn = lib().orc_uref_f64 if self.dt is np.float64 else lib().orc_uref_f32
        fn(_ptr(self.ux), _ptr(self.uy), _ptr(self.solid), _ptr(bw), _ptr(self.uref), self.nx, self.ny, self.nz)

    def pops(self):
        return self.f.reshape(19, self.nz, self.ny, self.nx)

    def shape3(self, a):
        return a.reshape(self.nz, self.ny, self.nx)


def drive(solver, bedwall, nsteps, cadence=20):
    """The solver-level part of Run3D.advance (reach3d.py:429-441): before every
    chunk of at most ``cadence`` steps, refresh uref from the current velocity
    (when a bed/bank mask exists), then step the chunk."""
    k = 0
    while k < nsteps:
        m = min(cadence, nsteps - k)
        if bedwall is not None:
            solver.update_uref(bedwall)
        solver.step(m)
        k += m
    return solver


class OracleRun:
    """Independent numpy restatement of Run3D (floodsim/reach3d.py:381-453):
    thrust ramp ``F0*s^2(3-2s)`` per chunk, uref refresh per chunk, and float32
    running sums of ux, uy, uz, rho, |u|, ux^2 once the averaging window opens.

    ``inputs`` holds the case arrays produced by a case builder: solid, Fx, Fy,
    Fz (lattice units), uin_lb, tau0, cs, cf, bedwall, and the scalars dt,
    t_total, t_avg, t_ramp.

    ``rho_sum``: "reference" sums rho in float32 as the reference does
    (reach3d.py:447); "deviation" sums rho - 1 in float32, as the library
    does, and ``sum["rho"]`` then reads n + sum(rho - 1), formed in double and
    rounded to float32 once (the library's readout).
    """

    cadence = 20   # steps between ramp/uref updates and averaging samples (reach3d.py:427)

    def __init__(self, inputs, precision="f32", rho_sum="deviation"):
        if rho_sum not in ("reference", "deviation"):
            raise ValueError("rho_sum: 'reference' or 'deviation'")
        self.rho_sum = rho_sum
        self.i = inputs
        self.lbm = OracleLBM(inputs["solid"], inputs["Fx"], inputs["Fy"], inputs["Fz"], inputs["uin_lb"],
                             inputs["tau0"], inputs["cs"], inputs["cf"], precision=precision)
        self.bedwall = inputs.get("bedwall")
        self.F0 = (self.lbm.Fx.copy(), self.lbm.Fy.copy(), self.lbm.Fz.copy())
        self.t = 0.0
        self.dt = inputs["dt"]
        self.t_total = inputs["t_total"]
        self.t_avg = inputs["t_avg"]
        self.t_ramp = inputs["t_ramp"]
        self.ramp_done = self.t_ramp <= 0
        if not self.ramp_done:
            for a in (self.lbm.Fx, self.lbm.Fy, self.lbm.Fz):
                a *= 0
        self._sum = {k: np.zeros(self.lbm.N, np.float32) for k in ("ux", "uy", "uz", "rho", "sp", "uxx")}
        self.navg = 0

    @property
    def sum(self):
        if self.rho_sum == "reference":
            return self._sum
        out = dict(self._sum)
        out["rho"] = (np.float64(self.navg) + self._sum["rho"].astype(np.float64)).astype(np.float32)
        return out

    def advance(self, nsteps):
        L = self.lbm
        k = 0
        while k < nsteps:
            m = min(self.cadence, nsteps - k)
            if not self.ramp_done:
                f = min(self.t / self.t_ramp, 1.0)
                f = np.float32(f * f * (3 - 2 * f))
                L.Fx = (self.F0[0] * f).astype(L.dt)
                L.Fy = (self.F0[1] * f).astype(L.dt)
                L.Fz = (self.F0[2] * f).astype(L.dt)
                self.ramp_done = f >= 1.0
            if self.bedwall is not None:
                L.update_uref(self.bedwall)
            L.step(m)
            k += m
            self.t += m * self.dt
            if self.t >= self.t_avg:
                ux, uy, uz, rho = (a.astype(np.float32) for a in (L.ux, L.uy, L.uz, L.rho))
                s = self._sum
                s["ux"] += ux
                s["uy"] += uy
                s["uz"] += uz
                s["rho"] += rho if self.rho_sum == "reference" else rho - np.float32(1)
                s["sp"] += np.sqrt(ux * ux + uy * uy + uz * uz)
                s["uxx"] += ux * ux
                self.navg += 1
