"""Many reaches stepped together on one GPU (the boat-test sweep).

``BatchRun`` takes built reaches, sets each up as a normal ``Run3D`` (solver
upload, bed/bank mask, initial uref, ramp state), then hands all of them to
one libwfs batch: every step is a single launch over all cases, and the run
driver's 20-step cadence (ramp, uref, time-mean sums; floodsim/reach3d.py
:424-453) runs on the device from per-case schedules, replayed from a HIP
graph. Results are read per case with the ``Run3D`` readouts (``field``,
``profiles``). This replaces the one-case-after-another loop of
``run_test`` (floodsim/boat_tests.py:79-110) for cases that run the same
number of steps.
"""
from __future__ import annotations

import ctypes
import time

from . import _native as nat
from .reach3d import Run3D


class BatchRun:
    def __init__(self, reaches, storage="f32", block_x=0, device=0, be=None):
        self.runs = [Run3D(r, be, storage=storage, block_x=block_x, device=device) for r in reaches]
        n = len(self.runs)
        handles = (ctypes.c_void_p * n)(*[run.lbm.handle.value for run in self.runs])
        self._h = ctypes.c_void_p()
        nat.check(nat.lib().wfs_batch_create(handles, n, ctypes.byref(self._h)), "wfs_batch_create")
        for k, run in enumerate(self.runs):
            nat.check(nat.lib().wfs_batch_set_schedule(self._h, k, run.r.dt, run.t_ramp, run.t_avg, run.t))
        self.steps_wall = 0.0
        self.nstep = 0

    def advance(self, nsteps, use_graph=True):
        """Advance every case by nsteps (asynchronous until synchronize())."""
        t0 = time.time()
        nat.check(nat.lib().wfs_batch_run(self._h, int(nsteps), 1 if use_graph else 0), "wfs_batch_run")
        self.nstep += int(nsteps)
        self.synchronize()
        self.steps_wall += time.time() - t0

    def synchronize(self):
        nat.check(nat.lib().wfs_batch_synchronize(self._h))
        t, navg = ctypes.c_double(0.0), ctypes.c_int32(0)
        for k, run in enumerate(self.runs):
            nat.check(nat.lib().wfs_batch_time(self._h, k, ctypes.byref(t), ctypes.byref(navg)))
            run.t = t.value
            run.navg = navg.value
            run.lbm.nstep = self.nstep
            run._ramp_done = run.t_ramp <= 0 or run.t >= run.t_ramp

    def mlups(self):
        cells = sum(run.r.cells for run in self.runs)
        return self.nstep * cells / max(self.steps_wall, 1e-9) / 1e6

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            nat.lib().wfs_batch_destroy(self._h)
            self._h = None
        for run in self.runs:
            run.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
