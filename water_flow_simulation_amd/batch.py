"""Many reaches stepped together on one GPU (the boat-test sweep).

``BatchRun`` takes built reaches, sets each up as a normal ``Run3D`` (solver
upload, bed/bank mask, initial uref, ramp state), then hands them to libwfs
batches: every step is a single launch over all cases of a batch, and the run
driver's 20-step cadence (ramp, uref, time-mean sums; floodsim/reach3d.py
:424-453) runs on the device from per-case schedules, replayed from a HIP
graph. With ``group`` cases per batch the cases are split into several
batches, each on its own stream, which the GPU runs side by side; the
autotuner picks ``group`` and the work-group width. Results are read per case
with the ``Run3D`` readouts (``field``, ``profiles``). This replaces the
one-case-after-another loop of ``run_test`` (floodsim/boat_tests.py:79-110):
``run_to_end`` gives every case a step budget equal to the steps its own
``Run3D.run_to_end`` would take (each case stops at its own end time), and
the library stops each case when its budget is spent while the others go on.
"""
from __future__ import annotations

import ctypes
import time

from . import _native as nat
from .reach3d import Run3D


class BatchRun:
    def __init__(self, reaches, storage="f32", block_x=0, device=0, be=None, group=0, pop_stagger=0,
                 streaming="ab"):
        """block_x: work-group width (0: library default, "auto": autotuned on
        the largest case); group: cases per batched launch (0: all, "auto":
        autotuned); pop_stagger: bytes between each case's population
        direction arrays (a layout choice, results unchanged; "auto": tuned on
        the largest case); streaming: "ab" (two population buffers per case)
        or "inplace" (one; half the population memory, the same results bit
        for bit)."""
        bx, ps = block_x, pop_stagger
        if block_x == "auto":
            from .autotune import block_x_for
            bx = block_x_for(max(reaches, key=lambda r: r.cells), storage)
        if pop_stagger == "auto":
            from .autotune import pop_stagger_for
            ps = pop_stagger_for(max(reaches, key=lambda r: r.cells), storage)
        self.runs = [Run3D(r, be, storage=storage, block_x=bx, device=device, pop_stagger=ps, streaming=streaming)
                     for r in reaches]
        self._hs = []
        self.group = 0
        self.steps_wall = 0.0
        self.nstep = 0          # batch steps run (each case takes at most this many)
        self.updates = 0        # cell updates over all cases (all cells counted, as MLUPS counts them)
        if group == "auto":
            from .autotune import tune_group
            tune_group(self)
        else:
            self.set_group(group or len(self.runs))

    def set_group(self, k):
        """Regroup the cases into batches of k (the last may be smaller)."""
        if self._hs:
            self.synchronize()
        self._destroy()
        n = len(self.runs)
        k = max(1, min(int(k), n))
        L = nat.lib()
        for start in range(0, n, k):
            sub = self.runs[start:start + k]
            handles = (ctypes.c_void_p * len(sub))(*[run.lbm.handle.value for run in sub])
            h = ctypes.c_void_p()
            nat.check(L.wfs_batch_create(handles, len(sub), ctypes.byref(h)), "wfs_batch_create")
            self._hs.append((h, start, len(sub)))
            for j, run in enumerate(sub):
                nat.check(L.wfs_batch_set_schedule(h, j, run.r.dt, run.t_ramp, run.t_avg, run.t))
        self.group = k

    def advance(self, nsteps, use_graph=True):
        """Advance every case by nsteps (fewer for a case whose budget runs
        out first); returns when all batches are done."""
        t0 = time.time()
        before = [run.lbm.nstep for run in self.runs]
        for h, _, _ in self._hs:          # queued on each batch's own stream: the batches overlap
            nat.check(nat.lib().wfs_batch_run(h, int(nsteps), 1 if use_graph else 0), "wfs_batch_run")
        self.nstep += int(nsteps)
        self.synchronize()
        self.steps_wall += time.time() - t0
        self.updates += sum((run.lbm.nstep - b) * run.r.cells for run, b in zip(self.runs, before))

    def _each(self):
        """(batch handle, index in that batch, index in self.runs) per case."""
        for h, start, n in self._hs:
            for j in range(n):
                yield h, j, start + j

    def set_budgets(self, steps):
        """Step budget per case from the next advance (None: no limit)."""
        if len(steps) != len(self.runs):
            raise ValueError("set_budgets: one budget per case")
        for h, j, k in self._each():
            b = steps[k]
            nat.check(nat.lib().wfs_batch_set_budget(h, j, -1 if b is None else int(b)), "wfs_batch_set_budget")

    def budgets_left(self):
        out = []
        v = ctypes.c_int64(0)
        for h, j, _ in self._each():
            nat.check(nat.lib().wfs_batch_budget(h, j, ctypes.byref(v)))
            out.append(None if v.value < 0 else int(v.value))
        return out

    def run_to_end(self, progress=None, cancel=None, chunk=400, use_graph=True):
        """Every case runs to its own end time exactly as Run3D.run_to_end
        would run it alone (floodsim/reach3d.py:455-467): the same number of
        steps in the same 20-step chunks, so the results are bit-identical,
        but all cases share each launch until their budgets are spent.
        ``progress(f)`` gets the fraction of the longest case done after
        every ``chunk`` steps; ``cancel()`` stops all cases there. Returns
        the per-case step counts."""
        from .reach3d import GPU_DUTY
        budgets = [run.steps_to_end(chunk) for run in self.runs]
        seg = self.runs[0].run_chunk(chunk) if self.runs else chunk
        duty = min(max(GPU_DUTY, 0.05), 1.0)
        self.set_budgets(budgets)
        total = max(budgets, default=0)
        done = 0
        try:
            while done < total:
                t0 = time.time()
                m = min(seg, total - done)
                self.advance(m, use_graph=use_graph)
                done += m
                if duty < 1:
                    time.sleep((time.time() - t0) * (1 / duty - 1))
                if progress:
                    progress(done / total)
                if cancel and cancel():
                    break
        finally:
            self.set_budgets([None] * len(self.runs))
        return budgets

    def synchronize(self):
        L = nat.lib()
        t, navg = ctypes.c_double(0.0), ctypes.c_int32(0)
        for h, start, n in self._hs:
            nat.check(L.wfs_batch_synchronize(h))
            for j in range(n):
                run = self.runs[start + j]
                nat.check(L.wfs_batch_time(h, j, ctypes.byref(t), ctypes.byref(navg)))
                run.t = t.value
                run.navg = navg.value
                run.lbm.nstep = nat.lib().wfs_nstep(run.lbm.handle)
                run._ramp_done = run.t_ramp <= 0 or run.t >= run.t_ramp

    def mlups(self):
        """Cell updates of all cases per second of batched stepping (millions)."""
        return self.updates / max(self.steps_wall, 1e-9) / 1e6

    def _destroy(self):
        for h, _, _ in getattr(self, "_hs", []):
            if h.value:
                nat.lib().wfs_batch_destroy(h)
        self._hs = []

    def close(self):
        self._destroy()
        for run in getattr(self, "runs", []):
            run.close()
        self.runs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
