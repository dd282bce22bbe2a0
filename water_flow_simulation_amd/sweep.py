"""Seeded boat-test ensembles, run one case after another or as one batch.

``ensemble(n, seed)`` draws reaches with bed widths 20-60 m, depths 1.5-5 m
and lengths 100-200 m (trapezoidal, 1:2 banks, Manning 0.03, discharge from
the Manning rule on the E3 slope) at dx = 0.2 m, each carrying 1-4 boats of
one of the seven presets, side by side or in line, with the shaft angle
drawn from 0-20 degrees. ``compare_modes`` runs the same cases for the same
number of steps with the same kernels twice: sequentially (one Run3D after
another, the loop of floodsim/boat_tests.py:79-110) and as one BatchRun,
and reports both wall times.

    python -m water_flow_simulation_amd.sweep [--cases 32] [--steps 10000] [--seed 0]
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import replace

import numpy as np

from . import reach3d as R
from .batch import BatchRun
from .geometry import Section
from .hydraulics1d import q_ref
from .workloads import E3_SLOPE

PRESETS = ("longtail", "outboard", "tug", "pontoon", "jet_hp", "jet_mp", "curtain")


def ensemble(n=32, seed=0, dx=0.2, scale=1.0):
    """n seeded reaches; ``scale`` shrinks the lengths and widths (tests)."""
    rng = np.random.default_rng(seed)
    reaches, labels = [], []
    for _ in range(n):
        width = float(rng.uniform(20, 60)) * scale
        depth = float(rng.uniform(1.5, 5.0))
        length = float(rng.uniform(100, 200)) * scale
        key = PRESETS[int(rng.integers(0, len(PRESETS)))]
        nboats = int(rng.integers(1, 5))
        inline = bool(rng.integers(0, 2))
        tilt = float(rng.uniform(0, 20))
        sec = Section(B=width, Hb=depth + 0.5, z=2.0, Bf=0.0, n=0.03)
        Q = float(q_ref(depth, 1.0, E3_SLOPE, sec))
        boat = replace(R.boat_type(key), tilt_deg=tilt)
        fleet = R.Fleet(boat, rows=nboats if inline else 1, per_row=1 if inline else nboats,
                        row_spacing=min(30.0, 0.15 * length), x_first=0.2 * length)
        cfg = R.ReachConfig(length=length, dx=dx)
        reaches.append(R.build_reach(sec, depth, Q, E3_SLOPE, fleet, cfg))
        labels.append(f"{key} x{nboats} {'in line' if inline else 'side by side'}, {width:.0f} m x {depth:.1f} m x "
                      f"{length:.0f} m")
    return reaches, labels


def run_sequential(reaches, steps, storage="f32"):
    """The cases one after another. Returns the runs, the wall time with the
    solvers' set-up and the wall time of the stepping alone."""
    t0 = time.perf_counter()
    runs = []
    step_s = 0.0
    for r in reaches:
        run = R.Run3D(r, storage=storage)
        t1 = time.perf_counter()
        run.advance(steps)
        step_s += time.perf_counter() - t1
        runs.append(run)
    return runs, time.perf_counter() - t0, step_s


def run_batched(reaches, steps, storage="f32", use_graph=True, streaming="ab"):
    """All cases as one batch (streaming: "ab" two population buffers per
    case, "inplace" one). Returns the batch, the wall time with the set-up
    and the wall time of the stepping alone."""
    t0 = time.perf_counter()
    b = BatchRun(reaches, storage=storage, streaming=streaming)
    t1 = time.perf_counter()
    b.advance(steps, use_graph=use_graph)
    return b, time.perf_counter() - t0, time.perf_counter() - t1


def compare_modes(n=32, steps=10000, seed=0, storage="f32"):
    tb = time.perf_counter()
    reaches, labels = ensemble(n, seed)
    build_s = time.perf_counter() - tb
    cells = sum(r.cells for r in reaches)
    runs, seq_s, seq_run_s = run_sequential(reaches, steps, storage)
    for run in runs:
        run.close()
    b, bat_s, bat_run_s = run_batched(reaches, steps, storage)
    b.close()
    return {"cases": n, "seed": seed, "steps": steps, "storage": storage, "cells_total": cells,
            "cells_min": min(r.cells for r in reaches), "cells_max": max(r.cells for r in reaches),
            "build_s": round(build_s, 2), "sequential_s": round(seq_s, 3), "batched_s": round(bat_s, 3),
            "sequential_run_s": round(seq_run_s, 3), "batched_run_s": round(bat_run_s, 3),
            "speedup": round(seq_s / bat_s, 3), "speedup_stepping": round(seq_run_s / bat_run_s, 3),
            "sequential_mlups": round(steps * cells / seq_s / 1e6, 1),
            "batched_mlups": round(steps * cells / bat_s / 1e6, 1), "labels": labels}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--cases", type=int, default=32)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--storage", default="f32")
    a = ap.parse_args()
    print(json.dumps(compare_modes(a.cases, a.steps, a.seed, a.storage)), flush=True)


if __name__ == "__main__":
    main()
