#!/usr/bin/env python3
"""Throughput of the river-reach solver on one MI355X.

A step is one solver step of the whole reach (fused collide-stream kernel,
inlet/outlet rows) inside the run driver's 20-step cadence: every chunk of 20
steps also refreshes the wall-model reference speed (column reduction) and
adds the time-mean sums, as a run does in its averaging window. The
workload is the 1500 x 522 x 30 trapezoidal reach with one pusher tug
(23.49 M cells, 20.64 M fluid), FP32 populations, synthetic inputs built by
the case builder; all state is resident in GPU memory before timing starts.

Prints one JSON line: MLUPS (all cells, the reference's convention), the
step kernel's HBM roofline fraction from HIP events around every launch on
the solver's stream, and a CPU baseline (the OpenMP oracle in its
Numba-like precision, on a bounded sample of the same reach). Beside it:
the FP32-fast, 16-bit, learned-closure, per-step column speed and in-place
(one population buffer) variants on the same reach, the batched sweep
against the same cases run one after another, and the meandering 1 km reach
(1.56 G-cell box, 0.67 G fluid cells) on the tile-sparse layout, FP16
storage, with and without the learned closure (``--meander`` also takes
``+inplace`` variants).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "MLUPS & % HBM roofline, plain FP32 vs FP16-store/batched/learned SGS"
PEAK_HBM_GBS = 8000.0                       # MI355X HBM3E
# algorithmic bytes per fluid-cell update: 19 populations read + 19 written + 1 flag byte
B_ALG = {"f32": 19 * 4 * 2 + 1, "f32fast": 19 * 4 * 2 + 1, "f16dev": 19 * 2 * 2 + 1, "bf16dev": 19 * 2 * 2 + 1,
         "f64": 19 * 8 * 2 + 1}
PUBLISHED_MLUPS = 900.0                     # reference CUDA path, RTX 5070 Laptop (top of 550-900)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_step_f32_trapezoid_tug_traffic.txt")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_case(name):
    from water_flow_simulation_amd import workloads
    return getattr(workloads, name)()


def solver_args(reach):
    nu_lb = max(1e-6 * reach.dt / reach.dx ** 2, 2e-4)
    return (reach.solid, reach.Fx, reach.Fy, reach.Fz, (reach.uin_phys / reach.vel).astype(np.float32),
            0.5 + 3 * nu_lb, reach.cfg.cs, reach.cf)


def chunks(lbm, nsteps, cadence=20, per_step_column_u=False):
    k = 0
    while k < nsteps:
        m = min(cadence, nsteps - k)
        if not per_step_column_u:
            lbm.update_uref()
        lbm.step(m)
        lbm.accumulate()
        k += m


def closure_weights():
    """The committed trained closure, or an untrained network (timing does not depend on the weights)."""
    from water_flow_simulation_amd import closure as CL
    return CL.load_params()[0] if os.path.exists(CL.WEIGHTS) else CL.init_params(0)


# Untimed steps before each variant line at least: a compute-bound step (the
# learned closure) runs about 12 % slow for its first 20-odd steps after the
# GPU sat idle during the solver set-up; the HBM-bound plain step does not.
VARIANT_WARMUP = 50


def time_gpu(reach, storage, steps, warmup, block_x, learned=False, column_u=False, inplace=False, pop_stagger=0):
    """column_u: the wall model's column speed refreshed every step inside the
    step kernel instead of once per 20-step chunk; inplace: one population
    buffer streamed in place; pop_stagger: bytes between the population
    direction arrays."""
    from water_flow_simulation_amd.lbm3d import LBM3D
    lbm = LBM3D(None, *solver_args(reach), storage=storage, block_x=block_x, pop_stagger=pop_stagger)
    if inplace:
        lbm.set_inplace(True)
    lbm.set_bedwall(reach.bedwall)
    if column_u:
        lbm.update_uref()
        lbm.set_column_u(True)
    if learned:
        from water_flow_simulation_amd import closure as CL
        geo, hcol = CL.geometry(reach.solid)
        uin = solver_args(reach)[4]
        lbm.set_closure("mlp", closure_weights(), geo, hcol, float(uin[uin > 0].mean()))
    chunks(lbm, warmup, per_step_column_u=column_u)
    lbm.synchronize()
    lbm.check_finite()
    lbm.reset_sums()
    lbm.timing(True)
    lbm.synchronize()
    t0 = time.perf_counter()
    chunks(lbm, steps, per_step_column_u=column_u)
    lbm.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms, launches = lbm.timing_read()
    lbm.timing(False)
    lbm.check_finite()
    mass = lbm.total("rho")
    lbm.close()
    return dict(wall=wall, kernel_ms=kernel_ms, launches=launches, mass=mass)


def cpu_baseline(reach, target_s=15.0):
    """The oracle's Numba-like instantiation (float storage, double arithmetic,
    OpenMP over cells) on a few steps of the same reach."""
    from oracle import oracle as O
    args = solver_args(reach)
    orc = O.OracleLBM(*args, precision="mixed")
    t0 = time.perf_counter()
    orc.step(1)
    one = time.perf_counter() - t0
    n = int(min(max(target_s / max(one, 1e-3), 2), 40))
    t0 = time.perf_counter()
    orc.step(n)
    dt = time.perf_counter() - t0
    # the reference's own CPU-runnable case, whole: 1000 steps of the 128x32x16 channel with one boat
    small = build_case("channel128")
    orc1 = O.OracleLBM(*solver_args(small), precision="mixed")
    t0 = time.perf_counter()
    orc1.step(1000)
    dt1 = time.perf_counter() - t0
    return dict(value=n * reach.cells / dt / 1e6, unit="MLUPS", cores=O.num_threads(), kind="port",
                sample=f"{n} steps of the full {reach.nx}x{reach.ny}x{reach.nz} reach (all cells counted)",
                cpu=cpu_model(),
                channel128={"value": round(1000 * small.cells / dt1 / 1e6, 1), "unit": "MLUPS", "steps": 1000,
                            "wall_s": round(dt1, 3),
                            "workload": "128x32x16 channel at dx 0.25 m, one boat (the reference's CPU case)"})


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) and its logical CPU count."""
    name = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name}, {os.cpu_count()} logical CPUs visible"


def time_batched(cases, steps, seed=0, scale=1.0):
    """The boat-test sweep: the same seeded ensemble (at scale 1 the
    32-reach boat-test ensemble of widths 20-60 m, depths 1.5-5 m, lengths
    100-200 m at dx 0.2 m) run one case after another and as one batch (one
    step launch over all cases, device-side cadence, HIP-graph replay)."""
    from water_flow_simulation_amd import sweep
    reaches, _ = sweep.ensemble(cases, seed, scale=scale)
    cells = sum(r.cells for r in reaches)
    runs, _, _ = sweep.run_sequential(reaches, 40)       # warm-up (allocation, first launches)
    for run in runs:
        run.close()
    runs, seq_all_s, seq_s = sweep.run_sequential(reaches, steps)
    for run in runs:
        run.close()
    b, bat_all_s, bat_s = sweep.run_batched(reaches, steps)
    b.close()
    # the same batch with every case streamed in place (one population buffer)
    b, _, aa_s = sweep.run_batched(reaches, steps, streaming="inplace")
    b.close()
    pop_gb = cells * 19 * 4 / 1e9
    # speedup: the stepping alone on both sides; speedup_with_setup: the
    # whole sweep, the solvers' set-up included on both sides
    return {"cases": cases, "seed": seed, "cells_total": cells, "steps": steps,
            "sequential_mlups": round(steps * cells / seq_s / 1e6, 1),
            "batched_mlups": round(steps * cells / bat_s / 1e6, 1), "speedup": round(seq_s / bat_s, 3),
            "speedup_with_setup": round(seq_all_s / bat_all_s, 3),
            "inplace": {"batched_mlups": round(steps * cells / aa_s / 1e6, 1), "speedup": round(seq_s / aa_s, 3),
                        "population_gb": round(pop_gb, 2), "population_gb_two_buffers": round(2 * pop_gb, 2)},
            "ensemble": f"seeded boat-test reaches at dx 0.2 m, lengths and widths x{scale}"}


def time_meander(variants, steps, scale=1.0):
    """The meandering-reach workload on the tile-sparse layout (1 km sine reach,
    80 m top width, 4 m deep, dx 0.08 m, 8 pushers): per variant the step
    kernel's average time, MLUPS over fluid and stored cells and the HBM
    fraction of the step's algorithmic bytes per fluid cell."""
    from water_flow_simulation_amd import meander
    out = {}
    for name in variants:
        st, opts = name.split("+")[0], name.split("+")[1:]
        learned = "mlp" in opts
        log(f"meander {name}")
        r = meander.run(scale, steps, warmup=VARIANT_WARMUP, storage=st, learned=learned, log=log,
                        streaming="inplace" if "inplace" in opts else "ab")
        k_s = r["kernel_ms"] / 1e3
        b_alg = B_ALG[st] + (2 if learned else 0)
        out[name] = {"kernel_ms_avg": r["kernel_ms"], "mlups_fluid": r["mlups_fluid"],
                     "mlups_stored": r["mlups_stored"],
                     "roofline_frac": round(b_alg * r["fluid_cells"] / k_s / 1e9 / PEAK_HBM_GBS, 4),
                     "device_gb": r["device_gb"], "setup_s": round(r["build_s"] + r["upload_s"], 1)}
        box = {k: r[k] for k in ("box", "box_cells", "stored_cells", "fluid_cells", "steps")}
    out["case"] = box
    return out


def traffic_per_launch():
    """HBM bytes per step-kernel launch from the committed rocprofv3 counter summary, if any."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    m = re.search(r"hbm_bytes_per_launch\s*[=:]\s*([0-9.eE+]+)", open(TRAFFIC_FILE).read())
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="trapezoid_tug")
    ap.add_argument("--storage", default="f32", choices=sorted(B_ALG))
    ap.add_argument("--block-x", default="auto",
                    help="work-group width of the step (64, 128, 192, 256; auto: tuned per variant on this GPU)")
    ap.add_argument("--pop-stagger", default="auto",
                    help="bytes between the population direction arrays of the reach's solvers (0: packed; "
                         "auto: tuned per storage format on this GPU before timing)")
    ap.add_argument("--variants", default="f32fast,f16dev,bf16dev,f32+mlp,f16dev+mlp,f16dev+colu,f32+inplace,"
                                          "f16dev+inplace",
                    help="extra variants timed (comma list; storage format, '+mlp' for the learned closure, "
                         "'+colu' for the per-step column speed, '+inplace' for one population buffer)")
    ap.add_argument("--batch-cases", type=int, default=32, help="cases of the batched-sweep line (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=400)
    ap.add_argument("--batch-scale", type=float, default=1.0, help="length and width factor of the ensemble")
    ap.add_argument("--meander", default="f16dev+inplace,f16dev+mlp+inplace",
                    help="variants timed on the meandering reach, tile-sparse layout ('' to skip); the default is "
                         "the configuration as specified: FP16 streamed in place, without and with the learned "
                         "closure")
    ap.add_argument("--meander-steps", type=int, default=200)
    ap.add_argument("--meander-scale", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path in this build "
              "(x-slab decomposition is design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from water_flow_simulation_amd import _native
    copy_gbs = _native.bandwidth_probe(2 << 30, 20)
    log(f"device copy bandwidth {copy_gbs:.0f} GB/s")
    log(f"building workload {a.workload}")
    reach = build_case(a.workload)
    n_fluid = int((~reach.solid).sum())
    # the step's own access pattern without its arithmetic, on the reach's padded box
    px = -(-reach.nx // 64) * 64
    ceiling = {eb: _native.stream_probe(px, reach.ny, reach.nz, eb, 0, 20) for eb in (4, 2)}
    log(f"19-stream pull ceiling {ceiling[4]:.0f} GB/s (4-byte), {ceiling[2]:.0f} GB/s (2-byte)")
    elem = {"f32": 4, "f32fast": 4, "f16dev": 2, "bf16dev": 2}

    def of_ceiling(st, gbs):
        return round(gbs / ceiling[elem[st]], 4) if st in elem else None
    log(f"{reach.nx}x{reach.ny}x{reach.nz} = {reach.cells} cells ({n_fluid} fluid); timing {a.steps} steps")
    def stagger(st):
        if a.pop_stagger != "auto":
            return int(a.pop_stagger)
        from water_flow_simulation_amd import autotune
        return autotune.pop_stagger_for(reach, st)

    def width(st, learned=False):
        if a.block_x != "auto":
            return int(a.block_x)
        from water_flow_simulation_amd import autotune
        return autotune.block_x_for(reach, st, closure_weights() if learned else None)
    r = time_gpu(reach, a.storage, a.steps, a.warmup, width(a.storage), pop_stagger=stagger(a.storage))
    log(f"timed: {r['wall']:.3f} s wall, step kernel {r['kernel_ms'] / max(r['launches'], 1):.4f} ms avg")
    mlups = a.steps * reach.cells / r["wall"] / 1e6
    k_avg_s = r["kernel_ms"] / max(r["launches"], 1) / 1e3
    achieved = B_ALG[a.storage] * n_fluid / k_avg_s / 1e9
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(r["wall"] / a.steps * 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": round(mlups / PUBLISHED_MLUPS, 2), "dtype": "fp32" if a.storage != "f64"
        else "fp64", "data": "synthetic",
        "config": {"workload": f"{a.workload}: trapezoidal reach 40 m bed, 1:2 banks, 3 m deep, 150 m at dx 0.1 m, "
                               f"{reach.nx}x{reach.ny}x{reach.nz} cells, one pusher tug (2 x 1.2 m props, 300 kW)",
                   "storage": a.storage, "cells": reach.cells, "fluid_cells": n_fluid, "cadence": 20,
                   "block_x": width(a.storage), "pop_stagger": stagger(a.storage)},
        "mlups_fluid": round(a.steps * n_fluid / r["wall"] / 1e6, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic_per_launch(),
                     "kernel": "k_step", "kernel_ms_avg": round(k_avg_s * 1e3, 5),
                     "bytes_alg_per_launch": B_ALG[a.storage] * n_fluid,
                     "copy_bw_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4),
                     "stream_ceiling_gbs": round(ceiling[elem[a.storage]], 1) if a.storage in elem else None,
                     "frac_of_stream_ceiling": of_ceiling(a.storage, achieved)},
        "mass_fp64": r["mass"],
    }
    variants = {}
    for name in [v for v in a.variants.split(",") if v and v != a.storage]:
        st, opts = name.split("+")[0], name.split("+")[1:]
        learned, colu, inplace = "mlp" in opts, "colu" in opts, "inplace" in opts
        log(f"timing variant {name}")
        rv = time_gpu(reach, st, a.steps, max(a.warmup, VARIANT_WARMUP), width(st, learned), learned, colu, inplace,
                      stagger(st))
        kv = rv["kernel_ms"] / max(rv["launches"], 1) / 1e3
        # + the 2-byte geometry code of the closure; + ux, uy written for the per-step column speed
        b_alg = B_ALG[st] + (2 if learned else 0) + (8 if colu else 0)
        variants[name] = {"mlups": round(a.steps * reach.cells / rv["wall"] / 1e6, 1),
                          "kernel_ms_avg": round(kv * 1e3, 5),
                          "roofline_frac": round(b_alg * n_fluid / kv / 1e9 / PEAK_HBM_GBS, 4),
                          "frac_of_stream_ceiling": of_ceiling(st, b_alg * n_fluid / kv / 1e9),
                          "speedup_vs_f32": round(r["wall"] / rv["wall"], 3), "pop_stagger": stagger(st),
                          "block_x": width(st, learned)}
    if variants:
        out["variants"] = variants
        out["variant_warmup"] = max(a.warmup, VARIANT_WARMUP)
    if a.batch_cases > 0:
        log(f"timing the batched sweep ({a.batch_cases} cases)")
        out["batched"] = time_batched(a.batch_cases, a.batch_steps, scale=a.batch_scale)
    if a.meander:
        out["meander"] = time_meander([v for v in a.meander.split(",") if v], a.meander_steps, a.meander_scale)
    if not a.no_cpu_baseline:
        log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(reach)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
