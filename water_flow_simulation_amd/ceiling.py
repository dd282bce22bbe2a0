"""Streaming ceilings of the device for the lattice step's access pattern.

    python -m water_flow_simulation_amd.ceiling [--box 1536,522,30] [--stagger 0,64,4096] [--repeat 3]
    python -m water_flow_simulation_amd.ceiling --sizes 128x64x32,512x256x32,1024x512x32 [--iters 50]

prints the plain copy bandwidth (wfs_bandwidth_probe) and, for 4-byte and
2-byte populations and each stagger, the rate of the step's 19-population
pull read and 19-population write without its arithmetic
(wfs_stream_probe): GB/s of algorithmic bytes, the best of --repeat runs of
--iters launches each, and every run. The box is the padded lattice (rows a
multiple of 64 cells); the default is the 1500x522x30 trapezoid reach's.
--sizes sweeps boxes of different sizes instead (stagger 0), to show where
the working set of the two population buffers stops fitting the last-level
cache.
"""
from __future__ import annotations

import argparse

from . import _native


def sweep(box=(1536, 522, 30), staggers=(0,), elems=(4, 2), repeat=3, iters=20):
    px, ny, nz = box
    rows = []
    for eb in elems:
        for st in staggers:
            if st % eb:
                continue
            runs = [_native.stream_probe(px, ny, nz, eb, st, iters) for _ in range(repeat)]
            rows.append({"elem_bytes": eb, "stagger_bytes": st, "best_gbs": max(runs), "runs_gbs": runs})
    return rows


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--box", default="1536,522,30")
    ap.add_argument("--stagger", default="0")
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--sizes", default="", help="boxes pxXnyXnz, comma separated (px a multiple of 64)")
    a = ap.parse_args()
    print(f"copy (float4, 2 GiB, 20 launches): {_native.bandwidth_probe(2 << 30, 20):.1f} GB/s", flush=True)
    if a.sizes:
        for spec in a.sizes.split(","):
            box = tuple(int(v) for v in spec.lower().split("x"))
            cells = box[0] * box[1] * box[2]
            for r in sweep(box, (0,), repeat=a.repeat, iters=a.iters):
                ws = 2 * 19 * r["elem_bytes"] * cells / 1e6
                print(f"19 streams, {r['elem_bytes']}-byte populations, box {spec} ({cells / 1e6:.2f} M cells, "
                      f"two buffers {ws:.0f} MB): {r['best_gbs']:7.1f} GB/s "
                      f"(runs {', '.join(f'{x:.0f}' for x in r['runs_gbs'])})", flush=True)
        return
    box = tuple(int(v) for v in a.box.split(","))
    for r in sweep(box, tuple(int(v) for v in a.stagger.split(",")), repeat=a.repeat, iters=a.iters):
        print(f"19 streams, {r['elem_bytes']}-byte populations, stagger {r['stagger_bytes']:8d} B: "
              f"{r['best_gbs']:7.1f} GB/s (runs {', '.join(f'{x:.0f}' for x in r['runs_gbs'])})", flush=True)


if __name__ == "__main__":
    main()
