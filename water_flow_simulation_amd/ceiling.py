"""Streaming ceilings of the device for the lattice step's access pattern.

    python -m water_flow_simulation_amd.ceiling [--box 1536,522,30] [--stagger 0,64,4096] [--repeat 3]

prints the plain copy bandwidth (wfs_bandwidth_probe) and, for 4-byte and
2-byte populations and each stagger, the rate of the step's 19-population
pull read and 19-population write without its arithmetic
(wfs_stream_probe): GB/s of algorithmic bytes, the best of --repeat runs of
20 launches each, and every run. The box is the padded lattice (rows a
multiple of 64 cells); the default is the 1500x522x30 trapezoid reach's.
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
    a = ap.parse_args()
    box = tuple(int(v) for v in a.box.split(","))
    print(f"copy (float4, 2 GiB, 20 launches): {_native.bandwidth_probe(2 << 30, 20):.1f} GB/s", flush=True)
    for r in sweep(box, tuple(int(v) for v in a.stagger.split(",")), repeat=a.repeat):
        print(f"19 streams, {r['elem_bytes']}-byte populations, stagger {r['stagger_bytes']:8d} B: "
              f"{r['best_gbs']:7.1f} GB/s (runs {', '.join(f'{x:.0f}' for x in r['runs_gbs'])})", flush=True)


if __name__ == "__main__":
    main()
