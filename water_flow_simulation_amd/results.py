"""Sweep outputs in the reference's on-disk formats.

- ``write_results_csv``: the boat-test table (floodsim/gui.py:1223-1231):
  columns in that order, numbers as %.5g, UTF-8 with a byte-order mark.
- ``write_cases_txt``: ``caseN: label`` lines (floodsim/__main__.py:77-79).
- ``export_case``: one case of the web viewer's cache
  (floodsim/web3d.py:44-86). The time-mean velocity is averaged over 2 x 2
  horizontal blocks of fluid cells and stored as int8 scaled by umax / 126,
  with -128 for blocks without fluid; the bed shear is the 2 x 2 block maximum
  in 0.01 N/m^2 as uint16; bed level, hulls, disks and a level/peak-speed
  profile complete the record. Arrays travel as base64.
- ``case_tag`` / ``cache_path``: the cache file names
  (floodsim/web3d.py:186-196).
- ``save_base_profiles``: the no-boat profiles, numpy .npz with the
  reference's dtypes (floodsim/web3d.py:236-242).
- ``write_vtk``: the velocity field as VTK legacy binary structured points
  for ParaView (floodsim/gui.py:1234-1248): velocity vectors, speed and the
  solid mask, big-endian float32, x fastest.

The records read the same as the reference's, so results of this build and
the ones in the reference's results/ folder can be set side by side.
"""
from __future__ import annotations

import base64
import csv
import json
import os

import numpy as np

STRIDE = 2   # horizontal pooling of the velocity field in a web-viewer record

RESULT_COLUMNS = ["label", "h", "Q", "V", "n", "T_kN", "P_kW", "T_per_kW", "F_net_kN", "d_eta_cm", "jet_reach",
                  "reach_capped", "us_gain", "us_gain_pct", "umax", "tau_max", "tau_base", "tau_area",
                  "tau_area_base", "dQ_pct", "dQ_net_pct", "dQ_per_100kW", "drop_cm_per_100kW", "kW_per_cm"]

# dtypes of the base-profile arrays as the reference stores them
BASE_PROFILE_DTYPES = {"x": np.float64, "eta": np.float64, "us": np.float64, "umax": np.float32,
                       "q": np.float32, "tau_bed": np.float32, "tau_max_x": np.float32, "mom": np.float32}


def write_results_csv(path, results):
    """One row per case result (compare() dicts), the reference's columns.
    Results that carry FP64 errors (run_test(fp64_errors=True)) add one
    column per error after the reference's, named err_fp64_<quantity>."""
    errs = sorted({k for c in results for k in c.get("err_fp64", {})}, key=_err_order)
    with open(path, "w", newline="", encoding="utf-8-sig") as f:
        w = csv.writer(f)
        w.writerow(RESULT_COLUMNS + [f"err_fp64_{k}" for k in errs])
        for c in results:
            row = [c[k] if isinstance(c[k], str) else f"{c[k]:.5g}" for k in RESULT_COLUMNS]
            e = c.get("err_fp64", {})
            w.writerow(row + [f"{e[k]:.5g}" if k in e else "" for k in errs])


def _err_order(k):
    from .boat_tests import FP64_ERROR_KEYS
    return (FP64_ERROR_KEYS.index(k), k) if k in FP64_ERROR_KEYS else (len(FP64_ERROR_KEYS), k)


def write_cases_txt(path, results):
    with open(path, "w", encoding="utf-8") as f:
        for i, c in enumerate(results, 1):
            f.write(f"case{i}: {c['label']}\n")


def _b64(a: np.ndarray) -> str:
    return base64.b64encode(np.ascontiguousarray(a).tobytes()).decode("ascii")


def pool(a, fl, s=STRIDE):
    """Mean over the fluid cells of s x s horizontal blocks. a: (nz, ny, nx) or
    (nz, ny, nx, k), fl: (nz, ny, nx) fluid weights. Returns the block means,
    (nz, ny // s, nx // s[, k]), and the mask of blocks holding fluid. Rows and
    columns past the last whole block are dropped."""
    nz, ny, nx = fl.shape
    ny2, nx2 = ny // s, nx // s
    w = fl[:, :ny2 * s, :nx2 * s].reshape(nz, ny2, s, nx2, s)
    n = w.sum((2, 4))
    tail = a.shape[3:]
    blocks = a[:, :ny2 * s, :nx2 * s].reshape((nz, ny2, s, nx2, s) + tail)
    tot = (blocks * w.reshape(w.shape + (1,) * len(tail))).sum((2, 4))
    nn = n.reshape(n.shape + (1,) * len(tail))
    return np.where(nn > 0, tot / np.maximum(nn, 1), 0.0), n > 0


def _quantized_velocity(mean_u, fluid):
    """int8 velocity, scale = largest block-mean component / 126, -128 where a
    block holds no fluid; returns the bytes' array, the scale and the speed."""
    peak = float(np.abs(mean_u).max()) or 1.0
    scale = peak / 126
    q = np.clip(np.round(mean_u / scale), -126, 126).astype(np.int8)
    q[~fluid] = -128
    return q, scale, np.sqrt((mean_u * mean_u).sum(-1))


def _bed_levels(r, ny2):
    """Per 2-row block: the lowest bed level (height of the lowest fluid cell,
    or the water depth for a dry column) of the inlet section."""
    wet = ~r.solid[:, :, 0]
    zb = np.where(wet.any(0), wet.argmax(0) * r.dx, r.h)
    return zb[:ny2 * STRIDE].reshape(ny2, STRIDE).min(1)


def _bed_shear_blocks(tau_bed, ny2, nx2):
    """Block maximum of the bed shear, its 99.5th percentile (at least 1) and
    median over wetted blocks, and the uint16 code in 0.01 N/m^2."""
    t = tau_bed[:ny2 * STRIDE, :nx2 * STRIDE].reshape(ny2, STRIDE, nx2, STRIDE).max((1, 3))
    wet = t > 0
    top = float(max(np.percentile(t[wet], 99.5) if wet.any() else 1.0, 1.0))
    med = float(np.median(t[wet])) if wet.any() else 1.0
    return np.clip(np.round(t * 100), 0, 65535).astype(np.uint16), top, med


def export_case(run, label: str, c: dict | None, res: dict) -> dict:
    """The web-viewer record of a finished run: ``run`` gives the reach
    (``run.r``) and the time-mean fields (``run.field(name, mean=True)``),
    ``res`` its profiles(), ``c`` its compare() dict (None for the base)."""
    r = run.r
    u = np.stack([run.field(k, mean=True) for k in ("ux", "uy", "uz")], -1)
    mean_u, fluid = pool(u, (~r.solid).astype(np.float32))
    q, scale, sp = _quantized_velocity(mean_u.astype(np.float32), fluid)
    nz, ny2, nx2 = fluid.shape
    tq, tau_top, tau_med = _bed_shear_blocks(res["tau_bed"], ny2, nx2)
    every = max(len(res["x"]) // 300, 1)
    fleet = r.fleet
    rec = {"label": label, "nx": int(nx2), "ny": int(ny2), "nz": int(nz), "dxs": float(r.dx * STRIDE),
           "dz": float(r.dx), "y0": float(r.y[0] - 0.5 * r.dx), "L": float(nx2 * r.dx * STRIDE),
           "W": float(ny2 * r.dx * STRIDE), "h": float(r.h), "V": float(r.V), "Q": float(r.Q), "S": float(r.S),
           "scale": float(scale), "u": _b64(q), "sp99": float(np.percentile(sp[fluid], 99.7)),
           "tau": _b64(tq), "tau_top": tau_top, "tau_med": tau_med,
           "zb": [round(float(v), 3) for v in _bed_levels(r, ny2)],
           "hulls": [[round(float(v), 3) for v in hb] for hb in r.hulls],
           "props": [{"x": p["x"], "y": p["y"], "z": p["z"], "D": p["D"], "tilt": p["tilt"],
                      "yaw": p.get("yaw", 0.0)} for p in r.props],
           "boat": fleet.boat.name if fleet.enabled else "",
           "jet": bool(fleet.boat.is_jet) if fleet.enabled else False,
           "prof": {"x": [round(float(v), 2) for v in res["x"][::every]],
                    "eta": [round(float(v) * 100, 3) for v in res["eta"][::every]],
                    "umax": [round(float(v), 3) for v in res["umax"][::every]]}}
    if c:
        stats = {}
        for k, v in c.items():
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                continue
            stats[k] = None if v != v else round(float(v), 4)
        rec["stats"] = stats
    return rec


def decode_velocity(rec: dict) -> tuple[np.ndarray, np.ndarray]:
    """(nz, ny, nx, 3) block-mean velocity (m/s) of a record and its fluid-block mask."""
    q = np.frombuffer(base64.b64decode(rec["u"]), np.int8).reshape(rec["nz"], rec["ny"], rec["nx"], 3)
    fluid = q[..., 0] != -128
    return np.where(fluid[..., None], q.astype(np.float64) * rec["scale"], 0.0), fluid


def decode_tau(rec: dict) -> np.ndarray:
    """(ny, nx) block-max bed shear (N/m^2) of a record."""
    return np.frombuffer(base64.b64decode(rec["tau"]), np.uint16).reshape(rec["ny"], rec["nx"]) / 100.0


def case_tag(exp: int, sec, h: float, max_cells: int, length: float, wall_model: bool = True,
             slope: float | None = None, x_first: float | None = None) -> str:
    """The cache tag of a river, level, resolution and wall model."""
    shape = f"B{sec.B:g}_Hb{sec.Hb:g}_z{sec.z:g}_Bf{sec.Bf:g}_n{sec.n:g}" + ("" if not slope else f"_S{slope:g}")
    if x_first is not None:
        shape += f"_x{x_first:g}"
    return f"e{exp}_{shape}_h{h:g}_n{max_cells}_L{length:g}{'' if wall_model else '_nowm'}"


def cache_path(cache_dir: str, key: str, tag: str) -> str:
    """``key`` is "base" or "t<test>c<case>"."""
    return os.path.join(cache_dir, f"{key}_{tag}.json")


def save_case(cache_dir: str, key: str, tag: str, rec: dict) -> str:
    os.makedirs(cache_dir, exist_ok=True)
    p = cache_path(cache_dir, key, tag)
    with open(p, "w", encoding="utf-8") as f:
        json.dump(rec, f, ensure_ascii=False)
    return p


def save_base_profiles(path: str, res: dict) -> None:
    """The no-boat profiles of a level (numpy .npz, the reference's dtypes)."""
    np.savez_compressed(path, **{k: np.asarray(res[k], dtype=BASE_PROFILE_DTYPES.get(k)) for k in res})


def load_base_profiles(path: str) -> dict:
    with np.load(path) as d:
        return {k: d[k] for k in d.files}


def write_sweep(out_dir: str, results) -> list[str]:
    """The files a boat-test sweep leaves in its output folder:
    boat_results.csv and cases.txt."""
    os.makedirs(out_dir, exist_ok=True)
    paths = [os.path.join(out_dir, "boat_results.csv"), os.path.join(out_dir, "cases.txt")]
    write_results_csv(paths[0], results)
    write_cases_txt(paths[1], results)
    return paths


def vtk_header(r) -> str:
    """Header of the VTK file of reach r, up to the point data."""
    return (f"# vtk DataFile Version 3.0\nFlood Sim 3D\nBINARY\nDATASET STRUCTURED_POINTS\n"
            f"DIMENSIONS {r.nx} {r.ny} {r.nz}\nORIGIN {0.5 * r.dx} {r.y[0]} {0.5 * r.dx}\n"
            f"SPACING {r.dx} {r.dx} {r.dx}\nPOINT_DATA {r.cells}\n")


def write_vtk(path: str, run, mean: bool = False) -> None:
    """The run's velocity (instantaneous or time mean) as a VTK file."""
    r = run.r
    u = np.stack([run.field(k, mean) for k in ("ux", "uy", "uz")], -1)
    speed = np.sqrt((u * u).sum(-1))
    with open(path, "wb") as f:
        f.write(vtk_header(r).encode())
        f.write(b"VECTORS velocity float\n")
        f.write(u.astype(">f4").tobytes())
        f.write(b"\nSCALARS speed float 1\nLOOKUP_TABLE default\n")
        f.write(speed.astype(">f4").tobytes())
        f.write(b"\nSCALARS solid float 1\nLOOKUP_TABLE default\n")
        f.write(r.solid.astype(">f4").tobytes())

