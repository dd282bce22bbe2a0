"""The recommended boat tests and their sequential runner.

``BOAT_TESTS`` lists the same six test definitions as the reference
(floodsim/boat_tests.py:9-60): boat type, rows x boats per row, optional row
spacing, water level ('low'/'mid'/'high') and per-boat overrides. ``run_test``
runs every variant plus one no-boat base per water level and compares them
(floodsim/boat_tests.py:73-113), one GPU solve after another.
"""
from __future__ import annotations

from dataclasses import replace

from .reach3d import Fleet, ReachConfig, Run3D, boat_type, build_reach, compare, profiles

BOAT_TESTS = [
    dict(key="B1", title="1) Boat type: which pushes water best",
         objective="four boat types, three boats each, same water level",
         variants=[dict(label="long-tail x3", boat="longtail", rows=1, per_row=3),
                   dict(label="outboard x3", boat="outboard", rows=1, per_row=3),
                   dict(label="tug x3", boat="tug", rows=1, per_row=3),
                   dict(label="pontoon pusher x3", boat="pontoon", rows=1, per_row=3)]),
    dict(key="B2", title="2) Number of boats: does the effect add up",
         objective="one to six tugs",
         variants=[dict(label="tug x1", boat="tug", rows=1, per_row=1),
                   dict(label="tug x2", boat="tug", rows=1, per_row=2),
                   dict(label="tug x3", boat="tug", rows=1, per_row=3),
                   dict(label="tug 2 rows x3", boat="tug", rows=2, per_row=3, row_spacing=40)]),
    dict(key="B3", title="3) Water level: low, middle, high",
         objective="four long-tail boats at three levels",
         variants=[dict(label="low water", boat="longtail", rows=1, per_row=4, level="low"),
                   dict(label="middle water", boat="longtail", rows=1, per_row=4, level="mid"),
                   dict(label="high water", boat="longtail", rows=1, per_row=4, level="high")]),
    dict(key="B4", title="4) Shaft angle: how far to tilt the propeller down",
         objective="four long-tail boats at 0, 12 and 25 degrees",
         variants=[dict(label="0 deg", boat="longtail", rows=1, per_row=4, over=dict(tilt_deg=0)),
                   dict(label="12 deg", boat="longtail", rows=1, per_row=4, over=dict(tilt_deg=12)),
                   dict(label="25 deg", boat="longtail", rows=1, per_row=4, over=dict(tilt_deg=25))]),
    dict(key="B5", title="5) Arrangement: across the river or in line",
         objective="six long-tail boats as 1x6, 2x3, 3x2 (rows 25 m apart)",
         variants=[dict(label="1 row x 6", boat="longtail", rows=1, per_row=6),
                   dict(label="2 rows x 3", boat="longtail", rows=2, per_row=3, row_spacing=25),
                   dict(label="3 rows x 2", boat="longtail", rows=3, per_row=2, row_spacing=25)]),
    dict(key="B6", title="6) Water jets and curtains against propellers at equal power",
         objective="three pontoons across the river, 300 kW in total in every case",
         variants=[dict(label="large propeller 100 kW x3", boat="pontoon", rows=1, per_row=3,
                        over=dict(power_kW=100)),
                   dict(label="100 bar jets straight x3", boat="jet_hp", rows=1, per_row=3),
                   dict(label="100 bar jets spread +-30 deg x3", boat="jet_hp", rows=1, per_row=3,
                        over=dict(spread_deg=30)),
                   dict(label="10 bar jets x3", boat="jet_mp", rows=1, per_row=3),
                   dict(label="3 m/s water curtain x3", boat="curtain", rows=1, per_row=3)]),
]


def pick_level(levels, which, Hb):
    """'low' = lowest level, 'high' = highest, 'mid' = middle of the in-bank
    levels (all levels if none is in bank); anything else is returned as is."""
    ordered = sorted(levels)
    in_bank = [h for h in ordered if h <= Hb] or ordered
    if which == "low":
        return ordered[0]
    if which == "high":
        return ordered[-1]
    return in_bank[len(in_bank) // 2] if which == "mid" else which


def variant_fleet(v, x_first=40.0, boat_edits=None):
    """The Fleet of one test variant (boat preset, edits, overrides)."""
    b = boat_type(v["boat"])
    if boat_edits and v["boat"] in boat_edits:
        b = replace(b, **boat_edits[v["boat"]])
    if "over" in v:
        b = replace(b, **v["over"])
    return Fleet(b, rows=v["rows"], per_row=v["per_row"], row_spacing=v.get("row_spacing", 30.0), x_first=x_first)


def run_test(test, river, sec, levels, h_sel, cfg: ReachConfig, be, flow_fn, x_first=40.0,
             progress=None, cancel=None, on_result=None, boat_edits=None, storage="f32"):
    """Run every variant of a test and its no-boat base per level; returns
    the comparison dicts with the profiles attached. ``flow_fn(h)`` gives
    (Q, S) for a level; ``river`` is accepted for signature compatibility."""
    base_cache = {}
    results = []
    n = len(test["variants"])
    for k, v in enumerate(test["variants"]):
        if cancel and cancel():
            break
        h = pick_level(levels, v["level"], sec.Hb) if "level" in v else h_sel
        Q, S = flow_fn(h)
        fleet = variant_fleet(v, x_first, boat_edits)

        def prog(f, k=k, part=0):
            if progress:
                progress((k + 0.5 * part + 0.5 * f) / n, v["label"])

        if h not in base_cache:
            rb = build_reach(sec, h, Q, S, replace(fleet, enabled=False), cfg)
            run_b = Run3D(rb, be, storage=storage)
            run_b.run_to_end(progress=lambda f: prog(f, part=0), cancel=cancel)
            base_cache[h] = profiles(run_b)
            run_b.close()
        rr = build_reach(sec, h, Q, S, fleet, cfg)
        run = Run3D(rr, be, storage=storage)
        run.run_to_end(progress=lambda f: prog(f, part=1), cancel=cancel)
        res = profiles(run)
        c = compare(res, base_cache[h], rr, cfg)
        c["label"] = v["label"]
        c["profiles"] = res
        c["base_profiles"] = base_cache[h]
        c["notes"] = rr.notes
        results.append(c)
        if on_result:
            on_result(c, run)
        run.close()
    return results
