"""The recommended boat tests and their runner.

``BOAT_TESTS`` lists the same six test definitions as the reference
(floodsim/boat_tests.py:9-60): boat type, rows x boats per row, optional row
spacing, water level ('low'/'mid'/'high') and per-boat overrides.

``run_test`` gives what the reference's runner gives (floodsim/boat_tests.py
:73-113): every variant compared with a no-boat base at its water level, the
profiles of both attached. It is organised differently. The whole test is
planned first (``plan_test``: one reach per variant, one base per level the
variants use), then every reach is simulated at once as one ``BatchRun`` in
which each case stops at its own end time (``simulate``), and the readouts
come last. ``batched=False`` simulates the same plan one case after another,
the reference's order of work; both give the same numbers bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

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


@dataclass
class TestPlan:
    """The reaches of one boat test: ``cases[k]`` is (label, level, Reach,
    index of its base in ``cases`` or None for a base itself)."""
    cases: list = field(default_factory=list)

    @property
    def reaches(self) -> list:
        return [c[2] for c in self.cases]

    def variants(self):
        return [(k, c) for k, c in enumerate(self.cases) if c[3] is not None]


def plan_test(test, sec, levels, h_sel, cfg: ReachConfig, flow_fn, x_first=40.0, boat_edits=None) -> TestPlan:
    """Build every reach the test needs. A variant's level is its 'level'
    entry resolved by ``pick_level`` (else ``h_sel``); ``flow_fn(h)`` gives
    (Q, S) once per level. The base of a level is built with the first
    variant's fleet switched off, so it carries no hull, disk or force."""
    plan = TestPlan()
    flows, base_at = {}, {}
    for v in test["variants"]:
        h = pick_level(levels, v["level"], sec.Hb) if "level" in v else h_sel
        if h not in flows:
            flows[h] = flow_fn(h)
        Q, S = flows[h]
        fleet = variant_fleet(v, x_first, boat_edits)
        if h not in base_at:
            base_at[h] = len(plan.cases)
            plan.cases.append((f"no boats, h = {h:g} m", h, build_reach(sec, h, Q, S, replace(fleet, enabled=False),
                                                                        cfg), None))
        plan.cases.append((v["label"], h, build_reach(sec, h, Q, S, fleet, cfg), base_at[h]))
    return plan


def simulate(reaches, be=None, storage="f32", batched=True, progress=None, cancel=None, group=0):
    """Run every reach to its own end time (Run3D.run_to_end); returns the
    runs, open for readout. Batched, all reaches share each launch until
    their step budgets are spent; otherwise one after another. ``progress``
    gets the fraction done (of the longest case's steps when batched, of all
    cell updates otherwise)."""
    if batched:
        from .batch import BatchRun
        b = BatchRun(reaches, storage=storage, be=be, group=group)
        b.run_to_end(progress=progress, cancel=cancel)
        b._destroy()           # the cases stay open for readout
        runs, b.runs = b.runs, []
    else:
        runs = [Run3D(r, be, storage=storage) for r in reaches]
        cost = [run.steps_to_end() * run.r.cells for run in runs]
        total = max(sum(cost), 1)
        done = 0
        for run, c in zip(runs, cost):
            if cancel and cancel():
                break
            run.run_to_end(progress=(lambda f, d=done, c=c: progress((d + f * c) / total)) if progress else None,
                           cancel=cancel)
            done += c
    return runs


# The reported quantities that get an FP64 error: water-level change, surface
# velocity, bed shear, discharge and the momentum budget (compare() keys).
FP64_ERROR_KEYS = ("d_eta_cm", "us_gain", "us_gain_pct", "umax", "tau_max", "tau_area", "jet_reach", "F_net_kN",
                   "dQ_net_pct", "dQ_per_100kW", "drop_cm_per_100kW")


def fp64_error(c, c64, prof, prof64, Q):
    """|value - FP64 value| for every reported quantity of a comparison,
    and the largest differences of its profiles from the FP64 run's."""
    err = {k: abs(float(c[k]) - float(c64[k])) for k in FP64_ERROR_KEYS}
    err["eta_x_max_mm"] = float(np.abs(prof["eta"] - prof64["eta"]).max() * 1e3)
    err["us_x_max"] = float(np.abs(prof["us"] - prof64["us"]).max())
    err["q_x_max_rel"] = float(np.abs(prof["q"] - prof64["q"]).max() / Q)
    err["tau_bed_max"] = float(np.abs(prof["tau_bed"] - prof64["tau_bed"]).max())
    return err


def run_test(test, river, sec, levels, h_sel, cfg: ReachConfig, be, flow_fn, x_first=40.0,
             progress=None, cancel=None, on_result=None, boat_edits=None, storage="f32", batched=True, group=0,
             fp64_errors=False):
    """One boat test, with the reference's signature and result records: a
    list with one dict per variant, ``compare`` of the variant against the
    base at its level plus 'label', 'profiles', 'base_profiles' and 'notes'
    (floodsim/boat_tests.py:73-113). ``progress(f, label)`` reports the
    fraction done; after ``cancel()`` turns true the runs stop where they
    are and the results are read from that state. ``on_result(result,
    run)`` is called per variant while its run is still open. ``river`` is
    accepted for signature compatibility.

    fp64_errors: the same plan runs a second time with FP64 storage, and
    every result carries 'err_fp64', the absolute difference of each
    reported quantity (FP64_ERROR_KEYS) and of the level, surface-velocity,
    discharge and bed-shear profiles from that FP64 run, and 'fp64', the
    FP64 run's comparison."""
    if cancel and cancel():
        return []
    plan = plan_test(test, sec, levels, h_sel, cfg, flow_fn, x_first, boat_edits)
    twin = fp64_errors and storage != "f64"
    tag = f"{len(plan.cases)} runs" + (" together" if batched else "")
    if progress:
        prog = (lambda f: progress(0.5 * f, tag)) if twin else (lambda f: progress(f, tag))
    else:
        prog = None
    runs = simulate(plan.reaches, be, storage, batched, prog, cancel, group)
    try:
        prof = [profiles(run) for run in runs]
        prof64 = prof
        if twin:
            runs64 = simulate(plan.reaches, be, "f64", batched,
                              (lambda f: progress(0.5 + 0.5 * f, tag + " (FP64)")) if progress else None, cancel, group)
            prof64 = [profiles(run) for run in runs64]
            for run in runs64:
                run.close()
        results = []
        for k, (label, h, reach, kb) in plan.variants():
            c = compare(prof[k], prof[kb], reach, cfg)
            c.update(label=label, profiles=prof[k], base_profiles=prof[kb], notes=reach.notes)
            if fp64_errors:
                c64 = compare(prof64[k], prof64[kb], reach, cfg)
                c["fp64"] = c64
                c["err_fp64"] = fp64_error(c, c64, prof[k], prof64[k], reach.Q)
            results.append(c)
            if on_result:
                on_result(c, runs[k])
        return results
    finally:
        for run in runs:
            run.close()
