"""Learned sub-grid closure: geometry codes, network weights, training samples
from fine runs of this solver, training with libwfs's own kernels, and the
judgement of the closure against fine-grid runs.

Model (csrc/closure.h). An 8-32-32-1 network (ReLU hidden layers, softplus
output) gives every cell a Smagorinsky coefficient

    C^2 = 0.01 + 0.0289 * softplus(y),   nu_t = C^2 |S|,

from eight features: strain-rate magnitude and shape from the
non-equilibrium moments, the share of vertical shear, speed and vertical
velocity relative to the mean inlet speed, wall distance, height in the
column and column depth. The step kernel evaluates it with BF16 MFMA on each
64-cell wavefront between the moments and the relaxation. Plain Smagorinsky
(the reference's constant Cs = 0.17, floodsim/lbm3d.py:120-123) is the
default and the fallback.

Training data. A reach is run at a fine spacing dx_f with the constant-Cs
model; snapshots are box-filtered 2x2x2 onto the grid of dx_c = 2 dx_f. For a
coarse cell the filter gives the resolved strain S (central differences of the
filtered velocity) and the subfilter stress T = mean(u u) - mean(u) mean(u);
the network is fitted to the relative error of the eddy-viscosity stress,

    sum |T_dev + 2 C^2 |S| S|^2 / sum |T_dev|^2,

plus (objective "mixed", the default) the relative error of the subfilter
energy transfer, sum (C^2 |S|^3 - Pi)^2 / sum Pi^2 with Pi = -T:S, by Adam on
the GPU with the library's forward/backward kernels (wfs_mlp_grad,
wfs_mlp_adam). PyTorch holds
the device buffers and shuffles indices; it computes nothing of the model.

Judgement. On held-out cases: the same relative stress error for the
network, for Cs = 0.17 and for the best constant coefficient (a priori), and
coarse runs with either closure against the filtered fine run's time means
(a posteriori).

    python -m water_flow_simulation_amd.closure [--train 10 --held-out 4 --dx 0.05 --length 20 ...]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time
from dataclasses import replace

import numpy as np
from scipy import ndimage

from . import _native as nat
from . import reach3d as R
from .geometry import Section
from .hydraulics1d import q_ref
from .workloads import E3_SLOPE

NF, NH = nat.MLP_FEATURES, 32
NPARAMS = nat.MLP_PARAMS
C2_MIN, C2_SCALE = nat.C2_MIN, nat.C2_SCALE
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "closure_weights.json")


# ---------------------------------------------------------------- geometry codes
def geometry(solid, chunk=256, workers=min(16, os.cpu_count() or 1)):
    """Per-cell code geo = round(255 / d) | round(255 z_rel) << 8 and per-column
    fluid-cell count hcol (ny, nx), for a solid mask (nz, ny, nx).

    d is the distance from a fluid cell centre to the nearest wall cell centre
    in cells, with walls as the step sees them: solid cells, the rows beyond
    the y edges and the layer below z = 0 (the lid and the x ends are open).
    z_rel = (z - z_bed + 1/2) / h, z_bed the lowest fluid cell of the column
    and h its fluid-cell count.

    The wall layer below z = 0 puts a wall within nz cells of every cell, so
    the distance transform runs on x-slabs of ``chunk`` cells with nz + 1
    cells of overlap on each side, exactly, on ``workers`` threads. Each slab
    is cut in y to its fluid rows and one solid row either side: a wall
    beyond that row is farther than the row's own cell."""
    solid = np.asarray(solid, bool)
    nz, ny, nx = solid.shape
    geo = np.zeros((nz, ny, nx), np.uint16)
    ncol = np.empty((ny, nx), np.int64)
    m = nz + 1
    z = np.arange(nz)[:, None, None]

    def slab(x0):
        x1 = min(nx, x0 + chunk)
        s = solid[:, :, x0:x1]
        fluid = ~s
        nc = fluid.sum(0)
        ncol[:, x0:x1] = nc
        a, b = max(0, x0 - m), min(nx, x1 + m)
        rows = np.flatnonzero(~solid[:, :, a:b].all(axis=(0, 2)))
        if rows.size == 0:
            return
        y0, y1 = max(int(rows[0]) - 1, 0), min(int(rows[-1]) + 2, ny)
        pad = np.ones((nz + 1, y1 - y0 + 2, b - a), bool)
        pad[1:, 1:-1, :] = solid[:, y0:y1, a:b]
        d = ndimage.distance_transform_edt(~pad)[1:, 1:-1, x0 - a:x1 - a]
        sc, nc = s[:, y0:y1], nc[y0:y1]
        inv = np.where(sc, 0.0, np.clip(np.rint(255.0 / np.maximum(d, 1.0)), 0, 255))
        zb = np.argmax(fluid[:, y0:y1], axis=0)
        zrel = np.clip((z - zb[None] + 0.5) / np.maximum(nc, 1)[None], 0.0, 1.0)
        zq = np.where(sc, 0.0, np.rint(255 * zrel))
        geo[:, y0:y1, x0:x1] = inv.astype(np.uint16) | (zq.astype(np.uint16) << 8)

    starts = range(0, nx, chunk)
    if workers > 1 and nx > chunk:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(slab, starts))
    else:
        for x0 in starts:
            slab(x0)
    return geo, np.minimum(ncol, 255).astype(np.uint8)


def coarsen_solid(solid, ncx, ncy, ncz):
    """Coarse mask of a fine one: solid where any of the 2x2x2 fine cells is."""
    s = np.asarray(solid, bool)[:2 * ncz, :2 * ncy, :2 * ncx]
    return s.reshape(ncz, 2, ncy, 2, ncx, 2).any(axis=(1, 3, 5))


# ---------------------------------------------------------------- weights
def smagorinsky_params(cs=0.17):
    """Weights for which the network gives C^2 = cs^2 in every cell."""
    p = np.zeros(NPARAMS, np.float32)
    s = (cs * cs - C2_MIN) / C2_SCALE
    if s <= 0:
        raise ValueError("cs^2 below the floor of the learned coefficient")
    p[-1] = math.log(math.expm1(s))
    return p


def init_params(seed=0, cs=0.17):
    """He-initialised hidden layers, a small output layer, output bias at the
    constant-Cs value."""
    rng = np.random.default_rng(seed)
    p = smagorinsky_params(cs)
    p[0:NH * NF] = rng.normal(0, math.sqrt(2 / NF), NH * NF)
    p[NH * NF + NH:NH * NF + NH + NH * NH] = rng.normal(0, math.sqrt(2 / NH), NH * NH)
    o = NH * NF + NH + NH * NH + NH
    p[o:o + NH] = rng.normal(0, 0.01, NH)
    return p.astype(np.float32)


def save_params(path, p, meta=None):
    with open(path, "w") as f:
        json.dump({"params": [float(v) for v in np.asarray(p, np.float32)], "meta": meta or {}}, f)


def load_params(path=WEIGHTS):
    with open(path) as f:
        d = json.load(f)
    p = np.asarray(d["params"], np.float32)
    if p.size != NPARAMS:
        raise ValueError(f"{path}: expected {NPARAMS} weights")
    return p, d.get("meta", {})


# ---------------------------------------------------------------- cases
def case_pair(seed, width=(20.0, 40.0), depth=3.0, length=60.0, dx_c=0.1, run_flows=1.0, avg_frac=0.5,
              with_boat=True):
    """One seeded channel reach built twice: at dx_c (the coarse run) and at
    dx_c / 2 with half the time step (the fine run), so that lattice
    velocities agree. Bed width drawn from ``width`` (a multiple of 2 dx_c),
    trapezoid with 1:2 banks, Manning 0.03, Q from the Manning rule on the E3
    slope; one or two propeller boats of a random preset at a random place,
    with the preset's power scaled by the flow area over 100 m^2 (capped at
    1) so that small test channels carry proportionate thrust."""
    rng = np.random.default_rng(seed)
    B = float(rng.uniform(*width))
    B = round(B / (2 * dx_c)) * 2 * dx_c
    sec = Section(B=B, Hb=depth + 0.5, z=2.0, Bf=0.0, n=0.03)
    Q = float(q_ref(depth, 1.0, E3_SLOPE, sec))
    key = ("longtail", "outboard", "tug", "pontoon")[int(rng.integers(0, 4))]
    nb = int(rng.integers(1, 3))
    boat = replace(R.boat_type(key), tilt_deg=float(rng.uniform(0, 15)))
    area = (B + sec.z * depth) * depth
    draft = min(boat.draft, 0.4 * depth)
    D = min(boat.prop_D, 0.4 * depth)
    # propeller disc below the keel and clear of the bed
    zp = min(max(boat.prop_depth, draft + 0.6 * D), depth - 0.6 * D)
    boat = replace(boat, prop_D=D, prop_depth=zp, draft=draft, hull_L=min(boat.hull_L, 0.2 * length),
                   hull_B=min(boat.hull_B, 0.3 * B), power_kW=boat.power_kW * min(1.0, area / 100.0))
    fleet = R.Fleet(boat, rows=1, per_row=nb, x_first=float(rng.uniform(0.2, 0.35)) * length, enabled=with_boat)
    cfg_c = R.ReachConfig(length=length, dx=dx_c, run_flows=run_flows, avg_frac=avg_frac)
    coarse = R.build_reach(sec, depth, Q, E3_SLOPE, fleet, cfg_c)
    cfg_f = replace(cfg_c, dx=dx_c / 2)
    fine = R.build_reach(sec, depth, Q, E3_SLOPE, fleet, cfg_f, dt=coarse.dt / 2)
    label = f"{key} x{nb}, bed {B:g} m, {depth:g} m deep, {length:g} m"
    return fine, coarse, label


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("closure training needs a GPU (torch.cuda) for its sample buffers")
    return torch


def _u0(reach):
    u = reach.uin_phys / reach.vel
    return float(u[u > 0].mean())


# ---------------------------------------------------------------- samples
def collect(run, coarse_solid, u0c, kv=1.0):
    """Training samples of the fine run's current state on the coarse grid of
    ``coarse_solid`` (nz_c, ny_c, nx_c): device tensors X (n, 8), ABP (n, 3)."""
    torch = _torch()
    ncz, ncy, ncx = coarse_solid.shape
    geo, hcol = geometry(coarse_solid)
    nc = ncx * ncy * ncz
    X = torch.empty((nc, NF), dtype=torch.float32, device="cuda")
    ABP = torch.empty((nc, 3), dtype=torch.float32, device="cuda")
    valid = torch.empty(nc, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    nat.check(nat.lib().wfs_closure_samples(run.lbm.handle, nat.ptr(geo), nat.ptr(hcol), ncx, ncy, ncz, float(kv),
                                            float(u0c), X.data_ptr(), ABP.data_ptr(), valid.data_ptr()),
              "wfs_closure_samples")
    keep = valid.bool()
    return X[keep].contiguous(), ABP[keep].contiguous()


def fine_run(fine, coarse, snapshots=8, storage="f32", log=None):
    """Runs the fine case to its end with the constant-Cs model and collects
    samples at ``snapshots`` evenly spaced times of the averaging window.
    Returns (X, ABP, run) with the run kept open for its time means, or None
    when the fine run does not stay finite (the case is then left out)."""
    torch = _torch()
    run = R.Run3D(fine, storage=storage)
    ncx, ncy, ncz = fine.nx // 2, fine.ny // 2, fine.nz // 2
    csol = coarsen_solid(fine.solid, ncx, ncy, ncz)
    u0c = _u0(coarse)
    kv = (fine.vel / coarse.vel)
    steps_total = int(math.ceil(run.t_total / fine.dt))
    steps_avg0 = int(math.ceil(run.t_avg / fine.dt))
    marks = sorted({steps_avg0 + int(round((k + 1) * (steps_total - steps_avg0) / snapshots)) for k in range(snapshots)})
    Xs, Ps = [], []
    done = 0
    for mk in marks:
        mk = max(mk - mk % R.Run3D.cadence, done + R.Run3D.cadence)
        run.advance(mk - done)
        done = mk
        try:
            run.lbm.check_finite()
        except nat.NonFiniteState as e:
            if log:
                log(f"  fine run left out: {e}")
            run.close()
            return None
        X, ABP = collect(run, csol, u0c, kv)
        Xs.append(X)
        Ps.append(ABP)
        if log:
            log(f"  fine step {done}/{steps_total}: {X.shape[0]} samples")
    while not run.done:
        run.advance(R.Run3D.cadence * 20)
    try:
        run.lbm.check_finite()
    except nat.NonFiniteState as e:
        if log:
            log(f"  fine run left out: {e}")
        run.close()
        return None
    return torch.cat(Xs), torch.cat(Ps), run


# ---------------------------------------------------------------- training
def evaluate(X, ABP, p, chunk=1 << 20):
    """Relative stress error sum(loss) / sum(P) of the network on device samples."""
    torch = _torch()
    n = X.shape[0]
    dev = X.device.index or 0
    pt = torch.as_tensor(np.asarray(p, np.float32), device=X.device)
    grad = torch.empty(NPARAMS, dtype=torch.float32, device=X.device)
    sums = torch.zeros(2, dtype=torch.float64, device=X.device)
    work = torch.empty(int(nat.lib().wfs_mlp_work_floats(min(chunk, n))), dtype=torch.float32, device=X.device)
    st = torch.cuda.current_stream().cuda_stream
    tot = np.zeros(2)
    for b0 in range(0, n, chunk):
        m = min(chunk, n - b0)
        nat.check(nat.lib().wfs_mlp_grad(dev, X[b0:].data_ptr(), ABP[b0:].data_ptr(), m, pt.data_ptr(),
                                         grad.data_ptr(), sums.data_ptr(), work.data_ptr(), work.numel(), st))
        tot += sums.cpu().numpy()
    return float(tot[0] / tot[1])


OBJECTIVES = ("stress", "dissipation", "mixed")


def objective_coeffs(ABP, objective="stress"):
    """Per-sample coefficients (A, B, P) of loss = P + 4 C^2 B + 4 C^4 A for a
    training objective:
      stress       |T + 2 C^2 |S| S|^2 (the samples as collected);
      dissipation  (C^2 |S|^3 - Pi)^2, Pi = -T:S the subfilter energy transfer,
                   i.e. A = |S|^6 / 4, B = |S|^2 B_stress / 2, P = Pi^2;
      mixed        the two, each divided by its sum of P over the samples."""
    if objective == "stress":
        return ABP
    if objective not in OBJECTIVES:
        raise ValueError(f"unknown objective {objective!r}")
    torch = _torch()
    A, B, P = ABP.double().unbind(1)
    s2 = torch.sqrt(2 * A)                  # |S|^2
    smag = torch.sqrt(s2)
    pi = torch.where(smag > 0, -B / torch.clamp(smag, min=1e-30), torch.zeros_like(B))
    diss = torch.stack([s2 ** 3 / 4, s2 * B / 2, pi * pi], 1)
    if objective == "dissipation":
        return diss.float().contiguous()
    st = ABP.double() / ABP[:, 2].double().sum()
    return (st + diss / diss[:, 2].sum()).float().contiguous()


def constant_errors(ABP, cs=0.17):
    """Relative stress errors of constant coefficients: Cs = cs, and the best
    constant C^2 = max(-sum B / (2 sum A), 0) with its value."""
    s = ABP.double().sum(0).cpu().numpy()
    A, B, P = s
    e = lambda c2: (P + 4 * c2 * B + 4 * c2 * c2 * A) / P  # noqa: E731
    best = max(-B / (2 * A), 0.0) if A > 0 else 0.0
    return {"smagorinsky": float(e(cs * cs)), "best_constant": float(e(best)), "best_c2": float(best)}


def train(X, ABP, p0=None, epochs=6, batch=65536, lr=3e-3, seed=0, log=None):
    """Adam on the stress-space loss with libwfs's gradient and update kernels;
    cosine learning-rate decay. Returns the float32 weights."""
    torch = _torch()
    n = X.shape[0]
    if n < batch:
        batch = n
    dev = X.device.index or 0
    p = torch.as_tensor(np.asarray(p0 if p0 is not None else init_params(seed), np.float32), device=X.device).clone()
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    grad = torch.empty_like(p)
    sums = torch.zeros(2, dtype=torch.float64, device=X.device)
    work = torch.empty(int(nat.lib().wfs_mlp_work_floats(batch)), dtype=torch.float32, device=X.device)
    st = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device=X.device)
    g.manual_seed(seed)
    L = nat.lib()
    per_epoch = max(n // batch, 1)
    total = epochs * per_epoch
    t = 0
    for ep in range(epochs):
        perm = torch.randperm(n, device=X.device, generator=g)
        for k in range(per_epoch):
            sel = perm[k * batch:(k + 1) * batch]
            xb = X.index_select(0, sel)
            ab = ABP.index_select(0, sel)
            t += 1
            lr_t = lr * 0.5 * (1 + math.cos(math.pi * (t - 1) / total))
            nat.check(L.wfs_mlp_grad(dev, xb.data_ptr(), ab.data_ptr(), xb.shape[0], p.data_ptr(), grad.data_ptr(),
                                     sums.data_ptr(), work.data_ptr(), work.numel(), st))
            nat.check(L.wfs_mlp_adam(dev, p.data_ptr(), grad.data_ptr(), m.data_ptr(), v.data_ptr(), t, lr_t, 0.9,
                                     0.999, 1e-8, st))
        if log:
            log(f"  epoch {ep + 1}/{epochs}: train error {evaluate(X, ABP, p.cpu().numpy()):.5f}")
    return p.cpu().numpy()


# ---------------------------------------------------------------- a posteriori
def filtered_means(run, coarse):
    """Time-mean velocity of the fine run, box-filtered onto the coarse run's
    grid (3, nz_c, ny_c, nx_c), and the mask of coarse cells it covers."""
    f = run.r
    u = np.stack([run.field(k, mean=True) for k in ("ux", "uy", "uz")])   # m/s
    fl = ~f.solid
    off = (2 * coarse.ny - f.ny) // 2            # fine cell k <-> coarse (k + off) // 2
    out = np.zeros((3, coarse.nz, coarse.ny, coarse.nx))
    cover = np.zeros((coarse.nz, coarse.ny, coarse.nx), bool)
    j0 = (off + 1) // 2
    j1 = min(coarse.ny, (f.ny + off) // 2)
    k0 = 2 * j0 - off
    nyc = j1 - j0
    ub = u[:, :2 * coarse.nz, k0:k0 + 2 * nyc, :2 * coarse.nx]
    fb = fl[:2 * coarse.nz, k0:k0 + 2 * nyc, :2 * coarse.nx]
    ub = ub.reshape(3, coarse.nz, 2, nyc, 2, coarse.nx, 2).mean(axis=(2, 4, 6))
    fb = fb.reshape(coarse.nz, 2, nyc, 2, coarse.nx, 2).all(axis=(1, 3, 5))
    out[:, :, j0:j1] = ub
    cover[:, j0:j1] = fb
    return out, cover


def _pairs(a):
    a = np.asarray(a, np.float64)
    return a[:2 * (len(a) // 2)].reshape(-1, 2).mean(1)


def profile_errors(pc, pf):
    """Coarse-run profiles pc against the fine run's pf (two fine x-cells per
    coarse one): level max |d eta| (mm, both relative to the outlet), surface
    velocity relative L2, discharge max relative error; inlet/outlet cells
    excluded."""
    sl = slice(2, -2)
    eta_f = _pairs(pf["eta"])
    eta_f = eta_f - eta_f[-1]
    eta_c = np.asarray(pc["eta"]) - pc["eta"][-1]
    us_f, q_f = _pairs(pf["us"]), _pairs(pf["q"])
    return {"eta_max_mm": float(np.abs(eta_c - eta_f)[sl].max() * 1e3),
            "us_rel_l2": float(np.linalg.norm((pc["us"] - us_f)[sl]) / np.linalg.norm(us_f[sl])),
            "q_max_rel": float(np.abs(pc["q"] - q_f)[sl].max() / np.abs(q_f[sl]).max())}


def aposteriori(coarse, fine_means, cover, closures, fine_profiles=None, storage="f32"):
    """Coarse runs, one per entry of ``closures`` (name -> weights, None for
    the constant-Cs model), against the fine run: relative L2 of the time-mean
    velocity over cells fluid in both grids away from the inlet and outlet
    rows, and with ``fine_profiles`` the errors of the reported profiles."""
    res = {}
    mask = cover & ~coarse.solid
    mask[:, :, :2] = False
    mask[:, :, -2:] = False
    ref = fine_means[:, mask]
    for name, cl in closures.items():
        run = R.Run3D(coarse, storage=storage, closure=cl)
        unstable = None
        while not run.done:
            run.advance(R.Run3D.cadence * 500)
            try:
                run.lbm.check_finite()
            except nat.NonFiniteState:
                unstable = run.lbm.nstep
                break
        if unstable is not None:        # the run blew up: record when
            res[name] = {"unstable_at_step": int(unstable)}
            run.close()
            continue
        u = np.stack([run.field(k, mean=True) for k in ("ux", "uy", "uz")])[:, mask]
        row = {"u_rel_l2": float(np.linalg.norm(u - ref) / np.linalg.norm(ref))}
        if fine_profiles is not None:
            row.update(profile_errors(R.profiles(run), fine_profiles))
        res[name] = row
        run.close()
    return res


# ---------------------------------------------------------------- study
def study(n_train=10, held_out=4, seed=0, width=(20.0, 40.0), depth=3.0, length=60.0, dx_c=0.1, snapshots=8,
          epochs=6, batch=65536, lr=3e-3, run_flows=1.0, avg_frac=0.5, posteriori=True, log=print,
          objective="mixed", spare=4):
    """Fine runs of seeded reaches (seeds seed, seed + 1, ...): the first
    ``n_train`` that stay finite give the training samples, the next
    ``held_out`` are kept for the judgement; a fine run that does not stay
    finite is left out and the next seed takes its place (at most ``spare``
    such replacements). Then training, and the a-priori and a-posteriori
    judgement."""
    torch = _torch()
    t0 = time.time()
    train_X, train_P, test = [], [], []
    seeds = {"train": [], "held_out": [], "left_out": []}
    cells = 0
    k = 0
    while (len(seeds["train"]) < n_train or len(seeds["held_out"]) < held_out) and \
            k < n_train + held_out + spare:
        fine, coarse, label = case_pair(seed + k, width, depth, length, dx_c, run_flows, avg_frac)
        k += 1
        cells += fine.cells
        if log:
            log(f"case seed {seed + k - 1}: {label}; fine {fine.nx}x{fine.ny}x{fine.nz}, "
                f"coarse {coarse.nx}x{coarse.ny}x{coarse.nz}")
        got = fine_run(fine, coarse, snapshots, log=log)
        if got is None:
            seeds["left_out"].append(seed + k - 1)
            continue
        X, ABP, run = got
        if len(seeds["train"]) < n_train:
            seeds["train"].append(seed + k - 1)
            train_X.append(X)
            train_P.append(ABP)
            run.close()
        else:
            seeds["held_out"].append(seed + k - 1)
            means, cover = filtered_means(run, coarse) if posteriori else (None, None)
            prof = R.profiles(run) if posteriori else None
            run.close()
            test.append((X, ABP, coarse, means, cover, prof, label))
    X = torch.cat(train_X)
    ABP = torch.cat(train_P)
    del train_X, train_P
    t_data = time.time() - t0
    Y = objective_coeffs(ABP, objective)
    consts = constant_errors(Y)
    if log:
        log(f"{X.shape[0]} training samples ({objective}); constant-Cs error {consts['smagorinsky']:.4f}, "
            f"best constant C^2 = {consts['best_c2']:.4f} error {consts['best_constant']:.4f}")
    t1 = time.time()
    p = train(X, Y, init_params(seed), epochs, batch, lr, seed, log)
    t_train = time.time() - t1
    out = {"objective": objective, "samples": int(X.shape[0]), "fine_cells_total": int(cells), "seeds": seeds,
           "data_seconds": round(t_data, 1), "train_seconds": round(t_train, 1),
           "train": {**consts, "learned": evaluate(X, Y, p)}, "held_out": []}
    del Y
    for Xh, Ph, coarse, means, cover, prof, label in test:
        out["held_out"].append(judge(Xh, Ph, p, consts["best_c2"], label, coarse if posteriori else None, means,
                                     cover, prof, objective))
        if log:
            log(f"held out {label}: {out['held_out'][-1]}")
    return p, out


def judge(Xh, Ph, p, best_c2, label, coarse=None, means=None, cover=None, prof=None, objective="stress"):
    """A-priori errors of the network, Cs = 0.17 and the training set's best
    constant on held-out samples (stress and dissipation metrics), and with a
    coarse case the a-posteriori comparison."""
    row = {"case": label, "samples": int(Xh.shape[0])}
    for obj in ("stress", "dissipation"):
        Y = objective_coeffs(Ph, obj)
        c = constant_errors(Y)
        s = Y.double().sum(0).cpu().numpy()
        row[obj] = {"smagorinsky": c["smagorinsky"], "learned": evaluate(Xh, Y, p)}
        if obj == objective or objective == "mixed":
            row[obj]["train_best_constant"] = float((s[2] + 4 * best_c2 * s[1] + 4 * best_c2 ** 2 * s[0]) / s[2])
    if coarse is not None:
        row["aposteriori"] = aposteriori(coarse, means, cover, {"smagorinsky": None, "learned": p}, prof)
    return row


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--train", type=int, default=10, help="training reaches (fine runs that stay finite)")
    ap.add_argument("--held-out", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--width", type=float, nargs=2, default=(20.0, 40.0))
    ap.add_argument("--depth", type=float, default=3.0)
    ap.add_argument("--length", type=float, default=60.0)
    ap.add_argument("--dx", type=float, default=0.05, help="fine spacing (coarse = 2 dx)")
    ap.add_argument("--snapshots", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--run-flows", type=float, default=1.0)
    ap.add_argument("--no-posteriori", action="store_true")
    ap.add_argument("--objective", default="mixed", choices=OBJECTIVES)
    ap.add_argument("--out", default=None, help="write the trained weights (JSON) here")
    a = ap.parse_args()
    p, res = study(a.train, a.held_out, a.seed, tuple(a.width), a.depth, a.length, 2 * a.dx, a.snapshots, a.epochs,
                   a.batch, a.lr, a.run_flows, posteriori=not a.no_posteriori,
                   log=lambda m: print(m, flush=True), objective=a.objective)
    print(json.dumps(res), flush=True)
    if a.out:
        save_params(a.out, p, {"study": res, "args": vars(a)})


if __name__ == "__main__":
    main()
