"""The learned sub-grid closure on the GPU against its CPU oracle
(oracle/closure_oracle.py): the in-step MFMA network, the training gradient
and Adam kernels, the box-filtered training samples, one learned LBM step,
the constant-network special case, and what a trained network gains over
plain Smagorinsky on held-out fine-grid data."""
import ctypes

import numpy as np
import pytest

from oracle import closure_oracle as C
from oracle import oracle as O
from tests.helpers import channel128, rel_l2, solver_inputs
from water_flow_simulation_amd import _native as nat
from water_flow_simulation_amd import closure as CL
from water_flow_simulation_amd.lbm3d import LBM3D

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _d(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _features(rng, n):
    return np.column_stack([rng.uniform(-1.5, 1.5, n), rng.uniform(-1, 1, n), rng.normal(0, 0.4, n),
                            rng.uniform(0, 1, n), rng.normal(0, 0.2, n), rng.uniform(0, 1, n), rng.uniform(0, 1, n),
                            rng.uniform(0.1, 1, n)]).astype(np.float32)


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = CL.init_params(seed).astype(np.float64)
    p[C.OFF_B1:C.OFF_W2] = rng.normal(0, 0.3, C.NH)
    p[C.OFF_B2:C.OFF_W3] = rng.normal(0, 0.3, C.NH)
    p[C.OFF_W3:C.OFF_B3] = rng.normal(0, 0.3 * scale, C.NH)
    return p.astype(np.float32)


def test_mfma_network_matches_oracle():
    print("case: wfs_mlp_eval, 5000 samples", flush=True)
    rng = np.random.default_rng(0)
    n = 5000                                       # not a multiple of 64: idle lanes in the last wavefront
    X = _features(rng, n)
    p = _params(1)
    Xd, pd = _d(X), _d(p)
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    work = torch.empty(int(nat.lib().wfs_mlp_work_floats(n)), dtype=torch.float32, device="cuda")
    nat.check(nat.lib().wfs_mlp_eval(0, Xd.data_ptr(), n, pd.data_ptr(), out.data_ptr(), work.data_ptr(),
                                     work.numel(), _stream()))
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.float64)
    ref = C.c2_of(C.forward(p, X)[0])
    rel = np.abs(got - ref) / ref
    print("c2 range", ref.min(), ref.max(), "rel err median", np.median(rel), "max", rel.max(), flush=True)
    assert ref.max() / ref.min() > 2          # the network really varies over the samples
    assert np.quantile(rel, 0.999) < 1e-4 and rel.max() < 5e-3


def test_training_gradient_matches_oracle():
    print("case: wfs_mlp_grad, 3000 samples", flush=True)
    rng = np.random.default_rng(2)
    n = 3000
    X = _features(rng, n)
    S, T = C.deviatoric(rng.normal(size=(n, 6))) * 1e-3, C.deviatoric(rng.normal(size=(n, 6))) * 2e-7
    ss = C.ddot(S, S)
    smag = np.sqrt(2 * ss)
    ABP = np.column_stack([smag * smag * ss, smag * C.ddot(T, S), C.ddot(T, T)]).astype(np.float32)
    p = _params(3)
    Xd, Pd, pd = _d(X), _d(ABP), _d(p)
    grad = torch.empty(C.NPARAMS, dtype=torch.float32, device="cuda")
    sums = torch.zeros(2, dtype=torch.float64, device="cuda")
    work = torch.empty(int(nat.lib().wfs_mlp_work_floats(n)), dtype=torch.float32, device="cuda")
    nat.check(nat.lib().wfs_mlp_grad(0, Xd.data_ptr(), Pd.data_ptr(), n, pd.data_ptr(), grad.data_ptr(),
                                     sums.data_ptr(), work.data_ptr(), work.numel(), _stream()))
    torch.cuda.synchronize()
    loss, psum, g = C.loss_grad(p, X, ABP.astype(np.float64))
    s = sums.cpu().numpy()
    print("loss", s[0], loss, "psum", s[1], psum, "grad rel l2", rel_l2(grad.cpu().numpy(), g), flush=True)
    assert s[1] == pytest.approx(psum, rel=1e-6)
    assert s[0] == pytest.approx(loss, rel=1e-5)
    assert rel_l2(grad.cpu().numpy(), g) < 1e-3
    # deterministic: a second call gives the same bits
    g1 = grad.clone()
    nat.check(nat.lib().wfs_mlp_grad(0, Xd.data_ptr(), Pd.data_ptr(), n, pd.data_ptr(), grad.data_ptr(),
                                     sums.data_ptr(), work.data_ptr(), work.numel(), _stream()))
    torch.cuda.synchronize()
    assert torch.equal(g1, grad)


def test_adam_step():
    rng = np.random.default_rng(4)
    p, g = rng.normal(size=C.NPARAMS).astype(np.float32), rng.normal(size=C.NPARAMS).astype(np.float32)
    m, v = rng.normal(size=C.NPARAMS).astype(np.float32), rng.uniform(0, 1, C.NPARAMS).astype(np.float32)
    pd, gd, md, vd = _d(p), _d(g), _d(m), _d(v)
    t, lr, b1, b2, eps = 7, 1e-3, 0.9, 0.999, 1e-8
    nat.check(nat.lib().wfs_mlp_adam(0, pd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(), t, lr, b1, b2,
                                     eps, _stream()))
    torch.cuda.synchronize()
    m1 = b1 * m.astype(np.float64) + (1 - b1) * g
    v1 = b2 * v.astype(np.float64) + (1 - b2) * g.astype(np.float64) ** 2
    p1 = p - lr * (m1 / (1 - b1 ** t)) / (np.sqrt(v1 / (1 - b2 ** t)) + eps)
    np.testing.assert_allclose(md.cpu().numpy(), m1, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(vd.cpu().numpy(), v1, rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(pd.cpu().numpy(), p1, rtol=1e-5, atol=1e-7)


def _developed(storage="f32", steps=300):
    r = channel128()
    kw = solver_inputs(r)
    L = LBM3D(None, storage=storage, **kw)
    L.step(steps)
    return r, kw, L


def test_learned_step_matches_oracle():
    print("case: channel128, 300 constant-Cs steps then learned steps vs FP64 oracle", flush=True)
    r, kw, L = _developed()
    geo, hcol = CL.geometry(r.solid)
    u0 = float(kw["uin"][kw["uin"] > 0].mean())
    fl = ~r.solid.ravel()
    # (1) a varying network: the C^2 field of one step
    p = _params(5)
    orc = O.OracleLBM(precision="f64", **kw)
    orc.f = L.populations().astype(np.float64)
    orc.rho, orc.ux, orc.uy, orc.uz = (a.astype(np.float64) for a in L.macro())
    L.set_closure("mlp", p, geo, hcol, u0)
    L.step(1)
    c2_gpu = L.closure_field().astype(np.float64)
    c2_orc = C.learned_step(orc, p, geo, hcol, 1.0 / u0)
    rel = np.abs(c2_gpu - c2_orc)[fl] / c2_orc[fl]
    print("c2 range", c2_orc[fl].min(), c2_orc[fl].max(), "rel err median", np.median(rel), "max", rel.max(),
          flush=True)
    assert c2_orc[fl].max() / c2_orc[fl].min() > 1.5
    # FP32 forms Pi_neq by cancellation of O(1/3) populations (relative error ~1e-4 in the
    # strain features), which now and then moves a feature across a bfloat16 rounding edge
    assert np.median(rel) < 1e-6
    assert np.quantile(rel, 0.999) < 5e-3 and rel.max() < 2e-2
    f19 = np.tile(fl, 19)                      # solid cells are never written (as in the reference)
    f_gpu = L.populations().astype(np.float64)
    assert np.abs(f_gpu - orc.f)[f19].max() < 5e-7
    # (2) a strongly dissipative network: five steps, with the plain model as the sensitivity check
    p = _params(6)
    p[-1] = 3.0
    plain = O.OracleLBM(precision="f64", **kw)
    plain.f, plain.rho, plain.ux, plain.uy, plain.uz = (a.copy() for a in (orc.f, orc.rho, orc.ux, orc.uy, orc.uz))
    L.set_closure("mlp", p, geo, hcol, u0)
    for _ in range(5):
        C.learned_step(orc, p, geo, hcol, 1.0 / u0)
    plain.step(5)
    L.step(5)
    f_gpu = L.populations().astype(np.float64)
    err, sens = np.abs(f_gpu - orc.f)[f19].max(), np.abs(plain.f - orc.f)[f19].max()
    print("5 learned steps: max |GPU - oracle|", err, " max |plain - learned| (oracle)", sens, flush=True)
    assert err < 1e-6 and sens > 20 * err
    L.close()


def test_constant_network_equals_constant_cs():
    print("case: channel128, learned constant network vs f32fast Smagorinsky, 200 steps", flush=True)
    r = channel128()
    kw = solver_inputs(r)
    a = LBM3D(None, storage="f32fast", **kw)
    b = LBM3D(None, storage="f32fast", **kw)
    geo, hcol = CL.geometry(r.solid)
    u0 = float(kw["uin"][kw["uin"] > 0].mean())
    b.set_closure("mlp", CL.smagorinsky_params(r.cfg.cs), geo, hcol, u0)
    a.step(200)
    b.step(200)
    d = rel_l2(b.populations(), a.populations())
    du = rel_l2(b.macro()[1], a.macro()[1])
    print("rel l2 populations", d, "ux", du, flush=True)
    assert d < 1e-6 and du < 1e-4
    np.testing.assert_allclose(b.closure_field()[~r.solid.ravel()], r.cfg.cs ** 2, rtol=1e-5)
    b.set_closure("smagorinsky")           # the fallback switch
    b.step(1)
    a.close()
    b.close()


def test_training_samples_match_oracle():
    print("case: channel128 developed state, samples on the 64x16x8 coarse grid", flush=True)
    r, kw, L = _developed()
    ncx, ncy, ncz = r.nx // 2, r.ny // 2, r.nz // 2
    csol = CL.coarsen_solid(r.solid, ncx, ncy, ncz)
    geo, hcol = CL.geometry(csol)
    u0 = 0.05
    nc = ncx * ncy * ncz
    X = torch.empty((nc, 8), dtype=torch.float32, device="cuda")
    ABP = torch.empty((nc, 3), dtype=torch.float32, device="cuda")
    valid = torch.empty(nc, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    nat.check(nat.lib().wfs_closure_samples(L.handle, nat.ptr(geo), nat.ptr(hcol), ncx, ncy, ncz, 1.0, u0,
                                            X.data_ptr(), ABP.data_ptr(), valid.data_ptr()))
    _, ux, uy, uz = L.macro(np.float64)
    u = np.stack([a.reshape(r.solid.shape) for a in (ux, uy, uz)])
    Xo, Po, vo = C.samples(u, r.solid, ncx, ncy, ncz, 1.0, geo, hcol, 1 / u0)
    vg = valid.cpu().numpy().astype(bool)
    np.testing.assert_array_equal(vg, vo)
    assert vg.sum() > 0.3 * nc
    Xg, Pg = X.cpu().numpy()[vg].astype(np.float64), ABP.cpu().numpy()[vg].astype(np.float64)
    Xo, Po = Xo[vo], Po[vo]
    mism = np.abs(Xg - Xo) > 1e-12
    print("samples", vg.sum(), "feature entries off by rounding", mism.mean(), flush=True)
    assert mism.mean() < 1e-3
    assert np.all(np.abs(Xg - Xo) <= 2 ** -7 * np.abs(Xo) + 1e-6)
    for k in range(3):
        scale = np.abs(Po[:, k]).max()
        assert np.abs(Pg[:, k] - Po[:, k]).max() < 1e-3 * scale, k
    L.close()


def test_trained_network_beats_constant_coefficients():
    """End to end on small reaches: fine runs at dx 0.05 m filtered onto dx
    0.1 m, training on three cases, a priori judgement on a fourth."""
    print("case: closure study, 4 small reaches (3 train, 1 held out)", flush=True)
    p, res = CL.study(n_train=3, held_out=1, seed=10, width=(5.0, 8.0), depth=1.5, length=16.0, dx_c=0.1,
                      snapshots=6, epochs=8, batch=16384, lr=3e-3, posteriori=True, log=lambda m: print(m, flush=True))
    print(res, flush=True)
    tr = res["train"]
    assert tr["learned"] < tr["smagorinsky"] and tr["best_constant"] <= tr["smagorinsky"]
    # the gain over the model without the network: lower subfilter stress and
    # energy-transfer errors than Cs = 0.17 on the held-out case
    for metric in ("stress", "dissipation"):
        h = res["held_out"][0][metric]
        assert h["learned"] < h["smagorinsky"], metric
    # a posteriori (reported): coarse runs against the filtered fine run; both
    # the constant-Cs and the learned run must complete
    post = res["held_out"][0]["aposteriori"]
    assert all(np.isfinite(v) for v in post["smagorinsky"].values())
    assert "u_rel_l2" in post["smagorinsky"] and "eta_max_mm" in post["smagorinsky"]
    assert "unstable_at_step" not in post["learned"], post["learned"]
    assert all(np.isfinite(v) for v in post["learned"].values())


def test_outlet_sponge_in_the_kernel():
    """The learned step's C^2 field with a network at the floor: Cs^2 (the
    float the kernel holds) in the last 64 cells before the outlet row, the
    floor elsewhere, on both layouts."""
    print("case: channel128, floor network, outlet sponge", flush=True)
    from dataclasses import replace
    r = channel128()
    kw = solver_inputs(r)
    geo, hcol = CL.geometry(r.solid)
    u0 = float(kw["uin"][kw["uin"] > 0].mean())
    p = CL.smagorinsky_params(r.cfg.cs)
    p[-1] = -40.0
    fl = ~r.solid
    x = np.broadcast_to(np.arange(r.nx), r.solid.shape)
    for layout in ("dense", "tiles"):
        L = LBM3D(None, storage="f16dev", layout=layout, **kw)
        L.set_closure("mlp", p, geo, hcol, u0)
        L.step(3)
        c2 = L.closure_field().reshape(r.solid.shape)
        sponge = fl & (x >= r.nx - 64)
        assert np.all(c2[sponge] == np.float32(r.cfg.cs ** 2)), layout
        assert np.all(c2[fl & ~sponge] < 0.0101) and np.all(c2[fl & ~sponge] >= 0.01), layout
        L.close()
