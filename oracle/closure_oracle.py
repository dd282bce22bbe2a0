"""CPU oracle for the learned sub-grid closure.

TEST INFRASTRUCTURE ONLY: the checker that the closure kernels of libwfs are
compared against (features, the 8-32-32-1 network with its BF16 rounding, the
stress-space loss and its gradient, the box-filtered training samples, and one
LBM step with a per-cell coefficient). Only tests/ imports it.

The closure has no counterpart in the reference, whose only sub-grid model is
the constant-coefficient Smagorinsky relaxation time of floodsim/lbm3d.py:
120-123; the learned model replaces Cs^2 there by a per-cell C^2. Everything
here is float64 numpy restating the mathematics; the tests pin it against
independent checks (torch's own bfloat16 rounding, numpy matrix products for
the invariants, central finite differences for the gradient, the direct
tensor form of the loss, brute-force loops and a linear field's closed form
for the box filter, and the constant-coefficient special case that must give
the plain Smagorinsky step).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import oracle as _orc

NF, NH = 8, 32
OFF_W1, OFF_B1 = 0, NH * NF
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAMS = OFF_B3 + 1
C2_MIN = float(np.float32(0.01))
C2_SCALE = float(np.float32(0.0289))


def split(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NH, NF), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], float(p[OFF_B3]))


def join(W1, b1, W2, b2, w3, b3, dtype=np.float32):
    return np.concatenate([np.ravel(W1), np.ravel(b1), np.ravel(W2), np.ravel(b2), np.ravel(w3),
                           np.atleast_1d(b3)]).astype(dtype)


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even) and back to float32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def softplus(y):
    return np.logaddexp(0.0, y)


def sigmoid(y):
    return 0.5 * (1.0 + np.tanh(0.5 * y))


def c2_of(y):
    return C2_MIN + C2_SCALE * softplus(y)


def trace_s3(S):
    """tr(S^3) of symmetric tensors given as (..., 6) [xx, yy, zz, xy, xz, yz]."""
    a, b, c, d, e, f = (S[..., k] for k in range(6))
    return a ** 3 + b ** 3 + c ** 3 + 3 * d * d * (a + b) + 3 * e * e * (a + c) + 3 * f * f * (b + c) + 6 * d * e * f


def ddot(S, T):
    """S:T of symmetric tensors in (..., 6) form."""
    return (S[..., 0] * T[..., 0] + S[..., 1] * T[..., 1] + S[..., 2] * T[..., 2] +
            2 * (S[..., 3] * T[..., 3] + S[..., 4] * T[..., 4] + S[..., 5] * T[..., 5]))


def deviatoric(S):
    S = np.array(S, np.float64, copy=True)
    t = (S[..., 0] + S[..., 1] + S[..., 2]) / 3
    for k in range(3):
        S[..., k] -= t
    return S


def features(S, u, geo, hcol, inv_u0):
    """The eight network inputs (csrc/closure.h, closure_features) of cells with
    deviatoric strain S (..., 6), velocity u (..., 3), geometry code geo
    (uint16) and column depth hcol (uint8)."""
    S = np.asarray(S, np.float64)
    u = np.asarray(u, np.float64)
    ss = ddot(S, S)
    smag = np.sqrt(2 * ss)
    live = ss > 1e-20
    x = np.zeros(S.shape[:-1] + (NF,))
    x[..., 0] = 0.25 * np.log(smag + 1e-7) + 1.5
    with np.errstate(divide="ignore", invalid="ignore"):
        x[..., 1] = np.where(live, np.clip(np.sqrt(6.0) * trace_s3(S) / (ss * np.sqrt(ss)), -1, 1), 0.0)
        x[..., 3] = np.where(live, 2 * (S[..., 4] ** 2 + S[..., 5] ** 2) / ss, 0.0)
    x[..., 2] = np.sqrt((u * u).sum(-1)) * inv_u0 - 1
    x[..., 4] = u[..., 2] * inv_u0
    geo = np.asarray(geo).astype(np.int64)
    x[..., 5] = (geo & 255) / 255.0
    x[..., 6] = (geo >> 8) / 255.0
    x[..., 7] = np.asarray(hcol).astype(np.float64) / 64.0
    return x


def strain_from_moments(mom, tau0, csmag2):
    """Resolved strain S = -3 Pi_dev / (2 rho tau_s) from the collision moments
    mom (10, N) = rho, u, Pi_neq (xx yy zz xy xz yz), with tau_s the
    constant-Cs relaxation time of floodsim/lbm3d.py:120-123."""
    r = mom[0]
    pi = np.stack(mom[4:10], -1)
    qn = np.sqrt(ddot(pi, pi))
    tau_s = 0.5 * (tau0 + np.sqrt(tau0 * tau0 + 18 * 1.41421356 * csmag2 * qn / r))
    k = -1.5 / (r * tau_s)
    return deviatoric(pi) * k[:, None], np.stack(mom[1:4], -1)


def forward(p, X, quantize=True):
    """Network output y and the intermediates. With quantize, the inputs, W1,
    W2 and the layer-1 activations are rounded to bfloat16 as in the MFMA path;
    arithmetic is float64."""
    W1, b1, W2, b2, w3, b3 = split(p)
    X = np.asarray(X, np.float64)
    if quantize:
        Xq, W1q, W2q = bf16(X).astype(np.float64), bf16(W1).astype(np.float64), bf16(W2).astype(np.float64)
        b1 = np.float32(b1).astype(np.float64)
    else:
        Xq, W1q, W2q = X, W1, W2
    z1 = Xq @ W1q.T + b1
    h1 = np.maximum(z1, 0)
    if quantize:
        h1 = bf16(h1).astype(np.float64)
    z2 = h1 @ W2q.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    return y, (Xq, W1q, W2q, z1, h1, z2, h2)


def loss_grad(p, X, ABP, quantize=True):
    """Loss sum(P + 4 C^2 B + 4 C^4 A), sum(P), and the gradient of the loss
    divided by sum(P) (straight-through for the rounding)."""
    y, (Xq, W1q, W2q, z1, h1, z2, h2) = forward(p, X, quantize)
    _, _, _, _, w3, _ = split(p)
    ABP = np.asarray(ABP, np.float64)
    A, B, P = ABP[:, 0], ABP[:, 1], ABP[:, 2]
    c2 = c2_of(y)
    loss = P + 4 * c2 * B + 4 * c2 * c2 * A
    dy = (4 * B + 8 * c2 * A) * C2_SCALE * sigmoid(y)
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = dy[:, None] * w3[None, :] * (z2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2q) * (z1 > 0)
    gW1 = d1.T @ Xq
    gb1 = d1.sum(0)
    psum = P.sum()
    g = join(gW1, gb1, gW2, gb2, gw3, gb3, np.float64) / psum
    return float(loss.sum()), float(psum), g


def stress_loss(c2, S, T):
    """|T + 2 C^2 |S| S|^2 per sample, written out on full 3x3 tensors."""
    def full(V):
        M = np.empty(V.shape[:-1] + (3, 3))
        M[..., 0, 0], M[..., 1, 1], M[..., 2, 2] = V[..., 0], V[..., 1], V[..., 2]
        M[..., 0, 1] = M[..., 1, 0] = V[..., 3]
        M[..., 0, 2] = M[..., 2, 0] = V[..., 4]
        M[..., 1, 2] = M[..., 2, 1] = V[..., 5]
        return M
    Sm, Tm = full(np.asarray(S, np.float64)), full(np.asarray(T, np.float64))
    smag = np.sqrt(2 * np.einsum("...ij,...ij->...", Sm, Sm))
    E = Tm + 2 * (c2 * smag)[..., None, None] * Sm
    return np.einsum("...ij,...ij->...", E, E)


# ---------------------------------------------------------------- training samples

def coarsen(u, solid, ncx, ncy, ncz, kv):
    """2x2x2 box filter of the fine velocity u (3, nz, ny, nx): the coarse mean,
    the subfilter stress mean((u - ubar)(u - ubar)) (..., 6), and the coarse
    fluid mask (no solid fine cell)."""
    u = np.asarray(u, np.float64)[:, :2 * ncz, :2 * ncy, :2 * ncx] * kv
    b = u.reshape(3, ncz, 2, ncy, 2, ncx, 2)
    m = b.mean(axis=(2, 4, 6))
    d = b - m[:, :, None, :, None, :, None]
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    T = np.stack([(d[i] * d[j]).mean(axis=(1, 3, 5)) for i, j in pairs], -1)
    sb = np.asarray(solid)[:2 * ncz, :2 * ncy, :2 * ncx].reshape(ncz, 2, ncy, 2, ncx, 2)
    fluid = ~sb.any(axis=(1, 3, 5))
    m = np.where(fluid[None], m, 0.0)
    T = np.where(fluid[..., None], T, 0.0)
    return m, T, fluid


def samples(u, solid, ncx, ncy, ncz, kv, geo, hcol, inv_u0):
    """Training samples of a fine state (the k_coarsen + k_samples pair):
    features X (nc, 8, rounded to bfloat16), ABP (nc, 3) and valid (nc,), in
    dense coarse order. Central differences of the filtered velocity; the lid
    above the top layer mirrors the cell (ghost (ux, uy, -uz))."""
    m, T, fluid = coarsen(u, solid, ncx, ncy, ncz, kv)
    nz, ny, nx = ncz, ncy, ncx
    valid = np.zeros((nz, ny, nx), bool)
    valid[1:, 1:-1, 1:-1] = fluid[1:, 1:-1, 1:-1]
    valid[1:, 1:-1, 1:-1] &= fluid[1:, 1:-1, :-2] & fluid[1:, 1:-1, 2:]
    valid[1:, 1:-1, 1:-1] &= fluid[1:, :-2, 1:-1] & fluid[1:, 2:, 1:-1]
    valid[1:, 1:-1, 1:-1] &= fluid[:-1, 1:-1, 1:-1]
    up_ok = np.ones((nz, ny, nx), bool)
    up_ok[:-1] = fluid[1:]
    valid &= up_ok
    ghost = m[:, -1:].copy()
    ghost[2] *= -1
    mz = np.concatenate([m, ghost], axis=1)   # z index nz is the mirror ghost
    g = np.zeros((3, 3, nz, ny, nx))
    for i in range(3):
        g[i, 0, :, :, 1:-1] = 0.5 * (m[i, :, :, 2:] - m[i, :, :, :-2])
        g[i, 1, :, 1:-1, :] = 0.5 * (m[i, :, 2:, :] - m[i, :, :-2, :])
        g[i, 2, 1:, :, :] = 0.5 * (mz[i, 2:, :, :] - mz[i, :-2, :, :])
    S = np.stack([g[0, 0], g[1, 1], g[2, 2], 0.5 * (g[0, 1] + g[1, 0]), 0.5 * (g[0, 2] + g[2, 0]),
                  0.5 * (g[1, 2] + g[2, 1])], -1)
    S = deviatoric(S)
    Td = deviatoric(T)
    ss = ddot(S, S)
    smag = np.sqrt(2 * ss)
    A = smag * smag * ss
    B = smag * ddot(Td, S)
    P = ddot(Td, Td)
    hc = np.broadcast_to(np.asarray(hcol)[None], (nz, ny, nx))
    X = features(S, np.moveaxis(m, 0, -1), geo, hc, inv_u0)
    X = bf16(X).astype(np.float64)
    X[~valid] = 0
    ABP = np.stack([A, B, P], -1)
    ABP[~valid] = 0
    return X.reshape(-1, NF), ABP.reshape(-1, 3), valid.ravel()


# ---------------------------------------------------------------- learned step

def _step_field(lbm, fd, cs2f, mom, rho, ux, uy, uz):
    L = _orc.lib()
    fn = L.orc_step_field_f64
    fn.argtypes = [ctypes.c_void_p] * 12 + [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + [ctypes.c_void_p] * 2
    fn.restype = ctypes.c_int
    P = _orc._ptr
    fn(P(lbm.f), P(fd), P(lbm.solid), P(lbm.Fx), P(lbm.Fy), P(lbm.Fz), P(rho), P(ux), P(uy), P(uz), P(lbm.cfw),
       P(lbm.uref), lbm.nx, lbm.ny, lbm.nz, lbm.tau0, lbm.csmag2,
       P(cs2f) if cs2f is not None else None, P(mom) if mom is not None else None)


def moments(lbm):
    """Collision moments (10, N) of the FP64 oracle's next step, without
    advancing it."""
    mom = np.zeros(10 * lbm.N)
    scratch = np.zeros_like(lbm.f)
    copies = [a.copy() for a in (lbm.rho, lbm.ux, lbm.uy, lbm.uz)]
    _step_field(lbm, scratch, None, mom, *copies)
    return mom.reshape(10, lbm.N)


OUTLET_SPONGE = 64   # cells before the outlet row where C^2 >= Cs^2


def outlet_sponge(c2, nx, ny, nz, csmag2):
    """C^2 of the learned closure with the outlet sponge: in the last
    OUTLET_SPONGE cells along x (x >= nx - OUTLET_SPONGE) at least the
    constant model's Cs^2 as the float kernel holds it."""
    x = np.tile(np.arange(nx), ny * nz)
    return np.where(x >= nx - OUTLET_SPONGE, np.maximum(c2, float(np.float32(csmag2))), c2)


def learned_step(lbm, p, geo, hcol, inv_u0):
    """One step of the FP64 oracle (precision "f64") with the learned closure:
    moments, features, network, then the step with the per-cell C^2 field and
    the inlet/outlet rows. Returns the C^2 field (N,)."""
    if lbm.precision != "f64":
        raise ValueError("learned_step needs the FP64 oracle")
    mom = moments(lbm)
    S, u = strain_from_moments(mom, lbm.tau0, lbm.csmag2)
    hc = np.broadcast_to(np.asarray(hcol).reshape(1, lbm.ny, lbm.nx), (lbm.nz, lbm.ny, lbm.nx)).ravel()
    X = features(S, u, np.asarray(geo).ravel(), hc, inv_u0)
    y, _ = forward(p, X)
    c2 = np.ascontiguousarray(outlet_sponge(c2_of(y), lbm.nx, lbm.ny, lbm.nz, lbm.csmag2))
    _step_field(lbm, lbm.f2, c2, None, lbm.rho, lbm.ux, lbm.uy, lbm.uz)
    _orc.lib().orc_bc_f64(_orc._ptr(lbm.f2), _orc._ptr(lbm.solid), _orc._ptr(lbm.uin), lbm.nx, lbm.ny, lbm.nz, 1.0)
    lbm.f, lbm.f2 = lbm.f2, lbm.f
    lbm.nstep += 1
    return c2
