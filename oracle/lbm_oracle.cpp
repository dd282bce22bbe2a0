// CPU oracle for the river-reach D3Q19 solver.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the HIP product is
// compared against and the CPU baseline that bench.py times beside it. Nothing
// in the product package links, loads or calls it.
//
// It restates, statement by statement, the reference solver in
// floodsim/lbm3d.py:
//   * lbm_step  (CUDA source lbm3d.py:31-138, Numba twin lbm3d.py:186-294)
//   * lbm_bc    (CUDA source lbm3d.py:140-175, Numba twin lbm3d.py:296-328)
//   * _equilibrium initial state (lbm3d.py:377-384)
// and the device-side pieces of the run driver in floodsim/reach3d.py:
//   * Run3D._update_uref column reduction (reach3d.py:412-418)
//
// Three arithmetic instantiations share one template:
//   f64   : double storage, double arithmetic              (the FP64 oracle)
//   f32   : float storage, float arithmetic, no FMA contraction,
//           evaluation order of the reference CUDA kernel   (CUDA-like FP32)
//   mixed : float storage, double scalar arithmetic, float
//           population/equilibrium temporaries             (Numba-like)
// Build flags keep -ffp-contract=off so that f32 follows IEEE single
// precision operation by operation.
//
// Layout (lbm3d.py:9): f[i*N + cell], cell = x + nx*(y + ny*z), x fastest.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <omp.h>

namespace {

// D3Q19 velocity set in the reference order (lbm3d.py:16-21).
const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
const int MIRZ[19] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16};

template <typename Real>
struct Weights {
    Real w[19];
    Weights() {
        w[0] = Real(1) / Real(3);
        for (int i = 1; i < 7; ++i) w[i] = Real(1) / Real(18);
        for (int i = 7; i < 19; ++i) w[i] = Real(1) / Real(36);
    }
};

// Real = storage / local population type, Acc = scalar arithmetic type.
template <typename Real, typename Acc>
struct Lbm {
    static const Weights<Real>& W() { static Weights<Real> w; return w; }

    // lbm3d.py:140-143 (feq_i) / lbm3d.py:297-299 (feq1)
    static Acc feq_i(int i, Acc r, Acc ux, Acc uy, Acc uz) {
        const Real* w = W().w;
        const Acc cu = Acc(3) * (CX[i] * ux + CY[i] * uy + CZ[i] * uz);
        return Acc(w[i]) * r * (Acc(1) + cu + Acc(0.5) * cu * cu - Acc(1.5) * (ux * ux + uy * uy + uz * uz));
    }

    // One cell of lbm_step (lbm3d.py:38-137; Numba lbm3d.py:189-294).
    static void cell(const Real* fs, Real* fd, const uint8_t* solid,
                     const Real* Fx, const Real* Fy, const Real* Fz,
                     Real* rho_o, Real* ux_o, Real* uy_o, Real* uz_o,
                     const Real* cfw, const Real* uref,
                     int nx, int ny, int nz, int64_t N, int x, int y, int z,
                     Acc tau0, Acc csmag2, const Acc* cs2f, Acc* mom) {
        const Real* w = W().w;
        const int64_t c = x + (int64_t)nx * (y + (int64_t)ny * z);
        if (solid[c]) {  // lbm3d.py:44
            rho_o[c] = Real(1); ux_o[c] = Real(0); uy_o[c] = Real(0); uz_o[c] = Real(0);
            if (mom) for (int k = 0; k < 10; ++k) mom[(int64_t)k * N + c] = Acc(k == 0 ? 1 : 0);
            return;
        }
        Real f[19];
        uint32_t wmask = 0u;
        Acc nwx = 0, nwy = 0, nwz = 0;
        // Pull streaming with x-edge copy, rigid-lid mirror and bounce-back (lbm3d.py:50-60).
        for (int i = 0; i < 19; ++i) {
            const int xs = x - CX[i];
            int ys = y - CY[i], zs = z - CZ[i];
            int ii = i;
            if (xs < 0 || xs >= nx) { f[i] = fs[(int64_t)i * N + c]; continue; }
            if (zs >= nz) { zs = z; ii = MIRZ[i]; }
            bool wall = (ys < 0 || ys >= ny || zs < 0);
            int64_t s = 0;
            if (!wall) { s = xs + (int64_t)nx * (ys + (int64_t)ny * zs); wall = solid[s] != 0; }
            if (wall) { wmask |= (1u << i); nwx -= CX[i]; nwy -= CY[i]; nwz -= CZ[i]; }
            f[i] = wall ? fs[(int64_t)OPP[i] * N + c] : fs[(int64_t)ii * N + s];
        }
        // Wall function: slip-velocity bounce-back delivering rho*cf*U^2*area (lbm3d.py:61-101).
        const Acc cf = Acc(cfw[c]);
        if (wmask && cf > Acc(0)) {
            const Acc nn = std::sqrt(nwx * nwx + nwy * nwy + nwz * nwz);
            const Acc afx = Acc(((wmask >> 1) & 1u) + ((wmask >> 2) & 1u));
            const Acc afy = Acc(((wmask >> 3) & 1u) + ((wmask >> 4) & 1u));
            const Acc afz = Acc(((wmask >> 5) & 1u) + ((wmask >> 6) & 1u));
            const Acc area = std::sqrt(afx * afx + afy * afy + afz * afz);
            if (nn > Acc(0) && area > Acc(0)) {
                const Acc ex = nwx / nn, ey = nwy / nn, ez = nwz / nn;
                const Acc px = Acc(ux_o[c]), py = Acc(uy_o[c]), pz = Acc(uz_o[c]);
                const Acc un = px * ex + py * ey + pz * ez;
                Acc tx = px - un * ex, ty = py - un * ey, tz = pz - un * ez;
                const Acc ut = std::sqrt(tx * tx + ty * ty + tz * tz);
                if (ut > Acc(1e-7)) {
                    tx /= ut; ty /= ut; tz /= ut;
                    const Acc rw = Acc(rho_o[c]);
                    Acc D = 0, K = 0;
                    for (int i = 0; i < 19; ++i)
                        if (wmask & (1u << i)) {
                            const Acc ct = CX[i] * tx + CY[i] * ty + CZ[i] * tz;
                            D += ct * Acc(f[i]);
                            K += Acc(w[i]) * ct * ct;
                        }
                    if (K > Acc(0)) {
                        const Acc ur = Acc(uref[c]);
                        const Acc U = ur > Acc(0) ? ur : ut;
                        const Acc us0 = (Acc(-2) * D - rw * cf * U * U * area) / (Acc(6) * rw * K);
                        const Acc us = std::fmin(std::fmax(us0, -ut), ut);
                        for (int i = 0; i < 19; ++i)
                            if (wmask & (1u << i))
                                f[i] = Real(Acc(f[i]) + Acc(6) * Acc(w[i]) * rw * (CX[i] * tx + CY[i] * ty + CZ[i] * tz) * us);
                    }
                }
            }
        }
        // Moments with the Guo half-force shift (lbm3d.py:103-108).
        Acc r = 0, mx = 0, my = 0, mz = 0;
        for (int i = 0; i < 19; ++i) {
            r += Acc(f[i]); mx += Acc(f[i]) * CX[i]; my += Acc(f[i]) * CY[i]; mz += Acc(f[i]) * CZ[i];
        }
        const Acc fx = Acc(Fx[c]), fy = Acc(Fy[c]), fz = Acc(Fz[c]);
        const Acc ux = (mx + Acc(0.5) * fx) / r, uy = (my + Acc(0.5) * fy) / r, uz = (mz + Acc(0.5) * fz) / r;
        const Acc usq = ux * ux + uy * uy + uz * uz;
        // Non-equilibrium stress (lbm3d.py:110-119).
        Acc pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
        Real feq[19];
        for (int i = 0; i < 19; ++i) {
            const Acc cu = Acc(3) * (CX[i] * ux + CY[i] * uy + CZ[i] * uz);
            feq[i] = Real(Acc(w[i]) * r * (Acc(1) + cu + Acc(0.5) * cu * cu - Acc(1.5) * usq));
            const Real fn = Real(f[i] - feq[i]);
            pxx += (CX[i] * CX[i]) * Acc(fn); pyy += (CY[i] * CY[i]) * Acc(fn); pzz += (CZ[i] * CZ[i]) * Acc(fn);
            pxy += (CX[i] * CY[i]) * Acc(fn); pxz += (CX[i] * CZ[i]) * Acc(fn); pyz += (CY[i] * CZ[i]) * Acc(fn);
        }
        // Optional export of the collision's moments (checks of the learned closure):
        // rho, u (with the half-force shift), Pi_neq xx yy zz xy xz yz.
        if (mom) {
            const Acc v[10] = {r, ux, uy, uz, pxx, pyy, pzz, pxy, pxz, pyz};
            for (int k = 0; k < 10; ++k) mom[(int64_t)k * N + c] = v[k];
        }
        // Smagorinsky relaxation time (lbm3d.py:120-123); with cs2f, a per-cell
        // coefficient in place of Cs^2 (the learned closure's C^2 field).
        const Acc cs2 = cs2f ? cs2f[c] : csmag2;
        const Acc qn = std::sqrt(pxx * pxx + pyy * pyy + pzz * pzz + Acc(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
        const Acc tau = Acc(0.5) * (tau0 + std::sqrt(tau0 * tau0 + Acc(18) * Acc(1.41421356) * cs2 * qn / r));
        const Acc om = Acc(1) / tau;
        const Acc sf = Acc(1) - Acc(0.5) * om;
        // Regularized post-collision state with Guo source (lbm3d.py:124-136).
        const Acc third = Acc(1) / Acc(3);
        for (int i = 0; i < 19; ++i) {
            const Acc cxx = CX[i] * CX[i] - third, cyy = CY[i] * CY[i] - third, czz = CZ[i] * CZ[i] - third;
            const Acc fneq = Acc(4.5) * Acc(w[i]) * (cxx * pxx + cyy * pyy + czz * pzz
                               + Acc(2) * ((CX[i] * CY[i]) * pxy + (CX[i] * CZ[i]) * pxz + (CY[i] * CZ[i]) * pyz));
            const Acc cu = CX[i] * ux + CY[i] * uy + CZ[i] * uz;
            const Acc cF = CX[i] * fx + CY[i] * fy + CZ[i] * fz;
            const Acc src = sf * Acc(w[i]) * (Acc(3) * ((CX[i] - ux) * fx + (CY[i] - uy) * fy + (CZ[i] - uz) * fz)
                                              + Acc(9) * cu * cF);
            const Acc fneq1 = Acc(-1.5) * Acc(w[i]) * cF;
            fd[(int64_t)i * N + c] = Real(Acc(feq[i]) + (Acc(1) - om) * (fneq + fneq1) + src);
        }
        rho_o[c] = Real(r); ux_o[c] = Real(ux); uy_o[c] = Real(uy); uz_o[c] = Real(uz);  // lbm3d.py:137
    }

    static void step(const Real* fs, Real* fd, const uint8_t* solid,
                     const Real* Fx, const Real* Fy, const Real* Fz,
                     Real* rho_o, Real* ux_o, Real* uy_o, Real* uz_o,
                     const Real* cfw, const Real* uref, int nx, int ny, int nz, Acc tau0, Acc csmag2,
                     const Acc* cs2f = nullptr, Acc* mom = nullptr) {
        const int64_t N = (int64_t)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x)
                    cell(fs, fd, solid, Fx, Fy, Fz, rho_o, ux_o, uy_o, uz_o, cfw, uref,
                         nx, ny, nz, N, x, y, z, tau0, csmag2, cs2f, mom);
    }

    // lbm_bc (lbm3d.py:145-175; Numba lbm3d.py:301-328): inlet velocity and outlet
    // pressure by non-equilibrium extrapolation on the post-collision field.
    static void bc(Real* fd, const uint8_t* solid, const Real* uin, int nx, int ny, int nz, Acc rho_out) {
        const int64_t N = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < ny * nz; ++j) {
            const int y = j % ny, z = j / ny;
            const int64_t row = (int64_t)nx * (y + (int64_t)ny * z);
            for (int side = 0; side < 2; ++side) {
                const int64_t c0 = side == 0 ? row : row + nx - 1;
                const int64_t c1 = side == 0 ? row + 1 : row + nx - 2;
                if (solid[c0] || solid[c1]) continue;
                Acc r = 0, mx = 0, my = 0, mz = 0;
                for (int i = 0; i < 19; ++i) {
                    const Acc v = Acc(fd[(int64_t)i * N + c1]);
                    r += v; mx += v * CX[i]; my += v * CY[i]; mz += v * CZ[i];
                }
                const Acc ux = mx / r, uy = my / r, uz = mz / r;
                for (int i = 0; i < 19; ++i) {
                    const Acc e = side == 0 ? feq_i(i, r, Acc(uin[j]), Acc(0), Acc(0))
                                            : feq_i(i, rho_out, ux, uy, uz);
                    fd[(int64_t)i * N + c0] = Real(e + (Acc(fd[(int64_t)i * N + c1]) - feq_i(i, r, ux, uy, uz)));
                }
            }
        }
    }
};

}  // namespace

extern "C" {

// Equilibrium initial state at rho = 1, u = (u0, 0, 0) (lbm3d.py:362-364, 377-384).
// The reference evaluates this in numpy float64 and stores float32; all three
// instantiations here do the same and round once to the storage type.
int orc_init_f64(double* f, const double* u0, int64_t N) {
    const double W[3] = {1.0 / 3, 1.0 / 18, 1.0 / 36};
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < N; ++c)
        for (int i = 0; i < 19; ++i) {
            const double w = i == 0 ? W[0] : (i < 7 ? W[1] : W[2]);
            const double ux = u0[c];
            const double cu = 3.0 * CX[i] * ux;
            f[(int64_t)i * N + c] = w * (1.0 + cu + 0.5 * cu * cu - 1.5 * ux * ux);
        }
    return 0;
}

int orc_init_f32(float* f, const float* u0, int64_t N) {
    const double W[3] = {1.0 / 3, 1.0 / 18, 1.0 / 36};
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < N; ++c)
        for (int i = 0; i < 19; ++i) {
            const double w = i == 0 ? W[0] : (i < 7 ? W[1] : W[2]);
            const double ux = (double)u0[c];
            const double cu = 3.0 * CX[i] * ux;
            f[(int64_t)i * N + c] = (float)(w * (1.0 + cu + 0.5 * cu * cu - 1.5 * ux * ux));
        }
    return 0;
}

int orc_step_f64(const double* fs, double* fd, const uint8_t* solid, const double* Fx, const double* Fy,
                 const double* Fz, double* rho, double* ux, double* uy, double* uz, const double* cfw,
                 const double* uref, int nx, int ny, int nz, double tau0, double csmag2) {
    Lbm<double, double>::step(fs, fd, solid, Fx, Fy, Fz, rho, ux, uy, uz, cfw, uref, nx, ny, nz, tau0, csmag2);
    return 0;
}

// FP64 step with a per-cell coefficient field cs2f (NULL: csmag2 everywhere) and
// an optional export mom[10][N] of each cell's collision moments.
int orc_step_field_f64(const double* fs, double* fd, const uint8_t* solid, const double* Fx, const double* Fy,
                       const double* Fz, double* rho, double* ux, double* uy, double* uz, const double* cfw,
                       const double* uref, int nx, int ny, int nz, double tau0, double csmag2, const double* cs2f,
                       double* mom) {
    Lbm<double, double>::step(fs, fd, solid, Fx, Fy, Fz, rho, ux, uy, uz, cfw, uref, nx, ny, nz, tau0, csmag2,
                              cs2f, mom);
    return 0;
}

int orc_step_f32(const float* fs, float* fd, const uint8_t* solid, const float* Fx, const float* Fy,
                 const float* Fz, float* rho, float* ux, float* uy, float* uz, const float* cfw,
                 const float* uref, int nx, int ny, int nz, float tau0, float csmag2) {
    Lbm<float, float>::step(fs, fd, solid, Fx, Fy, Fz, rho, ux, uy, uz, cfw, uref, nx, ny, nz, tau0, csmag2);
    return 0;
}

int orc_step_mixed(const float* fs, float* fd, const uint8_t* solid, const float* Fx, const float* Fy,
                   const float* Fz, float* rho, float* ux, float* uy, float* uz, const float* cfw,
                   const float* uref, int nx, int ny, int nz, float tau0, float csmag2) {
    Lbm<float, double>::step(fs, fd, solid, Fx, Fy, Fz, rho, ux, uy, uz, cfw, uref, nx, ny, nz,
                             (double)tau0, (double)csmag2);
    return 0;
}

int orc_bc_f64(double* fd, const uint8_t* solid, const double* uin, int nx, int ny, int nz, double rho_out) {
    Lbm<double, double>::bc(fd, solid, uin, nx, ny, nz, rho_out);
    return 0;
}

int orc_bc_f32(float* fd, const uint8_t* solid, const float* uin, int nx, int ny, int nz, float rho_out) {
    Lbm<float, float>::bc(fd, solid, uin, nx, ny, nz, rho_out);
    return 0;
}

int orc_bc_mixed(float* fd, const uint8_t* solid, const float* uin, int nx, int ny, int nz, float rho_out) {
    Lbm<float, double>::bc(fd, solid, uin, nx, ny, nz, (double)rho_out);
    return 0;
}

// Run3D._update_uref (reach3d.py:412-418): depth-averaged column speed written to
// the bed/bank-adjacent cells, zero elsewhere. Column sums run z = 0..nz-1.
int orc_uref_f32(const float* ux, const float* uy, const uint8_t* solid, const uint8_t* bedwall, float* uref,
                 int nx, int ny, int nz) {
    const int64_t plane = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int64_t col = 0; col < plane; ++col) {
        float sx = 0.f, sy = 0.f, nfl = 0.f;
        for (int z = 0; z < nz; ++z) {
            const int64_t c = col + plane * z;
            const float fl = solid[c] ? 0.f : 1.f;
            sx += ux[c] * fl; sy += uy[c] * fl; nfl += fl;
        }
        const float ncol = nfl > 1.f ? nfl : 1.f;
        const float U = std::sqrt(sx * sx + sy * sy) / ncol;
        for (int z = 0; z < nz; ++z) {
            const int64_t c = col + plane * z;
            uref[c] = bedwall[c] ? U : 0.f;
        }
    }
    return 0;
}

int orc_uref_f64(const double* ux, const double* uy, const uint8_t* solid, const uint8_t* bedwall, double* uref,
                 int nx, int ny, int nz) {
    const int64_t plane = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int64_t col = 0; col < plane; ++col) {
        double sx = 0.0, sy = 0.0, nfl = 0.0;
        for (int z = 0; z < nz; ++z) {
            const int64_t c = col + plane * z;
            const double fl = solid[c] ? 0.0 : 1.0;
            sx += ux[c] * fl; sy += uy[c] * fl; nfl += fl;
        }
        const double ncol = nfl > 1.0 ? nfl : 1.0;
        const double U = std::sqrt(sx * sx + sy * sy) / ncol;
        for (int z = 0; z < nz; ++z) {
            const int64_t c = col + plane * z;
            uref[c] = bedwall[c] ? U : 0.0;
        }
    }
    return 0;
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
