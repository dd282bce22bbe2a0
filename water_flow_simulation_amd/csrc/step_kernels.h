// Fused collide-stream step for the river-reach D3Q19 solver (gfx950).
//
// One work-item per lattice cell, x fastest. A work-group is one x-run of a
// single (y, z) row, so every population load and store of a wavefront is a
// contiguous 64-cell run (pull streaming: the x-shift of a neighbour row only
// moves the run by one element). Populations are structure-of-arrays
// f[i*N + cell], read from `fs` and written to `fd` (A-B ping-pong), with
// cell = x + px*(y + ny*z): rows are padded to a pitch px that is a multiple
// of 64 cells, so every row starts on a 256-byte boundary (padding cells are
// flagged solid and never touched by the step).
//
// Per cell (floodsim/lbm3d.py:38-137 describes the same update):
//   1. pull the 19 populations; x-edge cells keep their own value for sources
//      outside [0, nx); sources above the lid take the mirrored direction from
//      the same layer; wall links (out of the y/z range or a solid source)
//      bounce back the cell's own opposite population;
//   2. wall model: moving-wall bounce-back whose tangential slip makes the
//      momentum given to the wall per step equal rho*cf*U^2*area;
//   3. moments with the Guo half-force shift; non-equilibrium stress;
//      Smagorinsky relaxation time; regularized BGK with the Guo source.
//
// Floating-point contract: every expression keeps the operand order of the
// update it restates, and the library is compiled with -ffp-contract=off, so
// the FP32 instantiation rounds exactly like a scalar IEEE single-precision
// evaluation of the same formulas. Products with a zero lattice coefficient
// are dropped (x + 0*y == x), which changes no result bit apart from the sign
// of an exact zero.
//
// Only what each cell needs is read: the flag byte always; the 32-bit
// wall-link mask only next to walls; forces only inside actuator disks; the
// previous-step macroscopic values and uref only where the wall model acts.
// Macroscopic outputs are written where the wall model needs them next step,
// and everywhere on steps that end a host call (write_all).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include "lattice.h"

namespace wfs {

// ---------------------------------------------------------------- storage codecs
// Populations are stored as T and computed in R.
struct CodecF32 {
    using T = float;
    using R = float;
    template <int I>
    __device__ static R ld(const T* p) { return *p; }
    template <int I>
    __device__ static void st(T* p, R v) { *p = v; }
};

struct CodecF64 {
    using T = double;
    using R = double;
    template <int I>
    __device__ static R ld(const T* p) { return *p; }
    template <int I>
    __device__ static void st(T* p, R v) { *p = v; }
};

// 16-bit deviation storage: T holds (f_i - w_i) * 2^SHIFT, decoded as
// w_i + t * 2^-SHIFT. The power-of-two scale is exact; it lifts the small
// deviations out of the FP16 subnormal range. Arithmetic stays FP32.
constexpr int kDevShift = 8;

struct CodecF16Dev {
    using T = __half;
    using R = float;
    template <int I>
    __device__ static R ld(const T* p) {
        return weight<float>(I) + __half2float(*p) * (1.0f / float(1 << kDevShift));
    }
    template <int I>
    __device__ static void st(T* p, R v) {
        *p = __float2half_rn((v - weight<float>(I)) * float(1 << kDevShift));
    }
};

struct CodecBF16Dev {
    using T = __hip_bfloat16;
    using R = float;
    template <int I>
    __device__ static R ld(const T* p) {
        return weight<float>(I) + __bfloat162float(*p) * (1.0f / float(1 << kDevShift));
    }
    template <int I>
    __device__ static void st(T* p, R v) {
        *p = __float2bfloat16((v - weight<float>(I)) * float(1 << kDevShift));
    }
};

// ---------------------------------------------------------------- helpers
// cx*a + cy*b + cz*c for lattice coefficients in {-1, 0, 1}, summing the
// nonzero terms left to right (bitwise equal to the full expression).
template <int CXv, int CYv, int CZv, typename R>
__device__ __forceinline__ R dot3(R a, R b, R c) {
    const R ta = CXv > 0 ? a : -a;
    const R tb = CYv > 0 ? b : -b;
    const R tc = CZv > 0 ? c : -c;
    if constexpr (CXv != 0 && CYv != 0 && CZv != 0) return ta + tb + tc;
    else if constexpr (CXv != 0 && CYv != 0) return ta + tb;
    else if constexpr (CXv != 0 && CZv != 0) return ta + tc;
    else if constexpr (CYv != 0 && CZv != 0) return tb + tc;
    else if constexpr (CXv != 0) return ta;
    else if constexpr (CYv != 0) return tb;
    else if constexpr (CZv != 0) return tc;
    else return R(0);
}

// s + k*v for k in {-1, 0, 1}.
template <int K, typename R>
__device__ __forceinline__ R accum(R s, R v) {
    if constexpr (K > 0) return s + v;
    else if constexpr (K < 0) return s - v;
    else return s;
}

template <int... Is, typename F>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <typename R>
struct StepArgs {
    const uint8_t* __restrict__ flag;
    const uint32_t* __restrict__ wmask;
    const R* __restrict__ Fx;
    const R* __restrict__ Fy;
    const R* __restrict__ Fz;
    R* __restrict__ rho;
    R* __restrict__ ux;
    R* __restrict__ uy;
    R* __restrict__ uz;
    const R* __restrict__ cfw;
    const R* __restrict__ uref;
    R fscale;      // forces used are F * fscale (thrust ramp)
    R tau0;
    R tau0sq;      // tau0 * tau0
    R ksmag;       // (18 * 1.41421356) * Cs^2
    int nx, ny, nz;
    int px;        // row pitch in cells (nx rounded up to a multiple of 64)
    int rows;      // ny * nz
    int tiles;     // x-tiles per row (ceil(nx / blockDim.x))
    long long N;   // cells per population array, px * ny * nz
    int write_all;
};

// Work-group -> (row, x-tile) map. Work-groups are dispatched round-robin over
// the 8 XCDs (group L runs on XCD L % 8), each with its own L2. All x-tiles of
// a row get the same L % 8, so the lines a row's tiles share at their x-edges
// (the +-1 pull shift) and its partial-line writes meet in one L2.
__device__ __forceinline__ bool tile_of(int tiles, int rows, int& row, int& tile) {
    const unsigned L = blockIdx.x;
    const unsigned q = L >> 3;
    tile = int(q % unsigned(tiles));
    row = int((q / unsigned(tiles)) * 8u + (L & 7u));
    return row < rows;
}

// ---------------------------------------------------------------- the step
template <class Codec>
__global__ void __launch_bounds__(256) k_step(const typename Codec::T* __restrict__ fs,
                                              typename Codec::T* __restrict__ fd,
                                              StepArgs<typename Codec::R> a) {
    using R = typename Codec::R;
    using T = typename Codec::T;
    int row, tile;
    if (!tile_of(a.tiles, a.rows, row, tile)) return;
    const int x = tile * int(blockDim.x) + int(threadIdx.x);
    if (x >= a.nx) return;
    const int z = row / a.ny;
    const int px = a.px;
    const int plane = px * a.ny;
    const int c = x + px * row;
    const uint8_t fl = a.flag[c];
    if (fl & FLAG_SOLID) return;
    const uint32_t wm = (fl & FLAG_NEARWALL) ? a.wmask[c] : 0u;
    const bool xlo = x == 0, xhi = x == a.nx - 1, top = z == a.nz - 1;
    const long long N = a.N;

    // 1. pull streaming
    R f[Q];
    if (!(fl & FLAG_NEARWALL) && !xlo && !xhi && !top) {
        // interior: every source is the plain upstream neighbour
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            const int s = c - kCX[i] - px * kCY[i] - plane * kCZ[i];
            f[i] = Codec::template ld<i>(fs + (long long)i * N + s);
        });
    } else {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
            const bool xout = (cx == 1 && xlo) || (cx == -1 && xhi);
            const bool wall = (wm >> i) & 1u;
            int dir = i, s = c;
            if (xout) {
                dir = i;
                s = c;
            } else if (wall) {
                dir = kOPP[i];
                s = c;
            } else if (cz == -1 && top) {
                dir = kMIRZ[i];
                s = c - cx - px * cy;
            } else {
                s = c - cx - px * cy - plane * cz;
            }
            // i, OPP[i] and MIRZ[i] share one lattice weight, so one decoder serves all three.
            f[i] = Codec::template ld<i>(fs + (long long)dir * N + s);
        });
    }

    // 2. wall function
    if (wm != 0u && (fl & FLAG_CF)) {
        const R cf = a.cfw[c];
        int inx = 0, iny = 0, inz = 0;
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            if ((wm >> i) & 1u) { inx -= kCX[i]; iny -= kCY[i]; inz -= kCZ[i]; }
        });
        const R nwx = R(inx), nwy = R(iny), nwz = R(inz);
        const R nn = sqrt(nwx * nwx + nwy * nwy + nwz * nwz);
        const R afx = R(((wm >> 1) & 1u) + ((wm >> 2) & 1u));
        const R afy = R(((wm >> 3) & 1u) + ((wm >> 4) & 1u));
        const R afz = R(((wm >> 5) & 1u) + ((wm >> 6) & 1u));
        const R area = sqrt(afx * afx + afy * afy + afz * afz);
        if (nn > R(0) && area > R(0)) {
            const R ex = nwx / nn, ey = nwy / nn, ez = nwz / nn;
            const R px = a.ux[c], py = a.uy[c], pz = a.uz[c];
            const R un = px * ex + py * ey + pz * ez;
            R tx = px - un * ex, ty = py - un * ey, tz = pz - un * ez;
            const R ut = sqrt(tx * tx + ty * ty + tz * tz);
            if (ut > R(1e-7)) {
                tx /= ut; ty /= ut; tz /= ut;
                const R rw = a.rho[c];
                R D = R(0), K = R(0);
                static_for<Q>([&](auto I) {
                    constexpr int i = I;
                    if ((wm >> i) & 1u) {
                        const R ct = dot3<kCX[i], kCY[i], kCZ[i]>(tx, ty, tz);
                        D += ct * f[i];
                        K += weight<R>(i) * ct * ct;
                    }
                });
                if (K > R(0)) {
                    const R ur = a.uref[c];
                    const R U = ur > R(0) ? ur : ut;
                    const R us0 = (R(-2) * D - rw * cf * U * U * area) / (R(6) * rw * K);
                    const R us = fmin(fmax(us0, -ut), ut);
                    static_for<Q>([&](auto I) {
                        constexpr int i = I;
                        if ((wm >> i) & 1u) {
                            constexpr R c6w = R(6) * weight<R>(i);
                            f[i] = f[i] + c6w * rw * dot3<kCX[i], kCY[i], kCZ[i]>(tx, ty, tz) * us;
                        }
                    });
                }
            }
        }
    }

    // 3. moments (Guo half-force shift)
    R r = f[0], mx = R(0), my = R(0), mz = R(0);
    bool mx0 = true, my0 = true, mz0 = true;
    static_for<Q - 1>([&](auto J) {
        constexpr int i = J + 1;
        r = r + f[i];
        // first nonzero term starts the sum exactly (0 + v == v)
        if constexpr (kCX[i] != 0) { mx = mx0 ? (kCX[i] > 0 ? f[i] : -f[i]) : accum<kCX[i]>(mx, f[i]); mx0 = false; }
        if constexpr (kCY[i] != 0) { my = my0 ? (kCY[i] > 0 ? f[i] : -f[i]) : accum<kCY[i]>(my, f[i]); my0 = false; }
        if constexpr (kCZ[i] != 0) { mz = mz0 ? (kCZ[i] > 0 ? f[i] : -f[i]) : accum<kCZ[i]>(mz, f[i]); mz0 = false; }
    });
    const bool forced = (fl & FLAG_FORCED) != 0;
    R fx = R(0), fy = R(0), fz = R(0);
    R ux, uy, uz;
    if (forced) {
        fx = a.Fx[c] * a.fscale; fy = a.Fy[c] * a.fscale; fz = a.Fz[c] * a.fscale;
        ux = (mx + R(0.5) * fx) / r; uy = (my + R(0.5) * fy) / r; uz = (mz + R(0.5) * fz) / r;
    } else {
        ux = mx / r; uy = my / r; uz = mz / r;
    }
    const R usq = ux * ux + uy * uy + uz * uz;

    // non-equilibrium stress
    R feq[Q];
    R pxx = R(0), pyy = R(0), pzz = R(0), pxy = R(0), pxz = R(0), pyz = R(0);
    bool pxx0 = true, pyy0 = true, pzz0 = true, pxy0 = true, pxz0 = true, pyz0 = true;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
        const R cu = R(3) * dot3<cx, cy, cz>(ux, uy, uz);
        feq[i] = weight<R>(i) * r * (R(1) + cu + R(0.5) * cu * cu - R(1.5) * usq);
        const R fn = f[i] - feq[i];
        if constexpr (cx != 0) { pxx = pxx0 ? fn : pxx + fn; pxx0 = false; }
        if constexpr (cy != 0) { pyy = pyy0 ? fn : pyy + fn; pyy0 = false; }
        if constexpr (cz != 0) { pzz = pzz0 ? fn : pzz + fn; pzz0 = false; }
        if constexpr (cx * cy != 0) { pxy = pxy0 ? (cx * cy > 0 ? fn : -fn) : accum<cx * cy>(pxy, fn); pxy0 = false; }
        if constexpr (cx * cz != 0) { pxz = pxz0 ? (cx * cz > 0 ? fn : -fn) : accum<cx * cz>(pxz, fn); pxz0 = false; }
        if constexpr (cy * cz != 0) { pyz = pyz0 ? (cy * cz > 0 ? fn : -fn) : accum<cy * cz>(pyz, fn); pyz0 = false; }
    });

    // Smagorinsky relaxation time
    const R qn = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
    const R tau = R(0.5) * (a.tau0 + sqrt(a.tau0sq + a.ksmag * qn / r));
    const R om = R(1) / tau;
    const R om1 = R(1) - om;

    // regularized post-collision populations
    T* const out = fd + c;
    if (forced) {
        const R sf = R(1) - R(0.5) * om;
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
            constexpr R third = R(1) / R(3);
            constexpr R cxx = R(cx * cx) - third, cyy = R(cy * cy) - third, czz = R(cz * cz) - third;
            R sp = cxx * pxx + cyy * pyy + czz * pzz;
            if constexpr (cx * cy != 0) sp = sp + R(2) * (cx * cy > 0 ? pxy : -pxy);
            else if constexpr (cx * cz != 0) sp = sp + R(2) * (cx * cz > 0 ? pxz : -pxz);
            else if constexpr (cy * cz != 0) sp = sp + R(2) * (cy * cz > 0 ? pyz : -pyz);
            constexpr R c45w = R(4.5) * weight<R>(i);
            const R fneq = c45w * sp;
            const R cu = dot3<cx, cy, cz>(ux, uy, uz);
            const R cF = dot3<cx, cy, cz>(fx, fy, fz);
            const R src = sf * weight<R>(i) *
                          (R(3) * ((R(cx) - ux) * fx + (R(cy) - uy) * fy + (R(cz) - uz) * fz) + R(9) * cu * cF);
            constexpr R cm15w = R(-1.5) * weight<R>(i);
            const R fneq1 = cm15w * cF;
            Codec::template st<i>(out + (long long)i * N, feq[i] + om1 * (fneq + fneq1) + src);
        });
    } else {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
            constexpr R third = R(1) / R(3);
            constexpr R cxx = R(cx * cx) - third, cyy = R(cy * cy) - third, czz = R(cz * cz) - third;
            R sp = cxx * pxx + cyy * pyy + czz * pzz;
            if constexpr (cx * cy != 0) sp = sp + R(2) * (cx * cy > 0 ? pxy : -pxy);
            else if constexpr (cx * cz != 0) sp = sp + R(2) * (cx * cz > 0 ? pxz : -pxz);
            else if constexpr (cy * cz != 0) sp = sp + R(2) * (cy * cz > 0 ? pyz : -pyz);
            constexpr R c45w = R(4.5) * weight<R>(i);
            Codec::template st<i>(out + (long long)i * N, feq[i] + om1 * (c45w * sp));
        });
    }
    if (a.write_all || (wm != 0u && (fl & FLAG_CF))) {
        a.rho[c] = r; a.ux[c] = ux; a.uy[c] = uy; a.uz[c] = uz;
    }
}

// ---------------------------------------------------------------- inlet/outlet rows
// Non-equilibrium extrapolation on the post-collision field, one work-item
// per (y, z) row: inlet x = 0 takes rho from x = 1 and the prescribed
// velocity (uin, 0, 0); outlet x = nx-1 takes the velocity of x = nx-2 and
// density rho_out. Rows whose boundary or interior cell is solid are left.
template <class Codec>
__device__ __forceinline__ typename Codec::R feq_generic(int i, typename Codec::R r, typename Codec::R ux,
                                                         typename Codec::R uy, typename Codec::R uz) {
    using R = typename Codec::R;
    const R cu = R(3) * (R(kCX[i]) * ux + R(kCY[i]) * uy + R(kCZ[i]) * uz);
    return weight<R>(i) * r * (R(1) + cu + R(0.5) * cu * cu - R(1.5) * (ux * ux + uy * uy + uz * uz));
}

template <class Codec>
__global__ void __launch_bounds__(256) k_bc(typename Codec::T* __restrict__ fd, const uint8_t* __restrict__ flag,
                                            const typename Codec::R* __restrict__ uin, int nx, int px, int rows,
                                            long long N, typename Codec::R rho_out) {
    using R = typename Codec::R;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= rows) return;
    const long long row = (long long)px * j;  // j = y + ny*z
    for (int side = 0; side < 2; ++side) {
        const long long c0 = side == 0 ? row : row + nx - 1;
        const long long c1 = side == 0 ? row + 1 : row + nx - 2;
        if ((flag[c0] | flag[c1]) & FLAG_SOLID) continue;
        R v[Q];
        R r = R(0), mx = R(0), my = R(0), mz = R(0);
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            v[i] = Codec::template ld<i>(fd + (long long)i * N + c1);
            r += v[i];
            mx += v[i] * R(kCX[i]);
            my += v[i] * R(kCY[i]);
            mz += v[i] * R(kCZ[i]);
        });
        const R ux = mx / r, uy = my / r, uz = mz / r;
        const R uj = side == 0 ? uin[j] : R(0);
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            const R e = side == 0 ? feq_generic<Codec>(i, r, uj, R(0), R(0))
                                  : feq_generic<Codec>(i, rho_out, ux, uy, uz);
            Codec::template st<i>(fd + (long long)i * N + c0, e + (v[i] - feq_generic<Codec>(i, r, ux, uy, uz)));
        });
    }
}

}  // namespace wfs
