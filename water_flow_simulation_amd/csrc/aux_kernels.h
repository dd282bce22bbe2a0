// Set-up, cadence and readout kernels around the step (gfx950).
//
// Everything here is either once per case (initial state, cell flags), once
// per 20-step chunk (uref, time-mean sums) or once per readout (FP64
// reductions), so these kernels are simple and bandwidth-light; the step
// kernel in step_kernels.h is the hot path.
//
// Device arrays use the padded layout cell = x + px*(y + ny*z) (px >= nx, a
// multiple of 64); the host exchanges dense arrays (pitch nx) through
// k_pad / k_unpad.
#pragma once

#include <hip/hip_runtime.h>

#include "lattice.h"
#include "step_kernels.h"

namespace wfs {

// Where the host's dense cells (pitch nx, rows j = y + ny*z) live on the
// device. Dense layout (tmap = tcoord = null): cell = x + px*j, rows padded to
// px. Tile-sparse layout: tmap[(x / 64) + ntx*j] is the compact index of the
// 64-cell tile holding x (-1: an all-solid tile that is not stored) and
// tcoord[3t..3t+2] = (x-tile, y, z) of tile t; cell = t*64 + x % 64.
struct CellMap {
    int nx, px, ny, ntx;
    const int* tmap;
    const int* tcoord;
};

// Device cell of dense (x, row j), or -1 when its tile is not stored.
__device__ __forceinline__ long long cell_of(const CellMap& m, int x, long long j) {
    if (!m.tmap) return x + (long long)m.px * j;
    const int t = m.tmap[(x >> 6) + (long long)m.ntx * j];
    return t < 0 ? -1 : (long long)t * 64 + (x & 63);
}

// Dense (x, row j) of device cell c; x >= nx for padding cells.
__device__ __forceinline__ void dense_of(const CellMap& m, long long c, int& x, long long& j) {
    if (!m.tcoord) {
        j = c / m.px;
        x = int(c - j * m.px);
        return;
    }
    const long long t = c >> 6;
    x = m.tcoord[3 * t] * 64 + int(c & 63);
    j = m.tcoord[3 * t + 1] + (long long)m.ny * m.tcoord[3 * t + 2];
}

// In-place streaming: converts between the A-B layout (slot k holds f_k)
// and state S, both ways. The slots of each opposite pair are exchanged,
// except that a slot whose link enters from outside the box at this x-end row
// keeps its own value. S -> A-B -> S is exact. A-B -> S drops only the
// populations that leave the box through an x-end row (no step reads them);
// A-B -> S -> A-B then gives the entering population of the pair there.
// N cells, population direction stride Ns.
template <typename T>
__global__ void k_aa_exchange(T* __restrict__ f, long long N, long long Ns, CellMap m) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int x;
    long long j;
    dense_of(m, c, x, j);
    const bool xlo = x == 0, xhi = x == m.nx - 1;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (i % 2 == 1) {
            constexpr int o = kOPP[i];
            static_assert(o == i + 1, "opposite directions are adjacent pairs");
            const bool keep_i = (kCX[i] == 1 && xlo) || (kCX[i] == -1 && xhi);
            const bool keep_o = (kCX[o] == 1 && xlo) || (kCX[o] == -1 && xhi);
            T* pi = f + (long long)i * Ns + c;
            T* po = f + (long long)o * Ns + c;
            const T vi = *pi, vo = *po;
            *pi = keep_i ? vi : vo;
            *po = keep_o ? vo : vi;
        }
    });
}

// Host exchanges move a block of whole dense rows [j0, j1) at a time; the
// device cells of those rows are the range [c0, c1) (rows and, on the
// tile-sparse layout, tiles are stored in row order).
//
// dense rows j0.. -> device cells [c0, c1); padding cells get `fill`.
template <typename S, typename D>
__global__ void k_pad(const S* __restrict__ dense, D* __restrict__ dev, CellMap m, long long c0, long long c1,
                      long long j0, D fill) {
    const long long c = c0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    int x;
    long long j;
    dense_of(m, c, x, j);
    dev[c] = x < m.nx ? (D)dense[x + (long long)m.nx * (j - j0)] : fill;
}

// device cells -> dense cells [k0, k0 + n) (written from dense[0]); cells of
// tiles that are not stored get `fill`.
template <typename S, typename D>
__global__ void k_unpad(const S* __restrict__ dev, D* __restrict__ dense, CellMap m, long long k0, long long n,
                        D fill) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long k = k0 + i;
    const long long j = k / m.nx;
    const long long c = cell_of(m, int(k - j * m.nx), j);
    dense[i] = c < 0 ? fill : (D)dev[c];
}

// The same with `add` added in double before the one rounding to D (the
// time-mean sum of rho is kept as the sum of rho - 1; its readout adds the
// sample count back).
template <typename S, typename D>
__global__ void k_unpad_add(const S* __restrict__ dev, D* __restrict__ dense, CellMap m, long long k0, long long n,
                            double add) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long k = k0 + i;
    const long long j = k / m.nx;
    const long long c = cell_of(m, int(k - j * m.nx), j);
    dense[i] = (D)((c < 0 ? 0.0 : (double)dev[c]) + add);
}

// Equilibrium at rho = 1, u = (u0, 0, 0), evaluated in double and stored once
// through the codec; u0 = uin of the cell's row, 0 in solid cells
// (floodsim/lbm3d.py:362-364, 377-384). N cells, direction stride Ns.
template <class Codec>
__global__ void k_init(typename Codec::T* __restrict__ f, const uint8_t* __restrict__ flag,
                       const typename Codec::R* __restrict__ uin, CellMap m, long long N, long long Ns, Center ctr) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int x;
    long long j;  // row index y + ny*z
    dense_of(m, c, x, j);
    const double u = (flag[c] & FLAG_SOLID) ? 0.0 : (double)uin[j];
    const double W[3] = {1.0 / 3, 1.0 / 18, 1.0 / 36};
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        const double w = i == 0 ? W[0] : (i < 7 ? W[1] : W[2]);
        const double cu = 3.0 * kCX[i] * u;
        const double v = w * (1.0 + cu + 0.5 * cu * cu - 1.5 * u * u);
        Codec::template st<i>(f + (long long)i * Ns + c, (typename Codec::R)v, store_key<Codec>(ctr, i));
    });
}

// Cell flags and wall-link masks from the dense solid mask. A link i of fluid
// cell c is a wall link when its pull source lies inside [0, nx) in x but
// outside the y range, below the bed (z < 0) or on a solid cell; sources
// above the lid are mirrored into the same layer first
// (floodsim/lbm3d.py:50-58 applies the same rules per step). Padding cells
// (x >= nx) are flagged solid. One call covers the device cells [c0, c1) of
// whole z-planes; `solid` holds the dense planes from zlo on, one plane
// either side of them included.
template <typename R>
__global__ void k_flags(const uint8_t* __restrict__ solid, const R* __restrict__ cfw, uint8_t* __restrict__ flag,
                        uint32_t* __restrict__ wmask, CellMap m, int nz, long long c0, long long c1, int zlo) {
    const long long c = c0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    int x;
    long long j;
    dense_of(m, c, x, j);
    const int nx = m.nx, ny = m.ny;
    const int y = int(j % ny), z = int(j / ny);
    const long long s0 = (long long)nx * ny * zlo;   // dense cell of solid[0]
    uint8_t fl = 0;
    uint32_t wm = 0;
    if (x >= nx || solid[x + (long long)nx * j - s0]) {
        fl = FLAG_SOLID;
    } else {
        for (int i = 0; i < Q; ++i) {
            const int xs = x - kCX[i], ys = y - kCY[i];
            int zs = z - kCZ[i];
            if (xs < 0 || xs >= nx) continue;
            if (zs >= nz) zs = z;
            bool wall = ys < 0 || ys >= ny || zs < 0;
            if (!wall) wall = solid[xs + (long long)nx * (ys + (long long)ny * zs) - s0] != 0;
            if (wall) wm |= 1u << i;
        }
        if (wm) fl |= FLAG_NEARWALL;
        if (cfw[c] > R(0)) fl |= FLAG_CF;
    }
    flag[c] = fl;
    wmask[c] = wm;
}

// FLAG_FORCED where any force component of the group side table
// (StepAux::fslot / fval) is nonzero.
template <typename R>
__global__ void k_forced(uint8_t* __restrict__ flag, const int* __restrict__ fslot, const R* __restrict__ fval,
                         long long N) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    uint8_t fl = flag[c] & uint8_t(~FLAG_FORCED);
    const int s = fslot[c / kForceGroup];
    if (!(fl & FLAG_SOLID) && s >= 0) {
        const R* v = fval + (long long)s * kForceSlot + c % kForceGroup;
        if (v[0] != R(0) || v[kForceGroup] != R(0) || v[2 * kForceGroup] != R(0)) fl |= FLAG_FORCED;
    }
    flag[c] = fl;
}

// cell_of for the column kernels, whose work-groups are whole 64-lane
// x-tiles of one row (blockDim.x a multiple of 64, grid x-origin 0): every
// lane of a wavefront reads the same tile-map entry, so the entry is read
// once as a uniform value and a tile that is not stored is skipped by a
// scalar branch.
__device__ __forceinline__ long long column_cell(const CellMap& m, int x, long long j) {
    if (!m.tmap) return x + (long long)m.px * j;
    const int tx = __builtin_amdgcn_readfirstlane(x) >> 6;
    const int t = __builtin_amdgcn_readfirstlane(m.tmap[tx + (long long)m.ntx * j]);
    return t < 0 ? -1 : (long long)t * 64 + (x & 63);
}

// Depth-averaged column speed for the bed/bank wall model
// (floodsim/reach3d.py:412-418): one work-item per (x, y) column, sums over
// z = 0..nz-1 in order.
template <typename R>
__device__ __forceinline__ void uref_column(const R* __restrict__ ux, const R* __restrict__ uy,
                                            const uint8_t* __restrict__ flag, R* __restrict__ uref, CellMap m,
                                            int nz) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= m.nx) return;
    R sx = R(0), sy = R(0), nfl = R(0);
    for (int z = 0; z < nz; ++z) {
        const long long c = column_cell(m, x, y + (long long)m.ny * z);
        if (c < 0) continue;   // tile not stored: solid
        const R fl = (flag[c] & FLAG_SOLID) ? R(0) : R(1);
        sx += ux[c] * fl;
        sy += uy[c] * fl;
        nfl += fl;
    }
    const R ncol = nfl > R(1) ? nfl : R(1);
    const R U = sqrt(sx * sx + sy * sy) / ncol;
    for (int z = 0; z < nz; ++z) {
        const long long c = column_cell(m, x, y + (long long)m.ny * z);
        if (c >= 0) uref[c] = (flag[c] & FLAG_BEDWALL) ? U : R(0);
    }
}

template <typename R>
__global__ void k_uref(const R* __restrict__ ux, const R* __restrict__ uy, const uint8_t* __restrict__ flag,
                       R* __restrict__ uref, CellMap m, int nz) {
    uref_column(ux, uy, flag, uref, m, nz);
}

// FLAG_BEDWALL from a per-cell byte mask (nonzero: bed or bank cell of the
// wall model).
__global__ void k_bedwall_flag(uint8_t* __restrict__ flag, const uint8_t* __restrict__ bw, long long N) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const uint8_t fl = flag[c] & uint8_t(~FLAG_BEDWALL);
    flag[c] = bw[c] ? uint8_t(fl | FLAG_BEDWALL) : fl;
}

// Fluid cells per column (at least 1), for the per-step column velocity.
__global__ void k_colcount(const uint8_t* __restrict__ flag, uint8_t* __restrict__ count, CellMap m, int nz) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= m.nx) return;
    int n = 0;
    for (int z = 0; z < nz; ++z) {
        const long long c = cell_of(m, x, y + (long long)m.ny * z);
        if (c >= 0 && !(flag[c] & FLAG_SOLID)) ++n;
    }
    count[x + (long long)m.px * y] = uint8_t(n > 1 ? n : 1);
}

// The column velocity's first input: the current velocity of fluid cells,
// zero elsewhere.
template <typename R>
__global__ void k_colu_init(const uint8_t* __restrict__ flag, const R* __restrict__ ux, const R* __restrict__ uy,
                            R* __restrict__ cx, R* __restrict__ cy, long long N) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const bool fluid = !(flag[c] & FLAG_SOLID);
    cx[c] = fluid ? ux[c] : R(0);
    cy[c] = fluid ? uy[c] : R(0);
}

// Running sums for the time means (floodsim/reach3d.py:443-451).
template <typename R>
__global__ void k_accumulate(const R* __restrict__ rho, const R* __restrict__ ux, const R* __restrict__ uy,
                             const R* __restrict__ uz, R* __restrict__ s_ux, R* __restrict__ s_uy,
                             R* __restrict__ s_uz, R* __restrict__ s_rho, R* __restrict__ s_sp,
                             R* __restrict__ s_uxx, long long N) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const R a = ux[c], b = uy[c], d = uz[c];
    s_ux[c] += a;
    s_uy[c] += b;
    s_uz[c] += d;
    s_rho[c] += rho[c] - R(1);   // rho - 1: the level signal (1e-5..1e-4) survives the float sum
    s_sp[c] += sqrt(a * a + b * b + d * d);
    s_uxx[c] += a * a;
}

// ---------------------------------------------------------------- batched cadence
// Per-case run scalars kept on the device by a batch (the state of
// floodsim/reach3d.py Run3D.advance between chunks: time, ramp, averaging),
// and the case's step budget: each case stops when its own budget is spent,
// so cases of uneven length (each run to its own end time) share launches.
template <typename R>
struct CaseSched {
    double t, dt, t_ramp, t_avg;
    long long steps_left;   // step budget; negative: no limit
    long long steps_taken;  // steps taken since the host last read them
    R fscale;          // force scale used by the step
    int ramp_done;
    int navg;
    int accumulate;    // this chunk adds to the time-mean sums
    int active;        // steps this case takes in the current chunk (0: stopped)
    int flip;          // the case reads buffer f[parity ^ flip] (steps skipped while stopped)
};

// Before a chunk of m steps: how many of them the case takes (its budget),
// and for a running case the thrust ramp F = F0 * s, s = smoothstep(min(t /
// t_ramp, 1)) evaluated in double and rounded to float (reach3d.py:431-437).
// af[2k], af[2k+1] mirror case k's active and flip for the step kernels (a
// flat array they read beside the case's arguments).
template <typename R>
__global__ void k_batch_pre(CaseSched<R>* __restrict__ sch, int* __restrict__ af, int n, int m) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    CaseSched<R> c = sch[k];
    c.active = c.steps_left < 0 ? m : (int)(c.steps_left < m ? c.steps_left : m);
    af[2 * k] = c.active;
    af[2 * k + 1] = c.flip;
    if (c.active > 0 && !c.ramp_done) {
        double f = c.t / c.t_ramp;
        f = f < 1.0 ? f : 1.0;
        const float s = (float)(f * f * (3.0 - 2.0 * f));
        c.fscale = R(s);
        c.ramp_done = s >= 1.0f;
    }
    sch[k] = c;
}

// After a chunk of m steps, for a case that took a > 0 of them: t += a dt,
// and the time-mean sums take this state once t reaches the averaging window
// (reach3d.py:441-451). The m - a launches a stopping case sat out flipped
// the batch parity without it, so its buffer parity is corrected.
template <typename R>
__global__ void k_batch_post(CaseSched<R>* __restrict__ sch, int* __restrict__ af, int n, int m) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    CaseSched<R> c = sch[k];
    const int a = c.active;
    c.flip ^= (m - a) & 1;
    af[2 * k + 1] = c.flip;
    c.accumulate = 0;
    if (a > 0) {
        c.t += a * c.dt;
        c.accumulate = c.t >= c.t_avg;
        if (c.accumulate) c.navg += 1;
        c.steps_taken += a;
        if (c.steps_left > 0) c.steps_left -= a;
    }
    sch[k] = c;
}

// k_uref for case k of a batch, skipped when the case takes no step in
// this chunk (a stopped case keeps the uref of its last chunk, as a
// sequential run would).
template <typename R>
__global__ void k_uref_if(const CaseSched<R>* __restrict__ sch, int k, const R* __restrict__ ux,
                          const R* __restrict__ uy, const uint8_t* __restrict__ flag, R* __restrict__ uref,
                          CellMap m, int nz) {
    if (sch[k].active <= 0) return;
    uref_column(ux, uy, flag, uref, m, nz);
}

// k_accumulate for one case, skipped when its schedule says so.
template <typename R>
__global__ void k_accumulate_if(const CaseSched<R>* __restrict__ sch, int k, const R* __restrict__ rho,
                                const R* __restrict__ ux, const R* __restrict__ uy, const R* __restrict__ uz,
                                R* __restrict__ s_ux, R* __restrict__ s_uy, R* __restrict__ s_uz,
                                R* __restrict__ s_rho, R* __restrict__ s_sp, R* __restrict__ s_uxx, long long N) {
    if (!sch[k].accumulate) return;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const R a = ux[c], b = uy[c], d = uz[c];
    s_ux[c] += a;
    s_uy[c] += b;
    s_uz[c] += d;
    s_rho[c] += rho[c] - R(1);   // rho - 1: the level signal (1e-5..1e-4) survives the float sum
    s_sp[c] += sqrt(a * a + b * b + d * d);
    s_uxx[c] += a * a;
}

template <typename R>
__global__ void k_fill_macro(R* __restrict__ rho, R* __restrict__ ux, R* __restrict__ uy, R* __restrict__ uz,
                             long long N) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    rho[c] = R(1);
    ux[c] = R(0);
    uy[c] = R(0);
    uz[c] = R(0);
}

// Populations of dense cells [k0, k0 + n): padded storage (direction stride
// Npad) -> real array out[i*n + (k - k0)], and back.
template <class Codec, typename O>
__global__ void k_decode(const typename Codec::T* __restrict__ f, O* __restrict__ out, CellMap m,
                         long long Npad, long long k0, long long n, Center ctr) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const long long k = k0 + r;
    const long long j = k / m.nx;
    const long long c = cell_of(m, int(k - j * m.nx), j);
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        out[(long long)i * n + r] =
            c < 0 ? O(0) : (O)Codec::template ld<i>(f + (long long)i * Npad + c, ctr.c[center_class(i)]);
    });
}

template <class Codec>
__global__ void k_encode(typename Codec::T* __restrict__ f, const float* __restrict__ in, CellMap m,
                         long long Npad, long long k0, long long n, Center ctr) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const long long k = k0 + r;
    const long long j = k / m.nx;
    const long long c = cell_of(m, int(k - j * m.nx), j);
    if (c < 0) return;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        Codec::template st<i>(f + (long long)i * Npad + c, (typename Codec::R)in[(long long)i * n + r],
                              store_key<Codec>(ctr, i));
    });
}

// FP64 per-x sums over fluid cells of rows j in [j0, j1) (j = y + ny*z):
// pass 1 writes one partial per (x, chunk of rows), pass 2 adds the chunks
// in order. Deterministic (no atomics).
template <typename R>
__global__ void k_profile_partial(const R* __restrict__ v, const uint8_t* __restrict__ flag, CellMap m, int j0,
                                  int j1, int rows_per_chunk, double* __restrict__ partial) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= m.nx) return;
    const int chunk = blockIdx.y;
    const int a = j0 + chunk * rows_per_chunk;
    const int b = min(a + rows_per_chunk, j1);
    double s = 0.0;
    for (int j = a; j < b; ++j) {
        const long long c = cell_of(m, x, j);
        if (c >= 0 && !(flag[c] & FLAG_SOLID)) s += v ? (double)v[c] : 1.0;
    }
    partial[(long long)chunk * m.nx + x] = s;
}

__global__ void k_profile_final(const double* __restrict__ partial, int nx, int nchunks, double* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    double s = 0.0;
    for (int k = 0; k < nchunks; ++k) s += partial[(long long)k * nx + x];
    out[x] = s;
}

// Watchdog: counts fluid cells whose macroscopic state is not finite.
template <typename R>
__global__ void k_nonfinite(const R* __restrict__ rho, const R* __restrict__ ux, const R* __restrict__ uy,
                            const R* __restrict__ uz, const uint8_t* __restrict__ flag, long long N,
                            unsigned int* __restrict__ count) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    if (flag[c] & FLAG_SOLID) return;
    const bool bad = !isfinite(rho[c]) || !isfinite(ux[c]) || !isfinite(uy[c]) || !isfinite(uz[c]);
    if (bad) atomicAdd(count, 1u);
}

}  // namespace wfs

namespace wfs {

// ---------------------------------------------------------------- FP64 momentum budget
// Forces of one step, in lattice units, summed in double:
//   out[0..2] applied body force (thrust): F * fscale over forced fluid cells;
//   out[3..5] force of the fluid on the cells marked in `body` (hulls);
//   out[6..8] force on every other wall (bed, banks, domain edges).
// For a wall link i of a fluid cell (its pull source x - c_i is a wall), the
// next step returns the population that leaves toward the wall, g =
// f_opp(i)(x) as stored, as f_i = g + delta (bounce-back plus the wall
// function's slip correction, evaluated here by the step's own gather), so
// the wall receives the momentum c_opp(i) (g + f_i). Each block writes its
// nine partial sums; k_sum_parts adds the blocks in a fixed order.
template <class Codec, bool kSparse>
__device__ __forceinline__ void wall_force_cell(const typename Codec::T* __restrict__ fs,
                                                const StepArgs<typename Codec::R>& a, CellMap cm,
                                                const uint32_t* __restrict__ body, long long c, double* acc) {
    using R = typename Codec::R;
    if (c < a.N) {
        int x;
        long long j;
        dense_of(cm, c, x, j);
        const int row = int(j);
        const int z = row / a.ny, y = row - z * a.ny;
        const uint8_t fl = x < a.nx ? a.flag[c] : uint8_t(FLAG_SOLID);
        if (!(fl & FLAG_SOLID)) {
            if (fl & FLAG_FORCED) {
                R fx, fy, fz;
                body_force(*a.aux, int(c), a.fscale, fx, fy, fz);
                acc[0] += double(fx);
                acc[1] += double(fy);
                acc[2] += double(fz);
            }
            if (fl & FLAG_NEARWALL) {
                const uint32_t wm = a.wmask[c];
                R f[Q];
                gather<Codec, kSparse>(fs, a, int(c), x, y, z, fl, wm, f, kSparse ? a.nbr + 27 * (c >> 6) : nullptr,
                                       int(c & 63));
                static_for<Q>([&](auto I) {
                    constexpr int i = I;
                    if ((wm >> i) & 1u) {
                        constexpr int o = kOPP[i];
                        const R g = Codec::template ld<o>(fs + (long long)o * a.N + c, a.ctr.c[center_class(o)]);
                        const double m = double(g) + double(f[i]);
                        const int xs = x - kCX[i], ys = y - kCY[i], zs = z - kCZ[i];
                        const bool in = xs >= 0 && xs < a.nx && ys >= 0 && ys < a.ny && zs >= 0 && zs < a.nz;
                        // the wall cell's bit in the dense body mask
                        const long long d = xs + (long long)a.nx * (ys + (long long)a.ny * zs);
                        const bool hb = in && body && ((body[d >> 5] >> (d & 31)) & 1u);
                        const int k = hb ? 3 : 6;
                        acc[k] -= kCX[i] * m;
                        acc[k + 1] -= kCY[i] * m;
                        acc[k + 2] -= kCZ[i] * m;
                    }
                });
            }
        }
    }
}

// The nine sums of a 256-thread work-group in a fixed LDS tree; thread 0..8 write part[g * 9 + k].
__device__ __forceinline__ void block_sum9(const double* acc, double* __restrict__ part, long long g) {
    __shared__ double red[9][256];
#pragma unroll
    for (int k = 0; k < 9; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = int(blockDim.x) / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
#pragma unroll
            for (int k = 0; k < 9; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 9) part[g * 9 + threadIdx.x] = red[threadIdx.x][0];
}

// One 256-cell block per work-group, partials per block.
template <class Codec, bool kSparse>
__global__ void __launch_bounds__(256) k_wall_force(const typename Codec::T* __restrict__ fs,
                                                    StepArgs<typename Codec::R> a, CellMap cm,
                                                    const uint32_t* __restrict__ body, double* __restrict__ part) {
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    wall_force_cell<Codec, kSparse>(fs, a, cm, body, (long long)blockIdx.x * blockDim.x + threadIdx.x, acc);
    block_sum9(acc, part, blockIdx.x);
}

// The same sums with a fixed number of work-groups: group g takes the
// 256-cell blocks g, g + gridDim.x, ..., so the partials (gridDim.x x 9)
// do not grow with the lattice.
template <class Codec, bool kSparse>
__global__ void __launch_bounds__(256) k_wall_force_strided(const typename Codec::T* __restrict__ fs,
                                                            StepArgs<typename Codec::R> a, CellMap cm,
                                                            const uint32_t* __restrict__ body,
                                                            double* __restrict__ part) {
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const long long nblk = (a.N + 255) / 256;
    for (long long b = blockIdx.x; b < nblk; b += gridDim.x)
        wall_force_cell<Codec, kSparse>(fs, a, cm, body, b * 256 + threadIdx.x, acc);
    block_sum9(acc, part, blockIdx.x);
}

// ---------------------------------------------------------------- particle tracer
// Forward Euler on the cell velocity, as the reference's tracer
// (floodsim/reach3d.py:500-536): P (metres from the domain corner, x y z per
// particle, FP64) moves by float32(float32(u * vel) * dt), u the velocity of
// the cell holding P (indices truncated, clamped to the box; 0 where no tile
// is stored), a non-finite u * vel counting as 0. A particle that ends
// outside [0, nx-1) x [0, ny) x [0, nz) or in a solid cell restarts at a
// random inlet cell (rows of x = 1 in `inlet`): x in [0, 2 dx), y and z
// uniform within the cell. The random numbers come from a counter-based hash
// of (seed, particle, step), so a run repeats exactly.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double unit_of(uint64_t h) { return double(h >> 11) * 0x1.0p-53; }
__device__ __forceinline__ float finite_or_zero(float v) { return isfinite(v) ? v : 0.f; }

template <typename R>
__global__ void k_particles(double* __restrict__ P, long long n, const R* __restrict__ ux, const R* __restrict__ uy,
                            const R* __restrict__ uz, const uint8_t* __restrict__ flag, CellMap m, int nz, double dx,
                            float vel, float dt, const int* __restrict__ inlet, long long ninlet, uint64_t seed,
                            long long step, unsigned long long* __restrict__ respawned) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double p[3] = {P[3 * k], P[3 * k + 1], P[3 * k + 2]};
    const int dims[3] = {m.nx, m.ny, nz};
    int ic[3];
    for (int a = 0; a < 3; ++a) {
        const double q = p[a] / dx;
        const double t = isfinite(q) ? fmin(fmax(trunc(q), 0.0), double(dims[a] - 1)) : 0.0;
        ic[a] = int(t);
    }
    const long long c = cell_of(m, ic[0], ic[1] + (long long)m.ny * ic[2]);
    float u[3] = {0.f, 0.f, 0.f};
    if (c >= 0) {
        u[0] = float(ux[c]);
        u[1] = float(uy[c]);
        u[2] = float(uz[c]);
    }
    for (int a = 0; a < 3; ++a) p[a] += double(finite_or_zero(u[a] * vel) * dt);
    // leaving the box or entering a solid cell: restart at the inlet
    bool bad = false;
    long long jc[3];
    for (int a = 0; a < 3; ++a) {
        if (!isfinite(p[a])) { bad = true; break; }
        const double q = trunc(p[a] / dx);
        if (q < 0.0 || q >= double(a == 0 ? dims[a] - 1 : dims[a])) { bad = true; break; }
        jc[a] = (long long)q;
    }
    if (!bad) {
        const long long cc = cell_of(m, int(jc[0]), jc[1] + (long long)m.ny * jc[2]);
        bad = cc < 0 || (flag[cc] & FLAG_SOLID);
    }
    if (bad && ninlet > 0) {
        uint64_t h = mix64(seed ^ mix64(uint64_t(k) * 0x100000001b3ull + uint64_t(step)));
        const long long row = inlet[h % uint64_t(ninlet)];
        const long long y = row % m.ny, z = row / m.ny;
        h = mix64(h);
        p[0] = unit_of(h) * 2.0 * dx;
        h = mix64(h);
        p[1] = (double(y) + unit_of(h)) * dx;
        h = mix64(h);
        p[2] = (double(z) + unit_of(h)) * dx;
        atomicAdd(respawned, 1ull);
    }
    P[3 * k] = p[0];
    P[3 * k + 1] = p[1];
    P[3 * k + 2] = p[2];
}

// Fixed-order FP64 sum of per-block partials: out[g * NV + k] = sum of
// part[b * NV + k] over the b of chunk g, [g chunk, (g + 1) chunk). Thread t
// takes b = g chunk + t, t + 256, ...; the 256 thread sums meet in a fixed LDS
// tree. Two launches (chunks, then one work-group over the chunk sums) give a
// result that does not depend on timing.
template <int NV>
__global__ void __launch_bounds__(256) k_sum_parts(const double* __restrict__ part, long long blocks, long long chunk,
                                                   double* __restrict__ out) {
    __shared__ double red[NV][256];
    const int t = int(threadIdx.x);
    const long long b0 = (long long)blockIdx.x * chunk;
    const long long b1 = b0 + chunk < blocks ? b0 + chunk : blocks;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    for (long long b = b0 + t; b < b1; b += 256) {
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += part[b * NV + k];
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k][t] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int k = 0; k < NV; ++k) red[k][t] += red[k][t + w];
        }
        __syncthreads();
    }
    if (t < NV) out[(long long)blockIdx.x * NV + t] = red[t][0];
}
constexpr long long kSumChunks = 1024;   // work-groups of the first pass

}  // namespace wfs
