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

#include <type_traits>

#include "lattice.h"
#include "closure.h"

namespace wfs {

// ---------------------------------------------------------------- storage codecs
// Populations are stored as T and computed in R.
// Population stores carry the non-temporal hint: the next step reads them
// back only after the whole lattice (GBs) has streamed through the caches,
// so allocating them in L2 only evicts the pull halo lines still in use.
#ifndef WFS_NT_STORE
#define WFS_NT_STORE 1
#endif

template <typename T>
__device__ __forceinline__ void store_pop(T* p, T v) {
#if WFS_NT_STORE
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// Codec interface: dec decodes a stored value with the centre e of its
// direction (used by the 16-bit formats only), enc gives the value to store
// (its key: store_key below), ld/st are the same through a pointer, rt is the
// value a store followed by a load gives back.
struct CodecF32 {
    using T = float;
    using R = float;
    static constexpr bool kExact = true;   // operand order of the reference update, bit-exact
    template <int I>
    __device__ static R dec(T v, float) { return v; }
    template <int I>
    __device__ static T enc(R v, float) { return v; }
    template <int I>
    __device__ static R ld(const T* p, float) { return *p; }
    template <int I>
    __device__ static void st(T* p, R v, float) { store_pop(p, v); }
    template <int I>
    __device__ static R rt(R v, float, float) { return v; }
    static constexpr bool kStoreUsesN2 = false;
};

// FP32 storage with the pair-form collision (not bit-exact; judged by tolerance).
struct CodecF32Fast : CodecF32 {
    static constexpr bool kExact = false;
};

struct CodecF64 {
    using T = double;
    using R = double;
    static constexpr bool kExact = true;
    template <int I>
    __device__ static R dec(T v, float) { return v; }
    template <int I>
    __device__ static T enc(R v, float) { return v; }
    template <int I>
    __device__ static R ld(const T* p, float) { return *p; }
    template <int I>
    __device__ static void st(T* p, R v, float) { store_pop(p, v); }
    template <int I>
    __device__ static R rt(R v, float, float) { return v; }
    static constexpr bool kStoreUsesN2 = false;
};

// 16-bit deviation storage. FP16 holds 2 (f_i - e_i), decoded as
// e_i + t / 2; BF16 holds f_i - e_i. e_i is the equilibrium of the mean
// inlet flow, e_i = w_i (1 + 3 c_ix U0 + 4.5 (c_ix U0)^2 - 1.5 U0^2) (the
// lattice weight when U0 = 0). Centring on the mean flow leaves only the
// departures from it to the 11 (FP16) or 8 (BF16) significand bits. Small FP16
// departures fall into the subnormal range, whose absolute spacing is finer
// than that of the normal range the flow needs; the factor 2 (exact) only
// lets each FP16 decode and encode be one mixed-precision FMA whose constants
// are an inline 0.5 or 2 and the per-direction centre. Arithmetic is FP32.
struct CodecF16Dev {
    using T = __half;
    using R = float;
    static constexpr bool kExact = false;
    template <int I>
    __device__ static R dec(T v, float e) { return fmaf(__half2float(v), 0.5f, e); }
    // n2 = -2 e
    template <int I>
    __device__ static T enc(R v, float n2) { return __float2half_rn(fmaf(v, 2.0f, n2)); }
    template <int I>
    __device__ static R ld(const T* p, float e) { return dec<I>(*p, e); }
    template <int I>
    __device__ static void st(T* p, R v, float n2) {
        store_pop(reinterpret_cast<unsigned short*>(p), __half_as_ushort(enc<I>(v, n2)));
    }
    template <int I>
    __device__ static R rt(R v, float e, float n2) { return dec<I>(enc<I>(v, n2), e); }
    static constexpr bool kStoreUsesN2 = true;
};

struct CodecBF16Dev {
    using T = __hip_bfloat16;
    using R = float;
    static constexpr bool kExact = false;
    template <int I>
    __device__ static R dec(T v, float e) { return __bfloat162float(v) + e; }
    template <int I>
    __device__ static T enc(R v, float e) { return __float2bfloat16(v - e); }
    template <int I>
    __device__ static R ld(const T* p, float e) { return dec<I>(*p, e); }
    template <int I>
    __device__ static void st(T* p, R v, float e) {
        store_pop(reinterpret_cast<unsigned short*>(p), __bfloat16_as_ushort(enc<I>(v, e)));
    }
    template <int I>
    __device__ static R rt(R v, float e, float) { return dec<I>(enc<I>(v, e), e); }
    static constexpr bool kStoreUsesN2 = false;
};

// The per-direction constant a codec's store takes: -2 e_i for FP16, e_i otherwise.
template <class Codec>
__device__ __forceinline__ float store_key(const Center& ctr, int i) {
    if constexpr (Codec::kStoreUsesN2) return ctr.n2[center_class(i)];
    else return ctr.c[center_class(i)];
}

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

// Element idx of an array smaller than 4 GiB: a uniform base plus a 32-bit
// byte offset, the global-memory form with the base in scalar registers (no
// per-lane 64-bit address arithmetic). wfs_create keeps every per-direction
// population array and every cell array below 4 GiB.
template <typename T>
__device__ __forceinline__ T* elem(T* base, uint32_t idx) {
    using Byte = std::conditional_t<std::is_const<T>::value, const char, char>;
    return reinterpret_cast<T*>(reinterpret_cast<Byte*>(base) + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ T& at(T* base, int idx) {
    return *elem(base, (uint32_t)idx);
}

// Buffer access to one population direction: a resource in scalar registers
// (base, size in bytes; word 3 = 32-bit data format, the gfx9-family raw
// buffer) and a 32-bit byte offset per lane. Every direction of a cell then
// shares one lane offset, and the step spends no vector instruction on
// addresses. Offsets at or past the size read 0 and drop stores. The size is
// that of one direction array, below 4 GiB (wfs_create keeps cells * 8 B below
// 2^32), and every lane's offset is below it.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
// soff: a uniform byte offset added to the lane offset (scalar register)
// kAux: cache policy of the load (0 default, kPopLoadNT non-temporal)
constexpr int kPopLoadNT = 2;
// A non-temporal load through a pointer.
template <typename T>
__device__ __forceinline__ T load_nt(const T* p) {
    if constexpr (sizeof(T) == 2) {
        return __builtin_bit_cast(T, __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p)));
    } else if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_nontemporal_load(reinterpret_cast<const unsigned int*>(p)));
    } else {
        static_assert(sizeof(T) == 8, "2, 4 or 8 bytes");
        return __builtin_bit_cast(T, __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(p)));
    }
}
template <typename T, int kAux = 0>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff = 0) {
    if constexpr (sizeof(T) == 2) {
        return __builtin_bit_cast(T, (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(r, off, soff, kAux));
    } else if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, off, soff, kAux));
    } else {
        static_assert(sizeof(T) == 8, "2, 4 or 8 bytes");
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, kAux));
    }
}
// cache policy of population stores: non-temporal (nt) when WFS_NT_STORE
constexpr int kPopStoreAux = WFS_NT_STORE ? 2 : 0;
template <typename T>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, T v, uint32_t soff = 0) {
    if constexpr (sizeof(T) == 2) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, off, soff, kPopStoreAux);
    } else if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, soff, kPopStoreAux);
    } else {
        using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, off, soff, kPopStoreAux);
    }
}

// Stores the post-collision populations of cell c (all lanes that call it).
// pop32: one resource for the whole buffer, direction i at the uniform offset
// i N (fewer scalar instructions); otherwise the uniform pointer fd + i N and
// the cell's 32-bit offset.
template <class Codec>
__device__ __forceinline__ void store_cell(typename Codec::T* __restrict__ fd, long long N, const Center& ctr,
                                           const typename Codec::R* f, int c, bool pop32) {
    using T = typename Codec::T;
    const uint32_t off = (uint32_t)c * (uint32_t)sizeof(T);
    const uint32_t bytes = (uint32_t)N * (uint32_t)sizeof(T);
    if (pop32) {
        const auto r = buf_rsrc(fd, (uint32_t)Q * bytes);
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            buf_store<T>(r, off, Codec::template enc<i>(f[i], store_key<Codec>(ctr, i)), (uint32_t)i * bytes);
        });
    } else {
        // direction i at the uniform pointer fd + i N, the cell's 32-bit offset
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            Codec::template st<i>(elem(fd + (long long)i * N, (uint32_t)c), f[i], store_key<Codec>(ctr, i));
        });
    }
}

// Arrays the step reads or writes in some cells only: forces (forced cells),
// the macroscopic fields (wall-model cells, and every cell on the last step of
// a call), the wall model's inputs and the per-step column buffers. They sit
// in device memory behind one pointer, so a kernel loads each where it is
// used; as kernel arguments the compiler loads them all at entry and spills
// them to vector-register lanes.
// Body forces are few (actuator disks), so they are held per 64-cell group
// (cells 64 g .. 64 g + 63) rather than per cell: fslot[g] is the group's slot
// or -1, and fval holds per slot 64 values of Fx, then of Fy, then of Fz.
constexpr int kForceGroup = 64;
constexpr int kForceSlot = 3 * kForceGroup;

template <typename R>
struct StepAux {
    const int* fslot;
    const R* fval;
    R* rho;
    R* ux;
    R* uy;
    R* uz;
    const R* cfw;
    const R* uref;
    const uint8_t* colcount;  // per column (x + px y): fluid cells, at least 1
    // per-step column velocity for the wall model (when StepArgs::colu_on):
    // every step writes ux, uy of its fluid cells to colu_out_*, and a bed/bank
    // wall-model cell takes uref from its column of colu_in_* (the previous
    // step), as the run cadence does every 20 steps
    const R* colu_in_x;
    const R* colu_in_y;
    R* colu_out_x;
    R* colu_out_y;
};

// Force on cell c, one of FLAG_FORCED (its group has a slot), times s.
template <typename R>
__device__ __forceinline__ void body_force(const StepAux<R>& ax, int c, R s, R& fx, R& fy, R& fz) {
    const unsigned u = unsigned(c);
    const int b = ax.fslot[u / kForceGroup] * kForceSlot + int(u % kForceGroup);
    fx = ax.fval[b] * s;
    fy = ax.fval[b + kForceGroup] * s;
    fz = ax.fval[b + 2 * kForceGroup] * s;
}

template <typename R>
struct StepArgs {
    const uint8_t* __restrict__ flag;
    const uint32_t* __restrict__ wmask;
    const StepAux<R>* __restrict__ aux;
    int colu_on;   // per-step column velocity (StepAux::colu_*)
    R fscale;      // forces used are F * fscale (thrust ramp)
    R tau0;
    R tau0sq;      // tau0 * tau0
    R ksmag;       // (18 * 1.41421356) * Cs^2
    int nx, ny, nz;
    int px;        // row pitch in cells (nx rounded up to a multiple of 64)
    int rows;      // ny * nz
    int tiles;     // x-tiles per row (ceil(nx / blockDim.x))
    long long N;   // cells per population array, px * ny * nz
    const R* __restrict__ uin;  // inlet velocity per row (ny*nz)
    Center ctr;                 // per-direction centre of the 16-bit storage formats
    int bc_out_fused;           // outlet row inside the step ((nx-1) % 64 != 0)
    int pop32;                  // all 19 directions of a population buffer lie within 4 GiB
    int write_all;
    // learned closure (k_step_learned only)
    const uint16_t* __restrict__ geo;   // per cell: 255/d | 255 z_rel << 8
    const uint8_t* __restrict__ hcol;   // per column (x + px*y): fluid cells in the column
    const MlpPack* __restrict__ mlp;    // network weights in the wavefront layout
    float* __restrict__ c2out;          // C^2 per cell on write_all steps, when not null
    float inv_u0;                       // 1 / mean inlet speed (lattice units)
    float cs2;                          // the constant model's Cs^2 (float): the C^2 floor of the outlet sponge
    // tile-sparse layout (k_step_sparse only): a tile is one 64-cell x-run of a
    // row; cell c = tile * 64 + lane
    const int* __restrict__ wgrec;      // per work-group of the sparse step kernels: kWgRec ints (below)
    const int* __restrict__ nbr;        // per tile: 27 neighbour tiles (dx+1) + 3 ((dy+1) + 3 (dz+1)), -1 none
    const int* __restrict__ tcoord;     // per tile: x-tile, y, z
    const int* __restrict__ blk2tile;   // work-group -> tile (-1: idle), rows kept on one XCD
    const int* __restrict__ tmap;       // tile-sparse layout: x-tile + ntx * row -> stored tile or -1
    int ntx;
};

// Work-group -> (row, x-tile) map. Work-groups are dispatched round-robin over
// the 8 XCDs (group L runs on XCD L % 8), each with its own L2. All x-tiles of
// a row get the same L % 8, so the lines a row's tiles share at their x-edges
// (the +-1 pull shift) and its partial-line writes meet in one L2.
__device__ __forceinline__ bool tile_of(unsigned L, int tiles, int rows, int& row, int& tile) {
    const unsigned q = L >> 3;
    tile = int(q % unsigned(tiles));
    row = int((q / unsigned(tiles)) * 8u + (L & 7u));
    return row < rows;
}

// ---------------------------------------------------------------- the step
// Generic equilibrium with full lattice products, the form the boundary rows use.
template <typename R>
__device__ __forceinline__ R feq_full(int i, R r, R ux, R uy, R uz) {
    const R cu = R(3) * (R(kCX[i]) * ux + R(kCY[i]) * uy + R(kCZ[i]) * uz);
    return weight<R>(i) * r * (R(1) + cu + R(0.5) * cu * cu - R(1.5) * (ux * ux + uy * uy + uz * uz));
}

// Inlet (side 0) or outlet (side 1) row by non-equilibrium extrapolation from
// the neighbour's post-collision populations v (as stored):
//   inlet  f(0)    = feq(rho1, (uin, 0, 0)) + v - feq(rho1, u1)
//   outlet f(nx-1) = feq(rho_out, u1)      + v - feq(rho1, u1)
// with rho1, u1 the moments of v (floodsim/lbm3d.py:145-175 defines these rows).
template <typename R>
__device__ __forceinline__ void boundary_row(R* f, const R* v, int side, R uin, R rho_out) {
    R r = R(0), mx = R(0), my = R(0), mz = R(0);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        r += v[i];
        mx += v[i] * R(kCX[i]);
        my += v[i] * R(kCY[i]);
        mz += v[i] * R(kCZ[i]);
    }
    const R ux = mx / r, uy = my / r, uz = mz / r;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const R e = side == 0 ? feq_full<R>(i, r, uin, R(0), R(0)) : feq_full<R>(i, rho_out, ux, uy, uz);
        f[i] = e + (v[i] - feq_full<R>(i, r, ux, uy, uz));
    }
}


// Regularized BGK + Smagorinsky + Guo forcing in pair form, for the storage
// formats that are judged against an FP64 run rather than bit for bit.
// With Q_i = c_i c_i - I/3 and j = rho u (u with the Guo half-force shift),
// the post-collision population is
//   f_i = w_i rho + 4.5 w_i (Q_i : L) + 3 w_i (c_i . J)
//   L = rho u u + (1 - omega) Pi_neq + sf (u F + F u),   sf = 1 - omega / 2
//   J = j + (sf - (1 - omega) / 2) F
// which is feq + (1 - omega)(fneq + fneq1) + the Guo source of the
// per-direction form (floodsim/lbm3d.py:103-137) regrouped: w rho (1 - 1.5
// u^2 + 4.5 (c.u)^2) + 4.5 w (1 - omega) Q:Pi_neq + 9 w sf (c.u)(c.F) -
// 3 w sf u.F is w rho + 4.5 w Q:L, and 3 w rho (c.u) + w (c.F)(3 sf - 1.5 (1 -
// omega)) is 3 w c.J. Opposite directions share the even part and differ in
// the sign of the odd part, and only the deviatoric part of L enters
// (Q_i : L = c.L.c - tr L / 3):
//   axis x:        Q:L = L_xx - T                      (T = tr L / 3)
//   plane (1,+-1,0): Q:L = (T - (L_zz - T)) +- 2 L_xy
// Pi_neq comes from the second moment of f minus rho (u u + I/3), the second
// moment of the equilibrium, so no per-direction equilibrium is formed for
// the stress. Same mathematics as the per-direction form; different rounding
// (FMAs, one reciprocal).
//
// pair_moments: density, momentum and velocity, the non-equilibrium stress
// and its norm Q; pair_relax: the post-collision state.
// Reciprocal and square root of the pair form: the hardware instructions
// (about 1 ulp) in float, IEEE in double. The pair form is judged against
// FP64 runs, not bit for bit, and is VALU-bound in the 16-bit formats.
__device__ __forceinline__ float pair_rcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double pair_rcp(double v) { return 1.0 / v; }
__device__ __forceinline__ float pair_sqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
__device__ __forceinline__ double pair_sqrt(double v) { return sqrt(v); }

// qr = |Pi_neq| / rho, the Smagorinsky term's argument
template <typename R>
struct PairMoments {
    R r, ux, uy, uz, jx, jy, jz, pxx, pyy, pzz, pxy, pxz, pyz, qr;
};

// kForced: the wavefront holds a forced cell (fx, fy, fz zero on the others,
// where the forced formulas give the unforced values); otherwise the force
// terms are left out.
template <typename R, bool kForced>
__device__ __forceinline__ void pair_moments(const R* f, R fx, R fy, R fz, PairMoments<R>& m) {
    // pair sums and differences (i, i+1) for i = 1, 3, ..., 17
    R sp[9], dp[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        sp[k] = f[2 * k + 1] + f[2 * k + 2];
        dp[k] = f[2 * k + 1] - f[2 * k + 2];
    }
    // pairs: 0:x 1:y 2:z 3:(1,1,0) 4:(1,-1,0) 5:(1,0,1) 6:(1,0,-1) 7:(0,1,1) 8:(0,1,-1);
    // the four directions of each diagonal plane summed
    const R dxy = sp[3] + sp[4], dxz = sp[5] + sp[6], dyz = sp[7] + sp[8];
    const R r = f[0] + ((sp[0] + sp[1] + sp[2]) + (dxy + dxz + dyz));
    R jx = dp[0] + dp[3] + dp[4] + dp[5] + dp[6];
    R jy = dp[1] + dp[3] - dp[4] + dp[7] + dp[8];
    R jz = dp[2] + dp[5] - dp[6] + dp[7] - dp[8];
    if constexpr (kForced) {
        jx = fma(R(0.5), fx, jx); jy = fma(R(0.5), fy, jy); jz = fma(R(0.5), fz, jz);
    }
    const R ri = pair_rcp(r);
    const R ux = jx * ri, uy = jy * ri, uz = jz * ri;
    const R r3 = r * (R(1) / R(3));
    m.pxx = (sp[0] + dxy + dxz) - fma(jx, ux, r3);
    m.pyy = (sp[1] + dxy + dyz) - fma(jy, uy, r3);
    m.pzz = (sp[2] + dxz + dyz) - fma(jz, uz, r3);
    m.pxy = fma(-jx, uy, sp[3] - sp[4]);
    m.pxz = fma(-jx, uz, sp[5] - sp[6]);
    m.pyz = fma(-jy, uz, sp[7] - sp[8]);
    m.qr = ri * pair_sqrt(fma(m.pxx, m.pxx, fma(m.pyy, m.pyy, m.pzz * m.pzz)) +
                          R(2) * fma(m.pxy, m.pxy, fma(m.pxz, m.pxz, m.pyz * m.pyz)));
    m.r = r; m.ux = ux; m.uy = uy; m.uz = uz; m.jx = jx; m.jy = jy; m.jz = jz;
}

// Relaxation of the pair form with eddy-viscosity constant ksmag = 18 sqrt(2) C^2.
template <typename R, bool kForced>
__device__ __forceinline__ void pair_relax(R* f, R fx, R fy, R fz, const PairMoments<R>& m, R tau0, R tau0sq,
                                           R ksmag) {
    const R r = m.r, ux = m.ux, uy = m.uy, uz = m.uz;
    // omega = 1 / tau = 2 / (tau0 + sqrt(tau0^2 + ksmag Q / rho))
    const R om = R(2) * pair_rcp(tau0 + pair_sqrt(fma(ksmag, m.qr, tau0sq)));
    const R om1 = R(1) - om;
    R lxx = fma(om1, m.pxx, m.jx * ux), lyy = fma(om1, m.pyy, m.jy * uy), lzz = fma(om1, m.pzz, m.jz * uz);
    R lxy = fma(om1, m.pxy, m.jx * uy), lxz = fma(om1, m.pxz, m.jx * uz), lyz = fma(om1, m.pyz, m.jy * uz);
    // odd part 3 w (c.J) = (c.J) / 6 (axes), (c.J) / 12 (diagonals)
    R ax = m.jx, ay = m.jy, az = m.jz;
    if constexpr (kForced) {
        const R sf = fma(R(-0.5), om, R(1));
        const R gx = sf * fx, gy = sf * fy, gz = sf * fz;
        lxx = fma(R(2) * ux, gx, lxx); lyy = fma(R(2) * uy, gy, lyy); lzz = fma(R(2) * uz, gz, lzz);
        lxy = fma(ux, gy, fma(uy, gx, lxy));
        lxz = fma(ux, gz, fma(uz, gx, lxz));
        lyz = fma(uy, gz, fma(uz, gy, lyz));
        const R h = fma(R(-0.5), om1, sf);
        ax = fma(h, fx, ax); ay = fma(h, fy, ay); az = fma(h, fz, az);
    }
    const R T = (lxx + lyy + lzz) * (R(1) / R(3));
    const R dxx = lxx - T, dyy = lyy - T, dzz = lzz - T;
    // rest population: w0 = 1/3, Q_0 : L = -T
    f[0] = fma(R(-1.5), T, r * (R(1) / R(3)));
    const R wr1 = r * (R(1) / R(18)), wr2 = r * (R(1) / R(36));
    auto axis = [&](int k, R d, R a) {
        const R e = fma(R(0.25), d, wr1);          // 4.5 w1 = 0.25
        const R o = a * (R(1) / R(6));
        f[2 * k + 1] = e + o;
        f[2 * k + 2] = e - o;
    };
    axis(0, dxx, ax);
    axis(1, dyy, ay);
    axis(2, dzz, az);
    const R bx = ax * (R(1) / R(12)), by = ay * (R(1) / R(12)), bz = az * (R(1) / R(12));
    // plane of axes a, b (third axis n): Q:L = (T - d_n) +- 2 L_ab
    auto plane = [&](int k, R dn, R lab, R oa, R ob) {
        const R e0 = fma(R(0.125), T - dn, wr2);   // 4.5 w2 = 0.125
        const R ep = fma(R(0.25), lab, e0), em = fma(R(-0.25), lab, e0);
        const R op = oa + ob, on = oa - ob;
        f[2 * k + 1] = ep + op;
        f[2 * k + 2] = ep - op;
        f[2 * k + 3] = em + on;
        f[2 * k + 4] = em - on;
    };
    plane(3, dzz, lxy, bx, by);
    plane(5, dyy, lxz, bx, bz);
    plane(7, dxx, lyz, by, bz);
}

template <typename R, bool kForced>
__device__ __forceinline__ void collide_pairs(R* f, R fx, R fy, R fz, R tau0, R tau0sq, R ksmag, R& r, R& ux, R& uy,
                                              R& uz) {
    PairMoments<R> m;
    pair_moments<R, kForced>(f, fx, fy, fz, m);
    pair_relax<R, kForced>(f, fx, fy, fz, m, tau0, tau0sq, ksmag);
    r = m.r; ux = m.ux; uy = m.uy; uz = m.uz;
}

#ifndef WFS_STEP_WAVES_PER_EU
#define WFS_STEP_WAVES_PER_EU 1   // minimum resident waves per SIMD requested from the register allocator
#endif

// Pull streaming (x-edge own copy, rigid-lid mirror, bounce-back on wall
// links) and the wall function for one active cell: f gets the populations
// entering the collision.
//
// Pull source of direction (cx, cy, cz) for the cell: in the dense layout a
// fixed offset; in the tile-sparse layout the source tile comes from the
// tile's neighbour table (nb, wavefront-uniform) and only the first/last
// lane crosses into the x-neighbour tile.
template <bool kSparse, int cx, int cy, int cz>
__device__ __forceinline__ int pull_source(int c, int px, int plane, const int* nb, int lane) {
    if constexpr (!kSparse) {
        return c - cx - px * cy - plane * cz;
    } else {
        constexpr int kyz = 3 * ((1 - cy) + 3 * (1 - cz));
        const int ts = nb[1 + kyz];
        if constexpr (cx == 0) {
            return ts * 64 + lane;
        } else {
            const int td = nb[(cx > 0 ? 0 : 2) + kyz];
            const bool cross = cx > 0 ? lane == 0 : lane == 63;
            return cross ? td * 64 + ((lane - cx) & 63) : ts * 64 + lane - cx;
        }
    }
}

// Depth-averaged speed of column (x, y) from the previous step's velocity,
// the column reduction of the run cadence (floodsim/reach3d.py:412-418) in
// the order and arithmetic of k_uref: sums over z upward, solid cells
// contributing zero, |(sum ux, sum uy)| / fluid cells.
template <bool kSparse, typename R>
__device__ __forceinline__ R column_u(const StepArgs<R>& a, int x, int y) {
    R sx = R(0), sy = R(0);
    for (int zz = 0; zz < a.nz; ++zz) {
        const long long row = y + (long long)a.ny * zz;
        long long cc;
        if constexpr (kSparse) {
            const int t = a.tmap[(x >> 6) + (long long)a.ntx * row];
            if (t < 0) continue;   // tile not stored: solid
            cc = (long long)t * 64 + (x & 63);
        } else {
            cc = x + (long long)a.px * row;
        }
        sx += at(a.aux->colu_in_x, int(cc));
        sy += at(a.aux->colu_in_y, int(cc));
    }
    const R ncol = R(a.aux->colcount[x + (long long)a.px * y]);
    return sqrt(sx * sx + sy * sy) / ncol;
}

// The wall function (floodsim/lbm3d.py:67-101) on the populations f of a
// cell whose wall links are already bounced back: the slip correction of the
// links to the wall from the cell's previous-step velocity and the wall
// model's reference speed.
template <class Codec, bool kSparse>
__device__ __forceinline__ void wall_function(const StepArgs<typename Codec::R>& a, int c, int x, int y, uint8_t fl,
                                              uint32_t wm, typename Codec::R* f) {
    using R = typename Codec::R;
    if (wm != 0u && (fl & FLAG_CF)) {
        const StepAux<R>& ax = *a.aux;
        const R cf = at(ax.cfw, c);
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
            const R ox = at(ax.ux, c), oy = at(ax.uy, c), oz = at(ax.uz, c);   // previous step, own cell
            const R un = ox * ex + oy * ey + oz * ez;
            R tx = ox - un * ex, ty = oy - un * ey, tz = oz - un * ez;
            const R ut = sqrt(tx * tx + ty * ty + tz * tz);
            if (ut > R(1e-7)) {
                tx /= ut; ty /= ut; tz /= ut;
                const R rw = at(ax.rho, c);
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
                    const R ur = !a.colu_on ? at(ax.uref, c)
                                 : ((fl & FLAG_BEDWALL) ? column_u<kSparse>(a, x, y) : R(0));
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
}

// kLoadAux: cache policy of the population loads through the buffer resource.
template <class Codec, bool kSparse = false, int kLoadAux = 0>
__device__ __forceinline__ void gather(const typename Codec::T* __restrict__ fs, const StepArgs<typename Codec::R>& a,
                                       int c, int x, int y, int z, uint8_t fl, uint32_t wm, typename Codec::R* f,
                                       const int* nb = nullptr, int lane = 0) {
    const int px = a.px;
    const int plane = px * a.ny;
    const long long N = a.N;
    const bool xlo = x == 0, xhi = x == a.nx - 1, top = z == a.nz - 1;
    // 1. pull streaming
    if (!(fl & FLAG_NEARWALL)) {
        // No wall link: every source is the plain upstream neighbour, at a pull
        // base uniform over the wavefront plus the lane's own offset:
        //   dense:  (i N - (cx + px cy + plane cz)) + c
        //   sparse: (i N + 64 ts - cx) + lane, ts the tile of the (dy, dz) row;
        // stored tiles are numbered row by row with x ascending, so for a cell
        // with no wall link the x-neighbour tile of ts is ts -+ 1 and the lane
        // that crosses into it needs no special case. Under the lid (a row
        // property, uniform) a link with c_z = -1 takes direction MIRZ(i) from
        // the in-plane source, another uniform base. An x-end lane keeps its
        // own population for the links entering from outside the box (x-end
        // precedes the lid): in the wavefronts that hold an x-end (tested
        // uniformly first) those lanes add a per-direction offset to their
        // own cell, direction i.
        using T = typename Codec::T;
        constexpr uint32_t ts = (uint32_t)sizeof(T);
        const uint32_t cell = (uint32_t)(kSparse ? lane : c);
        const bool ends = __any(xlo || xhi);
        // pull base of direction I, elements (in [0, Q N))
        auto base_of = [&](auto I) -> long long {
            constexpr int i = I;
            constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
            const bool mir = cz == -1 && top;
            const long long d = mir ? kMIRZ[i] : i;
            if constexpr (kSparse) {
                const int rt = mir ? nb[1 + 3 * ((1 - cy) + 3 * 1)] : nb[1 + 3 * ((1 - cy) + 3 * (1 - cz))];
                return d * N + (long long)rt * 64 - cx;
            } else {
                return d * N - (cx + (long long)px * cy + (mir ? 0LL : (long long)plane * cz));
            }
        };
        auto xout_of = [&](auto I) -> bool {
            constexpr int cx = kCX[decltype(I)::value];
            return (cx == 1 && xlo) || (cx == -1 && xhi);
        };
        auto load_all = [&](auto with_ends) {
            constexpr bool kEnds = decltype(with_ends)::value;
            if (a.pop32) {
                // one resource for the buffer, the lowest candidate base as a
                // uniform byte offset; a lane adds its (non-negative) offset
                const auto rs = buf_rsrc(fs, (uint32_t)Q * (uint32_t)N * ts);
                const uint32_t vc = cell * ts;
                static_for<Q>([&](auto I) {
                    constexpr int i = I;
                    const uint32_t base = (uint32_t)base_of(I);
                    uint32_t soff = base, voff = vc;
                    if constexpr (kEnds && kCX[i] != 0) {
                        // own cell, direction i: i N + c = lo + (i N + c - cell - lo) + cell
                        const uint32_t own = (uint32_t)i * (uint32_t)N + ((uint32_t)c - cell);
                        const uint32_t lo = own < base ? own : base;
                        soff = lo;
                        voff = vc + ((xout_of(I) ? own : base) - lo) * ts;
                    }
                    f[i] = Codec::template dec<i>(buf_load<T, kLoadAux>(rs, voff, soff * ts), a.ctr.c[center_class(i)]);
                });
            } else {
                // a uniform pointer per direction at its pull base (>= 0) and the
                // lane's 32-bit offset; an x-end lane's own cell, direction i
                static_for<Q>([&](auto I) {
                    constexpr int i = I;
                    const T* p = elem(fs + base_of(I), cell);
                    if constexpr (kEnds && kCX[i] != 0) {
                        if (xout_of(I)) p = elem(fs + (long long)i * N, (uint32_t)c);
                    }
                    f[i] = Codec::template ld<i>(p, a.ctr.c[center_class(i)]);
                });
            }
        };
        if (ends) load_all(std::true_type{});
        else load_all(std::false_type{});
    } else {
        // Link by link (in this order of precedence): the x-end rows keep their
        // own population (xout: own cell, same direction), a wall link bounces
        // back (own cell, opposite direction), a link from above the lid
        // mirrors (in-plane source, z-mirrored direction; uniform, since top is
        // the row's), any other link pulls. The stored direction is one of
        // these constants, so a lane only selects among uniform offsets: the
        // lowest candidate direction is the uniform base and each lane adds a
        // non-negative multiple of N. No branches, no 64-bit products per lane.
        using T = typename Codec::T;
        constexpr uint32_t ts = (uint32_t)sizeof(T);
        auto link = [&](auto I, int& s, bool& xout, bool& wall, float& e) {
            constexpr int i = I;
            constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
            constexpr int io = kOPP[i], im = kMIRZ[i];
            xout = (cx == 1 && xlo) || (cx == -1 && xhi);
            wall = !xout && ((wm >> i) & 1u);
            const bool mirror = cz == -1 && top;
            s = (xout || wall) ? c
                : mirror ? pull_source<kSparse, cx, cy, 0>(c, px, plane, nb, lane)
                         : pull_source<kSparse, cx, cy, cz>(c, px, plane, nb, lane);
            (void)im;
            // decode with the centre of the direction actually read (the mirror
            // keeps c_x, so only the opposite direction has another centre); both
            // centres pinned in scalar registers by an empty asm, or a select of
            // the two loads becomes a per-lane load from a selected address
            constexpr int ki = center_class(i), ko = center_class(io);
            if constexpr (ki != ko) {
                float ei = a.ctr.c[ki], eo = a.ctr.c[ko];
                asm("" : "+s"(ei));
                asm("" : "+s"(eo));
                e = wall ? eo : ei;
            } else {
                e = a.ctr.c[ki];
            }
        };
        if (a.pop32) {
            const uint32_t nn = (uint32_t)N, nn2 = 2u * nn, nn3 = 3u * nn;
            const auto rs = buf_rsrc(fs, (uint32_t)Q * nn * ts);
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                constexpr int io = kOPP[i], im = kMIRZ[i];
                int s;
                bool xout, wall;
                float e;
                link(I, s, xout, wall, e);
                // direction read: i (xout), io (wall), else i or, under the lid,
                // im. The lowest candidate is the uniform base; the others are at
                // most 3 directions above it (opposites are adjacent pairs).
                constexpr int lo0 = io < i ? io : i;
                constexpr int lo1 = im < lo0 ? im : lo0;
                constexpr int di = i - lo0, dop = io - lo0;
                uint32_t adj = wall ? (dop ? nn : 0u) : (di ? nn : 0u);
                uint32_t base = (uint32_t)lo0 * nn;
                if constexpr (kCZ[i] == -1) {
                    static_assert(i - lo1 <= 3 && io - lo1 <= 3 && im - lo1 <= 3, "candidates within 3");
                    constexpr int ti = i - lo1, to = io - lo1, tm = im - lo1;
                    auto times = [&](int k) { return k == 0 ? 0u : (k == 1 ? nn : (k == 2 ? nn2 : nn3)); };
                    const uint32_t ai = times(ti), ao = times(to), am = times(tm);
                    if (top) {
                        adj = xout ? ai : (wall ? ao : am);
                        base = (uint32_t)lo1 * nn;
                    }
                }
                f[i] = Codec::template dec<i>(buf_load<T, kLoadAux>(rs, ((uint32_t)s + adj) * ts, base * ts), e);
            });
        } else {
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                int s;
                bool xout, wall;
                float e;
                link(I, s, xout, wall, e);
                const T* bd = fs + (long long)((kCZ[i] == -1 && top) ? kMIRZ[i] : i) * N;
                const T* b = xout ? fs + (long long)i * N : (wall ? fs + (long long)kOPP[i] * N : bd);
                f[i] = Codec::template dec<i>(*elem(b, (uint32_t)s), e);
            });
        }
    }

    wall_function<Codec, kSparse>(a, c, x, y, fl, wm, f);
}

// Fused inlet/outlet rows and the stores of one tile (every lane of the
// wavefront calls this; inactive lanes only take part in the shifts).
template <class Codec>
__device__ __forceinline__ void boundary_rows(const StepArgs<typename Codec::R>& a, typename Codec::R* f, int x,
                                              int row, int xw, bool active) {
    using R = typename Codec::R;
    // Inlet and outlet rows, fused: the boundary lane takes its neighbour's
    // post-collision populations (rounded through the storage format, as a
    // separate pass would read them) by a cross-lane shift. The outlet is done
    // here when x = nx-2 and nx-1 share a wavefront; otherwise by k_bc.
    // xw: first x of this wavefront.
    const bool in_wave = xw == 0;
    const bool out_wave = a.bc_out_fused && xw <= a.nx - 1 && a.nx - 1 < xw + 64;
    if (in_wave || out_wave) {
        R v[Q];
        if (in_wave) {
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                v[i] = __shfl_down(Codec::template rt<i>(f[i], a.ctr.c[center_class(i)], a.ctr.n2[center_class(i)]), 1);
            });
            const bool nb = __shfl_down(int(active), 1) != 0;
            if (x == 0 && active && nb) boundary_row<R>(f, v, 0, a.uin[row], R(1));
        }
        if (out_wave) {
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                v[i] = __shfl_up(Codec::template rt<i>(f[i], a.ctr.c[center_class(i)], a.ctr.n2[center_class(i)]), 1);
            });
            const bool nb = __shfl_up(int(active), 1) != 0;
            if (x == a.nx - 1 && active && nb) boundary_row<R>(f, v, 1, R(0), R(1));
        }
    }
}

template <class Codec>
__device__ __forceinline__ void finish(typename Codec::T* __restrict__ fd, const StepArgs<typename Codec::R>& a,
                                       typename Codec::R* f, int c, int x, int row, int xw, bool active) {
    boundary_rows<Codec>(a, f, x, row, xw, active);
    if (active) store_cell<Codec>(fd, a.N, a.ctr, f, c, a.pop32 != 0);
}

// Collision of one active cell whose incoming populations are in f (Guo
// half-force shift, regularized BGK with Smagorinsky relaxation time, Guo
// source; floodsim/lbm3d.py:103-137): f becomes the post-collision
// populations; the macroscopic fields are written where the next step or the
// caller reads them.
template <class Codec>
__device__ __forceinline__ void collide_cell(const StepArgs<typename Codec::R>& a, int c, uint8_t fl, uint32_t wm,
                                             typename Codec::R* f) {
    using R = typename Codec::R;
    R r, ux, uy, uz;
    if constexpr (Codec::kExact) {
        // 3. moments (Guo half-force shift)
        r = f[0];
        R mx = R(0), my = R(0), mz = R(0);
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
        if (forced) {
            body_force(*a.aux, int(c), a.fscale, fx, fy, fz);
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

        // regularized post-collision populations (kept in f until the store)
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
                f[i] = feq[i] + om1 * (fneq + fneq1) + src;
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
                f[i] = feq[i] + om1 * (c45w * sp);
            });
        }
    } else {
        const bool forced = (fl & FLAG_FORCED) != 0;
        if (__any(forced)) {   // uniform over the wavefront; forced cells are few
            R fx = R(0), fy = R(0), fz = R(0);
            if (forced) body_force(*a.aux, int(c), a.fscale, fx, fy, fz);
            collide_pairs<R, true>(f, fx, fy, fz, a.tau0, a.tau0sq, a.ksmag, r, ux, uy, uz);
        } else {
            collide_pairs<R, false>(f, R(0), R(0), R(0), a.tau0, a.tau0sq, a.ksmag, r, ux, uy, uz);
        }
    }
    if (a.write_all || (wm != 0u && (fl & FLAG_CF))) {
        const StepAux<R>& ax = *a.aux;
        at(ax.rho, c) = r; at(ax.ux, c) = ux; at(ax.uy, c) = uy; at(ax.uz, c) = uz;
    }
    if (a.colu_on) { at(a.aux->colu_out_x, c) = ux; at(a.aux->colu_out_y, c) = uy; }
}

// The update of one active cell: pull, wall function, collision; f gets the
// post-collision populations; the macroscopic fields are written where the
// next step or the caller reads them.
template <class Codec, bool kSparse>
__device__ __forceinline__ void cell_update(const typename Codec::T* __restrict__ fs,
                                            const StepArgs<typename Codec::R>& a, int c, int x, int y, int z,
                                            uint8_t fl, typename Codec::R* f, const int* nb, int lane) {
    const uint32_t wm = (fl & FLAG_NEARWALL) ? at(a.wmask, c) : 0u;
    gather<Codec, kSparse>(fs, a, c, x, y, z, fl, wm, f, nb, lane);
    collide_cell<Codec>(a, c, fl, wm, f);
}

// The wavefront holds the inlet row (x = 0) or the fused outlet row; xw is its
// first x. Those wavefronts exchange post-collision populations across lanes.
template <typename R>
__device__ __forceinline__ bool boundary_wave(const StepArgs<R>& a, int xw) {
    return xw == 0 || (a.bc_out_fused && xw <= a.nx - 1 && a.nx - 1 < xw + 64);
}

// One x-tile of one row: the whole step for its cells.
// Dense layout: tile = x-tile of blockDim.x cells of the row, t unused.
// Tile-sparse layout: a work-group is one wavefront, tile = its x-tile
// (64 cells) and t its index in the compact storage.
template <class Codec, bool kSparse = false>
__device__ __forceinline__ void step_tile(const typename Codec::T* __restrict__ fs, typename Codec::T* __restrict__ fd,
                                          const StepArgs<typename Codec::R>& a, int row, int tile, int t = 0,
                                          const int* nbp = nullptr) {
    using R = typename Codec::R;
    const int lane = int(threadIdx.x) & 63;
    const int x = kSparse ? tile * 64 + lane : tile * int(blockDim.x) + int(threadIdx.x);
    // first x of the wavefront, uniform (scalar register)
    const int xw = __builtin_amdgcn_readfirstlane(kSparse ? tile * 64 : tile * int(blockDim.x) + (int(threadIdx.x) & ~63));
    const bool inrow = x < a.nx;
    const int z = row / a.ny;
    const int c = kSparse ? t * 64 + lane : x + a.px * row;
    const int* nb = kSparse ? (nbp ? nbp : a.nbr + 27 * t) : nullptr;
    // Lanes outside the row or on solid cells do no work but stay in the wave
    // for the boundary-row shuffles.
    const uint8_t fl = inrow ? at(a.flag, c) : uint8_t(FLAG_SOLID);
    const bool active = !(fl & FLAG_SOLID);
    if (!boundary_wave(a, xw)) {
        // no cross-lane exchange: each active lane updates and stores its cell
        if (active) {
            R f[Q];
            cell_update<Codec, kSparse>(fs, a, c, x, row - a.ny * z, z, fl, f, nb, lane);
            store_cell<Codec>(fd, a.N, a.ctr, f, c, a.pop32 != 0);
        }
        return;
    }
    R f[Q];
    if (active) cell_update<Codec, kSparse>(fs, a, c, x, row - a.ny * z, z, fl, f, nb, lane);
    finish<Codec>(fd, a, f, c, x, row, xw, active);
}

// ---------------------------------------------------------------- in-place streaming
// One population buffer, two alternating steps (the AA pattern). Slot k of
// cell x is f[k * N + x]; a slot always holds a whole population p, decoded
// and encoded with p's centre.
//
// State S (after an even step, and the A-B layout with each pair of slots
// i / opp(i) exchanged): slot opp(i) of x holds the post-collision f*_i(x);
// at the x-end rows, where a link i enters from outside the box, slot i also
// holds f*_i(x) (the reference keeps the cell's own population there).
//
// Odd step, from S: every link reads where the pull step of the A-B scheme
// reads, with the slot exchanged: a plain link i from (x - c_i, opp(i)); a
// wall link from (x, i) (= f*_opp(i)(x), bounce-back); a link through the
// lid from (x - (cx, cy, 0), opp(MIRZ(i))); an x-end link from (x, i).
// After the collision f*_i(x) goes to where its reader of the next step
// looks: (x + c_i, i) for a plain link; (x, opp(i)) when x + c_i is a wall;
// (x + (cx, cy, 0), MIRZ(i)) when it leaves through the lid; (x, i) also at
// an x-end row. Each of these slots is read in this step by the writing cell
// alone, so the step needs no second buffer.
//
// Even step: every incoming population is in the cell's own slot i (state
// after an odd step); the results go back to the cell's own slots, giving S.
// Two steps equal two A-B steps, bit for bit.

// Either layout. Neighbour cells come from pull_source: the cell at x - c
// for a pull, at x + c (the offset negated) for a push. Interior cells of an
// odd step (no wall link, away from the x-ends and the lid) address every
// direction as a uniform pointer plus the cell's (dense) or lane's
// (tile-sparse) offset; other cells go link by link, a lane selecting
// between uniform slot pointers (f + k N) and adding its cell's 32-bit byte
// offset (each direction is below 4 GiB).
template <class Codec, bool kOdd, bool kSparse>
__device__ __forceinline__ void aa_gather(const typename Codec::T* f, const StepArgs<typename Codec::R>& a, int c,
                                          int x, int z, uint32_t wm, typename Codec::R* out, const int* nb,
                                          int lane) {
    using T = typename Codec::T;
    const int px = a.px;
    const int plane = px * a.ny;
    const long long N = a.N;
    const bool xlo = x == 0, xhi = x == a.nx - 1, top = z == a.nz - 1;
    if (kOdd && wm == 0u && !xlo && !xhi && !top) {
        // interior: slot opp(i) of x - c_i for every link, at a uniform pointer
        // plus the cell's (dense) or lane's (tile-sparse) offset
        const uint32_t off = (uint32_t)(kSparse ? lane : c);
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
            const T* pn = f + (long long)kOPP[i] * N +
                          (kSparse ? 64LL * nb[1 + 3 * ((1 - cy) + 3 * (1 - cz))] - cx
                                   : -(cx + (long long)px * cy + (long long)plane * cz));
            out[i] = Codec::template ld<i>(elem(pn, off), a.ctr.c[center_class(i)]);
        });
        return;
    }
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
        const bool xout = (cx == 1 && xlo) || (cx == -1 && xhi);
        const bool wall = !xout && ((wm >> i) & 1u);
        constexpr int ki = center_class(i), ko = center_class(kOPP[i]);
        const float e = (ki != ko && wall) ? a.ctr.c[ko] : a.ctr.c[ki];
        const T* own = f + (long long)i * N;   // x-end copy, bounce-back, or the even step
        if constexpr (kOdd) {
            const T* pn;
            int s;
            if (cz == -1 && top) {   // through the lid: slot opp(MIRZ(i)) of the in-plane source
                pn = f + (long long)kOPP[kMIRZ[i]] * N;
                s = pull_source<kSparse, cx, cy, 0>(c, px, plane, nb, lane);
            } else {
                pn = f + (long long)kOPP[i] * N;
                s = pull_source<kSparse, cx, cy, cz>(c, px, plane, nb, lane);
            }
            const bool mine = xout || wall;
            out[i] = Codec::template ld<i>(elem(mine ? own : pn, (uint32_t)(mine ? c : s)), e);
        } else {   // even step: own lines, non-temporal as in aa_gather32
            out[i] = Codec::template dec<i>(load_nt(elem(own, (uint32_t)c)), e);
        }
    });
}

template <class Codec, bool kOdd, bool kSparse>
__device__ __forceinline__ void aa_store(typename Codec::T* f, const StepArgs<typename Codec::R>& a, int c, int x,
                                         int z, uint32_t wm, const typename Codec::R* v, const int* nb, int lane) {
    using T = typename Codec::T;
    const int px = a.px;
    const int plane = px * a.ny;
    const long long N = a.N;
    const bool xlo = x == 0, xhi = x == a.nx - 1, top = z == a.nz - 1;
    if (kOdd && wm == 0u && !xlo && !xhi && !top) {
        // interior: slot i of x + c_i for every link
        const uint32_t off = (uint32_t)(kSparse ? lane : c);
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
            T* pt = f + (long long)i * N +
                    (kSparse ? 64LL * nb[1 + 3 * ((1 + cy) + 3 * (1 + cz))] + cx
                             : cx + (long long)px * cy + (long long)plane * cz);
            Codec::template st<i>(elem(pt, off), v[i], store_key<Codec>(a.ctr, i));
        });
        return;
    }
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
        const float key = store_key<Codec>(a.ctr, i);
        T* back = f + (long long)kOPP[i] * N;   // bounce-back, or the even step: own slot opp(i)
        if constexpr (!kOdd) {
            Codec::template st<i>(elem(back, (uint32_t)c), v[i], key);
        } else {
            const bool bounce = (wm >> kOPP[i]) & 1u;
            const bool inx = !((cx == 1 && xhi) || (cx == -1 && xlo));
            T* pt;
            int s;
            if (cz == 1 && top) {   // through the lid: slot MIRZ(i) of the in-plane target
                pt = f + (long long)kMIRZ[i] * N;
                s = pull_source<kSparse, -cx, -cy, 0>(c, px, plane, nb, lane);
            } else {
                pt = f + (long long)i * N;
                s = pull_source<kSparse, -cx, -cy, -cz>(c, px, plane, nb, lane);
            }
            if (bounce || inx) Codec::template st<i>(elem(bounce ? back : pt, (uint32_t)(bounce ? c : s)), v[i], key);
        }
    });
    // x-end rows: a link entering from outside the box keeps the cell's own population
    if (xlo || xhi) {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            if constexpr (kCX[i] != 0) {
                if ((kCX[i] == 1 && xlo) || (kCX[i] == -1 && xhi))
                    Codec::template st<i>(elem(f + (long long)i * N, (uint32_t)c), v[i], store_key<Codec>(a.ctr, i));
            }
        });
    }
}

// The same two steps with 32-bit addressing (Q * N * sizeof(T) < 4 GiB): one
// buffer resource for the whole buffer, the slot's offset uniform in a scalar
// register and only the cell's offset per lane. Where a link has two or three
// candidate slots, the lowest is the uniform base and a lane adds 0..3 N.
// Cells with no wall link away from the x-ends and the lid take the plain
// link of every direction.
template <class Codec, bool kOdd>
__device__ __forceinline__ void aa_gather32(const typename Codec::T* f, const StepArgs<typename Codec::R>& a, int c,
                                            int x, int z, uint32_t wm, typename Codec::R* out) {
    using T = typename Codec::T;
    constexpr uint32_t ts = (uint32_t)sizeof(T);
    const uint32_t nn = (uint32_t)a.N;
    const auto rs = buf_rsrc(f, (uint32_t)Q * nn * ts);
    const int px = a.px;
    const int plane = px * a.ny;
    const bool xlo = x == 0, xhi = x == a.nx - 1, top = z == a.nz - 1;
    const uint32_t vc = (uint32_t)c * ts;
    auto times = [&](int k) { return k == 0 ? 0u : (k == 1 ? nn : (k == 2 ? 2u * nn : 3u * nn)); };
    if (!kOdd || (wm == 0u && !xlo && !xhi && !top)) {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ki = center_class(i), ko = center_class(kOPP[i]);
            const float e = (ki != ko && ((wm >> i) & 1u)) ? a.ctr.c[ko] : a.ctr.c[ki];
            if constexpr (kOdd) {
                // slot opp(i) of the cell at x - c_i
                const int off = kCX[i] + px * kCY[i] + plane * kCZ[i];
                out[i] = Codec::template dec<i>(buf_load<T>(rs, vc, ((uint32_t)kOPP[i] * nn - (uint32_t)off) * ts), e);
            } else {
                // even step: the cell's own lines, read once and written back;
                // non-temporal loads measured 4 % (FP32) and 6 % (FP16) faster
                // here and 5 % on the tile-sparse meander; on the odd step's
                // neighbour loads they cost 13 %
                out[i] = Codec::template dec<i>(buf_load<T, kPopLoadNT>(rs, vc, (uint32_t)i * nn * ts), e);
            }
        });
        return;
    }
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
        constexpr int io = kOPP[i], mo = kOPP[kMIRZ[i]];
        const bool xout = (cx == 1 && xlo) || (cx == -1 && xhi);
        const bool wall = !xout && ((wm >> i) & 1u);
        const bool own = xout || wall;
        constexpr int ki = center_class(i), ko = center_class(io);
        const float e = (ki != ko && wall) ? a.ctr.c[ko] : a.ctr.c[ki];
        constexpr int lo0 = io < i ? io : i;
        uint32_t base = (uint32_t)lo0 * nn;
        int src = c - cx - px * cy - plane * cz;
        uint32_t adj = own ? times(i - lo0) : times(io - lo0);
        if constexpr (cz == -1) {
            if (top) {   // through the lid: slot opp(MIRZ(i)) of the in-plane source
                constexpr int lo1 = mo < lo0 ? mo : lo0;
                static_assert(i - lo1 <= 3 && mo - lo1 <= 3, "candidates within 3");
                base = (uint32_t)lo1 * nn;
                src = c - cx - px * cy;
                adj = own ? times(i - lo1) : times(mo - lo1);
            }
        }
        const int cell = own ? c : src;
        out[i] = Codec::template dec<i>(buf_load<T>(rs, ((uint32_t)cell + adj) * ts, base * ts), e);
    });
}

template <class Codec, bool kOdd>
__device__ __forceinline__ void aa_store32(typename Codec::T* f, const StepArgs<typename Codec::R>& a, int c, int x,
                                           int z, uint32_t wm, const typename Codec::R* v) {
    using T = typename Codec::T;
    constexpr uint32_t ts = (uint32_t)sizeof(T);
    const uint32_t nn = (uint32_t)a.N;
    const auto rs = buf_rsrc(f, (uint32_t)Q * nn * ts);
    const int px = a.px;
    const int plane = px * a.ny;
    const bool xlo = x == 0, xhi = x == a.nx - 1, top = z == a.nz - 1;
    const uint32_t vc = (uint32_t)c * ts;
    auto times = [&](int k) { return k == 0 ? 0u : (k == 1 ? nn : (k == 2 ? 2u * nn : 3u * nn)); };
    if (!kOdd || (wm == 0u && !xlo && !xhi && !top)) {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            const T w = Codec::template enc<i>(v[i], store_key<Codec>(a.ctr, i));
            if constexpr (kOdd) {   // slot i of the cell at x + c_i
                const int off = kCX[i] + px * kCY[i] + plane * kCZ[i];
                buf_store<T>(rs, vc, w, ((uint32_t)i * nn + (uint32_t)off) * ts);
            } else {
                buf_store<T>(rs, vc, w, (uint32_t)kOPP[i] * nn * ts);
            }
        });
    } else {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
            constexpr int io = kOPP[i], im = kMIRZ[i];
            const bool bounce = (wm >> io) & 1u;
            const bool inx = !((cx == 1 && xhi) || (cx == -1 && xlo));
            constexpr int lo0 = io < i ? io : i;
            uint32_t base = (uint32_t)lo0 * nn;
            int dst = c + cx + px * cy + plane * cz;
            uint32_t adj = bounce ? times(io - lo0) : times(i - lo0);
            if constexpr (cz == 1) {
                if (top) {   // through the lid: slot MIRZ(i) of the in-plane target
                    constexpr int lo1 = im < lo0 ? im : lo0;
                    static_assert(i - lo1 <= 3 && io - lo1 <= 3 && im - lo1 <= 3, "candidates within 3");
                    base = (uint32_t)lo1 * nn;
                    dst = c + cx + px * cy;
                    adj = bounce ? times(io - lo1) : times(im - lo1);
                }
            }
            if (bounce || inx) {
                const int cell = bounce ? c : dst;
                buf_store<T>(rs, ((uint32_t)cell + adj) * ts,
                             Codec::template enc<i>(v[i], store_key<Codec>(a.ctr, i)), base * ts);
            }
        });
    }
    // x-end rows: a link entering from outside the box keeps the cell's own population
    if (xlo || xhi) {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            if constexpr (kCX[i] != 0) {
                if ((kCX[i] == 1 && xlo) || (kCX[i] == -1 && xhi))
                    buf_store<T>(rs, vc, Codec::template enc<i>(v[i], store_key<Codec>(a.ctr, i)),
                                 (uint32_t)i * nn * ts);
            }
        });
    }
}

// One in-place step of one x-tile (dense: blockDim.x cells of a row; tile-
// sparse: the 64 cells of stored tile t). kPop32: the dense layout's 32-bit
// buffer addressing (whole buffer below 4 GiB).
template <class Codec, bool kOdd, bool kPop32, bool kSparse>
__device__ __forceinline__ void step_tile_aa(typename Codec::T* f, const StepArgs<typename Codec::R>& a, int row,
                                             int tile, int t = 0) {
    using R = typename Codec::R;
    static_assert(!(kPop32 && kSparse), "32-bit buffer addressing is for the dense layout");
    const int lane = int(threadIdx.x) & 63;
    const int x = kSparse ? tile * 64 + lane : tile * int(blockDim.x) + int(threadIdx.x);
    const int xw = __builtin_amdgcn_readfirstlane(kSparse ? tile * 64 : tile * int(blockDim.x) + (int(threadIdx.x) & ~63));
    const bool inrow = x < a.nx;
    const int z = row / a.ny;
    const int c = kSparse ? t * 64 + lane : x + a.px * row;
    const int* nb = kSparse ? a.nbr + 27 * t : nullptr;
    const uint8_t fl = inrow ? at(a.flag, c) : uint8_t(FLAG_SOLID);
    const bool active = !(fl & FLAG_SOLID);
    const uint32_t wm = (active && (fl & FLAG_NEARWALL)) ? at(a.wmask, c) : 0u;
    R v[Q];
    if (active) {
        if constexpr (kPop32)
            aa_gather32<Codec, kOdd>(f, a, c, x, z, wm, v);
        else
            aa_gather<Codec, kOdd, kSparse>(f, a, c, x, z, wm, v, nb, lane);
        wall_function<Codec, kSparse>(a, c, x, row - a.ny * z, fl, wm, v);
        collide_cell<Codec>(a, c, fl, wm, v);
    }
    if (boundary_wave(a, xw)) boundary_rows<Codec>(a, v, x, row, xw, active);
    if (active) {
        if constexpr (kPop32)
            aa_store32<Codec, kOdd>(f, a, c, x, z, wm, v);
        else
            aa_store<Codec, kOdd, kSparse>(f, a, c, x, z, wm, v, nb, lane);
    }
}

template <class Codec, bool kOdd, bool kPop32>
__global__ void __launch_bounds__(256) k_step_aa(typename Codec::T* __restrict__ f, StepArgs<typename Codec::R> a) {
    int row, tile;
    if (!tile_of(blockIdx.x, a.tiles, a.rows, row, tile)) return;   // whole work-group
    step_tile_aa<Codec, kOdd, kPop32, false>(f, a, row, tile);
}

// Tile-sparse layout: one stored tile per work-group (blk2tile; -1 idle).
template <class Codec, bool kOdd>
__global__ void __launch_bounds__(64) k_step_aa_sparse(typename Codec::T* __restrict__ f,
                                                       StepArgs<typename Codec::R> a) {
    const int t = __builtin_amdgcn_readfirstlane(a.blk2tile[blockIdx.x]);
    if (t < 0) return;
    const int tx = a.tcoord[3 * t], y = a.tcoord[3 * t + 1], z = a.tcoord[3 * t + 2];
    step_tile_aa<Codec, kOdd, false, true>(f, a, y + a.ny * z, tx, t);
}

// The step with the learned closure (float arithmetic, pair-form collision).
// The network runs on the whole wavefront between the moments and the
// relaxation, so every lane reaches it; inactive lanes feed zero features.
// Tile-sparse layout as in step_tile (one 64-cell tile per work-group).
template <class Codec, bool kSparse = false, int kStream = 0>
__device__ __forceinline__ void step_tile_learned(const typename Codec::T* __restrict__ fs,
                                                  typename Codec::T* __restrict__ fd, const StepArgs<float>& a,
                                                  int row, int tile, MlpPack* slots, int t = 0,
                                                  const int* nbp = nullptr) {
    using R = float;
    PackRegs pr;
    pack_load(a.mlp, pr);   // the network weights, issued before the population loads
    const int lane = int(threadIdx.x) & 63;
    const int x = kSparse ? tile * 64 + lane : tile * int(blockDim.x) + int(threadIdx.x);
    const int xw = kSparse ? tile * 64 : tile * int(blockDim.x) + (int(threadIdx.x) & ~63);
    const bool inrow = x < a.nx;
    const int z = row / a.ny;
    const int c = kSparse ? t * 64 + lane : x + a.px * row;
    const int* nb = kSparse ? (nbp ? nbp : a.nbr + 27 * t) : nullptr;
    const uint8_t fl = inrow ? at(a.flag, c) : uint8_t(FLAG_SOLID);
    const bool active = !(fl & FLAG_SOLID);
    const bool forced = active && (fl & FLAG_FORCED) != 0;
    const bool any_forced = __any(forced);   // uniform over the wavefront
    R f[Q];
    PairMoments<R> m;
    R fx = R(0), fy = R(0), fz = R(0);
    uint32_t wm = 0u;
    float feat[kNF] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (active) {
        wm = (fl & FLAG_NEARWALL) ? at(a.wmask, c) : 0u;
        if constexpr (kStream == 0) {
            // non-temporal population loads: faster here (FP32 0.79 -> 0.75 ms on
            // 23.5 M cells), slower in the plain step (0.617 -> 0.637 ms)
            gather<Codec, kSparse, kPopLoadNT>(fs, a, c, x, row - a.ny * z, z, fl, wm, f, nb, lane);
        } else {   // in place: fd is the one buffer
            aa_gather<Codec, kStream == 1, kSparse>(fd, a, c, x, z, wm, f, nb, lane);
            wall_function<Codec, kSparse>(a, c, x, row - a.ny * z, fl, wm, f);
        }
        if (forced) body_force(*a.aux, int(c), a.fscale, fx, fy, fz);
        if (any_forced) pair_moments<R, true>(f, fx, fy, fz, m);
        else pair_moments<R, false>(f, fx, fy, fz, m);
        // resolved strain rate S = k Pi_dev, k = -3 / (2 rho tau_s) = -3 / (rho (tau0 + sqrt(...)))
        // with tau_s the constant-Cs relaxation time
        const R k = R(-3) * __builtin_amdgcn_rcpf(m.r * (a.tau0 + pair_sqrt(fmaf(a.ksmag, m.qr, a.tau0sq))));
        const R t3 = (m.pxx + m.pyy + m.pzz) * (R(1) / R(3));
        closure_features(k, m.pxx - t3, m.pyy - t3, m.pzz - t3, m.pxy, m.pxz, m.pyz, m.ux, m.uy, m.uz, at(a.geo, c),
                         at(a.hcol, x + a.px * (row - a.ny * z)), a.inv_u0, feat);
    }
    // the network runs on wavefronts with a fluid cell (a uniform branch)
    float y = 0.f;
    const MlpPack* P = pack_store(pr, slots);   // all lanes (reconverged)
    if (__any(active)) y = mlp_wave(feat, P);
    // Less state across the network (more resident waves): the momentum is
    // formed again from rho and u, and the forces are read again, after it;
    // the empty asm keeps the compiler from carrying the first reads across.
    asm volatile("" ::: "memory");
    if (active) {
        m.jx = m.r * m.ux; m.jy = m.r * m.uy; m.jz = m.r * m.uz;   // u carries the half-force shift
        if (forced) body_force(*a.aux, int(c), a.fscale, fx, fy, fz);
        float c2 = kC2Min + kC2Scale * softplus_f(y);
        // outlet sponge: no less than the constant model's dissipation in the
        // last kOutletSponge cells before the outlet row
        if (x >= a.nx - kOutletSponge) c2 = fmaxf(c2, a.cs2);
        const R ks = R(18) * R(1.41421356) * c2;
        if (any_forced) pair_relax<R, true>(f, fx, fy, fz, m, a.tau0, a.tau0sq, ks);
        else pair_relax<R, false>(f, fx, fy, fz, m, a.tau0, a.tau0sq, ks);
        if (a.write_all || (wm != 0u && (fl & FLAG_CF))) {
            const StepAux<R>& ax = *a.aux;
            at(ax.rho, c) = m.r; at(ax.ux, c) = m.ux; at(ax.uy, c) = m.uy; at(ax.uz, c) = m.uz;
        }
        if (a.write_all && a.c2out) at(a.c2out, c) = c2;
        if (a.colu_on) { at(a.aux->colu_out_x, c) = m.ux; at(a.aux->colu_out_y, c) = m.uy; }
    }
    if constexpr (kStream == 0) {
        finish<Codec>(fd, a, f, c, x, row, xw, active);
    } else {
        if (boundary_wave(a, xw)) boundary_rows<Codec>(a, f, x, row, xw, active);
        if (active) aa_store<Codec, kStream == 1, kSparse>(fd, a, c, x, z, wm, f, nb, lane);
    }
}

// Resident waves per SIMD asked of the register allocator for the step. The
// 16-bit formats ask for 7 and get 8 (BF16 0.349 -> 0.340 ms on 23.5 M cells);
// FP32 gets 6 and does not change when asked for 7. The per-step column
// speed is a separate instance (kColu): at 8 waves its FP16 step lost 6 %.
template <class Codec, bool kColu>
constexpr int kStepWaves = (sizeof(typename Codec::T) == 2 && !(kColu && std::is_same<Codec, CodecF16Dev>::value) &&
                            WFS_STEP_WAVES_PER_EU < 7)
                               ? 7
                               : WFS_STEP_WAVES_PER_EU;

// kColu: the per-step column speed (StepArgs::colu_on); without it the
// column code drops out of the instance.
template <class Codec, bool kColu>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kStepWaves<Codec, kColu>)))
k_step(const typename Codec::T* __restrict__ fs, typename Codec::T* __restrict__ fd, StepArgs<typename Codec::R> a) {
    if constexpr (!kColu) a.colu_on = 0;
    int row, tile;
    if (!tile_of(blockIdx.x, a.tiles, a.rows, row, tile)) return;   // whole work-group
    step_tile<Codec>(fs, fd, a, row, tile);
}

// kStream: 0 two buffers (fs -> fd); 1 / 2 the odd / even in-place step on
// the one buffer fd (fs unused).
// Resident waves per SIMD asked for the learned step: the two-buffer FP32
// instances fit 6 without spills (0.683 -> 0.666 ms on 23.5 M cells); the
// 16-bit formats keep 5 (at 6 they spill a register and run 1.5 % slower,
// the tile-sparse meander 3 %), and so do the in-place instances (at 6 the
// FP32 ones spill).
template <class Codec, int kStream>
constexpr int kLearnedWaves = (sizeof(typename Codec::T) == 4 && kStream == 0) ? 6 : 5;

template <class Codec, int kStream = 0>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kLearnedWaves<Codec, kStream>))) k_step_learned(const typename Codec::T* __restrict__ fs,
                                                      typename Codec::T* __restrict__ fd, StepArgs<float> a) {
    extern __shared__ MlpPack slots[];   // one network pack per wavefront (dynamic LDS, blockDim / 64 of them)
    int row, tile;
    if (!tile_of(blockIdx.x, a.tiles, a.rows, row, tile)) return;   // whole work-group
    step_tile_learned<Codec, false, kStream>(fs, fd, a, row, tile, slots);
}

// Per-work-group record of the tile-sparse step kernels, in block order: the
// tile (-1: idle work-group), its (x-tile, y, z) and its 27 neighbour tiles,
// padded to 128 bytes. A work-group finds all it needs to start its loads in
// one round of scalar loads from blockIdx, where blk2tile -> (tcoord, nbr)
// took two dependent rounds: the two-buffer FP16 step on the 1 km meander
// 13.15 -> 12.43 ms, its in-place learned step 17.16 -> 17.01 ms. (The
// in-place constant-Cs step ran 4 % slower with it and keeps the per-tile
// tables.)
constexpr int kWgRec = 32;

// Tile-sparse layout: one stored tile per work-group of one wavefront; the
// block map keeps every tile of a row on one XCD. (Four tiles per
// work-group, which shared one staging of the network while it was staged
// per work-group, ran the 1 km meander's in-place learned step at 18.7 ms
// against 17.6 ms for one, with two at 17.9 ms.)
template <class Codec, int kStream = 0>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kLearnedWaves<Codec, kStream>))) k_step_learned_sparse(const typename Codec::T* __restrict__ fs,
                                                               typename Codec::T* __restrict__ fd,
                                                               StepArgs<float> a) {
    __shared__ MlpPack slots[1];   // the network pack of the wavefront
    // the tile is uniform over the wavefront: scalar registers, scalar loads of its record
    const int* rec = a.wgrec + kWgRec * (long long)blockIdx.x;
    const int t = __builtin_amdgcn_readfirstlane(rec[0]);
    if (t < 0) return;   // whole wavefront
    step_tile_learned<Codec, true, kStream>(fs, fd, a, rec[2] + a.ny * rec[3], rec[1], slots, t, rec + 4);
}

template <class Codec, int G>
__global__ void __launch_bounds__(64 * G) k_step_sparse(const typename Codec::T* __restrict__ fs,
                                                       typename Codec::T* __restrict__ fd,
                                                       StepArgs<typename Codec::R> a) {
    static_assert(G == 1, "one tile per work-group (the records are per work-group)");
    const int* rec = a.wgrec + kWgRec * (long long)blockIdx.x;
    const int t = __builtin_amdgcn_readfirstlane(rec[0]);
    if (t < 0) return;   // whole wavefront
    step_tile<Codec, true>(fs, fd, a, rec[2] + a.ny * rec[3], rec[1], t, rec + 4);
}

// ---------------------------------------------------------------- many reaches in one launch
// Per-case view for the batched step: the case's step arguments, its two
// population buffers, the device scalar that holds its current force scale
// (written by the cadence kernel), and the first work-group of the case in
// the flat launch (a multiple of 8, so the XCD map of each case is kept).
template <class Codec>
struct CaseArgs {
    StepArgs<typename Codec::R> a;
    typename Codec::T* f[2];
    const typename Codec::R* fscale;
    unsigned group_begin;
};

// Step s of a chunk for all cases: work-group g belongs to case
// group_case[g] and covers one (row, x-tile) of it. af[2k] is how many
// steps of the chunk case k takes (its budget; the cadence kernel writes
// it) and af[2k+1] its buffer-parity correction. A case whose budget ends
// before step s sits the launch out; the batch parity, corrected by the
// case's flip, selects its source buffer; the case's last step of the
// chunk writes the macroscopic fields of every cell (the chunk end of a
// sequential run).
// Resident waves per SIMD asked for the batched step. FP32 asks for 6,
// which k_step reaches by itself: the per-case schedule reads would
// otherwise take the batched FP32 step to 82 VGPRs and 5 waves. The 16-bit
// instances reach 7 unasked (65-66 VGPRs); asked for 7 they spill scalars.
template <class Codec>
constexpr int kBatchWaves = (sizeof(typename Codec::T) == 4 && WFS_STEP_WAVES_PER_EU < 6) ? 6 : WFS_STEP_WAVES_PER_EU;

template <class Codec>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kBatchWaves<Codec>)))
k_step_batch(const CaseArgs<Codec>* __restrict__ cases, const unsigned short* __restrict__ group_case,
             const int* __restrict__ af, int parity, int s) {
    const unsigned g = blockIdx.x;
    const int k = group_case[g];
    const int act = af[2 * k];
    if (s >= act) return;   // whole work-group: the case has stopped
    const CaseArgs<Codec>& cs = cases[k];
    const int par = parity ^ af[2 * k + 1];
    StepArgs<typename Codec::R> a = cs.a;
    a.fscale = *cs.fscale;
    a.write_all = s == act - 1 ? 1 : 0;
    int row, tile;
    if (!tile_of(g - cs.group_begin, a.tiles, a.rows, row, tile)) return;   // whole work-group
    step_tile<Codec>(cs.f[par], cs.f[par ^ 1], a, row, tile);
}

// The batched step for cases streamed in place, each on its one population
// buffer (f[0] == f[1]): the batch parity corrected by the case's flip is its
// streaming phase, 0 for state S (the odd step) and 1 for the even step, so
// cases stopped for an odd number of steps still take the right one. The
// branch is uniform over the work-group. In-place cases have their outlet
// row fused into the step. kPop32: every case's buffer is below 4 GiB.
// (Asked for 6 waves per SIMD at FP32, from 5 at 84 VGPRs, neither this
// kernel nor the single-reach in-place step ran faster.)
template <class Codec, bool kPop32>
__global__ void __launch_bounds__(256)
k_step_batch_aa(const CaseArgs<Codec>* __restrict__ cases, const unsigned short* __restrict__ group_case,
                const int* __restrict__ af, int parity, int s) {
    const unsigned g = blockIdx.x;
    const int k = group_case[g];
    const int act = af[2 * k];
    if (s >= act) return;   // whole work-group: the case has stopped
    const CaseArgs<Codec>& cs = cases[k];
    const int phase = (parity ^ af[2 * k + 1]) & 1;
    StepArgs<typename Codec::R> a = cs.a;
    a.fscale = *cs.fscale;
    a.write_all = s == act - 1 ? 1 : 0;
    int row, tile;
    if (!tile_of(g - cs.group_begin, a.tiles, a.rows, row, tile)) return;   // whole work-group
    if (phase == 0)
        step_tile_aa<Codec, true, kPop32, false>(cs.f[0], a, row, tile);
    else
        step_tile_aa<Codec, false, kPop32, false>(cs.f[0], a, row, tile);
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
__device__ __forceinline__ void bc_row(typename Codec::T* __restrict__ fd, const uint8_t* __restrict__ flag,
                                       const typename Codec::R* __restrict__ uin, int nx, int px, long long N,
                                       typename Codec::R rho_out, int sides, Center ctr, int j) {
    using R = typename Codec::R;
    const long long row = (long long)px * j;  // j = y + ny*z
    for (int side = 0; side < 2; ++side) {
        if (!((sides >> side) & 1)) continue;
        const long long c0 = side == 0 ? row : row + nx - 1;
        const long long c1 = side == 0 ? row + 1 : row + nx - 2;
        if ((flag[c0] | flag[c1]) & FLAG_SOLID) continue;
        R v[Q];
        R r = R(0), mx = R(0), my = R(0), mz = R(0);
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            v[i] = Codec::template ld<i>(fd + (long long)i * N + c1, ctr.c[center_class(i)]);
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
            Codec::template st<i>(fd + (long long)i * N + c0, e + (v[i] - feq_generic<Codec>(i, r, ux, uy, uz)),
                                  store_key<Codec>(ctr, i));
        });
    }
}

template <class Codec>
__global__ void __launch_bounds__(256) k_bc(typename Codec::T* __restrict__ fd, const uint8_t* __restrict__ flag,
                                            const typename Codec::R* __restrict__ uin, int nx, int px, int rows,
                                            long long N, typename Codec::R rho_out, int sides, Center ctr) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= rows) return;
    bc_row<Codec>(fd, flag, uin, nx, px, N, rho_out, sides, ctr, j);
}

// The outlet row of case k of a batch after step s of a chunk (cases whose
// outlet row opens a wavefront), on the buffer that step wrote; skipped
// when the case has stopped.
template <class Codec>
__global__ void __launch_bounds__(256) k_bc_batch(const CaseArgs<Codec>* __restrict__ cases,
                                                  const int* __restrict__ af, int k, int parity, int s) {
    if (s >= af[2 * k]) return;
    const CaseArgs<Codec>& cs = cases[k];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cs.a.rows) return;
    const int par = parity ^ af[2 * k + 1];
    bc_row<Codec>(cs.f[par ^ 1], cs.a.flag, cs.a.uin, cs.a.nx, cs.a.px, cs.a.N, typename Codec::R(1), 2, cs.a.ctr, j);
}

}  // namespace wfs
