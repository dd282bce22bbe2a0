// Learned sub-grid closure: a local Smagorinsky coefficient C^2 predicted per
// cell by an 8-32-32-1 network.
//
// The model. With the resolved deviatoric strain rate S (lattice units, filter
// width one cell) and |S| = sqrt(2 S:S), the eddy viscosity is
//     nu_t = C^2 |S|,   C^2 = kC2Min + kC2Scale * softplus(y(x)),
// y the network output for the feature vector x of the cell. Inside the LBM
// step nu_t enters exactly like the constant-coefficient Smagorinsky model:
//     tau = (tau0 + sqrt(tau0^2 + 18 sqrt(2) C^2 Q / rho)) / 2,
// Q = |Pi_neq|, so C^2 = Cs^2 everywhere is plain Smagorinsky.
//
// Features (all cells, lattice units):
//   x0 = ln(|S| + 1e-7) / 4 + 1.5        strain-rate magnitude
//   x1 = sqrt(6) tr(S^3) / (S:S)^1.5     strain shape, in [-1, 1]
//   x2 = |u| / U0 - 1                    speed relative to the mean inlet speed U0
//   x3 = 2 (Sxz^2 + Syz^2) / S:S         share of vertical shear, in [0, 1]
//   x4 = u_z / U0                        vertical velocity
//   x5 = 1 / d                           d: distance to the nearest wall cell centre (cells)
//   x6 = z_rel                           height above the bed over the column depth
//   x7 = h / 64                          h: fluid cells in the column
// x5 and x6 are stored per cell as two bytes (value * 255, rounded), h per
// column as one byte. Inputs and the two hidden layers' weights are BF16 with
// FP32 accumulation; the output layer is FP32.
//
// Training fits C^2 to box-filtered fine-grid runs in stress space: for a
// sample with subfilter stress tau (deviatoric part T) and resolved strain S,
//     loss = |T + 2 C^2 |S| S|^2 = P + 4 C^2 B + 4 C^4 A,
//     A = |S|^2 S:S,  B = |S| (T:S),  P = T:T,
// so a sample carries (A, B, P) besides its features.
#pragma once
#include <cstddef>

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wfs {

constexpr int kNF = 8;    // features
constexpr int kNH = 32;   // hidden width
// Parameter vector, natural order: W1[32][8] (out, in), b1[32], W2[32][32] (out, in), b2[32], w3[32], b3.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kNH * kNF;
constexpr int kOffW2 = kOffB1 + kNH;
constexpr int kOffB2 = kOffW2 + kNH * kNH;
constexpr int kOffW3 = kOffB2 + kNH;
constexpr int kOffB3 = kOffW3 + kNH;
constexpr int kNParams = kOffB3 + 1;   // 1377

constexpr float kC2Min = 0.01f;     // floor of C^2 (Cs = 0.1): the near-inviscid lattice (tau0 ~ 0.5006) needs it
constexpr float kC2Scale = 0.0289f; // C^2 per unit softplus (0.17^2)
// Outlet sponge: in the last kOutletSponge cells before the outlet row
// (x >= nx - kOutletSponge) the learned C^2 is at least the constant model's
// Cs^2. The outlet fixes rho = 1 and copies the non-equilibrium part of
// x = nx - 2 (floodsim/lbm3d.py:164-173); where that row sees backflow
// (pressure waves in a closed-lid reach turn the outlet discharge negative),
// a coefficient at the 0.01 floor let a disturbance there grow until the
// 1 km meander went non-finite after 40,000 steps, which the constant model
// damps. The sponge keeps the constant model's dissipation there.
constexpr int kOutletSponge = 64;

typedef short s4v __attribute__((ext_vector_type(4)));
typedef short s2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef uint32_t u2v __attribute__((ext_vector_type(2)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));

// Weights in the register layout of the wavefront evaluation below.
//
// Both MFMA layers compute the transposed product out^T = W x^T, so the rows
// of a 32x32 tile are units and the columns are 32 cells. For the 32x32 FP32
// result, lane l holds column l % 32 and, in accumulator r = 0..15, the unit
//     row(r, h) = 8 (r / 4) + 4 h + (r % 4),   h = l / 32.
// A lane's 8 accumulators r = 8s..8s+7 are, after ReLU and rounding, its
// B-operand slots for k-step s of the next 32x32x16 product; the matching
// A-operand slots of W2 are packed in the same unit order, so no data moves
// between lanes from layer 1 to layer 2.
// The per-lane-row vectors come first, within reach of the LDS pair-read
// offsets (1020 bytes) from one base per lane; w2 stays 16-byte aligned.
struct alignas(16) MlpPack {
    float b1[2 * 16];          // [h][r]: b1[row(r, h)], the layer-1 accumulator start
    float b2[2 * 16];          // [h][r]: b2[row(r, h)]
    float w3[2 * 16];          // [h][r]: w3[row(r, h)]
    float b3;
    float pad[3];
    uint32_t w1[64 * 2];       // lane l: BF16 W1[l%32][4h .. 4h+3] (32x32x8 A operand)
    uint32_t w2[2 * 64 * 4];   // step s, lane l: BF16 W2[l%32][unit(s, h, 0..7)]
};
static_assert(offsetof(MlpPack, w2) % 16 == 0, "w2 is read as 16-byte vectors");
constexpr int kPackWords = (int)(sizeof(MlpPack) / 4);
static_assert(kPackWords % 4 == 0, "the pack is copied as 16-byte vectors");
constexpr int kPackVec = kPackWords / 4;   // 185 vectors of 16 bytes

__host__ __device__ constexpr int mfma_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }
// Hidden unit in B slot j (0..7) of k-step s for lane half h.
__host__ __device__ constexpr int l2_unit(int s, int h, int j) { return mfma_row(8 * s + j, h); }

// Two floats rounded to bfloat16 (nearest even) in one register, low half a:
// one v_cvt_pk_bf16_f32 (a vector conversion; element-wise conversions can
// become two conversions and a byte permute)
__device__ __forceinline__ uint32_t pk_bf16(float a, float b) {
    const f2v v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf2v));
}
__device__ __forceinline__ float bf16_round(float a) { return (float)(__bf16)a; }

// softplus(y) = ln(1 + e^y) = ln2 log2(1 + 2^(y log2 e)), on the hardware
// base-2 exponential and logarithm (about 1 ulp; for y below -20 the sum
// rounds to 1 and the result to 0, against e^y < 2e-9 exactly)
__device__ __forceinline__ float softplus_f(float y) {
    const float e = __builtin_amdgcn_exp2f(y * 1.44269504f);
    return y > 15.f ? y : 0.693147181f * __builtin_amdgcn_logf(1.f + e);
}

// The eight features of one cell (formulas in the header comment) from the
// strain rate S = k P, P trace-free (the step passes the deviatoric part of
// Pi_neq and k = -3 / (2 rho tau_s); the training samples pass S and k = 1).
// The invariants scale out of k: S:S = k^2 P:P, |S| = |k| sqrt(2 P:P),
// tr(S^3) / (S:S)^1.5 = sign(k) tr(P^3) / (P:P)^1.5, and for a trace-free
// P, tr(P^3) = 3 det P (Cayley-Hamilton), so x1 = sign(k) 3 sqrt(6) det P
// / (P:P)^1.5 and x3 = 2 (P_xz^2 + P_yz^2) / P:P. The network sees the
// features rounded to bfloat16, so the hardware square root, reciprocal
// square root and base-2 logarithm (about 1 ulp) serve here.
__device__ __forceinline__ void closure_features(float k, float pxx, float pyy, float pzz, float pxy, float pxz,
                                                 float pyz, float ux, float uy, float uz, uint16_t geo, uint8_t hcol,
                                                 float inv_u0, float* x) {
    const float vz = fmaf(pxz, pxz, pyz * pyz);
    const float pp = fmaf(pxx, pxx, fmaf(pyy, pyy, pzz * pzz)) + 2.f * fmaf(pxy, pxy, vz);
    const bool live = (k * k) * pp > 1e-20f;
    const float smag = fabsf(k) * __builtin_amdgcn_sqrtf(2.f * pp);
    x[0] = fmaf(0.25f * 0.693147181f, __builtin_amdgcn_logf(smag + 1e-7f), 1.5f);
    const float rs = __builtin_amdgcn_rsqf(pp);
    const float rs2 = rs * rs;
    const float det = pxx * fmaf(pyy, pzz, -pyz * pyz) - pxy * fmaf(pxy, pzz, -pyz * pxz) +
                      pxz * fmaf(pxy, pyz, -pyy * pxz);
    const float c1 = k < 0.f ? -7.34846923f : 7.34846923f;   // 3 sqrt(6)
    x[1] = live ? __builtin_amdgcn_fmed3f(c1 * det * (rs2 * rs), -1.f, 1.f) : 0.f;
    x[2] = __builtin_amdgcn_sqrtf(fmaf(ux, ux, fmaf(uy, uy, uz * uz))) * inv_u0 - 1.f;
    x[3] = live ? 2.f * vz * rs2 : 0.f;
    x[4] = uz * inv_u0;
    x[5] = float(geo & 255u) * (1.f / 255.f);
    x[6] = float(geo >> 8) * (1.f / 255.f);
    x[7] = float(hcol) * (1.f / 64.f);
}

// ReLU as one integer max on the bit pattern: negative floats (and -0) have
// negative int32 patterns and become +0; others are unchanged. fmaxf with 0
// costs two instructions here (the operand is first made a canonical NaN).
__device__ __forceinline__ float relu_f(float v) {
    return __int_as_float(max(__float_as_int(v), 0));
}

// ReLU of two packed bfloat16 values as one packed 16-bit integer max:
// negative values (and -0) have negative int16 patterns and become +0,
// others are unchanged. Rounding to bfloat16 keeps the sign, so ReLU after
// the rounding equals the rounding after ReLU.
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
    const s2v z = {0, 0};
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s2v, v), z));
}

// Network output y for the cell of this lane. All 64 lanes of the wavefront
// must call this together (it exchanges data between lanes and runs MFMA).
// P points to the pack in LDS.
__device__ __forceinline__ float mlp_wave(const float* x, const MlpPack* P) {
    const int lane = int(threadIdx.x) & 63;
    const int h = lane >> 5;
    const uint32_t xp0 = pk_bf16(x[0], x[1]), xp1 = pk_bf16(x[2], x[3]);
    const uint32_t xp2 = pk_bf16(x[4], x[5]), xp3 = pk_bf16(x[6], x[7]);
    // Tile b holds cells 32b .. 32b+31 of the wavefront; lane l supplies
    // features 4h .. 4h+3 of cell 32b + l%32. Each lane sends the half it does
    // not use itself to its partner lane l ^ 32.
    const uint32_t r0 = __shfl_xor(h ? xp0 : xp2, 32);
    const uint32_t r1 = __shfl_xor(h ? xp1 : xp3, 32);
    const u2v own = h ? u2v{xp2, xp3} : u2v{xp0, xp1};
    const u2v other = u2v{r0, r1};
    const s4v a1 = __builtin_bit_cast(s4v, u2v{P->w1[2 * lane], P->w1[2 * lane + 1]});
    const bf8v a2s0 = __builtin_bit_cast(bf8v, *reinterpret_cast<const u4v*>(&P->w2[(0 * 64 + lane) * 4]));
    const bf8v a2s1 = __builtin_bit_cast(bf8v, *reinterpret_cast<const u4v*>(&P->w2[(1 * 64 + lane) * 4]));
    float y = 0.f;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        // this lane's rows of the biases and output weights (one LDS base,
        // constant offsets), read again for each tile: the row offset is made
        // opaque per tile, so the 48 values are not held in registers across
        // both tiles (126 -> 91 VGPRs in the learned step, 4 -> 5 waves per
        // SIMD: FP16 0.659 -> 0.596 ms on 23.5 M cells)
        int hb = 16 * h;
        asm volatile("" : "+v"(hb));
        const float* b1 = P->b1 + hb;
        const float* b2 = P->b2 + hb;
        const float* w3 = P->w3 + hb;
        const u2v xb = (b == h) ? own : other;
        f16v acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = b1[r];
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, __builtin_bit_cast(s4v, xb), acc, 0, 0, 0);
        u4v hq0, hq1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            hq0[q] = relu_bf16x2(pk_bf16(acc[2 * q], acc[2 * q + 1]));
            hq1[q] = relu_bf16x2(pk_bf16(acc[8 + 2 * q], acc[8 + 2 * q + 1]));
        }
        f16v acc2;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc2[r] = b2[r];
        acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2s0, __builtin_bit_cast(bf8v, hq0), acc2, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2s1, __builtin_bit_cast(bf8v, hq1), acc2, 0, 0, 0);
        float part = 0.f;
#pragma unroll
        for (int r = 0; r < 16; ++r) part = fmaf(w3[r], relu_f(acc2[r]), part);
        const float yb = part + __shfl_xor(part, 32) + P->b3;
        y = (b == h) ? yb : y;
    }
    return y;
}

// Copies the pack (global) into LDS; every thread of the block (at least 64)
// calls it. All loads are issued before the first LDS write.
__device__ __forceinline__ void stage_pack(const MlpPack* __restrict__ g, MlpPack* s) {
    const u4v* src = reinterpret_cast<const u4v*>(g);
    u4v* dst = reinterpret_cast<u4v*>(s);
    const int t = int(threadIdx.x), nt = int(blockDim.x);
    u4v v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (t + i * nt < kPackVec) v[i] = src[t + i * nt];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (t + i * nt < kPackVec) dst[t + i * nt] = v[i];
    __syncthreads();
}

// The step stages the pack per wavefront, in two parts: pack_load issues the
// global loads of the 64 lanes' share (3 x 16 bytes) into registers early,
// before the population loads, and pack_store writes them into the
// wavefront's own LDS slot once the populations are in flight. No
// work-group barrier is involved (each wavefront reads only its own slot,
// and one wavefront's LDS accesses complete in order), so no wavefront
// waits for another, and the weights' load latency hides behind the
// population loads. Staging once per work-group behind a barrier held every
// wavefront's population loads until the weights had arrived: the learned
// FP16 step took 0.552 ms on 23.5 M cells that way, 0.452 ms this way.
struct PackRegs {
    u4v v[3];
};
__device__ __forceinline__ void pack_load(const MlpPack* __restrict__ g, PackRegs& r) {
    const int lane = int(threadIdx.x) & 63;
    const u4v* src = reinterpret_cast<const u4v*>(g);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (lane + 64 * i < kPackVec) r.v[i] = src[lane + 64 * i];
}
// All 64 lanes call it together; returns the wavefront's slot.
__device__ __forceinline__ const MlpPack* pack_store(const PackRegs& r, MlpPack* slots) {
    const int lane = int(threadIdx.x) & 63;
    MlpPack* s = slots + (int(threadIdx.x) >> 6);
    u4v* dst = reinterpret_cast<u4v*>(s);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (lane + 64 * i < kPackVec) dst[lane + 64 * i] = r.v[i];
    // the compiler must not move the lanes' reads of other lanes' words above the writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return s;
}

// ---------------------------------------------------------------- kernels

// Natural parameter vector -> pack. One block of 256 threads.
__global__ void __launch_bounds__(256) k_mlp_pack(const float* __restrict__ p, MlpPack* __restrict__ out) {
    const int t = int(threadIdx.x);
    if (t < 64) {
        const int i = t & 31, h = t >> 5;
        const float* w = p + kOffW1 + i * kNF + 4 * h;
        out->w1[2 * t] = pk_bf16(w[0], w[1]);
        out->w1[2 * t + 1] = pk_bf16(w[2], w[3]);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float* w2 = p + kOffW2 + i * kNH;
                out->w2[(s * 64 + t) * 4 + q] = pk_bf16(w2[l2_unit(s, h, 2 * q)], w2[l2_unit(s, h, 2 * q + 1)]);
            }
    } else if (t < 96) {
        const int h = (t - 64) >> 4, r = (t - 64) & 15;
        const int u = mfma_row(r, h);
        out->b1[h * 16 + r] = p[kOffB1 + u];
        out->b2[h * 16 + r] = p[kOffB2 + u];
        out->w3[h * 16 + r] = p[kOffW3 + u];
    } else if (t == 96) {
        out->b3 = p[kOffB3];
        out->pad[0] = out->pad[1] = out->pad[2] = 0.f;
    }
}

// C^2 of n feature vectors (row-major n x 8) through the wavefront MFMA path.
__global__ void __launch_bounds__(256) k_mlp_eval(const float* __restrict__ X, long long n,
                                                  const MlpPack* __restrict__ pack, float* __restrict__ c2) {
    __shared__ MlpPack P;
    stage_pack(pack, &P);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float x[kNF];
#pragma unroll
    for (int j = 0; j < kNF; ++j) x[j] = k < n ? X[k * kNF + j] : 0.f;
    const float y = mlp_wave(x, &P);
    if (k < n) c2[k] = kC2Min + kC2Scale * softplus_f(y);
}

// Training: loss and gradient of a batch, one sample per thread, FP32 with
// the forward pass rounded like the MFMA path (BF16 inputs, W1, W2 and
// layer-1 activations); the gradient treats the rounding as the identity.
// Each block writes its partial gradient and loss sums (kNParams + 2 floats)
// to part[blockIdx.x]; k_mlp_reduce adds the blocks in a fixed order.
constexpr int kGradBlock = 64;
constexpr int kPartStride = kNParams + 3;   // grad, sum loss, sum P, pad

__global__ void __launch_bounds__(kGradBlock) k_mlp_grad(const float* __restrict__ X, const float* __restrict__ ABP,
                                                          long long n, const float* __restrict__ p,
                                                          float* __restrict__ part) {
    __shared__ float W[kNParams];
    __shared__ float sh1[kNH][kGradBlock];   // layer-1 activations (rounded)
    __shared__ float sd1[kNH][kGradBlock];   // layer-1 deltas
    __shared__ float sh2[kNH][kGradBlock];   // layer-2 activations
    __shared__ float sd2[kNH][kGradBlock];   // layer-2 deltas
    __shared__ float sx[kNF][kGradBlock];    // inputs (rounded)
    __shared__ float sdy[kGradBlock];
    __shared__ float sl[2][kGradBlock];
    const int t = int(threadIdx.x);
    for (int k = t; k < kNParams; k += kGradBlock) {
        const bool q = (k < kOffB1) || (k >= kOffW2 && k < kOffB2);
        W[k] = q ? bf16_round(p[k]) : p[k];
    }
    __syncthreads();
    const long long s = (long long)blockIdx.x * kGradBlock + t;
    const bool ok = s < n;
    float x[kNF];
#pragma unroll
    for (int j = 0; j < kNF; ++j) x[j] = ok ? bf16_round(X[s * kNF + j]) : 0.f;
    float z1[kNH], h1[kNH], z2[kNH];
#pragma unroll
    for (int i = 0; i < kNH; ++i) {
        float a = W[kOffB1 + i];
#pragma unroll
        for (int j = 0; j < kNF; ++j) a = fmaf(W[kOffW1 + i * kNF + j], x[j], a);
        z1[i] = a;
        h1[i] = bf16_round(fmaxf(a, 0.f));
    }
    float y = W[kOffB3];
#pragma unroll
    for (int i = 0; i < kNH; ++i) {
        float a = W[kOffB2 + i];
#pragma unroll
        for (int j = 0; j < kNH; ++j) a = fmaf(W[kOffW2 + i * kNH + j], h1[j], a);
        z2[i] = a;
        y = fmaf(W[kOffW3 + i], fmaxf(a, 0.f), y);
    }
    float dy = 0.f, loss = 0.f, pp = 0.f;
    if (ok) {
        const float A = ABP[s * 3], B = ABP[s * 3 + 1], P = ABP[s * 3 + 2];
        const float sp = y > 15.f ? y : log1pf(expf(y));
        const float sg = 1.f / (1.f + expf(-y));
        const float c2 = kC2Min + kC2Scale * sp;
        loss = P + 4.f * c2 * B + 4.f * c2 * c2 * A;
        pp = P;
        dy = (4.f * B + 8.f * c2 * A) * kC2Scale * sg;
    }
    float d2[kNH];
#pragma unroll
    for (int i = 0; i < kNH; ++i) d2[i] = z2[i] > 0.f ? dy * W[kOffW3 + i] : 0.f;
#pragma unroll
    for (int j = 0; j < kNH; ++j) {
        float a = 0.f;
#pragma unroll
        for (int i = 0; i < kNH; ++i) a = fmaf(W[kOffW2 + i * kNH + j], d2[i], a);
        sd1[j][t] = z1[j] > 0.f ? a : 0.f;
        sh1[j][t] = h1[j];
        sh2[j][t] = fmaxf(z2[j], 0.f);
        sd2[j][t] = d2[j];
    }
#pragma unroll
    for (int j = 0; j < kNF; ++j) sx[j][t] = x[j];
    sdy[t] = dy;
    sl[0][t] = loss;
    sl[1][t] = pp;
    __syncthreads();
    float* out = part + (long long)blockIdx.x * kPartStride;
    for (int k = t; k < kNParams + 2; k += kGradBlock) {
        float g = 0.f;
        if (k < kOffB1) {
            const int i = k / kNF, j = k % kNF;
            for (int m = 0; m < kGradBlock; ++m) g = fmaf(sd1[i][m], sx[j][m], g);
        } else if (k < kOffW2) {
            const int i = k - kOffB1;
            for (int m = 0; m < kGradBlock; ++m) g += sd1[i][m];
        } else if (k < kOffB2) {
            const int i = (k - kOffW2) / kNH, j = (k - kOffW2) % kNH;
            for (int m = 0; m < kGradBlock; ++m) g = fmaf(sd2[i][m], sh1[j][m], g);
        } else if (k < kOffW3) {
            const int i = k - kOffB2;
            for (int m = 0; m < kGradBlock; ++m) g += sd2[i][m];
        } else if (k < kOffB3) {
            const int i = k - kOffW3;
            for (int m = 0; m < kGradBlock; ++m) g = fmaf(sdy[m], sh2[i][m], g);
        } else if (k == kOffB3) {
            for (int m = 0; m < kGradBlock; ++m) g += sdy[m];
        } else {
            const int q = k - kNParams;
            for (int m = 0; m < kGradBlock; ++m) g += sl[q][m];
        }
        out[k] = g;
    }
}

// Sum of the block partials in block order (FP64). grad gets the mean over
// n samples divided by the mean P of the batch (the relative stress error's
// gradient); sums[0..1] get the loss and P sums.
__global__ void __launch_bounds__(256) k_mlp_reduce(const float* __restrict__ part, int blocks, long long n,
                                                    float* __restrict__ grad, double* __restrict__ sums) {
    __shared__ double psum;
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int b = 0; b < blocks; ++b) a += (double)part[(long long)b * kPartStride + kNParams + 1];
        psum = a;
    }
    __syncthreads();
    const double norm = psum > 0.0 ? 1.0 / psum : 0.0;
    for (int k = int(threadIdx.x); k < kNParams + 2; k += int(blockDim.x)) {
        double a = 0.0;
        for (int b = 0; b < blocks; ++b) a += (double)part[(long long)b * kPartStride + k];
        if (k < kNParams) grad[k] = (float)(a * norm);
        else sums[k - kNParams] = a;
    }
}

// Adam step (bias-corrected), t >= 1.
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + eps);
}

// ---------------------------------------------------------------- training samples
// A fine run is box-filtered 2x2x2 onto the coarse grid (coarse cell (X,Y,Z)
// covers fine cells 2X..2X+1, 2Y..2Y+1, 2Z..2Z+1). Velocities are scaled by
// kv into coarse lattice units.

// Pass 1: filtered velocity and subfilter stress mean((u - ubar)(u - ubar)).
template <typename R>
__global__ void __launch_bounds__(256) k_coarsen(const R* __restrict__ ux, const R* __restrict__ uy,
                                                 const R* __restrict__ uz, const uint8_t* __restrict__ flag,
                                                 int pxf, int nyf, int ncx, int ncy, int ncz, float kv,
                                                 float* __restrict__ ubar, float* __restrict__ tsub,
                                                 uint8_t* __restrict__ cfluid) {
    const long long nc = (long long)ncx * ncy * ncz;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int X = int(c % ncx), Y = int((c / ncx) % ncy), Z = int(c / ((long long)ncx * ncy));
    float u[8][3];
    bool fluid = true;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const long long f = (2 * X + (k & 1)) + (long long)pxf * ((2 * Y + ((k >> 1) & 1)) +
                                                                   (long long)nyf * (2 * Z + (k >> 2)));
        fluid = fluid && !(flag[f] & 1u);
        u[k][0] = kv * float(ux[f]);
        u[k][1] = kv * float(uy[f]);
        u[k][2] = kv * float(uz[f]);
    }
    float m[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) m[d] += u[k][d];
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] *= 0.125f;
    float t[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float a = u[k][0] - m[0], b = u[k][1] - m[1], e = u[k][2] - m[2];
        t[0] += a * a; t[1] += b * b; t[2] += e * e; t[3] += a * b; t[4] += a * e; t[5] += b * e;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) ubar[d * nc + c] = fluid ? m[d] : 0.f;
#pragma unroll
    for (int d = 0; d < 6; ++d) tsub[d * nc + c] = fluid ? t[d] * 0.125f : 0.f;
    cfluid[c] = fluid ? 1 : 0;
}

// Pass 2: resolved strain by central differences of the filtered velocity
// (the rigid lid mirrors the cell below it: ghost (ux, uy, -uz)), then the
// features and the loss coefficients. A sample is valid when the cell and its
// six face neighbours are fluid (or the lid above) and 0 < X < ncx-1.
__global__ void __launch_bounds__(256) k_samples(const float* __restrict__ ubar, const float* __restrict__ tsub,
                                                 const uint8_t* __restrict__ cfluid, const uint16_t* __restrict__ geo,
                                                 const uint8_t* __restrict__ hcol, int ncx, int ncy, int ncz,
                                                 float inv_u0, float* __restrict__ X, float* __restrict__ ABP,
                                                 uint8_t* __restrict__ valid) {
    const long long nc = (long long)ncx * ncy * ncz;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int x = int(c % ncx), y = int((c / ncx) % ncy), z = int(c / ((long long)ncx * ncy));
    const long long sx = 1, sy = ncx, sz = (long long)ncx * ncy;
    bool ok = cfluid[c] && x > 0 && x < ncx - 1 && y > 0 && y < ncy - 1 && z > 0;
    if (ok) ok = cfluid[c - sx] && cfluid[c + sx] && cfluid[c - sy] && cfluid[c + sy] && cfluid[c - sz] &&
                 (z == ncz - 1 || cfluid[c + sz]);
    valid[c] = ok ? 1 : 0;
    float feat[kNF] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float A = 0.f, B = 0.f, P = 0.f;
    if (ok) {
        float g[3][3];   // g[i][j] = d u_i / d x_j
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float* u = ubar + i * nc;
            g[i][0] = 0.5f * (u[c + sx] - u[c - sx]);
            g[i][1] = 0.5f * (u[c + sy] - u[c - sy]);
            const float up = z == ncz - 1 ? (i == 2 ? -u[c] : u[c]) : u[c + sz];
            g[i][2] = 0.5f * (up - u[c - sz]);
        }
        const float tr = (g[0][0] + g[1][1] + g[2][2]) * (1.f / 3.f);
        const float sxx = g[0][0] - tr, syy = g[1][1] - tr, szz = g[2][2] - tr;
        const float sxy = 0.5f * (g[0][1] + g[1][0]), sxz = 0.5f * (g[0][2] + g[2][0]),
                    syz = 0.5f * (g[1][2] + g[2][1]);
        const float* t = tsub + c;
        const float tt = (t[0] + t[nc] + t[2 * nc]) * (1.f / 3.f);
        const float Txx = t[0] - tt, Tyy = t[nc] - tt, Tzz = t[2 * nc] - tt;
        const float Txy = t[3 * nc], Txz = t[4 * nc], Tyz = t[5 * nc];
        const float ss = sxx * sxx + syy * syy + szz * szz + 2.f * (sxy * sxy + sxz * sxz + syz * syz);
        const float smag = sqrtf(2.f * ss);
        const float ts = Txx * sxx + Tyy * syy + Tzz * szz + 2.f * (Txy * sxy + Txz * sxz + Tyz * syz);
        A = smag * smag * ss;
        B = smag * ts;
        P = Txx * Txx + Tyy * Tyy + Tzz * Tzz + 2.f * (Txy * Txy + Txz * Txz + Tyz * Tyz);
        closure_features(1.f, sxx, syy, szz, sxy, sxz, syz, ubar[c], ubar[nc + c], ubar[2 * nc + c], geo[c],
                         hcol[x + (long long)ncx * y], inv_u0, feat);
#pragma unroll
        for (int j = 0; j < kNF; ++j) feat[j] = bf16_round(feat[j]);
    }
#pragma unroll
    for (int j = 0; j < kNF; ++j) X[c * kNF + j] = feat[j];
    ABP[c * 3] = A;
    ABP[c * 3 + 1] = B;
    ABP[c * 3 + 2] = P;
}

}  // namespace wfs
