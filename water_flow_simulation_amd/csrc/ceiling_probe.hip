// Streaming ceilings of the GPU for the access shapes of the lattice step.
//
// Measures, with HIP events: device copies of several widths, read-only and
// write-only streams, and the 19-population read + 19-population write
// pattern of one lattice step without its arithmetic, in the layouts a step
// kernel could use (structure of arrays with and without the pull offsets,
// vector access of 2 or 4 cells per work-item, blocked array-of-structure-of-
// arrays, in-place two-phase streaming, 16-bit populations). The fastest
// 19-stream figure bounds what any pull-stream step kernel can reach here.
//
//   build: python -m water_flow_simulation_amd.build   (writes water_flow_simulation_amd/ceiling_probe)
//   run:   water_flow_simulation_amd/ceiling_probe     (needs ~4.3 GB of device memory)
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("ERR %s %s\n", #x, hipGetErrorString(e)); exit(1);} } while (0)

__global__ void copy4x4(const float4* __restrict__ s, float4* __restrict__ d, long long n) {
    long long b = (long long)blockIdx.x * blockDim.x * 4 + threadIdx.x;
    for (int u = 0; u < 4; ++u) { long long k = b + (long long)u * blockDim.x; if (k < n) d[k] = s[k]; }
}
__global__ void copy4(const float4* __restrict__ s, float4* __restrict__ d, long long n) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; if (k < n) d[k] = s[k];
}
__global__ void copy4nt(const float4* __restrict__ s, float4* __restrict__ d, long long n) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) { float4 v = s[k]; __builtin_nontemporal_store(v.x, &d[k].x); __builtin_nontemporal_store(v.y, &d[k].y);
                 __builtin_nontemporal_store(v.z, &d[k].z); __builtin_nontemporal_store(v.w, &d[k].w); }
}
__global__ void copy1(const float* __restrict__ s, float* __restrict__ d, long long n) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; if (k < n) d[k] = s[k];
}
__global__ void copy1nt(const float* __restrict__ s, float* __restrict__ d, long long n) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; if (k < n) __builtin_nontemporal_store(s[k], &d[k]);
}
__global__ void read4(const float4* __restrict__ s, float* __restrict__ out, long long n) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0.f; if (k < n) { float4 v = s[k]; acc = v.x + v.y + v.z + v.w; }
    if (acc == 12345.678f) out[0] = acc;  // never true for zeroed input; keeps the load
}
__global__ void write4(float4* __restrict__ d, long long n) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; if (k < n) d[k] = make_float4(1.f, 2.f, 3.f, 4.f);
}
// 19 streams, one element per thread per stream; shift = pull offsets (0 = aligned copy)
__global__ void streams19(const float* __restrict__ s, float* __restrict__ d, long long N, int px, int plane, int pull, int nt) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    float v[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        long long sidx = pull ? c - CX[i] - (long long)px * CY[i] - (long long)plane * CZ[i] : c;
        if (sidx < 0) sidx = 0;
        if (sidx >= N) sidx = N - 1;
        v[i] = s[(long long)i * N + sidx];
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        if (nt) __builtin_nontemporal_store(v[i] * 1.0001f, &d[(long long)i * N + c]);
        else d[(long long)i * N + c] = v[i] * 1.0001f;
    }
}

template <int V>
struct VecT;
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));
template <> struct VecT<2> { using T = vf2; };
template <> struct VecT<4> { using T = vf4; };
// 19 streams, V consecutive cells per thread, vector loads at pull offsets (4-byte aligned), aligned vector stores
template <int V>
__global__ void streams19v(const float* __restrict__ s, float* __restrict__ d, long long N, int px, int plane, int nt) {
    using VT = typename VecT<V>::T;
    long long c0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * V;
    if (c0 + V > N) return;
    const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    VT v[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        long long sidx = c0 - CX[i] - (long long)px * CY[i] - (long long)plane * CZ[i];
        if (sidx < 0) sidx = 0;
        if (sidx > N - V) sidx = N - V;
        __builtin_memcpy(&v[i], s + (long long)i * N + sidx, sizeof(VT));
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        VT w = v[i];
        if (nt) __builtin_nontemporal_store(w, (VT*)(d + (long long)i * N + c0));
        else *(VT*)(d + (long long)i * N + c0) = w;
    }
}

// AoSoA: population i of cell c at ((c / B) * 19 + i) * B + c % B
template <int B>
__global__ void streams19_aosoa(const float* __restrict__ s, float* __restrict__ d, long long N, int px, int plane, int nt) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    float v[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        long long sidx = c - CX[i] - (long long)px * CY[i] - (long long)plane * CZ[i];
        if (sidx < 0) sidx = 0;
        if (sidx >= N) sidx = N - 1;
        v[i] = s[((sidx / B) * 19 + i) * B + (sidx % B)];
    }
    const long long base = (c / B) * 19 * B + (c % B);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        if (nt) __builtin_nontemporal_store(v[i] * 1.0001f, &d[base + (long long)i * B]);
        else d[base + (long long)i * B] = v[i] * 1.0001f;
    }
}

// In-place patterns on one array: mode 0 = same-cell read/write (AA even step),
// mode 1 = read at c - off_i, write at c + off_i (AA odd step shape).
__global__ void streams19_inplace(float* __restrict__ a, long long N, int px, int plane, int mode) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    float v[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        long long off = CX[i] + (long long)px * CY[i] + (long long)plane * CZ[i];
        long long sidx = mode ? c - off : c;
        if (sidx < 0) sidx = 0;
        if (sidx >= N) sidx = N - 1;
        v[i] = a[(long long)(mode ? OPP[i] : i) * N + sidx];
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        long long off = CX[i] + (long long)px * CY[i] + (long long)plane * CZ[i];
        long long didx = mode ? c + off : c;
        if (didx < 0) didx = 0;
        if (didx >= N) didx = N - 1;
        a[(long long)(mode ? i : OPP[i]) * N + didx] = v[i] * 1.0001f;
    }
}

// 16-bit populations. mode 0: one half per thread (ushort loads/stores);
// mode 1: two cells per thread, 4-byte loads at the (possibly 2-byte aligned) pull address;
// mode 2: two cells per thread, aligned 8-byte load around the pull address, pick the pair.
__global__ void streams19_half(const unsigned short* __restrict__ s, unsigned short* __restrict__ d, long long N, int px, int plane, int mode) {
    const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 0) {
        long long c = t;
        if (c >= N) return;
        unsigned short v[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            long long sidx = c - CX[i] - (long long)px * CY[i] - (long long)plane * CZ[i];
            sidx = sidx < 0 ? 0 : (sidx >= N ? N - 1 : sidx);
            v[i] = s[(long long)i * N + sidx];
        }
#pragma unroll
        for (int i = 0; i < 19; ++i) __builtin_nontemporal_store((unsigned short)(v[i] + 1), &d[(long long)i * N + c]);
        return;
    }
    long long c0 = t * 2;
    if (c0 + 2 > N) return;
    unsigned int v[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        long long sidx = c0 - CX[i] - (long long)px * CY[i] - (long long)plane * CZ[i];
        if (mode == 1) {
            sidx = sidx < 0 ? 0 : (sidx > N - 2 ? N - 2 : sidx);
            __builtin_memcpy(&v[i], s + (long long)i * N + sidx, 4);
        } else {
            long long al = sidx & ~1ll;               // pair-aligned start
            long long q = al - 2;                      // 4 halves [al-2, al+2)
            q = q < 0 ? 0 : (q > N - 4 ? N - 4 : q);
            q &= ~1ll;
            unsigned long long w;
            __builtin_memcpy(&w, s + (long long)i * N + q, 8);
            const int sh = (int)((sidx - q) & 3) * 16;
            v[i] = (unsigned int)(w >> sh);
        }
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) __builtin_nontemporal_store(v[i] + 1u, (unsigned int*)(d + (long long)i * N + c0));
}

int main() {
    const long long bytes = 2ll << 30;
    const long long n4 = bytes / 16, n1 = bytes / 4;
    float *a, *b, *o;
    CK(hipMalloc(&a, bytes)); CK(hipMalloc(&b, bytes)); CK(hipMalloc(&o, 4096));
    CK(hipMemset(a, 0, bytes)); CK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    auto timeit = [&](const char* name, double moved, auto launch) {
        launch(); CK(hipDeviceSynchronize());
        const int it = 10;
        CK(hipEventRecord(e0));
        for (int k = 0; k < it; ++k) launch();
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        printf("%-28s %8.1f GB/s  (%.3f ms)\n", name, moved * it / (ms * 1e-3) / 1e9, ms / it);
        fflush(stdout);
    };
    for (int bs : {256, 512, 1024}) {
        char nm[64];
        snprintf(nm, 64, "copy float4x4 bs%d", bs);
        timeit(nm, 2.0 * bytes, [&] { hipLaunchKernelGGL(copy4x4, dim3((n4 + bs * 4 - 1) / (bs * 4)), dim3(bs), 0, 0, (const float4*)a, (float4*)b, n4); });
        snprintf(nm, 64, "copy float4 bs%d", bs);
        timeit(nm, 2.0 * bytes, [&] { hipLaunchKernelGGL(copy4, dim3((n4 + bs - 1) / bs), dim3(bs), 0, 0, (const float4*)a, (float4*)b, n4); });
    }
    timeit("copy float4 nt-store", 2.0 * bytes, [&] { hipLaunchKernelGGL(copy4nt, dim3((n4 + 255) / 256), dim3(256), 0, 0, (const float4*)a, (float4*)b, n4); });
    timeit("copy float", 2.0 * bytes, [&] { hipLaunchKernelGGL(copy1, dim3((n1 + 255) / 256), dim3(256), 0, 0, a, b, n1); });
    timeit("copy float nt-store", 2.0 * bytes, [&] { hipLaunchKernelGGL(copy1nt, dim3((n1 + 255) / 256), dim3(256), 0, 0, a, b, n1); });
    timeit("read float4", 1.0 * bytes, [&] { hipLaunchKernelGGL(read4, dim3((n4 + 255) / 256), dim3(256), 0, 0, (const float4*)a, o, n4); });
    timeit("write float4", 1.0 * bytes, [&] { hipLaunchKernelGGL(write4, dim3((n4 + 255) / 256), dim3(256), 0, 0, (float4*)b, n4); });
    // LBM-shaped: N = 1536*522*30 cells, 19 streams each way (fits in the 2 GiB buffers: 19*4*N = 1.83 GB)
    const int px = 1536, ny = 522, nz = 30;
    const long long N = (long long)px * ny * nz;
    if (19 * N * 4 > bytes) { printf("size\n"); return 1; }
    for (int pull = 0; pull < 2; ++pull)
        for (int nt = 0; nt < 2; ++nt) {
            char nm[64]; snprintf(nm, 64, "19 streams pull=%d nt=%d", pull, nt);
            timeit(nm, 2.0 * 19 * 4 * N, [&] { hipLaunchKernelGGL(streams19, dim3((N + 255) / 256), dim3(256), 0, 0, a, b, N, px, px * ny, pull, nt); });
        }
    for (int nt = 0; nt < 2; ++nt) {
        char nm[64];
        snprintf(nm, 64, "19 streams pull float2 nt=%d", nt);
        timeit(nm, 2.0 * 19 * 4 * N, [&] { hipLaunchKernelGGL(streams19v<2>, dim3((N / 2 + 255) / 256), dim3(256), 0, 0, a, b, N, px, px * ny, nt); });
        snprintf(nm, 64, "19 streams pull float4 nt=%d", nt);
        timeit(nm, 2.0 * 19 * 4 * N, [&] { hipLaunchKernelGGL(streams19v<4>, dim3((N / 4 + 255) / 256), dim3(256), 0, 0, a, b, N, px, px * ny, nt); });
    }
    for (int nt = 0; nt < 2; ++nt) {
        char nm[64];
        snprintf(nm, 64, "aosoa B32 nt=%d", nt);
        timeit(nm, 2.0 * 19 * 4 * N, [&] { hipLaunchKernelGGL(streams19_aosoa<32>, dim3((N + 255) / 256), dim3(256), 0, 0, a, b, N, px, px * ny, nt); });
        snprintf(nm, 64, "aosoa B64 nt=%d", nt);
        timeit(nm, 2.0 * 19 * 4 * N, [&] { hipLaunchKernelGGL(streams19_aosoa<64>, dim3((N + 255) / 256), dim3(256), 0, 0, a, b, N, px, px * ny, nt); });
        snprintf(nm, 64, "aosoa B128 nt=%d", nt);
        timeit(nm, 2.0 * 19 * 4 * N, [&] { hipLaunchKernelGGL(streams19_aosoa<128>, dim3((N + 255) / 256), dim3(256), 0, 0, a, b, N, px, px * ny, nt); });
        snprintf(nm, 64, "aosoa B256 nt=%d", nt);
        timeit(nm, 2.0 * 19 * 4 * N, [&] { hipLaunchKernelGGL(streams19_aosoa<256>, dim3((N + 255) / 256), dim3(256), 0, 0, a, b, N, px, px * ny, nt); });
        snprintf(nm, 64, "aosoa B512 nt=%d", nt);
        timeit(nm, 2.0 * 19 * 4 * N, [&] { hipLaunchKernelGGL(streams19_aosoa<512>, dim3((N + 255) / 256), dim3(256), 0, 0, a, b, N, px, px * ny, nt); });
    }
    for (int mode = 0; mode < 2; ++mode) {
        char nm[64];
        snprintf(nm, 64, "inplace mode=%d", mode);
        timeit(nm, 2.0 * 19 * 4 * N, [&] { hipLaunchKernelGGL(streams19_inplace, dim3((N + 255) / 256), dim3(256), 0, 0, a, N, px, px * ny, mode); });
    }
    for (int mode = 0; mode < 3; ++mode) {
        char nm[64];
        snprintf(nm, 64, "half streams mode=%d", mode);
        const long long nthr = mode == 0 ? N : N / 2;
        timeit(nm, 2.0 * 19 * 2 * N, [&] { hipLaunchKernelGGL(streams19_half, dim3((nthr + 255) / 256), dim3(256), 0, 0, (const unsigned short*)a, (unsigned short*)b, N, px, px * ny, mode); });
    }
    printf("done\n");
    return 0;
}
