// C-ABI of libwfs.so: solver handles, uploads, stepping and readouts.
// Declarations and the reference interface each entry point replaces are in
// include/wfs.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "wfs.h"
#include "aux_kernels.h"
#include "step_kernels.h"
#include "closure.h"

using namespace wfs;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define WFS_HIP(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(WFS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define WFS_LAUNCH_CHECK() WFS_HIP(hipGetLastError())

inline unsigned blocks_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// Streaming copy, four 16-byte elements per work-item (a 4 KB run per wave per pass).
__global__ void __launch_bounds__(256) k_copy16(const float4* __restrict__ src, float4* __restrict__ dst,
                                                long long n) {
    const long long base = ((long long)blockIdx.x * blockDim.x) * 4 + threadIdx.x;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long long k = base + (long long)u * blockDim.x;
        if (k < n) dst[k] = src[k];
    }
}

// The memory traffic of one pull step without its arithmetic: cell c of a
// px x ny x nz box reads population i from direction array i of src at
// c - (cx_i + px cy_i + plane cz_i), clamped into the box, and writes it at
// c of direction array i of dst with a non-temporal store. Direction arrays
// start `stride` elements apart (stride - n elements of stagger between
// them).
template <typename T>
__global__ void __launch_bounds__(256) k_streams19(const T* __restrict__ src, T* __restrict__ dst, long long n,
                                                  long long stride, int px, long long plane) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    T v[Q];
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        long long s = c - (kCX[i] + (long long)px * kCY[i] + plane * kCZ[i]);
        s = s < 0 ? 0 : (s >= n ? n - 1 : s);
        v[i] = src[(long long)i * stride + s];
    });
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        __builtin_nontemporal_store(v[i], &dst[(long long)i * stride + c]);
    });
}

__global__ void k_fill_zero_u32(uint32_t* p, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0u;
}

// 64-bit fingerprint of a byte range: 8-byte words through a multiply-rotate
// mix, the tail zero-padded, the length folded in last. Identifies the
// geometry a checkpoint was taken on; not a cryptographic hash.
uint64_t fingerprint(const void* p, size_t n, uint64_t h) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    auto mix = [&h](uint64_t w) {
        h ^= w * 0x9E3779B97F4A7C15ull;
        h = ((h << 31) | (h >> 33)) * 0xBF58476D1CE4E5B9ull;
    };
    size_t k = 0;
    for (; k + 8 <= n; k += 8) {
        uint64_t w;
        std::memcpy(&w, b + k, 8);
        mix(w);
    }
    uint64_t w = 0;
    if (k < n) std::memcpy(&w, b + k, n - k);
    mix(w);
    mix((uint64_t)n);
    return h;
}

// Scratch of wfs_forces_noalloc, the allocation-free variant of wfs_forces:
// kNoAllocGroups work-groups each write nine partial sums, and one more
// work-group writes their nine totals after them. Fixed size (147,528
// bytes), one copy per GPU for the life of the process, shared by every
// call on that GPU: the partials are overwritten by each call, nothing needs
// clearing, and two calls that overlap on one GPU race.
constexpr int kNoAllocGroups = 2048;
constexpr long long kScratchDoubles = (long long)(kNoAllocGroups + 1) * 9;
constexpr long long kNoAllocMaxCells = (1ll << 31) - 1;   // stored cells per call (the solver's own limit)
__device__ __align__(256) double scratch[kScratchDoubles];

// The device address of `scratch` on each GPU, taken once per device (with
// that device current) by hipGetSymbolAddress; host code uses only these
// addresses, never the array's name.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_of(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return fail(WFS_ENODEV, "scratch: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(WFS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = static_cast<double*>(ptr);
    }
    *out = g_scratch[device];
    return WFS_OK;
}

// Device buffer owned by a solver: moved, never copied; assigning an empty
// one (buf = DevBuf()) frees what it held.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

struct wfs_solver {
    wfs_desc d{};
    long long N = 0;    // dense cell count nx*ny*nz (host arrays)
    long long Np = 0;   // device cell count px*ny*nz (rows padded to px)
    long long Ns = 0;   // population direction stride in elements: Np plus the stagger (wfs_set_pop_stagger)
    int px = 0;         // row pitch: nx rounded up to a multiple of 64
    int64_t nstep = 0;
    int32_t navg = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    size_t bytes = 0;
    bool uploaded = false;
    bool have_bedwall = false;
    // step-kernel timing: event pairs recorded around each launch
    bool timing = false;
    std::vector<hipEvent_t> events;
    size_t ev_used = 0;
    double timed_ms = 0.0;
    int64_t timed_launches = 0;
    virtual ~wfs_solver() {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
    // Sums the recorded pairs into timed_ms and frees the pool for reuse.
    int harvest() {
        if (ev_used == 0) return WFS_OK;
        WFS_HIP(hipEventSynchronize(events[ev_used - 1]));
        for (size_t k = 0; k + 1 < ev_used; k += 2) {
            float ms = 0.f;
            WFS_HIP(hipEventElapsedTime(&ms, events[k], events[k + 1]));
            timed_ms += ms;
            ++timed_launches;
        }
        ev_used = 0;
        return WFS_OK;
    }
    // Records one event of a pair on the stream (grows the pool to 4096 events).
    int mark() {
        if (ev_used == events.size()) {
            if (events.size() >= 4096) {
                int rc = harvest();
                if (rc) return rc;
            } else {
                hipEvent_t e;
                WFS_HIP(hipEventCreate(&e));
                events.push_back(e);
            }
        }
        WFS_HIP(hipEventRecord(events[ev_used++], stream));
        return WFS_OK;
    }
    virtual int alloc() = 0;
    virtual int set_pop_stagger(long long bytes) = 0;
    virtual int upload(const uint8_t* solid, const float* cf, const float* uin) = 0;
    virtual int set_force(const float* Fx, const float* Fy, const float* Fz) = 0;
    virtual int set_force_scale(float s) = 0;
    virtual int set_uin(const float* uin) = 0;
    virtual int set_uref(const float* uref) = 0;
    virtual int set_bedwall(const uint8_t* bw) = 0;
    virtual int update_uref() = 0;
    virtual int step(int n) = 0;
    virtual int get_macro(void* rho, void* ux, void* uy, void* uz, bool f64) = 0;
    virtual int get_uref(float* uref) = 0;
    virtual int set_macro(const float* rho, const float* ux, const float* uy, const float* uz) = 0;
    virtual int get_pops(void* f, bool f64) = 0;
    virtual int set_pops(const float* f) = 0;
    virtual int accumulate() = 0;
    virtual int reset_sums() = 0;
    virtual int get_sums(void* const* out, bool f64) = 0;
    virtual int profile_x(int field, int z0, int z1, double* out) = 0;
    virtual int total(int field, double* out) = 0;
    virtual int check_finite() = 0;
    virtual int set_closure(int kind, const float* params, const uint16_t* geo, const uint8_t* hcol, double u0) = 0;
    virtual int set_body_mask(const uint8_t* body) = 0;
    virtual int set_column_u(int per_step) = 0;
    virtual int set_inplace(int on) = 0;
    virtual int state_bytes(long long* out) = 0;
    virtual int state_save(void* out, long long bytes) = 0;
    virtual int state_load(const void* in, long long bytes) = 0;
    virtual int particles_set(long long n, const double* xyz, long long ninlet, const int32_t* inlet,
                              uint64_t seed) = 0;
    virtual int particles_move(double dt, double vel, double dx, long long* respawned) = 0;
    virtual int particles_get(double* xyz) = 0;
    virtual int forces(double* out) = 0;
    virtual int forces_noalloc(double* out) = 0;
    virtual int get_closure_field(float* c2) = 0;
    virtual int set_closure_field(int on) = 0;
    virtual int closure_samples(const uint16_t* geo, const uint8_t* hcol, int ncx, int ncy, int ncz, double kv,
                                double u0, float* X, float* ABP, uint8_t* valid) = 0;
    int closure = WFS_CLOSURE_SMAGORINSKY;
    bool colu_on = false;   // per-step column velocity for the wall model
    int colu_par = 0;       // which column-velocity buffer the next step reads
    // in-place streaming: one population buffer, steps alternate odd / even;
    // aa_phase 0 = state S (an even number of in-place steps taken)
    bool inplace = false;
    int aa_phase = 0;
    int batch_refs = 0;     // batches that hold this solver's buffers
    uint64_t geo_hash = 0;  // fingerprint of the uploaded geometry (dims, solid mask, wall cf)
    int bx = 128;   // step-kernel work-group width (64..256)
    // tile-sparse layout: only the 64-cell x-runs of a row that hold fluid are stored
    bool tiled = false;
    int ntx = 0;                                         // x-tiles per row (px / 64)
    long long T = 0;                                     // stored tiles
    std::vector<int> h_tmap, h_tcoord, h_nbr, h_blk2tile, h_wgrec;   // block map: one tile per work-group
    std::vector<int> h_rowtile;                          // tiles stored before row j (rows + 1 entries)
    virtual int set_force_cells(long long n, const int64_t* idx, const float* fx, const float* fy,
                                const float* fz) = 0;

    // Frees a buffer and takes it off the device-memory count.
    void release(DevBuf& b) {
        bytes -= b.bytes;
        b = DevBuf();
    }

    // (Re)allocates b; what it held before is freed first.
    int alloc_buf(DevBuf& b, size_t bytes_) {
        if (b.p) release(b);
        bytes_ = (bytes_ + 255) & ~size_t(255);  // whole 256-byte blocks: fills and vector access stay in bounds
        b.bytes = bytes_;
        if (hipMalloc(&b.p, bytes_) != hipSuccess) {
            (void)hipGetLastError();
            b.p = nullptr;
            return fail(WFS_ENOMEM, "hipMalloc of " + std::to_string(bytes_) + " bytes failed");
        }
        bytes += bytes_;
        return WFS_OK;
    }
};

namespace {

// Largest staging buffer for host exchanges of dense fields (blocks of whole
// rows pass through it one after another).
constexpr size_t kStageBytes = size_t(256) << 20;

template <class Codec>
struct Solver final : wfs_solver {
    using T = typename Codec::T;
    using R = typename Codec::R;
    DevBuf f[2], flag, wmask, fslot, fval, rho, ux, uy, uz, cfw, uref, uin;
    DevBuf sums[6], partial, out_x, counter, stage;
    DevBuf geo, hcol, mlp_params, mlp_pack, c2buf;   // learned closure (allocated when first set)
    bool c2_keep = true;                              // c2buf kept (wfs_set_closure_field)
    DevBuf body;                                      // hull cells for the force budget (when set): one bit per dense cell
    DevBuf colu[2][2], colcount;                      // per-step column velocity: [parity][ux, uy]; fluid cells per column
    DevBuf part_xyz, part_inlet, part_count;          // particle tracer: positions, inlet rows, respawn counter
    long long part_n = 0, part_ninlet = 0, part_step = 0;
    uint64_t part_seed = 0;
    DevBuf tmap_d, tcoord_d, nbr_d, blk2tile_d, wgrec_d;   // tile-sparse tables (when tiled)
    DevBuf aux_d;                                     // StepAux for column-buffer parity 0 and 1
    StepAux<R> aux_h[2]{};                            // what aux_d holds
    bool aux_up = false;

    CellMap cmap() const {
        CellMap m;
        m.nx = d.nx;
        m.px = px;
        m.ny = d.ny;
        m.ntx = ntx;
        m.tmap = tiled ? static_cast<const int*>(tmap_d.p) : nullptr;
        m.tcoord = tiled ? static_cast<const int*>(tcoord_d.p) : nullptr;
        return m;
    }
    float inv_u0 = 0.f;
    int cur = 0;
    R fscale = R(1);
    Center ctr{};   // storage centres (16-bit formats), fixed at upload

    T* fp(int k) { return static_cast<T*>(f[k].p); }
    R* rp(DevBuf& b) { return static_cast<R*>(b.p); }
    const uint8_t* flags() { return static_cast<const uint8_t*>(flag.p); }

    int alloc() override {
        const size_t n = (size_t)Np;
        int rc;
        // the second population buffer comes with the upload, unless the
        // solver streams in place by then
        if ((rc = alloc_buf(f[0], (size_t)Ns * Q * sizeof(T)))) return rc;
        if ((rc = alloc_buf(flag, n))) return rc;
        if ((rc = alloc_buf(wmask, n * sizeof(uint32_t)))) return rc;
        for (DevBuf* b : {&rho, &ux, &uy, &uz, &cfw, &uref})
            if ((rc = alloc_buf(*b, n * sizeof(R)))) return rc;
        if ((rc = alloc_buf(uin, (size_t)rows() * sizeof(R)))) return rc;
        if ((rc = alloc_buf(out_x, (size_t)d.nx * sizeof(double)))) return rc;
        if ((rc = alloc_buf(partial, (size_t)d.nx * 256 * sizeof(double)))) return rc;
        if ((rc = alloc_buf(counter, 256))) return rc;
        // host <-> device exchanges of dense fields pass through a staging
        // buffer of at most kStageBytes, a block of whole rows at a time; it
        // holds at least three z-planes of bytes (the solid mask around one
        // plane) and one row of FP64 populations
        size_t cap = kStageBytes;
        if (const char* e = std::getenv("WFS_STAGE_BYTES")) {   // a smaller cap (tests use it to force many blocks)
            const long long v = std::atoll(e);
            if (v > 0) cap = std::min(cap, (size_t)v);
        }
        size_t sb = std::min((size_t)N * Q * sizeof(double), cap);
        sb = std::max({sb, (size_t)3 * d.nx * d.ny, (size_t)d.nx * Q * sizeof(double)});
        if ((rc = alloc_buf(stage, sb))) return rc;
        if ((rc = alloc_buf(aux_d, 2 * sizeof(StepAux<R>)))) return rc;     // StepAux pair (sync_aux fills it)
        if (tiled) {
            for (auto pr : {std::make_pair(&tmap_d, &h_tmap), std::make_pair(&tcoord_d, &h_tcoord),
                            std::make_pair(&nbr_d, &h_nbr), std::make_pair(&blk2tile_d, &h_blk2tile),
                            std::make_pair(&wgrec_d, &h_wgrec)}) {
                if ((rc = alloc_buf(*pr.first, pr.second->size() * sizeof(int)))) return rc;
                WFS_HIP(hipMemcpy(pr.first->p, pr.second->data(), pr.second->size() * sizeof(int),
                                  hipMemcpyHostToDevice));
            }
        }
        return WFS_OK;
    }

    int rows() const { return d.ny * d.nz; }

    // Device cell where dense row j starts (rows, and on the tile-sparse
    // layout their tiles, are stored in row order).
    long long row_cell(long long j) const { return tiled ? (long long)h_rowtile[(size_t)j] * 64 : (long long)px * j; }
    // Rows per staging block when one row takes row_bytes of it.
    long long block_rows(size_t row_bytes) const {
        return std::max<long long>(1, (long long)(stage.bytes / row_bytes));
    }

    // host dense S cells -> device padded D cells (padding cells: fill), a
    // block of rows at a time
    template <typename S, typename D>
    int put_dense(DevBuf& b, const S* src, D fill) {
        const long long nrows = rows(), step = block_rows((size_t)d.nx * sizeof(S));
        for (long long j0 = 0; j0 < nrows; j0 += step) {
            const long long j1 = std::min(nrows, j0 + step);
            WFS_HIP(hipMemcpyAsync(stage.p, src + j0 * d.nx, (size_t)((j1 - j0) * d.nx) * sizeof(S),
                                   hipMemcpyHostToDevice, stream));
            const long long c0 = row_cell(j0), c1 = row_cell(j1);
            if (c1 > c0) {
                hipLaunchKernelGGL((k_pad<S, D>), dim3(blocks_for(c1 - c0, 256)), dim3(256), 0, stream,
                                   static_cast<const S*>(stage.p), static_cast<D*>(b.p), cmap(), c0, c1, j0, fill);
                WFS_LAUNCH_CHECK();
            }
        }
        WFS_HIP(hipStreamSynchronize(stream));
        return WFS_OK;
    }

    // host dense float cells -> device padded R cells
    int put_cells(DevBuf& b, const float* src) { return put_dense<float, R>(b, src, R(0)); }

    // host dense u8 cells -> device padded u8 cells
    int put_bytes(DevBuf& b, const uint8_t* src) { return put_dense<uint8_t, uint8_t>(b, src, uint8_t(0)); }

    // host float rows (ny*nz) -> device R rows
    int put_rows(DevBuf& b, const float* src) {
        const size_t n = (size_t)rows();
        std::vector<R> tmp(src, src + n);
        WFS_HIP(hipMemcpyAsync(b.p, tmp.data(), n * sizeof(R), hipMemcpyHostToDevice, stream));
        WFS_HIP(hipStreamSynchronize(stream));
        return WFS_OK;
    }

    // device R cells -> host dense D cells, a block of rows at a time: cells
    // of tiles that are not stored get `fill`, or with `add` set every cell
    // gets its value plus `fill` (added in double before the rounding to D)
    template <typename D>
    int get_dense(DevBuf& b, D* dst, double fill, bool add) {
        const long long nrows = rows(), step = block_rows((size_t)d.nx * sizeof(D));
        for (long long j0 = 0; j0 < nrows; j0 += step) {
            const long long j1 = std::min(nrows, j0 + step);
            const long long k0 = j0 * d.nx, n = (j1 - j0) * d.nx;
            if (add)
                hipLaunchKernelGGL((k_unpad_add<R, D>), dim3(blocks_for(n, 256)), dim3(256), 0, stream, rp(b),
                                   static_cast<D*>(stage.p), cmap(), k0, n, fill);
            else
                hipLaunchKernelGGL((k_unpad<R, D>), dim3(blocks_for(n, 256)), dim3(256), 0, stream, rp(b),
                                   static_cast<D*>(stage.p), cmap(), k0, n, (D)fill);
            WFS_LAUNCH_CHECK();
            WFS_HIP(hipMemcpyAsync(dst + k0, stage.p, (size_t)n * sizeof(D), hipMemcpyDeviceToHost, stream));
        }
        WFS_HIP(hipStreamSynchronize(stream));
        return WFS_OK;
    }

    // device R cells -> host dense float or double cells (cells of tiles that are not stored: fill)
    int get_cells(DevBuf& b, void* dst, bool f64, double fill = 0.0) {
        if (!dst) return WFS_OK;
        return f64 ? get_dense(b, static_cast<double*>(dst), fill, false)
                   : get_dense(b, static_cast<float*>(dst), fill, false);
    }

    int zero(DevBuf& b) {
        // a fill kernel rather than hipMemsetAsync keeps the call graph-capturable
        const long long n = (long long)(b.bytes / 4);
        hipLaunchKernelGGL(k_fill_zero_u32, dim3(blocks_for(n, 256)), dim3(256), 0, stream,
                           static_cast<uint32_t*>(b.p), n);
        WFS_LAUNCH_CHECK();
        return WFS_OK;
    }

    // Population direction arrays `bytes` apart beyond their Np cells (a
    // multiple of 128 bytes, at most 1 MiB; 0 packs them). Before the upload
    // only: the population buffer is allocated again.
    int set_pop_stagger(long long stagger) override {
        if (uploaded) return fail(WFS_EINVAL, "wfs_set_pop_stagger: set it before the upload");
        if (stagger < 0 || stagger % 128 || stagger > (1ll << 20))
            return fail(WFS_EINVAL, "wfs_set_pop_stagger: the stagger must be a multiple of 128 bytes in [0, 1 MiB]");
        const long long ns = Np + stagger / (long long)sizeof(T);
        if ((unsigned long long)ns * sizeof(T) >= (1ull << 32))
            return fail(WFS_EINVAL, "wfs_set_pop_stagger: a population direction would exceed 4 GiB");
        if (ns == Ns) return WFS_OK;
        WFS_HIP(hipStreamSynchronize(stream));
        for (auto& b : f) release(b);
        Ns = ns;
        return alloc_buf(f[0], (size_t)Ns * Q * sizeof(T));
    }

    int upload(const uint8_t* solid, const float* cf, const float* uin_h) override {
        int rc;
        const int32_t dims[4] = {d.nx, d.ny, d.nz, tiled ? 1 : 0};
        geo_hash = fingerprint(dims, sizeof dims, 0x57465347656f6dull);
        geo_hash = fingerprint(solid, (size_t)N, geo_hash);
        if (cf) geo_hash = fingerprint(cf, (size_t)N * sizeof(float), geo_hash);
        for (int k = 0; k < 2; ++k)
            if (!inplace && !f[k].p && (rc = alloc_buf(f[k], (size_t)Ns * Q * sizeof(T)))) return rc;
        if (cf) {
            if ((rc = put_cells(cfw, cf))) return rc;
        } else if ((rc = zero(cfw))) {
            return rc;
        }
        if ((rc = put_rows(uin, uin_h))) return rc;
        set_center(uin_h);
        // flags and wall masks a block of z-planes at a time: the staging
        // buffer holds the block's dense solid mask and one plane either side
        const size_t plane = (size_t)d.nx * d.ny;
        const int pz = (int)std::max<long long>(1, (long long)(stage.bytes / plane) - 2);
        for (int z0 = 0; z0 < d.nz; z0 += pz) {
            const int z1 = std::min(d.nz, z0 + pz), zlo = std::max(0, z0 - 1), zhi = std::min(d.nz, z1 + 1);
            WFS_HIP(hipMemcpyAsync(stage.p, solid + (size_t)zlo * plane, (size_t)(zhi - zlo) * plane,
                                   hipMemcpyHostToDevice, stream));
            const long long c0 = row_cell((long long)d.ny * z0), c1 = row_cell((long long)d.ny * z1);
            if (c1 > c0) {
                hipLaunchKernelGGL(k_flags<R>, dim3(blocks_for(c1 - c0, 256)), dim3(256), 0, stream,
                                   static_cast<const uint8_t*>(stage.p), rp(cfw), static_cast<uint8_t*>(flag.p),
                                   static_cast<uint32_t*>(wmask.p), cmap(), d.nz, c0, c1, zlo);
                WFS_LAUNCH_CHECK();
            }
        }
        for (int k = 0; k < 2; ++k) {
            if (!f[k].p) continue;   // in-place streaming keeps one buffer
            if (Ns != Np && (rc = zero(f[k]))) return rc;   // the gaps between direction arrays hold zeros
            hipLaunchKernelGGL(k_init<Codec>, dim3(blocks_for(Np, 256)), dim3(256), 0, stream, fp(k), flags(),
                               rp(uin), cmap(), Np, Ns, ctr);
            WFS_LAUNCH_CHECK();
        }
        if (inplace) {
            cur = f[0].p ? 0 : 1;
            aa_exchange();
        }
        hipLaunchKernelGGL(k_fill_macro<R>, dim3(blocks_for(Np, 256)), dim3(256), 0, stream, rp(rho), rp(ux),
                           rp(uy), rp(uz), Np);
        WFS_LAUNCH_CHECK();
        if ((rc = zero(uref))) return rc;
        if ((rc = put_forces({}, {}))) return rc;   // no body forces
        if ((rc = reset_sums())) return rc;
        WFS_HIP(hipStreamSynchronize(stream));
        if (!inplace) cur = 0;
        aa_phase = 0;
        colu_par = 0;
        nstep = 0;
        fscale = R(1);
        have_bedwall = false;
        uploaded = true;
        return WFS_OK;
    }

    // Centre of the 16-bit formats: the equilibrium of the mean inlet velocity U0
    // (mean of the positive inlet entries), in double, rounded to float.
    void set_center(const float* uin_h) {
        double sum = 0.0;
        long long n = 0;
        for (int j = 0; j < rows(); ++j)
            if (uin_h[j] > 0.f) {
                sum += uin_h[j];
                ++n;
            }
#ifdef WFS_CENTER_AT_WEIGHTS
        const double u0 = 0.0;
#else
        const double u0 = Codec::kExact || n == 0 ? 0.0 : sum / (double)n;
#endif
        for (int i = 0; i < Q; ++i) {
            const double w = i == 0 ? 1.0 / 3 : (i < 7 ? 1.0 / 18 : 1.0 / 36);
            const double cu = kCX[i] * u0;
            const int k = center_class(i);
            ctr.c[k] = (float)(w * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u0 * u0));
            ctr.n2[k] = -2.0f * ctr.c[k];
        }
    }

    // Stored cell of dense (x, row j), or -1 where the tile-sparse layout
    // stores no tile (the host twin of cell_of).
    long long cell_host(int x, long long j) const {
        if (!tiled) return x + (long long)px * j;
        const int t = h_tmap[(size_t)((x >> 6) + (long long)ntx * j)];
        return t < 0 ? -1 : (long long)t * 64 + (x & 63);
    }

    // Replaces the body forces by the listed stored cells (fx, fy, fz each;
    // a later entry of the same cell wins) through the 64-cell group side
    // table: a slot per group with a listed cell, every other group none.
    int put_forces(const std::vector<long long>& cells, const std::vector<float>& v) {
        const long long ng = Np / kForceGroup;
        std::vector<int> slot((size_t)ng, -1);
        int ns = 0;
        for (long long c : cells) {
            int& sl = slot[(size_t)(c / kForceGroup)];
            if (sl < 0) sl = ns++;
        }
        std::vector<R> val((size_t)std::max(ns, 1) * kForceSlot, R(0));
        for (size_t k = 0; k < cells.size(); ++k) {
            R* dst = val.data() + (size_t)slot[(size_t)(cells[k] / kForceGroup)] * kForceSlot + cells[k] % kForceGroup;
            dst[0] = R(v[3 * k]);
            dst[kForceGroup] = R(v[3 * k + 1]);
            dst[2 * kForceGroup] = R(v[3 * k + 2]);
        }
        DevBuf ns_d, nv_d;
        int rc = alloc_buf(ns_d, slot.size() * sizeof(int));
        if (!rc) rc = alloc_buf(nv_d, val.size() * sizeof(R));
        if (rc) {
            release(ns_d);
            release(nv_d);
            return rc;
        }
        WFS_HIP(hipStreamSynchronize(stream));   // steps in flight read the old table
        WFS_HIP(hipMemcpy(ns_d.p, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice));
        WFS_HIP(hipMemcpy(nv_d.p, val.data(), val.size() * sizeof(R), hipMemcpyHostToDevice));
        release(fslot);
        release(fval);
        fslot = std::move(ns_d);
        fval = std::move(nv_d);
        hipLaunchKernelGGL(k_forced<R>, dim3(blocks_for(Np, 256)), dim3(256), 0, stream,
                           static_cast<uint8_t*>(flag.p), static_cast<const int*>(fslot.p), rp(fval), Np);
        WFS_LAUNCH_CHECK();
        return sync_aux();
    }

    int set_force(const float* a, const float* b, const float* c) override {
        std::vector<long long> cells;
        std::vector<float> v;
        for (long long k = 0; k < N; ++k)
            if (a[k] != 0.f || b[k] != 0.f || c[k] != 0.f) {
                const long long j = k / d.nx;
                const long long cell = cell_host(int(k - j * d.nx), j);
                if (cell < 0) continue;
                cells.push_back(cell);
                v.insert(v.end(), {a[k], b[k], c[k]});
            }
        return put_forces(cells, v);
    }

    int set_force_cells(long long n, const int64_t* idx, const float* fx, const float* fy,
                        const float* fz) override {
        std::vector<long long> cells;
        std::vector<float> v;
        for (long long k = 0; k < n; ++k) {
            const long long j = idx[k] / d.nx;
            const long long cell = cell_host(int(idx[k] - j * d.nx), j);
            if (cell < 0) continue;   // a cell of a tile that is not stored
            cells.push_back(cell);
            v.insert(v.end(), {fx[k], fy[k], fz[k]});
        }
        return put_forces(cells, v);
    }

    int set_force_scale(float s) override {
        fscale = R(s);
        return WFS_OK;
    }

    int set_uin(const float* u) override { return put_rows(uin, u); }
    int set_uref(const float* u) override { return put_cells(uref, u); }

    // The mask becomes the FLAG_BEDWALL bit of the flag byte (through a
    // transient per-cell byte buffer).
    int set_bedwall(const uint8_t* bw) override {
        DevBuf tmp;
        int rc = alloc_buf(tmp, (size_t)Np);
        if (!rc) rc = put_bytes(tmp, bw);
        if (!rc) {
            hipLaunchKernelGGL(k_bedwall_flag, dim3(blocks_for(Np, 256)), dim3(256), 0, stream,
                               static_cast<uint8_t*>(flag.p), static_cast<const uint8_t*>(tmp.p), Np);
            if (hipGetLastError() != hipSuccess) rc = fail(WFS_EHIP, "wfs_set_bedwall: launch");
            if (hipStreamSynchronize(stream) != hipSuccess && !rc) rc = fail(WFS_EHIP, "wfs_set_bedwall: synchronize");
        }
        release(tmp);
        if (rc) return rc;
        have_bedwall = true;
        return WFS_OK;
    }

    int update_uref() override {
        if (!have_bedwall) return fail(WFS_EINVAL, "wfs_update_uref: no bed/bank mask set (wfs_set_bedwall)");
        hipLaunchKernelGGL(k_uref<R>, dim3(blocks_for(d.nx, 64), d.ny), dim3(64), 0, stream, rp(ux), rp(uy), flags(),
                           rp(uref), cmap(), d.nz);
        WFS_LAUNCH_CHECK();
        return WFS_OK;
    }

    // The StepAux pair (column-buffer parity 0 and 1) in device memory,
    // uploaded again when one of its pointers changed (set-up,
    // wfs_set_column_u); running steps read it, so the stream drains first.
    int sync_aux() {
        StepAux<R> v[2]{};
        for (int p = 0; p < 2; ++p) {
            v[p].fslot = static_cast<const int*>(fslot.p);
            v[p].fval = rp(fval);
            v[p].rho = rp(rho);
            v[p].ux = rp(ux);
            v[p].uy = rp(uy);
            v[p].uz = rp(uz);
            v[p].cfw = rp(cfw);
            v[p].uref = rp(uref);
            v[p].colcount = static_cast<const uint8_t*>(colcount.p);
            if (colu_on) {
                v[p].colu_in_x = rp(colu[p][0]);
                v[p].colu_in_y = rp(colu[p][1]);
                v[p].colu_out_x = rp(colu[p ^ 1][0]);
                v[p].colu_out_y = rp(colu[p ^ 1][1]);
            }
        }
        if (aux_up && std::memcmp(v, aux_h, sizeof v) == 0) return WFS_OK;
        int rc;
        if (!aux_d.p && (rc = alloc_buf(aux_d, sizeof v))) return rc;
        WFS_HIP(hipStreamSynchronize(stream));
        WFS_HIP(hipMemcpy(aux_d.p, v, sizeof v, hipMemcpyHostToDevice));
        std::memcpy(aux_h, v, sizeof v);
        aux_up = true;
        return WFS_OK;
    }

    int fill_args(StepArgs<R>& a) {
        int rc = sync_aux();
        if (rc) return rc;
        a.flag = flags();
        a.wmask = static_cast<const uint32_t*>(wmask.p);
        a.fscale = fscale;
        // FP32 solvers see tau0 and Cs^2 rounded to float, like the float kernel arguments of
        // lbm3d.py:403; the FP64 solver keeps them in double.
        a.tau0 = std::is_same<R, float>::value ? R((float)d.tau0) : R(d.tau0);
        a.tau0sq = a.tau0 * a.tau0;
        const R cs2 = std::is_same<R, float>::value ? R((float)(d.cs * d.cs)) : R(d.cs * d.cs);
        a.ksmag = R(18) * R(1.41421356) * cs2;
        a.nx = d.nx;
        a.ny = d.ny;
        a.nz = d.nz;
        a.px = px;
        a.rows = rows();
        a.tiles = (d.nx + bx - 1) / bx;
        a.N = Ns;
        a.uin = rp(uin);
        a.bc_out_fused = (d.nx - 1) % 64 != 0 ? 1 : 0;
        a.pop32 = (unsigned long long)Q * (unsigned long long)Ns * sizeof(T) < (1ull << 32) ? 1 : 0;
        a.write_all = 0;
        a.ctr = ctr;
        a.geo = static_cast<const uint16_t*>(geo.p);
        a.hcol = static_cast<const uint8_t*>(hcol.p);
        a.mlp = static_cast<const MlpPack*>(mlp_pack.p);
        a.c2out = static_cast<float*>(c2buf.p);
        a.inv_u0 = inv_u0;
        a.cs2 = (float)(d.cs * d.cs);
        a.nbr = static_cast<const int*>(nbr_d.p);
        a.tcoord = static_cast<const int*>(tcoord_d.p);
        a.blk2tile = static_cast<const int*>(blk2tile_d.p);
        a.wgrec = static_cast<const int*>(wgrec_d.p);
        a.tmap = static_cast<const int*>(tmap_d.p);
        a.ntx = ntx;
        set_colu(a);
        return WFS_OK;
    }

    // The column velocity this step reads (written by the previous step) and
    // the one it writes: the StepAux of this step's parity; the pair
    // alternates every step.
    void set_colu(StepArgs<R>& a) {
        a.colu_on = colu_on ? 1 : 0;
        a.aux = static_cast<const StepAux<R>*>(aux_d.p) + (colu_on ? colu_par : 0);
    }

    // A-B layout <-> state S of the in-place buffer (the same exchange both ways).
    int aa_exchange() {
        hipLaunchKernelGGL(k_aa_exchange<T>, dim3(blocks_for(Np, 256)), dim3(256), 0, stream, fp(cur), Np, Ns, cmap());
        WFS_LAUNCH_CHECK();
        return WFS_OK;
    }

    // Readouts and uploads of populations see the A-B layout; in-place
    // streaming converts around them, which needs state S.
    int aa_readable(const char* what) {
        if (inplace && aa_phase != 0)
            return fail(WFS_EINVAL, std::string(what) +
                                        ": in-place streaming holds a readable population state only after an even "
                                        "number of steps");
        return WFS_OK;
    }

    // In-place streaming on: the population buffer not in use is freed and
    // the state converted to S; off: the second buffer comes back. Needs the
    // outlet row fused into the step (nx - 1 not a multiple of 64).
    int set_inplace(int on) override {
        if (!on == !inplace) return WFS_OK;
        if (batch_refs) return fail(WFS_EINVAL, "wfs_set_inplace: the solver is a case of a batch");
        int rc;
        if (on) {
            if ((d.nx - 1) % 64 == 0)
                return fail(WFS_EINVAL, "wfs_set_inplace: nx - 1 must not be a multiple of 64 (the outlet row is "
                                        "fused into the step)");
            WFS_HIP(hipStreamSynchronize(stream));
            if (uploaded && (rc = aa_exchange())) return rc;
            WFS_HIP(hipStreamSynchronize(stream));
            release(f[cur ^ 1]);
            inplace = true;
            aa_phase = 0;
            return WFS_OK;
        }
        if ((rc = aa_readable("wfs_set_inplace"))) return rc;
        if ((rc = alloc_buf(f[cur ^ 1], (size_t)Ns * Q * sizeof(T)))) return rc;
        if (Ns != Np && (rc = zero(f[cur ^ 1]))) return rc;
        if (uploaded && (rc = aa_exchange())) return rc;
        WFS_HIP(hipStreamSynchronize(stream));
        inplace = false;
        return WFS_OK;
    }

    // Dynamic LDS of the dense learned step: one network pack per wavefront.
    unsigned learned_lds() const { return (unsigned)((bx / 64) * sizeof(MlpPack)); }

    // One in-place step: odd (from state S) or even, by aa_phase.
    template <bool kOdd>
    void launch_step_aa(const dim3& grid, const StepArgs<R>& a) {
        constexpr int kStream = kOdd ? 1 : 2;
        if constexpr (std::is_same<R, float>::value) {
            if (closure == WFS_CLOSURE_MLP) {
                if (tiled) {
                    hipLaunchKernelGGL((k_step_learned_sparse<Codec, kStream>), dim3((unsigned)h_blk2tile.size()),
                                       dim3(64), 0, stream, nullptr, fp(cur), a);
                } else {
                    hipLaunchKernelGGL((k_step_learned<Codec, kStream>), grid, dim3(bx), learned_lds(), stream,
                                       nullptr, fp(cur), a);
                }
                return;
            }
        }
        if (tiled)
            hipLaunchKernelGGL((k_step_aa_sparse<Codec, kOdd>), dim3((unsigned)h_blk2tile.size()), dim3(64), 0, stream,
                               fp(cur), a);
        else if (a.pop32)
            hipLaunchKernelGGL((k_step_aa<Codec, kOdd, true>), grid, dim3(bx), 0, stream, fp(cur), a);
        else
            hipLaunchKernelGGL((k_step_aa<Codec, kOdd, false>), grid, dim3(bx), 0, stream, fp(cur), a);
    }

    // One step launch: the constant-Cs kernel or, with the learned closure, k_step_learned.
    void launch_step(const dim3& grid, const StepArgs<R>& a) {
        if (inplace) {
            if (aa_phase == 0)
                launch_step_aa<true>(grid, a);
            else
                launch_step_aa<false>(grid, a);
            return;
        }
        if (tiled) {
            if constexpr (std::is_same<R, float>::value) {
                if (closure == WFS_CLOSURE_MLP) {
                    hipLaunchKernelGGL((k_step_learned_sparse<Codec>), dim3((unsigned)h_blk2tile.size()), dim3(64), 0,
                                       stream, fp(cur), fp(cur ^ 1), a);
                    return;
                }
            }
            hipLaunchKernelGGL((k_step_sparse<Codec, 1>), dim3((unsigned)h_blk2tile.size()), dim3(64), 0, stream,
                               fp(cur), fp(cur ^ 1), a);
            return;
        }
        if constexpr (std::is_same<R, float>::value) {
            if (closure == WFS_CLOSURE_MLP) {
                hipLaunchKernelGGL(k_step_learned<Codec>, grid, dim3(bx), learned_lds(), stream, fp(cur), fp(cur ^ 1),
                                   a);
                return;
            }
        }
        if (a.colu_on)
            hipLaunchKernelGGL((k_step<Codec, true>), grid, dim3(bx), 0, stream, fp(cur), fp(cur ^ 1), a);
        else
            hipLaunchKernelGGL((k_step<Codec, false>), grid, dim3(bx), 0, stream, fp(cur), fp(cur ^ 1), a);
    }

    int step(int n) override {
        StepArgs<R> a;
        int rc0 = fill_args(a);
        if (rc0) return rc0;
        const long long groups = (long long)((a.rows + 7) / 8) * 8 * a.tiles;
        const dim3 grid((unsigned)groups);
        for (int k = 0; k < n; ++k) {
            a.write_all = (k == n - 1) ? 1 : 0;
            set_colu(a);
            if (timing) {
                if (ev_used % 2) return fail(WFS_EHIP, "step timing: unpaired event");
                int rc = mark();
                if (rc) return rc;
            }
            launch_step(grid, a);
            WFS_LAUNCH_CHECK();
            if (timing) {
                int rc = mark();
                if (rc) return rc;
            }
            if (!a.bc_out_fused) {   // x = nx-1 opens a wavefront: outlet row in its own pass
                hipLaunchKernelGGL(k_bc<Codec>, dim3(blocks_for(a.rows, 64)), dim3(64), 0, stream, fp(cur ^ 1),
                                   flags(), rp(uin), d.nx, px, a.rows, Ns, R(1), 2, ctr);
                WFS_LAUNCH_CHECK();
            }
            if (inplace)
                aa_phase ^= 1;
            else
                cur ^= 1;
            colu_par ^= 1;
        }
        nstep += n;
        return WFS_OK;
    }

    int get_macro(void* r, void* a, void* b, void* c, bool f64) override {
        int rc;
        if ((rc = get_cells(rho, r, f64, 1.0)) || (rc = get_cells(ux, a, f64)) || (rc = get_cells(uy, b, f64)) ||
            (rc = get_cells(uz, c, f64)))
            return rc;
        return WFS_OK;
    }

    int set_macro(const float* r, const float* a, const float* b, const float* c) override {
        int rc;
        if ((rc = put_cells(rho, r)) || (rc = put_cells(ux, a)) || (rc = put_cells(uy, b)) || (rc = put_cells(uz, c)))
            return rc;
        return WFS_OK;
    }

    int get_uref(float* u) override { return get_cells(uref, u, false); }

    int get_pops(void* out, bool f64) override {
        int rc = aa_readable("wfs_get_pops");
        if (rc) return rc;
        if (inplace && (rc = aa_exchange())) return rc;
        rc = get_pops_ab(out, f64);
        if (inplace) {
            int rc2 = aa_exchange();
            if (!rc) rc = rc2;
            if (hipStreamSynchronize(stream) != hipSuccess && !rc) rc = fail(WFS_EHIP, "wfs_get_pops: synchronize");
        }
        return rc;
    }

    // populations f[i*N + k] of the dense cells, through the staging buffer
    // a block of rows at a time
    int get_pops_ab(void* out, bool f64) {
        const size_t esz = f64 ? sizeof(double) : sizeof(float);
        const long long nrows = rows(), step = block_rows((size_t)d.nx * Q * esz);
        for (long long j0 = 0; j0 < nrows; j0 += step) {
            const long long j1 = std::min(nrows, j0 + step);
            const long long k0 = j0 * d.nx, n = (j1 - j0) * d.nx;
            if (f64)
                hipLaunchKernelGGL((k_decode<Codec, double>), dim3(blocks_for(n, 256)), dim3(256), 0, stream,
                                   fp(cur), static_cast<double*>(stage.p), cmap(), Ns, k0, n, ctr);
            else
                hipLaunchKernelGGL((k_decode<Codec, float>), dim3(blocks_for(n, 256)), dim3(256), 0, stream,
                                   fp(cur), static_cast<float*>(stage.p), cmap(), Ns, k0, n, ctr);
            WFS_LAUNCH_CHECK();
            for (int i = 0; i < Q; ++i)
                WFS_HIP(hipMemcpyAsync(static_cast<char*>(out) + ((size_t)i * N + k0) * esz,
                                       static_cast<const char*>(stage.p) + (size_t)i * n * esz, (size_t)n * esz,
                                       hipMemcpyDeviceToHost, stream));
        }
        WFS_HIP(hipStreamSynchronize(stream));
        return WFS_OK;
    }

    int set_pops(const float* in) override {
        int rc = aa_readable("wfs_set_pops");
        if (rc) return rc;
        if ((rc = set_pops_ab(in))) return rc;
        if (inplace) {
            if ((rc = aa_exchange())) return rc;
            WFS_HIP(hipStreamSynchronize(stream));
        }
        return WFS_OK;
    }

    int set_pops_ab(const float* in) {
        const long long nrows = rows(), step = block_rows((size_t)d.nx * Q * sizeof(float));
        for (long long j0 = 0; j0 < nrows; j0 += step) {
            const long long j1 = std::min(nrows, j0 + step);
            const long long k0 = j0 * d.nx, n = (j1 - j0) * d.nx;
            for (int i = 0; i < Q; ++i)
                WFS_HIP(hipMemcpyAsync(static_cast<float*>(stage.p) + (size_t)i * n, in + (size_t)i * N + k0,
                                       (size_t)n * sizeof(float), hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(k_encode<Codec>, dim3(blocks_for(n, 256)), dim3(256), 0, stream, fp(cur),
                               static_cast<const float*>(stage.p), cmap(), Ns, k0, n, ctr);
            WFS_LAUNCH_CHECK();
        }
        WFS_HIP(hipStreamSynchronize(stream));
        return WFS_OK;
    }

    // ---- checkpoints: the raw solver state. The population buffer in use (its
    // device layout and storage format as they are, in-place phase
    // included), the macroscopic fields (the next step's wall function reads
    // them), uref, the six time-mean sums, the column buffers when on, and
    // the counters. Loaded into a solver built the same way (dimensions,
    // storage, layout, streaming mode, column-speed mode), a run continues
    // bit for bit.
    static constexpr uint64_t kStateMagic = 0x3144335374534657ull;   // the bytes "WFStS3D1"
    // Version 2 adds the lattice size and the geometry fingerprint: a state of
    // another reach with the same stored-cell count is refused.
    static constexpr int32_t kStateVersion = 2;
    struct StateHead {
        uint64_t magic;
        int32_t version, storage, tiled, inplace, aa_phase, colu_on, colu_par, navg;
        int64_t Np, nstep;
        double fscale;
        int32_t nx, ny, nz, pad;
        uint64_t geo_hash;
    };
    // (buffer, byte offset, bytes) of each piece; the populations direction by
    // direction, without the gaps of a staggered layout, so a state loads
    // into a solver of any population stagger
    struct StatePart {
        DevBuf* b;
        size_t off, n;
    };
    std::vector<StatePart> state_parts() {
        std::vector<StatePart> v;
        for (int i = 0; i < Q; ++i) v.push_back({&f[cur], (size_t)i * Ns * sizeof(T), (size_t)Np * sizeof(T)});
        for (DevBuf* b : {&rho, &ux, &uy, &uz, &uref}) v.push_back({b, 0, (size_t)Np * sizeof(R)});
        for (auto& sb : sums) v.push_back({&sb, 0, (size_t)Np * sizeof(R)});
        if (colu_on)
            for (auto& pr : colu)
                for (auto& b : pr) v.push_back({&b, 0, (size_t)Np * sizeof(R)});
        return v;
    }
    int state_bytes(long long* out) override {
        size_t n = sizeof(StateHead);
        for (auto& pr : state_parts()) n += pr.n;
        *out = (long long)n;
        return WFS_OK;
    }
    int state_save(void* out, long long bytes) override {
        if (!uploaded) return fail(WFS_EINVAL, "wfs_state_save: no geometry uploaded");
        long long need;
        state_bytes(&need);
        if (bytes != need) return fail(WFS_EINVAL, "wfs_state_save: buffer size " + std::to_string(bytes) +
                                                       " != wfs_state_bytes " + std::to_string(need));
        StateHead h{kStateMagic, kStateVersion, d.storage, tiled ? 1 : 0, inplace ? 1 : 0, aa_phase,
                    colu_on ? 1 : 0, colu_par, navg, Np, (int64_t)nstep, (double)fscale, d.nx, d.ny, d.nz, 0,
                    geo_hash};
        std::memcpy(out, &h, sizeof h);
        WFS_HIP(hipStreamSynchronize(stream));
        char* o = static_cast<char*>(out) + sizeof h;
        for (auto& pr : state_parts()) {
            if (pr.b->p) WFS_HIP(hipMemcpy(o, static_cast<const char*>(pr.b->p) + pr.off, pr.n, hipMemcpyDeviceToHost));
            else std::memset(o, 0, pr.n);   // time-mean sums not allocated: all zero
            o += pr.n;
        }
        return WFS_OK;
    }
    int state_load(const void* in, long long bytes) override {
        if (!uploaded) return fail(WFS_EINVAL, "wfs_state_load: upload the geometry first");
        StateHead h;
        if (bytes < (long long)sizeof h) return fail(WFS_EINVAL, "wfs_state_load: not a solver state (too short)");
        std::memcpy(&h, in, sizeof h);
        if (h.magic != kStateMagic || h.version != kStateVersion)
            return fail(WFS_EINVAL, "wfs_state_load: not a solver state");
        if (h.storage != d.storage || h.tiled != (tiled ? 1 : 0) || h.Np != Np)
            return fail(WFS_EINVAL, "wfs_state_load: the state is of another storage format, layout or size");
        if (h.nx != d.nx || h.ny != d.ny || h.nz != d.nz || h.geo_hash != geo_hash)
            return fail(WFS_EINVAL, "wfs_state_load: the state is of another geometry (lattice size, solid mask or "
                                    "wall coefficients differ)");
        if (h.inplace != (inplace ? 1 : 0))
            return fail(WFS_EINVAL, "wfs_state_load: the state is of the other streaming mode");
        if (h.colu_on != (colu_on ? 1 : 0))
            return fail(WFS_EINVAL, "wfs_state_load: per-step column speed on in one and off in the other");
        long long need;
        state_bytes(&need);
        if (bytes != need) return fail(WFS_EINVAL, "wfs_state_load: size " + std::to_string(bytes) + " != " +
                                                       std::to_string(need));
        WFS_HIP(hipStreamSynchronize(stream));
        const char* p = static_cast<const char*>(in) + sizeof h;
        for (auto& pr : state_parts()) {
            if (!pr.b->p) {   // time-mean sums not allocated: only a state that holds some allocates them
                const bool nonzero = std::any_of(p, p + pr.n, [](char ch) { return ch != 0; });
                int rc = nonzero ? ensure_sums() : WFS_OK;
                if (rc) return rc;
                WFS_HIP(hipStreamSynchronize(stream));
            }
            if (pr.b->p) WFS_HIP(hipMemcpy(static_cast<char*>(pr.b->p) + pr.off, p, pr.n, hipMemcpyHostToDevice));
            p += pr.n;
        }
        nstep = h.nstep;
        navg = h.navg;
        aa_phase = h.aa_phase;
        colu_par = h.colu_par;
        fscale = R(h.fscale);
        return WFS_OK;
    }

    // The six time-mean sums (6 values per stored cell) are allocated, zeroed,
    // by the first accumulation or read: a run that never averages holds none.
    int ensure_sums() {
        if (sums[0].p) return WFS_OK;
        for (auto& sb : sums) {
            int rc = alloc_buf(sb, (size_t)Np * sizeof(R));
            if (!rc) rc = zero(sb);
            if (rc) {
                for (auto& b : sums) release(b);
                return rc;
            }
        }
        return WFS_OK;
    }

    int accumulate() override {
        int rc = ensure_sums();
        if (rc) return rc;
        hipLaunchKernelGGL(k_accumulate<R>, dim3(blocks_for(Np, 256)), dim3(256), 0, stream, rp(rho), rp(ux), rp(uy),
                           rp(uz), rp(sums[0]), rp(sums[1]), rp(sums[2]), rp(sums[3]), rp(sums[4]), rp(sums[5]), Np);
        WFS_LAUNCH_CHECK();
        ++navg;
        return WFS_OK;
    }

    int reset_sums() override {
        if (sums[0].p)
            for (auto& sb : sums) {
                int rc = zero(sb);
                if (rc) return rc;
            }
        navg = 0;
        return WFS_OK;
    }

    // sums[3] holds the sum of rho - 1; the readout is the sum of rho.
    int get_sums(void* const* out, bool f64) override {
        int rc0 = ensure_sums();
        if (rc0) return rc0;
        for (int k = 0; k < 6; ++k) {
            int rc = k == 3 ? get_cells_add(sums[k], out[k], f64, (double)navg) : get_cells(sums[k], out[k], f64);
            if (rc) return rc;
        }
        return WFS_OK;
    }

    int get_cells_add(DevBuf& b, void* dst, bool f64, double add) {
        if (!dst) return WFS_OK;
        return f64 ? get_dense(b, static_cast<double*>(dst), add, true)
                   : get_dense(b, static_cast<float*>(dst), add, true);
    }

    const R* field_ptr(int field) {
        switch (field) {
            case WFS_F_RHO: return rp(rho);
            case WFS_F_UX: return rp(ux);
            case WFS_F_UY: return rp(uy);
            case WFS_F_UZ: return rp(uz);
            case WFS_F_SUM_UX: return rp(sums[0]);
            case WFS_F_SUM_UY: return rp(sums[1]);
            case WFS_F_SUM_UZ: return rp(sums[2]);
            case WFS_F_SUM_RHO: return rp(sums[3]);
            case WFS_F_SUM_SP: return rp(sums[4]);
            case WFS_F_SUM_UXX: return rp(sums[5]);
            default: return nullptr;
        }
    }

    int profile_dev(int field, int z0, int z1) {
        if (field < 0 || field > WFS_F_ONE) return fail(WFS_EINVAL, "wfs_profile_x: unknown field");
        if (z0 < 0 || z1 > d.nz || z0 >= z1) return fail(WFS_EINVAL, "wfs_profile_x: bad z range");
        if (field >= WFS_F_SUM_UX && field <= WFS_F_SUM_UXX) {
            int rc = ensure_sums();
            if (rc) return rc;
        }
        const R* v = field_ptr(field);
        const int j0 = z0 * d.ny, j1 = z1 * d.ny;
        const int nrows = j1 - j0;
        const int nchunks = std::min(256, std::max(1, nrows / 16));
        const int per = (nrows + nchunks - 1) / nchunks;
        const int used = (nrows + per - 1) / per;
        hipLaunchKernelGGL(k_profile_partial<R>, dim3(blocks_for(d.nx, 64), used), dim3(64), 0, stream, v, flags(),
                           cmap(), j0, j1, per, static_cast<double*>(partial.p));
        WFS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_profile_final, dim3(blocks_for(d.nx, 64)), dim3(64), 0, stream,
                           static_cast<const double*>(partial.p), d.nx, used, static_cast<double*>(out_x.p));
        WFS_LAUNCH_CHECK();
        return WFS_OK;
    }

    int profile_x(int field, int z0, int z1, double* out) override {
        int rc = profile_dev(field, z0, z1);
        if (rc) return rc;
        WFS_HIP(hipMemcpyAsync(out, out_x.p, (size_t)d.nx * sizeof(double), hipMemcpyDeviceToHost, stream));
        WFS_HIP(hipStreamSynchronize(stream));
        return WFS_OK;
    }

    int total(int field, double* out) override {
        std::vector<double> pxs((size_t)d.nx);
        int rc = profile_x(field, 0, d.nz, pxs.data());
        if (rc) return rc;
        double s = 0.0;
        for (double v : pxs) s += v;
        *out = s;
        return WFS_OK;
    }

    int set_closure(int kind, const float* params, const uint16_t* geo_h, const uint8_t* hcol_h,
                    double u0) override {
        if (kind == WFS_CLOSURE_SMAGORINSKY) {
            closure = kind;
            return WFS_OK;
        }
        if (kind != WFS_CLOSURE_MLP) return fail(WFS_EINVAL, "wfs_set_closure: unknown closure kind");
        if constexpr (!std::is_same<R, float>::value) {
            return fail(WFS_EINVAL, "wfs_set_closure: the learned closure runs in float arithmetic (not FP64 storage)");
        } else {
            if (!params || !geo_h || !hcol_h) return fail(WFS_EINVAL, "wfs_set_closure: null argument");
            if (!(u0 > 0.0) || !std::isfinite(u0)) return fail(WFS_EINVAL, "wfs_set_closure: u0 must be positive");
            for (int k = 0; k < kNParams; ++k)
                if (!std::isfinite(params[k])) return fail(WFS_EINVAL, "wfs_set_closure: non-finite weight");
            int rc;
            if (!geo.p) {
                if ((rc = alloc_buf(geo, (size_t)Np * sizeof(uint16_t))) ||
                    (rc = alloc_buf(hcol, (size_t)px * d.ny)) || (rc = alloc_buf(mlp_params, kNParams * sizeof(float))) ||
                    (rc = alloc_buf(mlp_pack, sizeof(MlpPack))))
                    return rc;
            }
            if (c2_keep && !c2buf.p && (rc = alloc_buf(c2buf, (size_t)Np * sizeof(float)))) return rc;
            if ((rc = put_dense<uint16_t, uint16_t>(geo, geo_h, uint16_t(0)))) return rc;
            const long long ncol = (long long)px * d.ny;
            WFS_HIP(hipMemcpyAsync(stage.p, hcol_h, (size_t)d.nx * d.ny, hipMemcpyHostToDevice, stream));
            const CellMap cm{d.nx, px, d.ny, 0, nullptr, nullptr};   // per column: rows of the (x, y) plane
            hipLaunchKernelGGL((k_pad<uint8_t, uint8_t>), dim3(blocks_for(ncol, 256)), dim3(256), 0, stream,
                               static_cast<const uint8_t*>(stage.p), static_cast<uint8_t*>(hcol.p), cm, 0LL, ncol,
                               0LL, uint8_t(0));
            WFS_LAUNCH_CHECK();
            WFS_HIP(hipMemcpyAsync(mlp_params.p, params, kNParams * sizeof(float), hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(256), 0, stream, static_cast<const float*>(mlp_params.p),
                               static_cast<MlpPack*>(mlp_pack.p));
            WFS_LAUNCH_CHECK();
            if (c2buf.p && (rc = zero(c2buf))) return rc;
            WFS_HIP(hipStreamSynchronize(stream));
            inv_u0 = (float)(1.0 / u0);
            closure = kind;
            return WFS_OK;
        }
    }

    // Per-step column velocity: off, the wall model reads uref, which the run
    // cadence refreshes (wfs_update_uref); on, every step refreshes it from the
    // previous step inside the step kernel. Switching on starts from the
    // current velocity field (valid after a wfs_step call).
    int set_column_u(int per_step) override {
        if (!per_step) {
            colu_on = false;
            WFS_HIP(hipStreamSynchronize(stream));
            for (auto& p : colu)
                for (auto& b : p) release(b);
            release(colcount);
            return WFS_OK;
        }
        if (d.nz > 255) return fail(WFS_EINVAL, "wfs_set_column_u: more than 255 cells per column");
        int rc;
        for (auto& p : colu)
            for (auto& b : p)
                if (!b.p && (rc = alloc_buf(b, (size_t)Np * sizeof(R)))) return rc;
        if (!colcount.p && (rc = alloc_buf(colcount, (size_t)px * d.ny))) return rc;
        hipLaunchKernelGGL(k_colcount, dim3(blocks_for(d.nx, 64), d.ny), dim3(64), 0, stream, flags(),
                           static_cast<uint8_t*>(colcount.p), cmap(), d.nz);
        WFS_LAUNCH_CHECK();
        for (int k = 0; k < 2; ++k) {
            hipLaunchKernelGGL(k_colu_init<R>, dim3(blocks_for(Np, 256)), dim3(256), 0, stream, flags(), rp(ux),
                               rp(uy), rp(colu[k][0]), rp(colu[k][1]), Np);
            WFS_LAUNCH_CHECK();
        }
        WFS_HIP(hipStreamSynchronize(stream));
        colu_on = true;
        return WFS_OK;
    }

    // Particle tracer: positions (metres from the domain corner, x y z per
    // particle) live on the device; each move reads the velocity field in
    // place, so only positions cross to the host.
    int particles_set(long long n, const double* xyz, long long ninlet, const int32_t* inlet, uint64_t seed) override {
        if (n < 0 || ninlet < 0) return fail(WFS_EINVAL, "wfs_particles_set: negative count");
        WFS_HIP(hipStreamSynchronize(stream));
        release(part_xyz);
        release(part_inlet);
        part_n = 0;
        if (n == 0) return WFS_OK;
        int rc;
        if ((rc = alloc_buf(part_xyz, (size_t)n * 3 * sizeof(double))) ||
            (ninlet > 0 && (rc = alloc_buf(part_inlet, (size_t)ninlet * sizeof(int32_t)))) ||
            (!part_count.p && (rc = alloc_buf(part_count, sizeof(unsigned long long)))))
            return rc;
        for (long long k = 0; k < ninlet; ++k)
            if (inlet[k] < 0 || inlet[k] >= rows()) return fail(WFS_EINVAL, "wfs_particles_set: inlet row out of range");
        WFS_HIP(hipMemcpy(part_xyz.p, xyz, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice));
        if (ninlet > 0)
            WFS_HIP(hipMemcpy(part_inlet.p, inlet, (size_t)ninlet * sizeof(int32_t), hipMemcpyHostToDevice));
        part_n = n;
        part_ninlet = ninlet;
        part_seed = seed;
        part_step = 0;
        return WFS_OK;
    }

    int particles_move(double dt, double vel, double dx, long long* respawned) override {
        if (part_n == 0) return fail(WFS_EINVAL, "wfs_particles_move: no particles (wfs_particles_set)");
        if (!(dx > 0)) return fail(WFS_EINVAL, "wfs_particles_move: dx must be positive");
        WFS_HIP(hipMemsetAsync(part_count.p, 0, sizeof(unsigned long long), stream));
        hipLaunchKernelGGL(k_particles<R>, dim3(blocks_for(part_n, 256)), dim3(256), 0, stream,
                           static_cast<double*>(part_xyz.p), part_n, rp(ux), rp(uy), rp(uz), flags(), cmap(), d.nz,
                           dx, (float)vel, (float)dt, static_cast<const int*>(part_inlet.p), part_ninlet, part_seed,
                           part_step, static_cast<unsigned long long*>(part_count.p));
        WFS_LAUNCH_CHECK();
        ++part_step;
        unsigned long long cnt = 0;
        WFS_HIP(hipMemcpyAsync(&cnt, part_count.p, sizeof(cnt), hipMemcpyDeviceToHost, stream));
        WFS_HIP(hipStreamSynchronize(stream));
        if (respawned) *respawned = (long long)cnt;
        return WFS_OK;
    }

    int particles_get(double* xyz) override {
        if (part_n == 0) return WFS_OK;
        WFS_HIP(hipMemcpyAsync(xyz, part_xyz.p, (size_t)part_n * 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        WFS_HIP(hipStreamSynchronize(stream));
        return WFS_OK;
    }

    // The mask is kept as one bit per dense cell (x + nx (y + ny z)), so that
    // hull cells stay known where the tile-sparse layout stores no tile (a run
    // of 64 solid cells): the budget looks up the wall cell of every link.
    int set_body_mask(const uint8_t* m) override {
        if (!m) {
            WFS_HIP(hipStreamSynchronize(stream));
            release(body);
            return WFS_OK;
        }
        const size_t words = (size_t)((N + 31) / 32);
        std::vector<uint32_t> bits(words, 0u);
        for (long long k = 0; k < N; ++k)
            if (m[k]) bits[(size_t)(k >> 5)] |= 1u << (k & 31);
        if (!body.p) {
            int rc = alloc_buf(body, words * sizeof(uint32_t));
            if (rc) return rc;
        }
        WFS_HIP(hipMemcpy(body.p, bits.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice));
        return WFS_OK;
    }

    int forces(double* out) override {
        int rc = aa_readable("wfs_forces");
        if (rc) return rc;
        if (inplace && (rc = aa_exchange())) return rc;
        rc = forces_ab(out);
        if (inplace) {
            int rc2 = aa_exchange();
            if (!rc) rc = rc2;
            if (hipStreamSynchronize(stream) != hipSuccess && !rc) rc = fail(WFS_EHIP, "wfs_forces: synchronize");
        }
        return rc;
    }

    // wfs_forces through the fixed scratch array: no device allocation.
    int forces_noalloc(double* out) override {
        int rc = aa_readable("wfs_forces_noalloc");
        if (rc) return rc;
        double* sc = nullptr;
        if ((rc = scratch_of(device, &sc))) return rc;
        if (inplace && (rc = aa_exchange())) return rc;
        StepArgs<R> a;
        rc = fill_args(a);
        if (!rc) {
            if (tiled)
                hipLaunchKernelGGL((k_wall_force_strided<Codec, true>), dim3(kNoAllocGroups), dim3(256), 0, stream,
                                   fp(cur), a, cmap(), static_cast<const uint32_t*>(body.p), sc);
            else
                hipLaunchKernelGGL((k_wall_force_strided<Codec, false>), dim3(kNoAllocGroups), dim3(256), 0, stream,
                                   fp(cur), a, cmap(), static_cast<const uint32_t*>(body.p), sc);
            if (hipGetLastError() != hipSuccess) rc = fail(WFS_EHIP, "wfs_forces_noalloc: launch");
        }
        if (!rc) {
            hipLaunchKernelGGL(k_sum_parts<9>, dim3(1), dim3(256), 0, stream, static_cast<const double*>(sc),
                               (long long)kNoAllocGroups, (long long)kNoAllocGroups, sc + (long long)kNoAllocGroups * 9);
            if (hipGetLastError() != hipSuccess) rc = fail(WFS_EHIP, "wfs_forces_noalloc: launch");
        }
        if (!rc && hipMemcpyAsync(out, sc + (long long)kNoAllocGroups * 9, 9 * sizeof(double), hipMemcpyDeviceToHost,
                                  stream) != hipSuccess)
            rc = fail(WFS_EHIP, "wfs_forces_noalloc: copy");
        if (inplace) {
            int rc2 = aa_exchange();
            if (!rc) rc = rc2;
        }
        if (hipStreamSynchronize(stream) != hipSuccess && !rc) rc = fail(WFS_EHIP, "wfs_forces_noalloc: synchronize");
        return rc;
    }

    int forces_ab(double* out) {
        StepArgs<R> a;
        int rc = fill_args(a);
        if (rc) return rc;
        const long long blocks = (Np + 255) / 256;
        DevBuf part, sum;
        if ((rc = alloc_buf(part, (size_t)(blocks + kSumChunks) * 9 * sizeof(double))) ||
            (rc = alloc_buf(sum, 9 * sizeof(double))))
            return rc;
        if (tiled)
            hipLaunchKernelGGL((k_wall_force<Codec, true>), dim3((unsigned)blocks), dim3(256), 0, stream, fp(cur), a,
                               cmap(), static_cast<const uint32_t*>(body.p), static_cast<double*>(part.p));
        else
            hipLaunchKernelGGL((k_wall_force<Codec, false>), dim3((unsigned)blocks), dim3(256), 0, stream, fp(cur),
                               a, cmap(), static_cast<const uint32_t*>(body.p), static_cast<double*>(part.p));
        WFS_LAUNCH_CHECK();
        // fixed-order two-pass sum: kSumChunks chunk sums, then their sum
        const long long chunk = (blocks + kSumChunks - 1) / kSumChunks;
        const long long nch = (blocks + chunk - 1) / chunk;
        hipLaunchKernelGGL(k_sum_parts<9>, dim3((unsigned)nch), dim3(256), 0, stream, static_cast<const double*>(part.p),
                           blocks, chunk, static_cast<double*>(part.p) + blocks * 9);
        WFS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sum_parts<9>, dim3(1), dim3(256), 0, stream,
                           static_cast<const double*>(part.p) + blocks * 9, nch, nch, static_cast<double*>(sum.p));
        WFS_LAUNCH_CHECK();
        WFS_HIP(hipMemcpyAsync(out, sum.p, 9 * sizeof(double), hipMemcpyDeviceToHost, stream));
        WFS_HIP(hipStreamSynchronize(stream));
        bytes -= part.bytes + sum.bytes;
        return WFS_OK;
    }

    int get_closure_field(float* out) override {
        if (!c2buf.p) return fail(WFS_EINVAL, "wfs_get_closure_field: no learned closure set, or its field is off "
                                              "(wfs_set_closure_field)");
        return get_cells(c2buf, out, false);
    }

    // The learned step records C^2 per cell (4 bytes per stored cell) only
    // while this is on; off frees the field. On again, it holds zeros until
    // the next step.
    int set_closure_field(int on) override {
        c2_keep = on != 0;
        if (!c2_keep) {
            WFS_HIP(hipStreamSynchronize(stream));
            release(c2buf);
            return WFS_OK;
        }
        if (closure != WFS_CLOSURE_MLP || c2buf.p) return WFS_OK;
        int rc = alloc_buf(c2buf, (size_t)Np * sizeof(float));
        return rc ? rc : zero(c2buf);
    }

    // Training samples from this (fine) run, box-filtered 2x2x2 onto the
    // coarse grid ncx x ncy x ncz; outputs are device arrays (dense coarse order).
    int closure_samples(const uint16_t* geo_h, const uint8_t* hcol_h, int ncx, int ncy, int ncz, double kv, double u0,
                        float* X, float* ABP, uint8_t* valid) override {
        if (tiled) return fail(WFS_EINVAL, "wfs_closure_samples: runs on the dense layout");
        if (ncx < 3 || ncy < 3 || ncz < 2 || 2 * ncx > d.nx || 2 * ncy > d.ny || 2 * ncz > d.nz)
            return fail(WFS_EINVAL, "wfs_closure_samples: coarse grid must be at most half the fine grid");
        if (!(u0 > 0.0) || !(kv > 0.0)) return fail(WFS_EINVAL, "wfs_closure_samples: u0 and kv must be positive");
        if (!geo_h || !hcol_h || !X || !ABP || !valid) return fail(WFS_EINVAL, "wfs_closure_samples: null argument");
        const long long nc = (long long)ncx * ncy * ncz;
        DevBuf ub, ts, cf, g, hc;
        int rc;
        if ((rc = alloc_buf(ub, (size_t)nc * 3 * sizeof(float))) || (rc = alloc_buf(ts, (size_t)nc * 6 * sizeof(float))) ||
            (rc = alloc_buf(cf, (size_t)nc)) || (rc = alloc_buf(g, (size_t)nc * sizeof(uint16_t))) ||
            (rc = alloc_buf(hc, (size_t)ncx * ncy)))
            return rc;
        WFS_HIP(hipMemcpyAsync(g.p, geo_h, (size_t)nc * sizeof(uint16_t), hipMemcpyHostToDevice, stream));
        WFS_HIP(hipMemcpyAsync(hc.p, hcol_h, (size_t)ncx * ncy, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_coarsen<R>, dim3(blocks_for(nc, 256)), dim3(256), 0, stream, rp(ux), rp(uy), rp(uz),
                           flags(), px, d.ny, ncx, ncy, ncz, (float)kv, static_cast<float*>(ub.p),
                           static_cast<float*>(ts.p), static_cast<uint8_t*>(cf.p));
        WFS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_samples, dim3(blocks_for(nc, 256)), dim3(256), 0, stream, static_cast<const float*>(ub.p),
                           static_cast<const float*>(ts.p), static_cast<const uint8_t*>(cf.p),
                           static_cast<const uint16_t*>(g.p), static_cast<const uint8_t*>(hc.p), ncx, ncy, ncz,
                           (float)(1.0 / u0), X, ABP, valid);
        WFS_LAUNCH_CHECK();
        WFS_HIP(hipStreamSynchronize(stream));
        bytes -= ub.bytes + ts.bytes + cf.bytes + g.bytes + hc.bytes;
        return WFS_OK;
    }

    int check_finite() override {
        int rc = zero(counter);
        if (rc) return rc;
        hipLaunchKernelGGL(k_nonfinite<R>, dim3(blocks_for(Np, 256)), dim3(256), 0, stream, rp(rho), rp(ux), rp(uy),
                           rp(uz), flags(), Np, static_cast<unsigned int*>(counter.p));
        WFS_LAUNCH_CHECK();
        unsigned int cnt = 0;
        WFS_HIP(hipMemcpyAsync(&cnt, counter.p, sizeof(cnt), hipMemcpyDeviceToHost, stream));
        WFS_HIP(hipStreamSynchronize(stream));
        if (cnt) return fail(WFS_ENONFINITE, std::to_string(cnt) + " fluid cells hold a non-finite state after step " +
                                                 std::to_string(nstep));
        return WFS_OK;
    }
};

// ---------------------------------------------------------------- batches
// A batch steps several uploaded solvers (same storage format, same device)
// together: one step launch covers every case, and the 20-step cadence of the
// run driver (ramp, uref, time-mean sums) runs on the device from per-case
// schedules, so a whole chunk can be captured once as a HIP graph and replayed.
struct BatchBase {
    int device = 0;
    int storage = 0;
    hipStream_t stream = nullptr;
    virtual ~BatchBase() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    virtual int set_schedule(int k, double dt, double t_ramp, double t_avg, double t0) = 0;
    virtual int set_budget(int k, long long steps) = 0;
    virtual int budget_of(int k, long long* steps_left) = 0;
    virtual int run(int nsteps, int use_graph) = 0;
    virtual int synchronize() = 0;
    virtual int time_of(int k, double* t, int32_t* navg) = 0;
};

template <class Codec>
struct Batch final : BatchBase {
    using R = typename Codec::R;
    std::vector<Solver<Codec>*> cases;
    std::vector<CaseArgs<Codec>> args;   // host copy of the per-case step views
    DevBuf d_args, d_group_case, d_sched, d_af;   // d_af: per case (active, flip) for the step kernels
    std::vector<CaseSched<R>> sched;     // host mirror, valid after synchronize()
    unsigned groups = 0;
    int parity = 0;                      // batch parity: case k reads f[parity ^ flip_k]
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool dirty_sched = true;
    int bx = 128;                        // work-group width fixed when the batch is made
    bool inplace = false;                // every case streams in place on one buffer
    bool pop32 = true;                   // every case's population buffer is below 4 GiB

    ~Batch() override {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto* c : cases) --c->batch_refs;
    }

    int init(wfs_solver* const* hs, int n) {
        int bx0 = -1;
        std::vector<unsigned short> gcase;
        for (int k = 0; k < n; ++k) {
            auto* c = dynamic_cast<Solver<Codec>*>(hs[k]);
            if (!c) return fail(WFS_EINVAL, "wfs_batch_create: cases must share one storage format");
            if (!c->uploaded) return fail(WFS_EINVAL, "wfs_batch_create: case " + std::to_string(k) + " has no geometry");
            if (c->device != device) return fail(WFS_EINVAL, "wfs_batch_create: cases must be on one device");
            if (bx0 < 0) bx0 = c->bx;
            if (c->bx != bx0) return fail(WFS_EINVAL, "wfs_batch_create: cases must share block_x");
            if (c->tiled) return fail(WFS_EINVAL, "wfs_batch_create: batched cases use the dense layout");
            if (c->colu_on) return fail(WFS_EINVAL, "wfs_batch_create: batched cases use the run cadence's uref");
            if (k == 0) inplace = c->inplace;
            if (c->inplace != inplace)
                return fail(WFS_EINVAL, "wfs_batch_create: cases must share the streaming mode (in place or two buffers)");
            if (c->closure != WFS_CLOSURE_SMAGORINSKY)
                return fail(WFS_EINVAL, "wfs_batch_create: batched cases use the constant-Cs closure");
            int rs = c->ensure_sums();   // the batch accumulates into them
            if (rs) return rs;
            WFS_HIP(hipStreamSynchronize(c->stream));
            cases.push_back(c);
            ++c->batch_refs;
            CaseArgs<Codec> ca{};
            int rc = c->fill_args(ca.a);
            if (rc) return rc;
            ca.f[0] = c->fp(c->cur);
            ca.f[1] = c->inplace ? c->fp(c->cur) : c->fp(c->cur ^ 1);
            pop32 = pop32 && ca.a.pop32;
            ca.fscale = nullptr;   // set below, once the schedule array exists
            ca.group_begin = (unsigned)gcase.size();
            const long long ng = (long long)((ca.a.rows + 7) / 8) * 8 * ca.a.tiles;
            if ((long long)gcase.size() + ng > (1ll << 31)) return fail(WFS_EINVAL, "wfs_batch_create: too many cells");
            gcase.insert(gcase.end(), (size_t)ng, (unsigned short)k);
            args.push_back(ca);
            CaseSched<R> sc{};
            sc.t = 0.0;
            sc.dt = 1.0;
            sc.t_ramp = 0.0;
            sc.t_avg = 1e300;
            sc.fscale = c->fscale;
            sc.ramp_done = 1;
            sc.navg = c->navg;
            sc.accumulate = 0;
            sc.steps_left = -1;
            sc.steps_taken = 0;
            sc.active = 0;
            sc.flip = c->inplace ? c->aa_phase : 0;   // in place: parity ^ flip is the streaming phase
            sched.push_back(sc);
        }
        groups = (unsigned)gcase.size();
        bx = bx0;
        int rc;
        if ((rc = cases[0]->alloc_buf(d_group_case, gcase.size() * sizeof(unsigned short)))) return rc;
        if ((rc = cases[0]->alloc_buf(d_sched, sched.size() * sizeof(CaseSched<R>)))) return rc;
        if ((rc = cases[0]->alloc_buf(d_args, args.size() * sizeof(CaseArgs<Codec>)))) return rc;
        if ((rc = cases[0]->alloc_buf(d_af, args.size() * 2 * sizeof(int)))) return rc;
        for (size_t k = 0; k < args.size(); ++k)
            args[k].fscale = &static_cast<CaseSched<R>*>(d_sched.p)[k].fscale;
        WFS_HIP(hipMemcpyAsync(d_group_case.p, gcase.data(), gcase.size() * sizeof(unsigned short),
                               hipMemcpyHostToDevice, stream));
        WFS_HIP(hipMemcpyAsync(d_args.p, args.data(), args.size() * sizeof(CaseArgs<Codec>), hipMemcpyHostToDevice,
                               stream));
        // the device schedule is valid from the start: a synchronize before
        // the first run reads back what the cases hold now
        WFS_HIP(hipMemcpyAsync(d_sched.p, sched.data(), sched.size() * sizeof(CaseSched<R>), hipMemcpyHostToDevice,
                               stream));
        WFS_HIP(hipStreamSynchronize(stream));
        dirty_sched = false;
        return WFS_OK;
    }

    int set_schedule(int k, double dt, double t_ramp, double t_avg, double t0) override {
        if (k < 0 || k >= (int)cases.size()) return fail(WFS_EINVAL, "wfs_batch_set_schedule: case out of range");
        if (!(dt > 0.0)) return fail(WFS_EINVAL, "wfs_batch_set_schedule: dt must be positive");
        CaseSched<R>& c = sched[(size_t)k];
        c.t = t0;
        c.dt = dt;
        c.t_ramp = t_ramp;
        c.t_avg = t_avg;
        c.ramp_done = t_ramp <= 0.0;
        if (!c.ramp_done) c.fscale = R(0);
        dirty_sched = true;
        return WFS_OK;
    }

    // Budget of case k in steps from now (negative: no limit). A case whose
    // budget is spent sits out the launches until it gets a new one.
    int set_budget(int k, long long steps) override {
        if (k < 0 || k >= (int)cases.size()) return fail(WFS_EINVAL, "wfs_batch_set_budget: case out of range");
        sched[(size_t)k].steps_left = steps < 0 ? -1 : steps;
        dirty_sched = true;
        return WFS_OK;
    }

    int budget_of(int k, long long* steps_left) override {
        if (k < 0 || k >= (int)cases.size()) return fail(WFS_EINVAL, "wfs_batch_budget: case out of range");
        *steps_left = sched[(size_t)k].steps_left;
        return WFS_OK;
    }

    // The kernels of one chunk of m steps, queued on the batch stream. Each
    // case takes min(m, its budget) of them (k_batch_pre), with the cadence
    // of a sequential chunk of that length.
    int enqueue_chunk(int m, int parity0) {
        const int n = (int)cases.size();
        CaseSched<R>* ds = static_cast<CaseSched<R>*>(d_sched.p);
        int* af = static_cast<int*>(d_af.p);
        hipLaunchKernelGGL(k_batch_pre<R>, dim3(blocks_for(n, 64)), dim3(64), 0, stream, ds, af, n, m);
        WFS_LAUNCH_CHECK();
        for (size_t k = 0; k < cases.size(); ++k) {
            Solver<Codec>* c = cases[k];
            if (!c->have_bedwall) continue;
            hipLaunchKernelGGL(k_uref_if<R>, dim3(blocks_for(c->d.nx, 64), c->d.ny), dim3(64), 0, stream,
                               static_cast<const CaseSched<R>*>(ds), (int)k, c->rp(c->ux), c->rp(c->uy), c->flags(),
                               c->rp(c->uref), c->cmap(), c->d.nz);
            WFS_LAUNCH_CHECK();
        }
        const CaseArgs<Codec>* da = static_cast<const CaseArgs<Codec>*>(d_args.p);
        int par = parity0;
        const unsigned short* gc = static_cast<const unsigned short*>(d_group_case.p);
        for (int s = 0; s < m; ++s) {
            if (inplace) {   // outlet rows fused into the step (required for in-place streaming)
                if (pop32)
                    hipLaunchKernelGGL((k_step_batch_aa<Codec, true>), dim3(groups), dim3(bx), 0, stream, da, gc,
                                       static_cast<const int*>(af), par, s);
                else
                    hipLaunchKernelGGL((k_step_batch_aa<Codec, false>), dim3(groups), dim3(bx), 0, stream, da, gc,
                                       static_cast<const int*>(af), par, s);
                WFS_LAUNCH_CHECK();
                par ^= 1;
                continue;
            }
            hipLaunchKernelGGL(k_step_batch<Codec>, dim3(groups), dim3(bx), 0, stream, da, gc,
                               static_cast<const int*>(af), par, s);
            WFS_LAUNCH_CHECK();
            for (size_t k = 0; k < cases.size(); ++k) {
                Solver<Codec>* c = cases[k];
                if ((c->d.nx - 1) % 64 != 0) continue;   // outlet row fused into the step
                hipLaunchKernelGGL(k_bc_batch<Codec>, dim3(blocks_for(c->rows(), 64)), dim3(64), 0, stream, da,
                                   static_cast<const int*>(af), (int)k, par, s);
                WFS_LAUNCH_CHECK();
            }
            par ^= 1;
        }
        hipLaunchKernelGGL(k_batch_post<R>, dim3(blocks_for(n, 64)), dim3(64), 0, stream, ds, af, n, m);
        WFS_LAUNCH_CHECK();
        for (size_t k = 0; k < cases.size(); ++k) {
            Solver<Codec>* c = cases[k];
            hipLaunchKernelGGL(k_accumulate_if<R>, dim3(blocks_for(c->Np, 256)), dim3(256), 0, stream,
                               static_cast<const CaseSched<R>*>(d_sched.p), (int)k, c->rp(c->rho), c->rp(c->ux),
                               c->rp(c->uy), c->rp(c->uz), c->rp(c->sums[0]), c->rp(c->sums[1]), c->rp(c->sums[2]),
                               c->rp(c->sums[3]), c->rp(c->sums[4]), c->rp(c->sums[5]), c->Np);
            WFS_LAUNCH_CHECK();
        }
        return WFS_OK;
    }

    int run(int nsteps, int use_graph) override {
        if (nsteps < 0) return fail(WFS_EINVAL, "wfs_batch_run: nsteps must be >= 0");
        if (dirty_sched) {
            WFS_HIP(hipMemcpyAsync(d_sched.p, sched.data(), sched.size() * sizeof(CaseSched<R>),
                                   hipMemcpyHostToDevice, stream));
            dirty_sched = false;
        }
        const int cadence = 20;
        const int full = nsteps / cadence, rest = nsteps % cadence;
        int rc;
        if (use_graph && full > 0) {
            if (!graph_exec) {
                // one chunk of 20 steps (an even count, so it starts and ends on the same parity)
                WFS_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                rc = enqueue_chunk(cadence, parity);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(stream, &g);
                if (rc) return rc;
                if (e != hipSuccess) return fail(WFS_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                graph = g;
                WFS_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
            }
            for (int k = 0; k < full; ++k) WFS_HIP(hipGraphLaunch(graph_exec, stream));
        } else {
            for (int k = 0; k < full; ++k)
                if ((rc = enqueue_chunk(cadence, parity))) return rc;
        }
        if (rest) {
            if ((rc = enqueue_chunk(rest, parity))) return rc;
            parity ^= rest & 1;
            if (graph_exec) {   // the captured chunk assumed the old parity
                (void)hipGraphExecDestroy(graph_exec);
                (void)hipGraphDestroy(graph);
                graph_exec = nullptr;
                graph = nullptr;
            }
        }
        return WFS_OK;
    }

    int synchronize() override {
        WFS_HIP(hipStreamSynchronize(stream));
        WFS_HIP(hipMemcpy(sched.data(), d_sched.p, sched.size() * sizeof(CaseSched<R>), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < cases.size(); ++k) {
            Solver<Codec>* c = cases[k];
            c->nstep += sched[k].steps_taken;
            sched[k].steps_taken = 0;
            c->navg = sched[k].navg;
            c->fscale = sched[k].fscale;
            if (inplace)
                c->aa_phase = (parity ^ sched[k].flip) & 1;
            else
                c->cur = (args[k].f[(parity ^ sched[k].flip) & 1] == c->fp(0)) ? 0 : 1;
        }
        dirty_sched = true;   // steps_taken cleared on the host copy
        return WFS_OK;
    }

    int time_of(int k, double* t, int32_t* navg) override {
        if (k < 0 || k >= (int)cases.size()) return fail(WFS_EINVAL, "wfs_batch_time: case out of range");
        *t = sched[(size_t)k].t;
        *navg = sched[(size_t)k].navg;
        return WFS_OK;
    }
};

wfs_solver* make_solver(int storage) {
    switch (storage) {
        case WFS_STORE_F32: return new Solver<wfs::CodecF32>();
        case WFS_STORE_F16DEV: return new Solver<wfs::CodecF16Dev>();
        case WFS_STORE_BF16DEV: return new Solver<wfs::CodecBF16Dev>();
        case WFS_STORE_F64: return new Solver<wfs::CodecF64>();
        case WFS_STORE_F32_FAST: return new Solver<wfs::CodecF32Fast>();
        default: return nullptr;
    }
}

int check_handle(const wfs_solver* h, bool need_upload = true) {
    if (!h) return fail(WFS_EINVAL, "null solver handle");
    if (need_upload && !h->uploaded) return fail(WFS_EINVAL, "solver has no geometry yet (call wfs_upload first)");
    return WFS_OK;
}

// A batch keeps its own pointers to a case's population buffers and its
// parity: while a batch holds a solver, calls that step it or replace its
// state directly would be overwritten by the batch's next run (or race its
// stream), so they are refused.
int check_not_batched(const wfs_solver* h, const char* what) {
    if (h->batch_refs) return fail(WFS_EINVAL, std::string(what) + ": the solver is a case of a batch (destroy the batch first)");
    return WFS_OK;
}

// Every call runs with the handle's device current.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int now = -1;
        if (prev >= 0 && hipGetDevice(&now) == hipSuccess && now != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" {

int wfs_abi_version(void) { return WFS_ABI_VERSION; }

int wfs_device_count(int32_t* count) {
    if (!count) return fail(WFS_EINVAL, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return WFS_OK;
}

const char* wfs_last_error(void) { return g_err.c_str(); }

int wfs_create(const wfs_desc* desc, wfs_solver** out) {
    if (!desc || !out) return fail(WFS_EINVAL, "wfs_create: null argument");
    *out = nullptr;
    const wfs_desc& d = *desc;
    if (d.nx < 3 || d.ny < 1 || d.nz < 1) return fail(WFS_EINVAL, "wfs_create: need nx >= 3, ny >= 1, nz >= 1");
    const long long N = (long long)d.nx * d.ny * d.nz;
    if (N >= (1ll << 31)) return fail(WFS_EINVAL, "wfs_create: nx*ny*nz must be below 2^31 cells");
    if (d.ny > 65535 || d.nz > 65535) return fail(WFS_EINVAL, "wfs_create: ny and nz must be at most 65535");
    if (!(d.tau0 > 0.5) || !std::isfinite(d.tau0)) return fail(WFS_EINVAL, "wfs_create: tau0 must exceed 0.5");
    if (!(d.cs >= 0.0) || !std::isfinite(d.cs)) return fail(WFS_EINVAL, "wfs_create: cs must be >= 0");
    const int bx = d.block_x == 0 ? 128 : d.block_x;
    if (bx < 64 || bx > 256 || bx % 64) return fail(WFS_EINVAL, "wfs_create: block_x must be 64, 128, 192 or 256");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(WFS_ENODEV, "wfs_create: no HIP device visible");
    }
    if (d.device < 0 || d.device >= ndev) return fail(WFS_ENODEV, "wfs_create: device ordinal out of range");
    std::unique_ptr<wfs_solver> h(make_solver(d.storage));
    if (!h) return fail(WFS_EINVAL, "wfs_create: unknown storage format");
    h->d = d;
    h->N = N;
    h->px = (d.nx + 63) / 64 * 64;
    h->Np = (long long)h->px * d.ny * d.nz;
    if (h->Np >= (1ll << 31)) return fail(WFS_EINVAL, "wfs_create: padded lattice exceeds 2^31 cells");
    // the step kernel addresses each field and each population direction with 32-bit byte offsets
    const unsigned long long rsz = d.storage == WFS_STORE_F64 ? 8ull : 4ull;
    if ((unsigned long long)h->Np * rsz >= (1ull << 32))
        return fail(WFS_EINVAL, "wfs_create: one field or population direction would exceed 4 GiB "
                                "(at most 2^30 padded cells with float fields, 2^29 with FP64 storage)");
    h->Ns = h->Np;
    h->device = d.device;
    DeviceGuard g(d.device);
    WFS_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    int rc = h->alloc();
    if (rc) return rc;
    h->bx = bx;
    *out = h.release();
    return WFS_OK;
}

// Tile tables of a solid mask: tiles are the 64-cell x-runs of each row that
// hold a fluid cell, numbered in row order (z, y, x-tile).
static void build_tiles(wfs_solver* h, const uint8_t* solid) {
    const wfs_desc& d = h->d;
    const int ntx = h->px / 64;
    h->ntx = ntx;
    const long long rows = (long long)d.ny * d.nz;
    h->h_tmap.assign((size_t)(ntx * rows), -1);
    h->h_tcoord.clear();
    std::vector<std::vector<int>> row_tiles;   // per non-empty row, its tiles
    h->h_rowtile.assign((size_t)rows + 1, 0);
    long long T = 0;
    for (long long j = 0; j < rows; ++j) {
        h->h_rowtile[(size_t)j] = (int)T;
        std::vector<int> rt;
        const uint8_t* r = solid + (size_t)j * d.nx;
        for (int tx = 0; tx < ntx; ++tx) {
            const int x0 = tx * 64, x1 = std::min(d.nx, x0 + 64);
            bool fluid = false;
            for (int x = x0; x < x1 && !fluid; ++x) fluid = r[x] == 0;
            if (!fluid) continue;
            h->h_tmap[(size_t)(tx + ntx * j)] = (int)T;
            h->h_tcoord.push_back(tx);
            h->h_tcoord.push_back((int)(j % d.ny));
            h->h_tcoord.push_back((int)(j / d.ny));
            rt.push_back((int)T);
            ++T;
        }
        if (!rt.empty()) row_tiles.push_back(std::move(rt));
    }
    h->T = T;
    h->h_rowtile[(size_t)rows] = (int)T;
    h->h_nbr.assign((size_t)T * 27, -1);
    for (long long t = 0; t < T; ++t) {
        const int tx = h->h_tcoord[3 * t], y = h->h_tcoord[3 * t + 1], z = h->h_tcoord[3 * t + 2];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int a = tx + dx, b = y + dy, c = z + dz;
                    if (a < 0 || a >= ntx || b < 0 || b >= d.ny || c < 0 || c >= d.nz) continue;
                    h->h_nbr[(size_t)(t * 27 + (dx + 1) + 3 * ((dy + 1) + 3 * (dz + 1)))] =
                        h->h_tmap[(size_t)(a + ntx * (b + (long long)d.ny * c))];
                }
    }
    // work-group -> tiles: rows dealt round-robin to the 8 XCD slots (work-group
    // L runs on XCD L % 8), so every tile of a row shares one L2; a work-group
    // takes G consecutive tiles of its slot
    std::vector<std::vector<int>> slot(8);
    for (size_t r = 0; r < row_tiles.size(); ++r)
        slot[r % 8].insert(slot[r % 8].end(), row_tiles[r].begin(), row_tiles[r].end());
    auto block_map = [&](size_t G, std::vector<int>& out) {
        size_t groups = 0;
        for (auto& s8 : slot) groups = std::max(groups, (s8.size() + G - 1) / G);
        out.assign(8 * groups * G, -1);
        for (int s8 = 0; s8 < 8; ++s8)
            for (size_t k = 0; k < slot[s8].size(); ++k) out[(8 * (k / G) + s8) * G + k % G] = slot[s8][k];
    };
    block_map(1, h->h_blk2tile);
    // the same map as per-work-group records (kWgRec ints: tile, x-tile, y, z, 27 neighbour tiles)
    h->h_wgrec.assign(h->h_blk2tile.size() * kWgRec, -1);
    for (size_t L = 0; L < h->h_blk2tile.size(); ++L) {
        const int t = h->h_blk2tile[L];
        if (t < 0) continue;
        int* r = &h->h_wgrec[(size_t)kWgRec * L];
        r[0] = t;
        for (int k = 0; k < 3; ++k) r[1 + k] = h->h_tcoord[3 * (size_t)t + k];
        for (int k = 0; k < 27; ++k) r[4 + k] = h->h_nbr[27 * (size_t)t + k];
    }
}

int wfs_create_tiled(const wfs_desc* desc, const uint8_t* solid, wfs_solver** out) {
    if (!desc || !out || !solid) return fail(WFS_EINVAL, "wfs_create_tiled: null argument");
    *out = nullptr;
    const wfs_desc& d = *desc;
    if (d.nx < 3 || (d.nx - 1) % 64 == 0)
        return fail(WFS_EINVAL, "wfs_create_tiled: need nx >= 3 with (nx - 1) % 64 != 0 (outlet row fused)");
    const long long N = (long long)d.nx * d.ny * d.nz;
    if (N >= (1ll << 40)) return fail(WFS_EINVAL, "wfs_create_tiled: lattice too large");
    if (!(d.tau0 > 0.5) || !std::isfinite(d.tau0)) return fail(WFS_EINVAL, "wfs_create_tiled: tau0 must exceed 0.5");
    if (!(d.cs >= 0.0) || !std::isfinite(d.cs)) return fail(WFS_EINVAL, "wfs_create_tiled: cs must be >= 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(WFS_ENODEV, "wfs_create_tiled: no HIP device visible");
    }
    if (d.device < 0 || d.device >= ndev) return fail(WFS_ENODEV, "wfs_create_tiled: device ordinal out of range");
    if (d.ny > 65535 || d.nz > 65535) return fail(WFS_EINVAL, "wfs_create_tiled: ny and nz must be at most 65535");
    std::unique_ptr<wfs_solver> h(make_solver(d.storage));
    if (!h) return fail(WFS_EINVAL, "wfs_create_tiled: unknown storage format");
    h->d = d;
    h->N = N;
    h->px = (d.nx + 63) / 64 * 64;
    h->tiled = true;
    h->bx = 64;
    build_tiles(h.get(), solid);
    h->Np = h->T * 64;
    if (h->Np == 0) return fail(WFS_EINVAL, "wfs_create_tiled: no fluid cell");
    const unsigned long long rsz = d.storage == WFS_STORE_F64 ? 8ull : 4ull;
    if (h->Np >= (1ll << 31) || (unsigned long long)h->Np * rsz >= (1ull << 32))
        return fail(WFS_EINVAL, "wfs_create_tiled: stored cells exceed 4 GiB per field or population direction");
    h->Ns = h->Np;
    h->device = d.device;
    DeviceGuard g(d.device);
    WFS_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    int rc = h->alloc();
    if (rc) return rc;
    *out = h.release();
    return WFS_OK;
}

int64_t wfs_stored_cells(const wfs_solver* h) { return h ? (int64_t)h->Np : -1; }

int wfs_set_force_cells(wfs_solver* h, int64_t n, const int64_t* cells, const float* fx, const float* fy,
                        const float* fz) {
    int rc = check_handle(h);
    if (rc) return rc;
    if ((rc = check_not_batched(h, "wfs_set_force_cells"))) return rc;
    if (n < 0 || (n > 0 && (!cells || !fx || !fy || !fz))) return fail(WFS_EINVAL, "wfs_set_force_cells: bad arguments");
    for (int64_t k = 0; k < n; ++k)
        if (cells[k] < 0 || cells[k] >= h->N) return fail(WFS_EINVAL, "wfs_set_force_cells: cell index out of range");
    DeviceGuard g(h->device);
    return h->set_force_cells(n, cells, fx, fy, fz);
}

int wfs_set_block_x(wfs_solver* h, int32_t block_x) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (block_x < 64 || block_x > 256 || block_x % 64)
        return fail(WFS_EINVAL, "wfs_set_block_x: block_x must be 64, 128, 192 or 256");
    if (h->tiled && block_x != 64) return fail(WFS_EINVAL, "wfs_set_block_x: the tile-sparse layout runs 64-cell tiles");
    h->bx = block_x;
    return WFS_OK;
}

int32_t wfs_block_x(const wfs_solver* h) { return h ? h->bx : -1; }

int wfs_destroy(wfs_solver* h) {
    if (!h) return WFS_OK;
    DeviceGuard g(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
    return WFS_OK;
}

int wfs_upload(wfs_solver* h, const uint8_t* solid, const float* cf, const float* uin) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!solid || !uin) return fail(WFS_EINVAL, "wfs_upload: solid and uin are required");
    DeviceGuard g(h->device);
    return h->upload(solid, cf, uin);
}

int wfs_set_force(wfs_solver* h, const float* Fx, const float* Fy, const float* Fz) {
    int rc = check_handle(h);
    if (rc) return rc;
    if ((rc = check_not_batched(h, "wfs_set_force"))) return rc;
    if (!Fx || !Fy || !Fz) return fail(WFS_EINVAL, "wfs_set_force: null force array");
    DeviceGuard g(h->device);
    return h->set_force(Fx, Fy, Fz);
}

int wfs_set_force_scale(wfs_solver* h, float scale) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!std::isfinite(scale)) return fail(WFS_EINVAL, "wfs_set_force_scale: scale must be finite");
    return h->set_force_scale(scale);
}

int wfs_set_uin(wfs_solver* h, const float* uin) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!uin) return fail(WFS_EINVAL, "wfs_set_uin: null array");
    DeviceGuard g(h->device);
    return h->set_uin(uin);
}

int wfs_set_uref(wfs_solver* h, const float* uref) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!uref) return fail(WFS_EINVAL, "wfs_set_uref: null array");
    DeviceGuard g(h->device);
    return h->set_uref(uref);
}

int wfs_set_bedwall(wfs_solver* h, const uint8_t* bedwall) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!bedwall) return fail(WFS_EINVAL, "wfs_set_bedwall: null array");
    DeviceGuard g(h->device);
    return h->set_bedwall(bedwall);
}

int wfs_update_uref(wfs_solver* h) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->update_uref();
}

int wfs_step(wfs_solver* h, int32_t n) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (n < 0) return fail(WFS_EINVAL, "wfs_step: n must be >= 0");
    if ((rc = check_not_batched(h, "wfs_step"))) return rc;
    if (n == 0) return WFS_OK;
    DeviceGuard g(h->device);
    return h->step(n);
}

int wfs_synchronize(wfs_solver* h) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    DeviceGuard g(h->device);
    WFS_HIP(hipStreamSynchronize(h->stream));
    return WFS_OK;
}

int64_t wfs_nstep(const wfs_solver* h) { return h ? h->nstep : -1; }

int wfs_get_macro(wfs_solver* h, float* rho, float* ux, float* uy, float* uz) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->get_macro(rho, ux, uy, uz, false);
}

int wfs_get_macro_f64(wfs_solver* h, double* rho, double* ux, double* uy, double* uz) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->get_macro(rho, ux, uy, uz, true);
}

int wfs_set_macro(wfs_solver* h, const float* rho, const float* ux, const float* uy, const float* uz) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!rho || !ux || !uy || !uz) return fail(WFS_EINVAL, "wfs_set_macro: null array");
    if ((rc = check_not_batched(h, "wfs_set_macro"))) return rc;
    DeviceGuard g(h->device);
    return h->set_macro(rho, ux, uy, uz);
}

int wfs_get_uref(wfs_solver* h, float* uref) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!uref) return fail(WFS_EINVAL, "wfs_get_uref: null array");
    DeviceGuard g(h->device);
    return h->get_uref(uref);
}

int wfs_get_populations(wfs_solver* h, float* f) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!f) return fail(WFS_EINVAL, "wfs_get_populations: null array");
    DeviceGuard g(h->device);
    return h->get_pops(f, false);
}

int wfs_get_populations_f64(wfs_solver* h, double* f) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!f) return fail(WFS_EINVAL, "wfs_get_populations_f64: null array");
    DeviceGuard g(h->device);
    return h->get_pops(f, true);
}

int wfs_set_populations(wfs_solver* h, const float* f) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!f) return fail(WFS_EINVAL, "wfs_set_populations: null array");
    if ((rc = check_not_batched(h, "wfs_set_populations"))) return rc;
    DeviceGuard g(h->device);
    return h->set_pops(f);
}

int wfs_accumulate(wfs_solver* h) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->accumulate();
}

int wfs_reset_sums(wfs_solver* h) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->reset_sums();
}

int wfs_get_sums(wfs_solver* h, float* ux, float* uy, float* uz, float* rho, float* sp, float* uxx, int32_t* navg) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    void* const out[6] = {ux, uy, uz, rho, sp, uxx};
    rc = h->get_sums(out, false);
    if (rc) return rc;
    if (navg) *navg = h->navg;
    return WFS_OK;
}

int wfs_get_sums_f64(wfs_solver* h, double* ux, double* uy, double* uz, double* rho, double* sp, double* uxx,
                     int32_t* navg) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    void* const out[6] = {ux, uy, uz, rho, sp, uxx};
    rc = h->get_sums(out, true);
    if (rc) return rc;
    if (navg) *navg = h->navg;
    return WFS_OK;
}

int wfs_profile_x(wfs_solver* h, int32_t field, int32_t z0, int32_t z1, double* out) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!out) return fail(WFS_EINVAL, "wfs_profile_x: null output");
    DeviceGuard g(h->device);
    return h->profile_x(field, z0, z1, out);
}

int wfs_total(wfs_solver* h, int32_t field, double* out) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!out) return fail(WFS_EINVAL, "wfs_total: null output");
    DeviceGuard g(h->device);
    return h->total(field, out);
}

int wfs_check_finite(wfs_solver* h) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->check_finite();
}

int wfs_timing_enable(wfs_solver* h, int32_t on) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    DeviceGuard g(h->device);
    if ((rc = h->harvest())) return rc;
    h->timed_ms = 0.0;
    h->timed_launches = 0;
    h->timing = on != 0;
    return WFS_OK;
}

int wfs_timing_read(wfs_solver* h, double* kernel_ms, int64_t* launches) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!kernel_ms || !launches) return fail(WFS_EINVAL, "wfs_timing_read: null output");
    DeviceGuard g(h->device);
    WFS_HIP(hipStreamSynchronize(h->stream));
    if ((rc = h->harvest())) return rc;
    *kernel_ms = h->timed_ms;
    *launches = h->timed_launches;
    h->timed_ms = 0.0;
    h->timed_launches = 0;
    return WFS_OK;
}

struct wfs_batch {
    std::unique_ptr<BatchBase> impl;
};

int wfs_batch_create(wfs_solver* const* cases, int32_t n, wfs_batch** out) {
    if (!cases || !out || n < 1) return fail(WFS_EINVAL, "wfs_batch_create: need at least one case");
    if (n > 65535) return fail(WFS_EINVAL, "wfs_batch_create: at most 65535 cases");
    *out = nullptr;
    for (int k = 0; k < n; ++k)
        if (!cases[k]) return fail(WFS_EINVAL, "wfs_batch_create: null case handle");
    const int storage = cases[0]->d.storage;
    std::unique_ptr<BatchBase> b;
    switch (storage) {
        case WFS_STORE_F32: b.reset(new Batch<wfs::CodecF32>()); break;
        case WFS_STORE_F16DEV: b.reset(new Batch<wfs::CodecF16Dev>()); break;
        case WFS_STORE_BF16DEV: b.reset(new Batch<wfs::CodecBF16Dev>()); break;
        case WFS_STORE_F64: b.reset(new Batch<wfs::CodecF64>()); break;
        case WFS_STORE_F32_FAST: b.reset(new Batch<wfs::CodecF32Fast>()); break;
        default: return fail(WFS_EINVAL, "wfs_batch_create: unknown storage");
    }
    b->device = cases[0]->device;
    b->storage = storage;
    DeviceGuard g(b->device);
    WFS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    int rc = 0;
    switch (storage) {
        case WFS_STORE_F32: rc = static_cast<Batch<wfs::CodecF32>*>(b.get())->init(cases, n); break;
        case WFS_STORE_F16DEV: rc = static_cast<Batch<wfs::CodecF16Dev>*>(b.get())->init(cases, n); break;
        case WFS_STORE_BF16DEV: rc = static_cast<Batch<wfs::CodecBF16Dev>*>(b.get())->init(cases, n); break;
        case WFS_STORE_F64: rc = static_cast<Batch<wfs::CodecF64>*>(b.get())->init(cases, n); break;
        case WFS_STORE_F32_FAST: rc = static_cast<Batch<wfs::CodecF32Fast>*>(b.get())->init(cases, n); break;
    }
    if (rc) return rc;
    *out = new wfs_batch{std::move(b)};
    return WFS_OK;
}

int wfs_batch_destroy(wfs_batch* b) {
    if (!b) return WFS_OK;
    DeviceGuard g(b->impl->device);
    (void)hipStreamSynchronize(b->impl->stream);
    delete b;
    return WFS_OK;
}

int wfs_batch_set_schedule(wfs_batch* b, int32_t k, double dt, double t_ramp, double t_avg, double t0) {
    if (!b) return fail(WFS_EINVAL, "null batch");
    return b->impl->set_schedule(k, dt, t_ramp, t_avg, t0);
}

int wfs_batch_set_budget(wfs_batch* b, int32_t k, int64_t steps) {
    if (!b) return fail(WFS_EINVAL, "null batch");
    return b->impl->set_budget(k, (long long)steps);
}

int wfs_batch_budget(wfs_batch* b, int32_t k, int64_t* steps_left) {
    if (!b || !steps_left) return fail(WFS_EINVAL, "wfs_batch_budget: null argument");
    long long v = 0;
    int rc = b->impl->budget_of(k, &v);
    *steps_left = v;
    return rc;
}

int wfs_batch_run(wfs_batch* b, int32_t nsteps, int32_t use_graph) {
    if (!b) return fail(WFS_EINVAL, "null batch");
    DeviceGuard g(b->impl->device);
    return b->impl->run(nsteps, use_graph);
}

int wfs_batch_synchronize(wfs_batch* b) {
    if (!b) return fail(WFS_EINVAL, "null batch");
    DeviceGuard g(b->impl->device);
    return b->impl->synchronize();
}

int wfs_batch_time(wfs_batch* b, int32_t k, double* t, int32_t* navg) {
    if (!b || !t || !navg) return fail(WFS_EINVAL, "wfs_batch_time: null argument");
    return b->impl->time_of(k, t, navg);
}

int wfs_stream_probe(int32_t device, int32_t px, int32_t ny, int32_t nz, int32_t elem_bytes, int64_t stagger_bytes,
                     int32_t iters, double* gbs) {
    if (!gbs || px < 64 || ny < 3 || nz < 3 || iters < 1 || (elem_bytes != 2 && elem_bytes != 4) ||
        stagger_bytes < 0 || stagger_bytes % elem_bytes)
        return fail(WFS_EINVAL, "wfs_stream_probe: bad arguments");
    const long long n = (long long)px * ny * nz;
    const long long stride = n + stagger_bytes / elem_bytes;
    if ((unsigned long long)Q * (unsigned long long)stride * (unsigned long long)elem_bytes >= (1ull << 36))
        return fail(WFS_EINVAL, "wfs_stream_probe: box too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return fail(WFS_ENODEV, "wfs_stream_probe: no such device");
    }
    DeviceGuard g(device);
    const size_t bytes = (size_t)Q * (size_t)stride * (size_t)elem_bytes;
    DevBuf a, b;
    if (hipMalloc(&a.p, bytes) != hipSuccess || hipMalloc(&b.p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(WFS_ENOMEM, "wfs_stream_probe: allocation failed");
    }
    hipStream_t st;
    WFS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    hipLaunchKernelGGL(k_fill_zero_u32, dim3(blocks_for((long long)(bytes / 4), 256)), dim3(256), 0, st,
                       static_cast<uint32_t*>(a.p), (long long)(bytes / 4));
    WFS_LAUNCH_CHECK();
    const long long plane = (long long)px * ny;
    auto launch = [&]() {
        if (elem_bytes == 4)
            hipLaunchKernelGGL(k_streams19<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                               static_cast<const float*>(a.p), static_cast<float*>(b.p), n, stride, px, plane);
        else
            hipLaunchKernelGGL(k_streams19<unsigned short>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                               static_cast<const unsigned short*>(a.p), static_cast<unsigned short*>(b.p), n, stride,
                               px, plane);
    };
    launch();   // warm-up
    WFS_LAUNCH_CHECK();
    hipEvent_t e0, e1;
    WFS_HIP(hipEventCreate(&e0));
    WFS_HIP(hipEventCreate(&e1));
    WFS_HIP(hipEventRecord(e0, st));
    for (int k = 0; k < iters; ++k) {
        launch();
        WFS_LAUNCH_CHECK();
    }
    WFS_HIP(hipEventRecord(e1, st));
    WFS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    WFS_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbs = 2.0 * Q * (double)elem_bytes * (double)n * iters / (ms * 1e-3) / 1e9;
    return WFS_OK;
}

int wfs_bandwidth_probe(int32_t device, int64_t bytes, int32_t iters, double* gbs) {
    if (!gbs || bytes < 16 * 1024 || iters < 1) return fail(WFS_EINVAL, "wfs_bandwidth_probe: bad arguments");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return fail(WFS_ENODEV, "wfs_bandwidth_probe: no such device");
    }
    DeviceGuard g(device);
    const long long n = bytes / 16;
    DevBuf a, b;
    if (hipMalloc(&a.p, (size_t)n * 16) != hipSuccess || hipMalloc(&b.p, (size_t)n * 16) != hipSuccess) {
        (void)hipGetLastError();
        return fail(WFS_ENOMEM, "wfs_bandwidth_probe: allocation failed");
    }
    hipStream_t st;
    WFS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    const unsigned grid = (unsigned)((n + 1023) / 1024);
    hipLaunchKernelGGL(k_fill_zero_u32, dim3(blocks_for(n * 4, 256)), dim3(256), 0, st, static_cast<uint32_t*>(a.p),
                       n * 4);
    WFS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, static_cast<const float4*>(a.p),
                       static_cast<float4*>(b.p), n);  // warm-up
    WFS_LAUNCH_CHECK();
    hipEvent_t e0, e1;
    WFS_HIP(hipEventCreate(&e0));
    WFS_HIP(hipEventCreate(&e1));
    WFS_HIP(hipEventRecord(e0, st));
    for (int k = 0; k < iters; ++k) {
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, static_cast<const float4*>(a.p),
                           static_cast<float4*>(b.p), n);
        WFS_LAUNCH_CHECK();
    }
    WFS_HIP(hipEventRecord(e1, st));
    WFS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    WFS_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbs = 2.0 * (double)n * 16 * iters / (ms * 1e-3) / 1e9;
    return WFS_OK;
}

int64_t wfs_device_bytes(const wfs_solver* h) { return h ? (int64_t)h->bytes : -1; }

// ---------------------------------------------------------------- learned closure

int wfs_set_closure(wfs_solver* h, int32_t kind, const float* params, const uint16_t* geo, const uint8_t* hcol,
                    double u0) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->set_closure(kind, params, geo, hcol, u0);
}

int wfs_particles_set(wfs_solver* h, int64_t n, const double* xyz, int64_t ninlet, const int32_t* inlet_rows,
                      uint64_t seed) {
    if (!h || (n > 0 && !xyz) || (ninlet > 0 && !inlet_rows)) return fail(WFS_EINVAL, "wfs_particles_set: null argument");
    DeviceGuard g(h->device);
    return h->particles_set(n, xyz, ninlet, inlet_rows, seed);
}

int wfs_particles_move(wfs_solver* h, double dt, double vel, double dx, int64_t* respawned) {
    if (!h) return fail(WFS_EINVAL, "wfs_particles_move: null handle");
    DeviceGuard g(h->device);
    long long r = 0;
    const int rc = h->particles_move(dt, vel, dx, &r);
    if (respawned) *respawned = r;
    return rc;
}

int wfs_particles_get(wfs_solver* h, double* xyz) {
    if (!h || !xyz) return fail(WFS_EINVAL, "wfs_particles_get: null argument");
    DeviceGuard g(h->device);
    return h->particles_get(xyz);
}

int wfs_set_column_u(wfs_solver* h, int32_t per_step) {
    if (!h) return fail(WFS_EINVAL, "wfs_set_column_u: null handle");
    DeviceGuard g(h->device);
    return h->set_column_u(per_step);
}

int wfs_set_pop_stagger(wfs_solver* h, int64_t bytes) {
    if (!h) return fail(WFS_EINVAL, "wfs_set_pop_stagger: null handle");
    DeviceGuard g(h->device);
    return h->set_pop_stagger((long long)bytes);
}

int64_t wfs_pop_stride(const wfs_solver* h) { return h ? (int64_t)h->Ns : -1; }

int wfs_set_inplace(wfs_solver* h, int32_t on) {
    if (!h) return fail(WFS_EINVAL, "wfs_set_inplace: null handle");
    DeviceGuard g(h->device);
    return h->set_inplace(on);
}

int wfs_state_bytes(wfs_solver* h, int64_t* bytes) {
    if (!h || !bytes) return fail(WFS_EINVAL, "wfs_state_bytes: null argument");
    long long n = 0;
    int rc = h->state_bytes(&n);
    *bytes = n;
    return rc;
}

int wfs_state_save(wfs_solver* h, void* out, int64_t bytes) {
    if (!h || !out) return fail(WFS_EINVAL, "wfs_state_save: null argument");
    DeviceGuard g(h->device);
    return h->state_save(out, bytes);
}

int wfs_state_load(wfs_solver* h, const void* in, int64_t bytes) {
    if (!h || !in) return fail(WFS_EINVAL, "wfs_state_load: null argument");
    int rc = check_not_batched(h, "wfs_state_load");
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->state_load(in, bytes);
}

int wfs_forces_noalloc(wfs_solver* h, double* out) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!out) return fail(WFS_EINVAL, "wfs_forces_noalloc: null output");
    if (h->Np >= kNoAllocMaxCells)
        return fail(WFS_EINVAL, "wfs_forces_noalloc: more than 2^31 - 1 stored cells (" + std::to_string(h->Np) + ")");
    DeviceGuard g(h->device);
    return h->forces_noalloc(out);
}

int wfs_set_body_mask(wfs_solver* h, const uint8_t* body) {
    int rc = check_handle(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->set_body_mask(body);
}

int wfs_forces(wfs_solver* h, double* out) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!out) return fail(WFS_EINVAL, "wfs_forces: null output");
    DeviceGuard g(h->device);
    return h->forces(out);
}

int wfs_set_closure_field(wfs_solver* h, int32_t on) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    DeviceGuard g(h->device);
    return h->set_closure_field(on);
}

int wfs_get_closure_field(wfs_solver* h, float* c2) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!c2) return fail(WFS_EINVAL, "wfs_get_closure_field: null output");
    DeviceGuard g(h->device);
    return h->get_closure_field(c2);
}

int wfs_closure_samples(wfs_solver* fine, const uint16_t* geo, const uint8_t* hcol, int32_t ncx, int32_t ncy,
                        int32_t ncz, double kv, double u0, float* X, float* ABP, uint8_t* valid) {
    int rc = check_handle(fine);
    if (rc) return rc;
    DeviceGuard g(fine->device);
    WFS_HIP(hipStreamSynchronize(fine->stream));
    return fine->closure_samples(geo, hcol, ncx, ncy, ncz, kv, u0, X, ABP, valid);
}

int64_t wfs_mlp_work_floats(int64_t n) {
    if (n < 0) return -1;
    const int64_t blocks = (n + kGradBlock - 1) / kGradBlock;
    const int64_t grad = blocks * kPartStride;
    const int64_t pack = kPackWords;
    return grad > pack ? grad : pack;
}

static int mlp_device(int32_t device) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return fail(WFS_ENODEV, "no such device");
    }
    return WFS_OK;
}

int wfs_mlp_eval(int32_t device, const float* X, int64_t n, const float* params, float* c2, float* work,
                 int64_t work_floats, void* stream) {
    int rc = mlp_device(device);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!X || !c2)) || !params || !work) return fail(WFS_EINVAL, "wfs_mlp_eval: bad arguments");
    if (work_floats < kPackWords) return fail(WFS_EINVAL, "wfs_mlp_eval: work buffer too small");
    DeviceGuard g(device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(256), 0, st, params, reinterpret_cast<MlpPack*>(work));
    WFS_LAUNCH_CHECK();
    if (n == 0) return WFS_OK;
    hipLaunchKernelGGL(k_mlp_eval, dim3(blocks_for(n, 256)), dim3(256), 0, st, X, (long long)n,
                       reinterpret_cast<const MlpPack*>(work), c2);
    WFS_LAUNCH_CHECK();
    return WFS_OK;
}

int wfs_mlp_grad(int32_t device, const float* X, const float* ABP, int64_t n, const float* params, float* grad,
                 double* sums, float* work, int64_t work_floats, void* stream) {
    int rc = mlp_device(device);
    if (rc) return rc;
    if (n < 1 || !X || !ABP || !params || !grad || !sums || !work) return fail(WFS_EINVAL, "wfs_mlp_grad: bad arguments");
    if (work_floats < wfs_mlp_work_floats(n)) return fail(WFS_EINVAL, "wfs_mlp_grad: work buffer too small");
    const long long blocks = (n + kGradBlock - 1) / kGradBlock;
    if (blocks > (1ll << 30)) return fail(WFS_EINVAL, "wfs_mlp_grad: batch too large");
    DeviceGuard g(device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)blocks), dim3(kGradBlock), 0, st, X, ABP, (long long)n, params, work);
    WFS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mlp_reduce, dim3(1), dim3(256), 0, st, work, (int)blocks, (long long)n, grad, sums);
    WFS_LAUNCH_CHECK();
    return WFS_OK;
}

int wfs_mlp_adam(int32_t device, float* params, const float* grad, float* m, float* v, int64_t t, double lr,
                 double beta1, double beta2, double eps, void* stream) {
    int rc = mlp_device(device);
    if (rc) return rc;
    if (!params || !grad || !m || !v || t < 1 || !(lr > 0.0)) return fail(WFS_EINVAL, "wfs_mlp_adam: bad arguments");
    DeviceGuard g(device);
    const float c1 = (float)(1.0 / (1.0 - std::pow(beta1, (double)t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(beta2, (double)t)));
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(kNParams, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), params,
                       grad, m, v, kNParams, (float)lr, (float)beta1, (float)beta2, (float)eps, c1, c2);
    WFS_LAUNCH_CHECK();
    return WFS_OK;
}

}  // extern "C"
