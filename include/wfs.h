/* C-ABI of the MI355X river-reach lattice-Boltzmann solver (libwfs.so).
 *
 * One opaque handle = one solver instance = the state of the reference's
 * `LBM3D` object (floodsim/lbm3d.py:333-418) held in GPU memory, plus the
 * device-side pieces of its driver `Run3D` (floodsim/reach3d.py:378-453).
 * The reference has no FFI of its own: its `LBM3D` methods launch CuPy
 * kernels directly. Each entry point below replaces the reference code named
 * in its comment; the Python class water_flow_simulation_amd.lbm3d.LBM3D
 * binds them with ctypes (see INTEGRATION.md for the binding).
 *
 * Conventions
 *  - Arrays are flat, C order, cell = x + nx*(y + ny*z) (lbm3d.py:9);
 *    per-row arrays (uin) are indexed j = y + ny*z.
 *  - Every function returns 0 on success or a negative WFS_E* code;
 *    wfs_last_error() gives the message of the calling thread's last failure.
 *  - The caller owns host buffers: inputs are copied in, outputs copied out.
 *  - Work is queued on the handle's own HIP stream; wfs_step() returns before
 *    the GPU has finished, every wfs_get_* call waits for it.
 *  - One handle must not be used from two threads at once; distinct handles
 *    are independent.
 */
#ifndef WFS_H
#define WFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WFS_ABI_VERSION 1

enum wfs_status {
    WFS_OK = 0,
    WFS_EINVAL = -1,   /* bad argument (maps to TypeError/ValueError) */
    WFS_ENOMEM = -2,   /* device allocation failed */
    WFS_EHIP = -3,     /* HIP runtime error */
    WFS_ENODEV = -4,   /* no usable GPU */
    WFS_ENONFINITE = -5 /* watchdog: non-finite state found */
};

/* Population storage. F32 is the reference's layout and precision, with the
 * reference's operation order (bit-exact against an IEEE FP32 evaluation);
 * the two 16-bit formats store f_i - e_i, e_i the equilibrium of the
 * mean inlet velocity given at wfs_upload, and compute in FP32; F64
 * stores and computes in double (the FP64 run of the same path); F32_FAST
 * stores FP32 and uses the same opposite-pair collision as the 16-bit
 * formats (agreement within rounding, not bit for bit). */
enum wfs_storage { WFS_STORE_F32 = 0, WFS_STORE_F16DEV = 1, WFS_STORE_BF16DEV = 2, WFS_STORE_F64 = 3,
                   WFS_STORE_F32_FAST = 4 };

typedef struct wfs_desc {
    int32_t nx, ny, nz;   /* lattice size; nx >= 3, nx*ny*nz < 2^31 */
    double tau0;          /* molecular relaxation time (reach3d.py:383-385); FP32/16-bit storage rounds it to float */
    double cs;            /* Smagorinsky constant; Cs^2 = cs*cs, rounded to float for FP32/16-bit storage */
    int32_t storage;      /* enum wfs_storage */
    int32_t block_x;      /* work-group size along x (64..256, multiple of 64); 0 = default */
    int32_t device;       /* HIP device ordinal */
} wfs_desc;

typedef struct wfs_solver wfs_solver;

/* Version of this ABI (WFS_ABI_VERSION). */
int wfs_abi_version(void);

/* Number of visible HIP devices (0 if none). */
int wfs_device_count(int32_t* count);

/* Message of the calling thread's last failure ("" if none). */
const char* wfs_last_error(void);

/* Creates a solver (replaces LBM3D.__init__ allocation, lbm3d.py:342-366). */
int wfs_create(const wfs_desc* desc, wfs_solver** out);

/* Work-group width of the step kernel (64, 128, 192 or 256 cells along x;
 * the x-tile of one work-group). Takes effect at the next wfs_step; batches
 * keep the width their cases had when the batch was made. No counterpart in
 * the reference (its launch shape is fixed, lbm3d.py:401-404); used by the
 * launch-shape autotuner. */
int wfs_set_block_x(wfs_solver* h, int32_t block_x);
int32_t wfs_block_x(const wfs_solver* h);

/* Creates a solver on the tile-sparse layout: a tile is one 64-cell x-run of
 * a row (y, z); only tiles holding a fluid cell of the dense solid mask
 * (u8, nx*ny*nz) are stored, so a meandering reach in a wide bounding box
 * keeps only its channel. Same physics and the same results as wfs_create
 * (bit for bit); wfs_upload must then be given the same solid mask. The work
 * of one step is one wavefront per stored tile. Requires (nx - 1) % 64 != 0.
 * The learned closure and wfs_forces work on this layout; training samples
 * and batches need the dense one. No reference counterpart (the reference's
 * reaches are straight prisms, reach3d.py:213-216). */
int wfs_create_tiled(const wfs_desc* desc, const uint8_t* solid, wfs_solver** out);
/* Cells held on the device (padded rows, or stored tiles x 64). */
int64_t wfs_stored_cells(const wfs_solver* h);

/* Body forces on a list of n dense cell indices (lattice units); every other
 * cell gets zero. Replaces wfs_set_force's dense arrays when only actuator
 * disks carry force. */
int wfs_set_force_cells(wfs_solver* h, int64_t n, const int64_t* cells, const float* fx, const float* fy,
                        const float* fz);

/* Frees all device memory of the solver. NULL is accepted. */
int wfs_destroy(wfs_solver* h);

/* Uploads the geometry and initial state (lbm3d.py:352-364): solid mask
 * (u8, N), wall friction coefficients cf (f32, N; NULL = all zero), inlet
 * velocity uin (f32, ny*nz, lattice units). Populations are set to the
 * equilibrium at rho = 1, u = (uin broadcast along x, 0, 0), zero velocity in
 * solid cells; rho = 1, u = 0 for the macroscopic fields; uref = 0; forces 0. */
int wfs_upload(wfs_solver* h, const uint8_t* solid, const float* cf, const float* uin);

/* Replaces the body-force fields (lattice units, f32, N each) (lbm3d.py:386-388,
 * and the caller reassignment of Fx/Fy/Fz in reach3d.py:434-436). */
int wfs_set_force(wfs_solver* h, const float* Fx, const float* Fy, const float* Fz);

/* Scales the forces used by the step: F_used = F * scale in FP32, i.e. the
 * thrust ramp F = F0 * s of reach3d.py:431-437 without re-uploading F. */
int wfs_set_force_scale(wfs_solver* h, float scale);

/* Replaces the inlet velocity profile (f32, ny*nz). */
int wfs_set_uin(wfs_solver* h, const float* uin);

/* Replaces the wall-model reference speed (f32, N) (reach3d.py:418 assignment). */
int wfs_set_uref(wfs_solver* h, const float* uref);

/* Sets the bed/bank wall mask used by wfs_update_uref (u8, N). */
int wfs_set_bedwall(wfs_solver* h, const uint8_t* bedwall);

/* Run3D._update_uref (reach3d.py:412-418) on the device: uref = depth-averaged
 * column speed |(sum_z ux, sum_z uy)| / n_fluid on bed/bank cells, 0 elsewhere. */
int wfs_update_uref(wfs_solver* h);

/* Per-step column velocity for the wall model. per_step != 0: every step
 * computes the same column speed as wfs_update_uref from the previous step's
 * velocity inside the step kernel (each step writes its ux, uy; each bed/bank
 * wall-model cell sums its column), so uref is never older than one step;
 * wfs_update_uref and wfs_set_uref then have no effect on the step. Starts from
 * the current velocity field (call after wfs_step). Needs wfs_set_bedwall and
 * at most 255 cells per column. per_step == 0: back to the run cadence.
 * No reference counterpart: the reference refreshes uref once per 20-step
 * chunk (reach3d.py:427-440); with a cadence of one step the two agree. */
int wfs_set_column_u(wfs_solver* h, int32_t per_step);

/* In-place streaming. on != 0: the step keeps one population buffer instead
 * of two (the second is freed) and alternates two kernels: an odd step pulls
 * from the neighbours' exchanged slots and pushes its results to where the
 * next step reads them; an even step reads and writes the cell's own slots.
 * Two steps equal two steps of the two-buffer scheme bit for bit (macros,
 * forces and populations). wfs_get_pops, wfs_set_pops, wfs_forces and
 * switching off need an even number of in-place steps since switching on
 * (WFS_EINVAL otherwise); in wfs_get_pops a population leaving the box
 * through an x-end row (x = 0 or nx - 1), which no step reads, reads as its
 * opposite. Called between wfs_create and wfs_upload, the second buffer is
 * never allocated. Both layouts and both closures; nx - 1 not a multiple of
 * 64; not for a case of a batch (WFS_EINVAL while a batch holds the solver).
 * No reference counterpart: the reference streams between two
 * arrays (floodsim/lbm3d.py:50-60); this halves the population memory. */
int wfs_set_inplace(wfs_solver* h, int32_t on);

/* Population layout: each of the 19 direction arrays starts `bytes` after
 * the end of the previous one (a multiple of 128 bytes, at most 1 MiB; 0,
 * the default, packs them). The 19-stream access of the step reaches
 * different HBM rates for different offsets between its streams (the
 * address interleave decides; 5.3-6.2 TB/s on one box), so a solver can
 * take the stagger an autotuner found fastest for its lattice. Between
 * wfs_create and wfs_upload only (WFS_EINVAL after): the population buffer
 * is allocated again. Results do not depend on it; checkpoints hold the
 * populations without the gaps and load across staggers.
 * No reference counterpart: the reference keeps f[i*N + c]
 * (floodsim/lbm3d.py:38-60). */
int wfs_set_pop_stagger(wfs_solver* h, int64_t bytes);

/* Population direction stride in cells (stored cells plus the stagger), or
 * -1 for a null handle. No reference counterpart. */
int64_t wfs_pop_stride(const wfs_solver* h);

/* Checkpoints: the raw solver state as one host buffer of wfs_state_bytes
 * bytes. It holds the population buffer in use (device layout and storage
 * format as they are, the in-place phase included), the macroscopic fields,
 * uref, the six time-mean sums, the per-step column buffers when on, the step
 * count, the sample count and the force scale. wfs_state_load restores it
 * into a solver created and uploaded the same way (dimensions, storage,
 * layout, streaming mode, column-speed mode, and the same geometry: the
 * state carries a fingerprint of the solid mask and the wall coefficients
 * given to wfs_upload; WFS_EINVAL otherwise), and the run then continues
 * bit for bit. wfs_state_load, like wfs_step, wfs_set_macro and
 * wfs_set_populations, is refused while a batch holds the solver. Inputs (geometry, forces, wall cf, inlet
 * profile, closure) come from the solver's own set-up; particles are not
 * part of the state. No reference counterpart: the reference never
 * checkpoints a run (it caches results, web3d.py:187-206). */
int wfs_state_bytes(wfs_solver* h, int64_t* bytes);
int wfs_state_save(wfs_solver* h, void* out, int64_t bytes);
int wfs_state_load(wfs_solver* h, const void* in, int64_t bytes);

/* Particle tracer (Run3D.init_particles / move_particles / particles_xyz,
 * reach3d.py:500-541). wfs_particles_set uploads n positions (f64, x y z per
 * particle, metres from the domain corner) and the rows y + ny*z of the inlet
 * cells particles restart in (the fluid cells of the plane x = 1). Each
 * wfs_particles_move advances every particle by float32(float32(u * vel) * dt)
 * with u the lattice velocity of the cell holding it (vel: m/s per lattice
 * unit, dt in seconds, dx the cell size in metres) and restarts the particles
 * that left the box or entered a solid cell at a random inlet cell (a hash
 * of seed, particle and move count); *respawned gets their number. Positions
 * stay on the device until wfs_particles_get. */
int wfs_particles_set(wfs_solver* h, int64_t n, const double* xyz, int64_t ninlet, const int32_t* inlet_rows,
                      uint64_t seed);
int wfs_particles_move(wfs_solver* h, double dt, double vel, double dx, int64_t* respawned);
int wfs_particles_get(wfs_solver* h, double* xyz);

/* n solver steps (lbm3d.py:390-412): fused collide-stream, inlet/outlet rows,
 * buffer swap. Asynchronous. The macroscopic fields hold the last step's
 * values for every cell when the call's work completes. */
int wfs_step(wfs_solver* h, int32_t n);

/* Waits for all queued work of the handle. */
int wfs_synchronize(wfs_solver* h);

/* Number of steps taken (LBM3D.nstep). */
int64_t wfs_nstep(const wfs_solver* h);

/* Copies the macroscopic fields of the last step (f32, N each; NULL skips). */
int wfs_get_macro(wfs_solver* h, float* rho, float* ux, float* uy, float* uz);
int wfs_get_macro_f64(wfs_solver* h, double* rho, double* ux, double* uy, double* uz);

/* Overwrites the macroscopic fields that the next step's wall model reads
 * as the previous step's values (f32, N each). */
int wfs_set_macro(wfs_solver* h, const float* rho, const float* ux, const float* uy, const float* uz);

/* Copies uref (f32, N). */
int wfs_get_uref(wfs_solver* h, float* uref);

/* Populations f[i*N + cell] decoded to f32 / f64 (19*N values), and the reverse. */
int wfs_get_populations(wfs_solver* h, float* f);
int wfs_get_populations_f64(wfs_solver* h, double* f);
int wfs_set_populations(wfs_solver* h, const float* f);

/* Run3D time-mean accumulation (reach3d.py:443-451): adds ux, uy, uz, rho,
 * |u| and ux^2 of the current state to running sums (storage precision).
 * Unlike the reference, the density enters as rho - 1: a float sum of
 * rho ~ 1 loses the level signal (rho - 1 ~ 1e-5..1e-4) once its ulp
 * exceeds it (about 40 % of it lost after 3000 samples, 90 % after 11000),
 * while a float sum of rho - 1 keeps it to its own relative precision. */
int wfs_accumulate(wfs_solver* h);
int wfs_reset_sums(wfs_solver* h);
/* Copies the six running sums (f32, N each; NULL skips) and the sample count.
 * The rho sum is returned as the sum of rho, n + sum(rho - 1) formed in
 * double and rounded once. */
int wfs_get_sums(wfs_solver* h, float* ux, float* uy, float* uz, float* rho, float* sp, float* uxx,
                 int32_t* navg);

int wfs_get_sums_f64(wfs_solver* h, double* ux, double* uy, double* uz, double* rho, double* sp, double* uxx,
                     int32_t* navg);

/* FP64 per-x reductions over fluid cells with z0 <= z < z1:
 * out[x] = sum_{y, z} field(x, y, z) for field in enum wfs_field
 * (WFS_F_SUM_RHO is the device sum of rho - 1). */
enum wfs_field {
    WFS_F_RHO = 0, WFS_F_UX = 1, WFS_F_UY = 2, WFS_F_UZ = 3,
    WFS_F_SUM_UX = 4, WFS_F_SUM_UY = 5, WFS_F_SUM_UZ = 6, WFS_F_SUM_RHO = 7, WFS_F_SUM_SP = 8, WFS_F_SUM_UXX = 9,
    WFS_F_ONE = 10
};
int wfs_profile_x(wfs_solver* h, int32_t field, int32_t z0, int32_t z1, double* out);

/* FP64 total over fluid cells of a field (e.g. mass = total of WFS_F_RHO). */
int wfs_total(wfs_solver* h, int32_t field, double* out);

/* FP64 momentum budget of the next step, lattice units (momentum per step):
 * out[0..2] the applied body force (thrust, F * scale summed over fluid
 * cells), out[3..5] the force of the fluid on the cells marked by
 * wfs_set_body_mask (hulls), out[6..8] the force on all other walls (bed,
 * banks, domain edges). Wall forces come from momentum exchange on the
 * bounce-back links, including the wall function's slip correction. Together
 * with the inlet/outlet fluxes these close the reach's momentum balance, the
 * budget that compare() (floodsim/reach3d.py:571-618) estimates from profiles.
 * No reference counterpart (the reference reduces nothing on the device). */
int wfs_set_body_mask(wfs_solver* h, const uint8_t* body);
int wfs_forces(wfs_solver* h, double* out);

/* wfs_forces without any device allocation: the same nine sums, through a
 * fixed scratch array in the library (147,528 bytes per GPU, held until the
 * process ends) instead of partial-sum buffers allocated per call. No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The result equals wfs_forces up to the
 * order of the FP64 summation (a relative difference of order 1e-15 of the
 * largest component). The scratch does not grow with the lattice: any solver
 * up to the library's 2^31 - 1 stored cells is accepted (WFS_EINVAL
 * above). Every call on one GPU uses that GPU's one scratch array: calls
 * that overlap on one GPU (from two threads) race and must be serialised by
 * the caller. Blocks until the sums are on the host. */
int wfs_forces_noalloc(wfs_solver* h, double* out);

/* Watchdog: WFS_ENONFINITE if any fluid cell's rho or u is NaN/Inf. */
int wfs_check_finite(wfs_solver* h);

/* Kernel timing: while enabled, every launch of the fused step kernel is
 * bracketed by HIP events on the handle's stream. wfs_timing_read waits for
 * the stream and returns the summed kernel time (ms) and the number of
 * launches since the last enable/read, then starts a new tally. */
int wfs_timing_enable(wfs_solver* h, int32_t on);
int wfs_timing_read(wfs_solver* h, double* kernel_ms, int64_t* launches);

/* ---- Batches: many reaches stepped together (the boat-test sweep).
 * A batch references solvers that were created and uploaded as usual (same
 * storage format, block_x, device and streaming mode: all two-buffer or all
 * in place, wfs_set_inplace). Each run step is one launch over every
 * case; the run driver's 20-step cadence of floodsim/reach3d.py:424-453
 * (thrust ramp, uref refresh, time-mean sums once t >= t_avg) runs on the
 * device from per-case schedules, and with use_graph a 20-step chunk is
 * captured once as a HIP graph and replayed. Replaces the sequential per-case
 * loop of floodsim/boat_tests.py:79-110 (run_test). While a batch runs, its
 * cases must not be used directly; wfs_batch_synchronize() hands them back
 * with nstep, the sample count and the force scale brought up to date. */
typedef struct wfs_batch wfs_batch;
int wfs_batch_create(wfs_solver* const* cases, int32_t n, wfs_batch** out);
int wfs_batch_destroy(wfs_batch* b);
/* Case k: time step dt (s), ramp length t_ramp (s; <= 0: no ramp), start of
 * the averaging window t_avg (s), and the case's current time t0 (s). */
int wfs_batch_set_schedule(wfs_batch* b, int32_t k, double dt, double t_ramp, double t_avg, double t0);
/* Step budget of case k, counted from the next wfs_batch_run (steps < 0: no
 * limit, the default). A case takes steps until its budget is spent and then
 * sits out the rest of the batch's launches: inside a chunk it stops after
 * its last step with the cadence of a sequential chunk of that length (the
 * time, the averaging sample and the fields of a run of exactly that many
 * steps), so cases with different end times (floodsim/reach3d.py:455-467,
 * run_to_end of each case of floodsim/boat_tests.py:97-103) share one
 * batched run. wfs_batch_budget reads the budget left as of the last
 * synchronize. */
int wfs_batch_set_budget(wfs_batch* b, int32_t k, int64_t steps);
int wfs_batch_budget(wfs_batch* b, int32_t k, int64_t* steps_left);
int wfs_batch_run(wfs_batch* b, int32_t nsteps, int32_t use_graph);
int wfs_batch_synchronize(wfs_batch* b);
/* Case k's time and sample count as of the last synchronize. */
int wfs_batch_time(wfs_batch* b, int32_t k, double* t, int32_t* navg);

/* Device copy bandwidth: copies `bytes` bytes between two fresh buffers
 * `iters` times with a grid-stride 16-byte copy kernel and reports
 * 2*bytes*iters / elapsed in GB/s (1e9 bytes/s). The practical HBM ceiling
 * that the step kernel's traffic is compared with; also the known-size copy
 * used to calibrate hardware byte counters. */
int wfs_bandwidth_probe(int32_t device, int64_t bytes, int32_t iters, double* gbs);

/* Streaming ceiling of the lattice step's access pattern: the 19-population
 * pull read and 19-population write of one step over a px x ny x nz box
 * without the arithmetic (each population read from x - c_i, written with a
 * non-temporal store), populations of elem_bytes (4: FP32, 2: the 16-bit
 * formats), direction arrays stagger_bytes apart beyond their length.
 * Reports 2 * 19 * elem_bytes * cells * iters / elapsed in GB/s: what any
 * pull-stream step kernel with this layout can reach on the device. No
 * reference counterpart (bench.py reports it beside the roofline). */
int wfs_stream_probe(int32_t device, int32_t px, int32_t ny, int32_t nz, int32_t elem_bytes, int64_t stagger_bytes,
                     int32_t iters, double* gbs);

/* Device memory held by the solver, in bytes. */
int64_t wfs_device_bytes(const wfs_solver* h);

/* ---- Learned sub-grid closure.
 * No counterpart in the reference, whose only closure is the constant-Cs
 * Smagorinsky model of floodsim/lbm3d.py:120-123 (kept as the default and the
 * fallback). A network 8 -> 32 -> 32 -> 1 (ReLU, softplus output) predicts per
 * cell the coefficient C^2 = WFS_C2_MIN + WFS_C2_SCALE * softplus(y) that
 * replaces Cs^2 in tau = (tau0 + sqrt(tau0^2 + 18 sqrt(2) C^2 Q / rho)) / 2.
 * Its inputs are strain-rate invariants from the non-equilibrium moments, the
 * velocity relative to U0, the wall distance, the height in the column and
 * the column depth (formulas: water_flow_simulation_amd/csrc/closure.h).
 * Parameters are WFS_MLP_PARAMS floats in the order
 *   W1[32][8] (out, in), b1[32], W2[32][32] (out, in), b2[32], w3[32], b3.
 * Hidden layers run as BF16 MFMA with FP32 accumulation inside the step. */
#define WFS_MLP_FEATURES 8
#define WFS_MLP_HIDDEN 32
#define WFS_MLP_PARAMS 1377
#define WFS_C2_MIN 0.01
#define WFS_C2_SCALE 0.0289
enum wfs_closure { WFS_CLOSURE_SMAGORINSKY = 0, WFS_CLOSURE_MLP = 1 };

/* Selects the closure. kind WFS_CLOSURE_SMAGORINSKY ignores the other
 * arguments. kind WFS_CLOSURE_MLP (not for FP64 storage) takes the weights,
 * the per-cell geometry code geo (nz, ny, nx) uint16: round(255/d) |
 * round(255 z_rel) << 8, the per-column fluid-cell count hcol (ny, nx) uint8,
 * and u0 > 0, the mean inlet speed in lattice units. */
int wfs_set_closure(wfs_solver* h, int32_t kind, const float* params, const uint16_t* geo, const uint8_t* hcol,
                    double u0);
/* C^2 of every cell as used in the last step of the previous wfs_step call
 * (learned closure; dense (nz, ny, nx) float). */
int wfs_get_closure_field(wfs_solver* h, float* c2);
/* on != 0 (the default): the learned step keeps the C^2 field that
 * wfs_get_closure_field reads (4 bytes per stored cell); 0 frees it, for runs
 * that do not read it. No reference counterpart (the reference holds no
 * learned closure). */
int wfs_set_closure_field(wfs_solver* h, int32_t on);

/* Training samples from a fine run (its state after the last wfs_step call):
 * the velocity is box-filtered 2x2x2 onto the coarse grid ncx x ncy x ncz
 * (coarse cell (X,Y,Z) covers fine cells 2X..2X+1, 2Y..2Y+1, 2Z..2Z+1) and
 * scaled by kv into coarse lattice units; geo/hcol (host) describe the coarse
 * grid as in wfs_set_closure and u0 is the coarse mean inlet speed. Outputs
 * are DEVICE arrays in dense coarse order: X (nc x 8) features, ABP (nc x 3)
 * loss coefficients A = |S|^2 S:S, B = |S| (T:S), P = T:T (T the deviatoric
 * subfilter stress, S the resolved strain by central differences), valid (nc)
 * 1 where the cell and its face neighbours are fluid. Blocks until done. */
int wfs_closure_samples(wfs_solver* fine, const uint16_t* geo, const uint8_t* hcol, int32_t ncx, int32_t ncy,
                        int32_t ncz, double kv, double u0, float* X, float* ABP, uint8_t* valid);

/* Training and evaluation kernels. All pointers are DEVICE pointers on
 * `device`; work is a scratch array of at least wfs_mlp_work_floats(n)
 * floats; stream is a hipStream_t (NULL: the null stream); calls are
 * asynchronous on it. */
int64_t wfs_mlp_work_floats(int64_t n);
/* c2[k] = C^2 of the features X[k][0..7] through the in-step MFMA path. */
int wfs_mlp_eval(int32_t device, const float* X, int64_t n, const float* params, float* c2, float* work,
                 int64_t work_floats, void* stream);
/* Loss sum(P + 4 C^2 B + 4 C^4 A) over the n samples and its gradient with
 * respect to params, divided by sum(P): grad (WFS_MLP_PARAMS floats),
 * sums[0] = loss sum, sums[1] = sum P (double). The forward pass rounds like
 * the MFMA path (BF16 inputs, W1, W2 and layer-1 activations); the gradient
 * treats the rounding as the identity. Deterministic. */
int wfs_mlp_grad(int32_t device, const float* X, const float* ABP, int64_t n, const float* params, float* grad,
                 double* sums, float* work, int64_t work_floats, void* stream);
/* One bias-corrected Adam step (t >= 1) on params with moments m, v. */
int wfs_mlp_adam(int32_t device, float* params, const float* grad, float* m, float* v, int64_t t, double lr,
                 double beta1, double beta2, double eps, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WFS_H */
