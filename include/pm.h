/*
 * pm.h -- C ABI of the MI355X particle-mesh gravity engine (libpm.so).
 *
 * Drop-in for the compute path of CosmosSimulationWithCuda's header-only
 * simulator (CosmosCuda.cuh, template struct Universe, cuh:830-1428): load
 * particle arrays, step, copy positions and velocities out.  Plain C types
 * and pointers only; every call returns a status (the reference exits the
 * process on error, cuh:19-27).  One handle = one simulation = one HIP stream
 * on one device.  A handle is not thread-safe; distinct handles may be used
 * from distinct threads.
 *
 * Each entry point names the reference code it replaces ("cuh" =
 * CosmosCuda.cuh).
 */
#ifndef COSMOS_PM_H
#define COSMOS_PM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_sim* pm_handle;

typedef enum pm_status {
    PM_OK = 0,
    PM_EINVAL = 1,      /* bad argument or configuration */
    PM_ENOMEM = 2,      /* device allocation failed */
    PM_EDEVICE = 3,     /* HIP runtime or kernel launch error */
    PM_EBUSY = 4,       /* resource held (serving: budget or slots exhausted) */
    PM_ECANCELLED = 5,  /* request cancelled */
    PM_ENODEV = 6       /* no usable GPU */
} pm_status;

/* mass assignment: the reference's lowAccuracy flag (cuh:881, 890) */
enum { PM_NGP = 0, PM_CIC = 1 };
/* wrap handling at the lattice border (cuh:288-303) */
enum { PM_WRAP_KEEP_VELOCITY = 0, PM_WRAP_REFERENCE = 1 };
/* force lattice: spectral interior + reference stencil on the edge bands, or
 * the reference's stencil (cuh:174-229) on every cell */
enum { PM_FORCE_SPECTRAL = 0, PM_FORCE_STENCIL = 1 };
/* FFT arithmetic: FP32 Stockham stages, or radix-16 stages on the matrix
 * cores in split BF16 with FP32 accumulation (256 <= n_lattice <= 8192):
 * BF16X3 splits each FP32 operand as bf16 hi + lo and forms hi*hi + hi*lo +
 * lo*hi; BF16X6 splits it as hi + mid + lo and forms hi*hi + hi*mid + mid*hi
 * + hi*lo + lo*hi + mid*mid.  The spectra and diagnostics stay FP64. */
enum { PM_FFT_FP32 = 0, PM_FFT_BF16X3 = 1, PM_FFT_BF16X6 = 2 };
/* lattices readable with pm_read_lattice */
enum {
    PM_LATTICE_MASS = 0,      /* int64 fixed point, N*N, index i + j*N       */
    PM_LATTICE_POTENTIAL = 1, /* float, N*N                                   */
    PM_LATTICE_FORCE = 2,     /* float2 (Fx, Fy) interleaved, N*N             */
    PM_LATTICE_TRANSFER = 3   /* float, (N/2+1)*N: T[kx*N+ky] incl. 1/N^2     */
};
/* single stages of one step, in order (cuh:1336-1341) */
enum { PM_STAGE_DEPOSIT = 0, PM_STAGE_FIELD = 1, PM_STAGE_PUSH = 2 };
/* timed kernels (pm_profile_read) */
enum {
    PM_K_CLEAR = 0, PM_K_DEPOSIT = 1, PM_K_ROW_R2C = 2, PM_K_COL_CONV = 3, PM_K_ROW_C2R = 4,
    PM_K_EDGE = 5, PM_K_STENCIL = 6, PM_K_PUSH = 7, PM_K_SORT = 8, PM_K_PUSH_DEPOSIT = 9,
    PM_K_PUSH_DEPOSIT_SORT = 10, PM_K_PUSH_DEPOSIT_COUNT = 11, PM_K_COUNT = 12
};
/* PM_K_PUSH_DEPOSIT: the steady-state step kernel; PM_K_PUSH_DEPOSIT_COUNT: the
 * same kernel on the step that also counts the new bins for the next re-sort;
 * PM_K_PUSH_DEPOSIT_SORT: the step that writes the re-sorted order. */

/* Replaces the compile-time namespace Constants (cuh:29-74, main.cu:5-25) and
 * the ctor arguments (cuh:881) with runtime fields. */
typedef struct pm_config {
    int32_t n_lattice;     /* N: power of two, 64..8192 (Constants::N)                */
    int32_t mode;          /* PM_NGP / PM_CIC                                          */
    int64_t n_particles;   /* capacity in particles                                    */
    float dt;              /* time step (Constants::dt)                                */
    float gravity;         /* force multiplier (Constants::gravityMultiplier)          */
    double green_origin;   /* G at r = 0: 3.0 in the reference (cuh:113-116)           */
    double deconv_eps;     /* regulariser of the deconvolution, 0.001 (cuh:139)        */
    int32_t wrap_mode;     /* PM_WRAP_*                                                */
    int32_t force_mode;    /* PM_FORCE_*                                               */
    int32_t precision;     /* PM_FFT_*                                                 */
    int32_t mass_shift;    /* fixed-point bits (rho = q * 2^-shift); 0 = from the masses */
    int32_t device;        /* HIP device ordinal                                       */
    int32_t resort_interval; /* steps between re-binning the particles by lattice tile;
                                0 = chosen at each load: 9, or 7 with more than 64
                                particles per cell, or 2 for CIC runs of at least
                                3e7 particles with more than 64 per cell; shortened
                                as the particles drift faster (tunable
                                "resort_drift_ppm").
                                Affects speed only, never results.                    */
    int32_t reserved[6];
} pm_config;

typedef struct pm_info {
    int32_t n_lattice;
    int32_t mass_shift;    /* in effect after the last pm_load */
    int64_t n_particles;   /* loaded */
    int64_t capacity;
    int64_t device_bytes;  /* device memory held by the handle */
    int32_t device;
    int32_t reserved;
    int64_t sorts;         /* particle re-binnings performed since creation            */
    int64_t migrants;      /* deposits that fell outside their tile window (global atomics) */
} pm_info;

/* Defaults of the reference: N 2048, CIC, dt 50, G 1, G(0) 3, eps 0.001. */
void pm_config_default(pm_config* cfg);

/* Device memory one simulation of this configuration holds (for admission
 * by memory budget).  -1 on an invalid configuration. */
int64_t pm_instance_bytes(const pm_config* cfg);

/* Number of HIP devices visible (0 without a GPU). */
int pm_device_count(void);

/* Allocate lattices and particle arrays, build the FP64 transfer function.
 * Replaces the device part of Universe::Universe (cuh:969-1035). */
pm_status pm_create(const pm_config* cfg, pm_handle* out);

/* Free everything.  Replaces Universe::~Universe (cuh:1396-1427). */
pm_status pm_destroy(pm_handle h);

/* Copy n <= capacity particles from host SoA arrays (caller keeps them).
 * Replaces the H2D uploads of the ctor / clear / addGalaxy
 * (cuh:1012-1016, 1053-1060, 1151-1155).  Picks the fixed-point shift from
 * sum|m| unless cfg.mass_shift is set. */
pm_status pm_load(pm_handle h, const float* x, const float* y, const float* vx, const float* vy,
                  const float* m, int64_t n);

/* Same, from the reference's Verlet state (x, xOld): v = (x - xOld)/dt. */
pm_status pm_load_verlet(pm_handle h, const float* x, const float* y, const float* x_old,
                         const float* y_old, const float* m, int64_t n);

/* Enqueue nsteps full steps on the handle's stream (asynchronous).  Replaces
 * the worker loop body (cuh:1336-1341): deposit, FFT convolution, gradient,
 * gather + update.  The deposit of each step after the first is fused into
 * the previous step's gather + update; the mass lattice of the positions a
 * step ends with is therefore already built when it returns.  With the
 * adaptive re-sort on ("resort_migrants_permille" > 0) a simulation of 2^20
 * or more particles keeps at most two steps queued ahead of the device. */
pm_status pm_step(pm_handle h, int32_t nsteps);

/* Enqueue one stage of a step (PM_STAGE_*), for stage-level checks. */
pm_status pm_stage(pm_handle h, int32_t stage);

/* Wait for all work on the handle's stream. */
pm_status pm_sync(pm_handle h);

/* Blocking copy of the state into caller buffers (any pointer may be NULL).
 * The reference has no read-back path (its device state never returns). */
pm_status pm_read(pm_handle h, float* x, float* y, float* vx, float* vy);

/* Same, as the reference's Verlet state: x_old = x - v*dt. */
pm_status pm_read_verlet(pm_handle h, float* x, float* y, float* x_old, float* y_old);

/* Blocking copy of a lattice (PM_LATTICE_*) into dst of at least bytes. */
pm_status pm_read_lattice(pm_handle h, int32_t which, void* dst, size_t bytes);

/* FP64 reductions on the device: out[0] mass, [1] px, [2] py, [3] kinetic
 * energy, [4] potential energy 1/2 sum rho*Phi (of the last field solve),
 * [5] max |v|, [6] non-finite particles, [7] particles in bounds. */
pm_status pm_diagnostics(pm_handle h, double out[8]);

pm_status pm_get_info(pm_handle h, pm_info* info);

/* Density frame of the current mass lattice, replacing the reference's render
 * stage (cuh:613-828, scheduled at 1249-1256): mass per cell -> blur_r x blur_r
 * filter with weights 1/(r+1) for r < (blur_r-1)/2 and 0 otherwise (the
 * identity for the reference's BLUR_R = 3), zero outside the lattice -> the
 * frame's min and max -> running extremes kept by the handle since its first
 * frame (reset_extremes != 0 restarts them) -> 1.8 * ((v - min)/(max - min))^0.25.
 * blur_r odd in [1, 15]; frame: N*N floats on the host, index i + j*N.
 * Blocking.  The reference renders the lattice deposited at the start of its
 * K-th step; this renders the lattice of the current positions. */
pm_status pm_frame(pm_handle h, int32_t blur_r, int32_t reset_extremes, float* frame);

/* Checkpoints (no counterpart in the reference, whose state never leaves the
 * device): a header (configuration, fixed-point shift, FFT plan, tunables,
 * step count, FNV-1a checksum) followed by x, y, vx, vy, m as little-endian
 * float32 in load order.  Restoring creates a handle on `device` whose next
 * steps are bit-identical to those of the saved handle.  pm_checkpoint_bytes
 * gives the buffer size for pm_checkpoint_write (-1 for a null handle). */
int64_t pm_checkpoint_bytes(pm_handle h);
pm_status pm_checkpoint_write(pm_handle h, void* dst, int64_t bytes);
pm_status pm_checkpoint_restore(const void* src, int64_t bytes, int32_t device, pm_handle* out);
pm_status pm_checkpoint_save(pm_handle h, const char* path);
pm_status pm_checkpoint_open(const char* path, int32_t device, pm_handle* out);

/* Runtime tunables: "resort_interval"
 * (steps between particle re-binnings), "resort_migrants_permille" (re-bin
 * early when more than this many per thousand particles per step land outside
 * their tile's window; 0 = off, default 10), "resort_drift_ppm" (with the
 * interval chosen at load and the adaptive re-sort on: shorten the interval
 * by one, down to 3, after a re-sort cycle whose peak rate of particles
 * outside their tile windows exceeded this many per million per step;
 * 0 = off, default 200), "deposit_chunk" (particles per
 * particle-kernel block, a multiple of 512; by default set at each load:
 * below 3e7 particles about n_particles / 256 within [2048, 8192]; above,
 * 7680 on lattices up to 2048 and 20480 above, or about n_particles / 8192
 * within [12288, 131072] for dense runs (> 64 particles per cell) whose
 * re-sorts run inside the step kernel; setting it fixes it),
 * "profile_mask" (kernels timed while pm_profile is on, bit k = PM_K_* id k),
 * "bin_max_chunks" (a dense tile is split into at most this many equal
 * blocks, default 0 = no cap), "tile_log" (bin edge 2^5 or
 * 2^6 cells; not below the lattice's default) change speed, never results;
 * the FFT plan "fft_rp", "fft_cb", "fft_radix" changes FP32 rounding only
 * (see pm_autotune; "fft_radix" 64 selects the wave-local passes, one
 * 64-lane wavefront per row pair / column, N = 256 .. 2048); "fft_it" (row
 * pairs per block of the C2R pass, handled in sequence with the next one's
 * inputs prefetched and the edge bands done in the pass, the R2C and column
 * passes then running one row pair / column per block: 1 (the one-per-block
 * passes and a separate edge pass), or 2 or 4 with fft_rp = fft_cb = 1 at
 * N = 256, 512 (2 only) and 1024 radix 4 and N = 2048, 4096 radix 8; default
 * 2 at those sizes, else 1; setting another batch or radix resets it to 1
 * where needed) changes speed only.  pm_get_param also reads "adaptive_sorts" (early
 * re-bins so far) and "interval_changes" (shortenings of the interval by the
 * drift rule so far).  No counterpart in the reference, whose
 * launch shape is the compile-time Constants::THREADS (cuh:31). */
pm_status pm_set_param(pm_handle h, const char* name, int64_t value);
pm_status pm_get_param(pm_handle h, const char* name, int64_t* value);

/* Autotuner (no counterpart in the reference).  PM_TUNE_FFT times every
 * compiled FFT plan ("fft_rp" row pairs per row-pass block in {1,2,4},
 * "fft_cb" columns per column-pass block in {1,2,4,8}, "fft_radix" 4 or 8,
 * "fft_it" 1, 2 or 4) on
 * the current mass lattice and keeps the fastest; PM_TUNE_PARTICLES times
 * tile size, chunk size and re-sort interval over `reps`-step trials on a
 * device snapshot of the state, which is restored afterwards.  The particle
 * parameters never change results; a different FFT plan changes the FP32
 * rounding of the field (every plan meets the 1e-5 relative L2 bar). */
enum { PM_TUNE_FFT = 1, PM_TUNE_PARTICLES = 2 };
pm_status pm_autotune(pm_handle h, int32_t flags, int32_t reps);

/* Per-kernel device timing with HIP events on the handle's stream.  Each timed
 * launch adds two event records to the stream (about 5 us per launch on
 * MI355X); pm_set_param "profile_mask" (bit k = kernel id k, default all)
 * limits the timing to the kernels of interest. */
pm_status pm_profile(pm_handle h, int32_t enable);
pm_status pm_profile_read(pm_handle h, int32_t kernel, double* total_ms, int64_t* launches);
pm_status pm_profile_reset(pm_handle h);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  pm_diagnostics' FP64 reductions over caller-owned
 * DEVICE arrays, written to a caller-owned DEVICE buffer: x, y, vx, vy, m
 * (np floats each), optionally the int64 mass lattice rho (fixed point
 * 2^-mass_shift) and the float potential phi of an n_lattice^2 lattice (both
 * NULL: out[4] = 0), out = 8 doubles in pm_diagnostics' order.  Allocates
 * nothing: the block partial sums use a fixed device array of the library,
 * PM_DIAG_NOALLOC_SCRATCH_BYTES per device, held until the process ends.
 * Asynchronous on `stream` (a hipStream_t; NULL = the null stream) of
 * `device`; returns PM_EINVAL for arguments the fixed array cannot hold.
 * Every call on one device shares that array: calls that overlap on one
 * device race.  Same inputs in the same order give pm_diagnostics' bits. */
#define PM_DIAG_NOALLOC_SCRATCH_BYTES 73728
pm_status pm_diagnostics_noalloc(int32_t device, void* stream, int32_t n_lattice, const float* x, const float* y,
                                 const float* vx, const float* vy, const float* m, int64_t np, const int64_t* rho,
                                 int32_t mass_shift, const float* phi, double* out);

/* The handle's hipStream_t, for callers that order their own work after it. */
void* pm_stream(pm_handle h);

const char* pm_last_error(pm_handle h);
const char* pm_status_string(pm_status s);

#ifdef __cplusplus
}
#endif
#endif /* COSMOS_PM_H */
