// Internal launcher interface between the C-ABI layer (pm_api.cpp) and the
// HIP kernels (pm_*.hip).  Not installed; the public interface is include/pm.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmk {

constexpr int kBlock = 256;   // threads per block for every streaming kernel (4 wave64s)

// Particle binning: square tiles of 2^tl cells; a particle kernel block works
// on one chunk of one tile with LDS windows covering the tile plus a
// kHalo-cell margin (plus the +1 column/row CIC reaches).  Tiles are 32 cells
// for lattices up to 2048 (at most 4096 bins, 24.2 KB per window) and 64
// above.  The margin is the largest with which the re-sorting step (two
// windows and two per-bin tables) still runs two blocks per CU at 32-cell
// tiles: fewer particles leave their windows between re-sorts (each such
// deposit is a global atomic), so the drift rule keeps longer intervals.
// Against 7 cells, ms per step: colliding disks 1e8 CIC over steps
// 5-25 / 25-61 / 61-200 0.868-0.883 / 0.850 / 0.863 against 0.870 / 0.854 /
// 0.903; single disk 2e7 over steps 325-500 0.290 against 0.328; 1e9 NGP
// over steps 15-35 9.36 against 9.92; 1e9 CIC over steps 35-55 13.10
// against 14.33; 12 cells: 0.986 / 0.947 / 0.948 (one re-sorting block per
// CU).  At 64-cell tiles (4096^2 Plummer binary, 20 steps) 1.204 against
// 1.193 ms.
constexpr int kHalo = 11;
constexpr int kDepositBlock = 512;
constexpr int kDepositChunkDefault = 12288;          // particles per particle-kernel block (tunable)
// Automatic chunk, chosen at each load:
//  * below kSmallRunParticles: about np / 256 within [2048, 8192], a multiple
//    of 2048 (a 12288 floor left most CUs idle: 1e6 particles, 82 blocks);
//    measured on exponential disks of 1.5e5..1e7 particles at 256^2..2048^2,
//    the step kernel took 0.55x to 1.0x the time it takes with 12288 (4e5
//    particles at 1024^2: 27.5 -> 15.7 us; 1e7: 0.96..0.98x);
//  * from there on 32-cell tiles (N <= 2048): kChunkTiled32.  Two colliding
//    disks of 1e8 on 2048^2 (ms per step / step kernel ms): CIC 0.925 / 0.771
//    at 12288, 0.907 / 0.742 at 8192, 0.899 / 0.728 at 7680, 0.905 / 0.728 at
//    6656, 0.975 / 0.791 at 4096; NGP 0.924 at 12288, 0.888 at 7680; 1e9 CIC
//    (separate re-sort every 2 steps) 14.22 at 122880, 13.95 at 32768, 13.77
//    at 16384, 13.46 at 7680;
//  * on 64-cell tiles (N >= 4096, windows 4x larger): kChunkTiled64.  Plummer
//    spheres of 1e8 on 4096^2: 1.224 ms per step at 12288, 1.205 at 16384,
//    1.191 at 20480, 1.197 at 24576, 1.32 at 7680;
//  * dense runs whose re-sorts are fused into the step kernel (1e9 NGP on
//    2048^2): about kChunkTargetBlocks blocks per launch within
//    [kDepositChunkDefault, kDepositChunkMax], a multiple of 4096 (11.13 ms
//    per step at 122880, 12.15 at 32768: the fused re-sort step slows with
//    more blocks flushing into the same dense cells).
constexpr int kChunkTargetBlocks = 8192;
constexpr int kDepositChunkMax = 131072;
constexpr int kChunkTiled32 = 7680;
constexpr int kChunkTiled64 = 20480;
constexpr int64_t kSmallRunParticles = 30000000;
// large_chunks: a dense run whose re-sorts are fused into the step kernel
inline int auto_chunk(int64_t np, int n, bool large_chunks) {
    if (np < kSmallRunParticles) {
        const int64_t c = (np / 256 + 2047) / 2048 * 2048;
        return (int)(c < 2048 ? 2048 : c > 8192 ? 8192 : c);
    }
    if (!large_chunks) return n <= 2048 ? kChunkTiled32 : kChunkTiled64;
    const int64_t c = (np / kChunkTargetBlocks + 4095) / 4096 * 4096;
    return (int)(c < kDepositChunkDefault ? kDepositChunkDefault : c > kDepositChunkMax ? kDepositChunkMax : c);
}

inline int tile_log_for(int n) { return n <= 2048 ? 5 : 6; }
template <int TL>
struct TileGeom {
    static constexpr int kEdge = 1 << TL;
    static constexpr int kWin = kEdge + 2 * kHalo + 1;
};

// Workgroup barrier that orders LDS only.  __syncthreads() is a full
// workgroup fence: before every barrier it waits for all of the thread's
// outstanding global loads and stores (vmcnt(0)).  Kernels whose threads
// exchange data through LDS only (FFT stages, the step kernels' count / rank /
// window phases) use an LDS-scope fence around s_barrier instead, so
// prefetched loads and earlier stores stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// cnt[b] += 1 for every active lane (b >= 0); returns the value before this
// lane's increment.  Lanes in one bin get consecutive values, so a block's
// values for a bin are a permutation of 0 .. count - 1.  The lanes that share
// the wave's first active bin are served by one LDS atomic (in a dense tile
// nearly every lane of a wave is in the same bin, where per-lane atomics to
// one address serialise); the others add individually.
__device__ __forceinline__ int wave_count_add(int* cnt, int b) {
    const int b0 = __builtin_amdgcn_readfirstlane(b);        // bin of the first active lane
    const unsigned long long same = __ballot(b == b0);
    if (b != b0) return atomicAdd(&cnt[b], 1);
    const int lane = (int)__lane_id();
    int base = 0;
    if (lane == __ffsll((long long)same) - 1) base = atomicAdd(&cnt[b0], __popcll(same));
    base = __builtin_amdgcn_readfirstlane(base);              // the first active lane is the leader
    return base + __popcll(same & ((1ull << lane) - 1ull));
}

// The same count without the ranks: no value returns from the atomics.
__device__ __forceinline__ void wave_count(int* cnt, int b) {
    const int b0 = __builtin_amdgcn_readfirstlane(b);
    const unsigned long long same = __ballot(b == b0);
    if (b != b0) {
        atomicAdd(&cnt[b], 1);
        return;
    }
    if ((int)__lane_id() == __ffsll((long long)same) - 1) atomicAdd(&cnt[b0], __popcll(same));
}

// Bin of a position: its tile when in bounds (1 <= x, y < N-1, the
// reference's rule, cuh:257, 497), else the parked bin (N >> tl)^2.
__device__ __forceinline__ int bin_of(float x, float y, int n, int tl) {
    const int tpr = n >> tl;
    const float hi = (float)(n - 1);
    if (!(x >= 1.0f && x < hi && y >= 1.0f && y < hi)) return tpr * tpr;
    return ((int)y >> tl) * tpr + ((int)x >> tl);
}

struct ParticleArrays {
    float *x, *y, *vx, *vy, *m;
    uint32_t* id;   // load index of the particle in each slot
};

// 0 = no cap: measured as fast or faster than a cap of 64 at 1e8 (2048^2
// disks, 4096^2 Plummer spheres) and, with the automatic chunk, at 1e9.
constexpr int kMaxChunksPerBinDefault = 0;
struct BinState {
    int* count;          // [nbins + 1] particles per bin (last = parked)
    int64_t* start;      // [nbins + 2] exclusive prefix of count
    int* chunk_start;    // [nbins + 1] exclusive prefix of deposit chunks per in-bounds bin
    int* cursor;         // [nbins + 1] scatter cursors
    int chunk;           // particles per deposit block (target)
    int tl;              // log2 of the tile edge in cells
    int max_chunks_per_bin = kMaxChunksPerBinDefault;   // a dense bin is split into at most this many
                                                        // equal chunks (fewer window flushes to the same cells)
};

// Lattice sizes with compiled kernels (powers of two).
inline bool lattice_supported(int n) {
    return n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096 ||
           n == 8192;
}

// ----- binning (pm_binning.hip) -----------------------------------------------
int bin_count_of(int n, int tl);   // in-bounds bins + 1 parked bin
// Counting sort of cur into alt by bin; fills bs.  Caller swaps cur and alt.
// counted: bs.count already holds the bins of every particle outside the parked
// range (filled by push_deposit), so the counting pass is skipped.
hipError_t sort_particles(int n, ParticleArrays cur, ParticleArrays alt, int64_t np, BinState bs, bool counted,
                          hipStream_t s);
hipError_t iota_ids(uint32_t* id, int64_t np, hipStream_t s);
hipError_t unpermute(const uint32_t* id, int64_t np, const float* a0, const float* a1, const float* a2,
                     const float* a3, float* o0, float* o1, float* o2, float* o3, hipStream_t s);

// ----- deposit (pm_particles.hip) ---------------------------------------------
// rho must be zero on entry; adds int64 fixed-point contributions round(w*m*scale)
// for the particles of every in-bounds bin (bs from the last sort_particles).
// max_chunks: upper bound of the chunk count (grid size).  migrants counts the
// contributions made outside the tile window.
int deposit_max_chunks(int n, int tl, int64_t capacity, int chunk);
// Re-sort fused into a step: push + deposit the chunks of `cur` (read from
// src), writing each particle to dst grouped by the bin of its start
// position (next.start / next.cursor from scan_next), and move the standing
// parked particles.  scatter_fits: the kernel's LDS fits this lattice/tile.
bool scatter_fits(int n, int tl);
hipError_t push_deposit_scatter(int n, int mode, int wrap_mode, const float2* force, ParticleArrays src,
                                ParticleArrays dst, BinState cur, BinState next, int max_chunks, float dt,
                                float scale, int64_t* rho_next, unsigned long long* migrants, hipStream_t s);
// next.count (counted by a push with count_bins) -> next.start, chunk_start,
// cursor; the parked bin also holds the particles standing in cur's parked
// range (old_start), and its cursor starts after them.
hipError_t scan_next(int n, BinState next, const int64_t* old_start, hipStream_t s);
hipError_t deposit(int n, int mode, const float* x, const float* y, const float* m, BinState bs, int max_chunks,
                   float scale, int64_t* rho, unsigned long long* migrants, hipStream_t s);
// Fused step tail: gather + kick-drift every particle of every in-bounds bin,
// then deposit its new position into rho_next (zero on entry).  count_bins:
// also histogram the new bins into bs.count (zero on entry) for the next sort.
// pipelined: the software-pipelined variant (faster below kPipelineMaxDensity
// particles per cell; slower above, where LDS atomics contend).
constexpr int kPipelineMaxDensity = 64;
hipError_t push_deposit(int n, int mode, int wrap_mode, const float2* force, ParticleArrays p, BinState bs,
                        int max_chunks, float dt, float scale, int64_t* rho_next, unsigned long long* migrants,
                        bool count_bins, bool pipelined, hipStream_t s);

// ----- FFT field solve (pm_fft.hip) -------------------------------------------
// Half spectrum layout: H[k * N + j], k in [0, N/2] (x frequency), j in [0, N)
// (y row before the column pass, y frequency after it).  Every pass takes an
// array of `count` descriptors in device memory (one per simulation, all of
// lattice size n) and launches grid.y = count.
struct FieldDesc {
    const int64_t* rho;   // mass lattice (int64 fixed point)
    float2* H;            // half spectrum
    float2* Hfy;          // half spectrum of the y derivative
    const float* T;       // transfer function incl. 1/N^2, [kx][ky]
    const float* dk;      // stencil symbol G*(8 sin t - sin 2t)/6
    float* phi;           // potential
    float2* force;        // (Fx, Fy)
    float inv_scale;      // 2^-shift
    float gravity;        // force multiplier
    int64_t* rho_zero = nullptr;   // if set, the row R2C pass zeroes its rows of this
                                   // lattice (the next deposit target) as it reads rho
};
// Tile batch and radix of the three passes: rp row pairs per row-pass block,
// cb columns per column-pass block, radix 4 or 8 (FP32 Stockham stages) or 16
// (split-BF16 matrix-core stages, rp = cb = 1, 256 <= n <= 8192); it > 1:
// it row pairs / columns per block in sequence with the next one's inputs
// prefetched (rp = cb = 1, pm_fft_pipe.hip).
struct FftPlan {
    int rp = 2, cb = 1, radix = 4, it = 1;
};
// radix kWaveRadix: the wave-local passes (pm_fft_wave.hip), one 64-lane
// wavefront per row pair / column, rp = cb = it = 1, N = 256 .. 2048
constexpr int kWaveRadix = 64;
bool wave_plan_compiled(int n);
hipError_t wave_pass(int pass, int n, const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s);
bool plan_valid(int n, FftPlan p);
bool pipe_plan_compiled(int n, int radix, int it);
// The prefetching C2R pass (it > 1) applies the reference's clamped stencil to
// the edge bands itself; with it = 1 the separate edge_force pass does.
// (With the matrix-core radix codes, 16 and 17, only the R2C and column
// passes take the prefetching kernels: the C2R pass's three-transform batch
// cuts their blocks in flight, 4096^2: 0.24 -> 0.41 ms.)
inline bool c2r_does_edges(const FftPlan& p) { return p.it != 1 && p.radix < 16; }
hipError_t row_r2c_pipe(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, hipStream_t s);
hipError_t col_conv_pipe(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                         hipStream_t s);
hipError_t row_c2r_pipe(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                        hipStream_t s);
FftPlan default_plan(int n);
hipError_t row_r2c(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, hipStream_t s);
hipError_t col_conv(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                    hipStream_t s);
hipError_t row_c2r(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                   hipStream_t s);
hipError_t edge_force(int n, const FieldDesc* d, int count, hipStream_t s);
hipError_t stencil_force(int n, const FieldDesc* d, int count, hipStream_t s);

// ----- batched small simulations (pm_particles.hip, pm_util.hip) ------------------
// One fused push + deposit launch over many simulations whose lattices are
// small enough to stay in L2: no tile windows, force gathered and mass
// deposited with global memory operations.  Particle order is irrelevant.
struct PartDesc {
    float *x, *y, *vx, *vy;
    const float* m;
    const float2* force;
    unsigned long long* rho_next;   // zero on entry
    int64_t np;
    int n;
    float dt;
    float scale;                    // 2^shift
    int wrap_mode;
};
// block_start[i]: first block of simulation i (exclusive prefix of
// ceil(np_i / (4 * kBlock)) blocks), block_start[count] = total blocks.
hipError_t push_deposit_direct(int mode, const PartDesc* d, const int* block_start, int count, int total_blocks,
                               hipStream_t s);
int direct_blocks(int64_t np);
// ----- persistent round kernel (pm_round.hip) ------------------------------------
// One batched member: field and particle descriptors per step parity (as the
// per-pass batched launches use them; fd[p].rho_zero set), its twiddles,
// lattice, deposit mode and items per step for the four passes (R2C, column,
// C2R, push).
struct RoundInst {
    FieldDesc fd[2];
    PartDesc pd[2];
    const float2* tw;
    int n, mode;
    int items[4];
};
// A run of queue items: pass `phase` of member `inst` in step `step`,
// starting at item `start`; segments are in queue order (step, pass, member).
struct RoundSeg {
    int start, inst, phase, step;
};
inline size_t round_ctl_bytes(int ninst) { return ((size_t)(2 + 4 * ninst) * 4 + 15) / 16 * 16; }
// lattices and plans the kernel runs: the default FP32 plans at N = 256..2048
bool round_supported(int n, const FftPlan& p);
void round_items(int n, int64_t np, int items[4]);
size_t round_lds_bytes(int max_n);
// blocks of the kernel co-resident on the device at this LDS size (sets the
// kernel's dynamic LDS limit; call outside stream capture)
int round_max_blocks(int max_n, size_t lds_bytes);
// wall-clock ticks of the device's constant-rate counter in `seconds`
long long round_wait_ticks(double seconds);
// steps x passes x members work queue; ctl: round_ctl_bytes(ninst) of device
// memory (zeroed here, by a fill kernel), ctl[1] != 0 afterwards if a
// dependency wait exceeded wait_ticks
hipError_t round_persistent(const RoundInst* inst, const RoundSeg* seg, int nseg, int total, int* ctl, int ninst,
                            int max_n, size_t lds_bytes, int blocks, long long wait_ticks, hipStream_t s);

// One simulation of a batched serving round that keeps its particles sorted
// by tile (more particles than the direct kernel suits): its step kernels,
// the fused re-sort's counting, scan and re-sorting steps and the parked-range
// copy run over arrays of these, one launch per group of equal lattice size,
// tile size, deposit mode and wrap mode (pm_particles.hip, pm_binning.hip).
struct TiledDesc {
    ParticleArrays cur, alt;              // current order; target of the re-sort
    const float2* force;
    unsigned long long* rho_next[2];      // deposit target by step parity
    const int64_t* start;                 // bins of the current order
    const int* chunk_start;
    int* next_count;                      // second bin set, filled by the counting step
    int64_t* next_start;
    int* next_chunk_start;
    int* next_cursor;
    unsigned long long* migrants;
    int n, nbins, chunk, kmax;            // lattice, bins incl. the parked one, deposit chunk, chunks per bin cap
    float dt, scale;                      // time step, 2^shift
};
// chunk_first[i]: chunks of simulations 0 .. i-1 in their current chunk maps,
// chunk_first[count] their total (count + 1 entries, written on the device;
// run again whenever the chunk maps change: after each re-sort).
hipError_t member_chunks_batched(const TiledDesc* d, int count, int* chunk_first, hipStream_t s);
// chunk_first: as member_chunks_batched left it for these descriptors;
// total_blocks: an upper bound of the chunks (the sum of the members'
// max_chunks), which bounds the grid.  p: parity of the step (rho_next[p] is
// the deposit target).  count_bins: the counting step (histogram of the new
// bins into next_count, zero on entry).  The loop is the one loading at the
// top of each pass (the software-pipelined one needs more registers than two
// blocks per CU leave once blocks stride over chunks).
hipError_t push_deposit_batched(int n, int tl, int mode, int wrap_mode, bool count_bins, const TiledDesc* d,
                                const int* chunk_first, int count, int total_blocks, int p, hipStream_t s);
// The re-sorting step (cur -> alt in the order of the counted bins, with the
// standing parked particles) of every simulation.
hipError_t push_deposit_scatter_batched(int n, int tl, int mode, int wrap_mode, const TiledDesc* d,
                                        const int* chunk_first, int count, int total_blocks, int p, hipStream_t s);
// next_count = 0 (nbins entries) of every simulation.
hipError_t zero_next_counts_batched(const TiledDesc* d, int count, int max_nbins, hipStream_t s);
// scan_next of every simulation (next bins from next_count; old starts = start).
hipError_t scan_next_batched(const TiledDesc* d, int count, hipStream_t s);

struct FillDesc {
    void* p;
    size_t bytes;   // multiple of 16
};
hipError_t fill_zero_batched(const FillDesc* d, int count, size_t max_bytes, hipStream_t s);

// ----- FP64 setup spectra (pm_spectra.hip) --------------------------------------
// T[k*N + ky] = Ghat(k,ky) * c/(c^2+eps) / N^2 as float, k in [0,N/2].
// work64 must hold 2 * N * N doubles (complex lattice).
hipError_t build_transfer(int n, int mode, double green_origin, double eps, double* work64,
                          float* T, hipStream_t s);

// ----- push (pm_push.hip) -----------------------------------------------------------
hipError_t push(int n, int mode, const float2* force, float* x, float* y, float* vx, float* vy,
                int64_t np, float dt, int wrap_mode, hipStream_t s);

// ----- density frames (pm_frame.hip) --------------------------------------------------
// out: N*N floats; keys: 2 unsigned; smooth: 2 floats (running extremes, kept
// between frames).  blur_r odd in [1, 15].
hipError_t frame(int n, const int64_t* rho, float inv_scale, int blur_r, bool reset, float* out, unsigned* keys,
                 float* smooth, hipStream_t s);

// ----- utilities (pm_util.hip) ------------------------------------------------------
hipError_t fill_zero(void* p, size_t bytes, hipStream_t s);
// Writes four field descriptors to device memory, in stream order.
hipError_t set_field_desc(FieldDesc* dst, const FieldDesc* src4, hipStream_t s);
// Per-block FP64 partial sums: out[b*8 + {0..7}] = {mass, px, py, ke, max|v|, nonfinite, n_inb, 0}
constexpr int kDiagBlocks = 1024;   // partial sums of the diagnostic reductions: at most this many blocks
int diag_blocks(int64_t np);
hipError_t diag_particles(int n, const float* x, const float* y, const float* vx, const float* vy,
                          const float* m, int64_t np, double* partial, hipStream_t s);
// out[b] = partial sums of rho*phi (mass units)
int diag_lattice_blocks(int n);
hipError_t diag_lattice(int n, const int64_t* rho, double inv_scale, const float* phi,
                        double* partial, hipStream_t s);

}  // namespace pmk
