// Particle kernels: mass deposit (a) and force gather + kick-drift (d).
//
// Deposit restates k_scatterMassOnAccumulator (CosmosCuda.cuh:446-595) with
// int64 fixed-point sums: each contribution w*m (FP32, separately rounded) is
// quantised to round(w*m*2^shift) and added with a 64-bit integer add, so the
// lattice is bit-exact and independent of particle order.
//
// Push restates k_forceMultiSampling (CosmosCuda.cuh:231-312) on (x, v) state:
// v' = v + a dt, x' = x + v' dt (the reference's position Verlet with
// v = (x - xOld)/dt), the same in-bounds rule and the same wrap thresholds.
//
// Particles are kept sorted by tile (pm_binning.hip).  The chunk-mapped
// kernels give each block one chunk of one tile and stage that tile's
// neighbourhood in LDS: an int64 window for the deposit and, in the fused
// step kernel, a float2 window of the force lattice for the gathers.  Cells
// outside the windows (particles that left their tile by more than kHalo
// cells since the last sort) are served from the lattice in global memory;
// the arithmetic is the same on both paths, so results never depend on it.
#include "pm_particle_core.h"

namespace pmk {
namespace {

// Block c -> (bin b, slots [s0, s1)): chunk_start is the exclusive prefix of
// chunks over the in-bounds bins, chunk_start[nin] their total; the k chunks
// of a bin of P particles split it into equal parts [j P/k, (j+1) P/k).
__device__ __forceinline__ bool chunk_of(const int* __restrict__ chunk_start, const int64_t* __restrict__ start,
                                         int nin, int c, int& b, int64_t& s0, int64_t& s1) {
    if (c >= chunk_start[nin]) return false;
    int lo = 0, hi = nin;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (chunk_start[mid] <= c) lo = mid + 1; else hi = mid;
    }
    b = lo - 1;
    const int64_t p = start[b + 1] - start[b], k = chunk_start[b + 1] - chunk_start[b], j = c - chunk_start[b];
    s0 = start[b] + j * p / k;
    s1 = start[b] + (j + 1) * p / k;
    return true;
}

template <int W, int BLK>
__device__ __forceinline__ void flush_window(const unsigned long long* win, int wx0, int wy0,
                                             unsigned long long* __restrict__ rho, int n) {
    for (int e = threadIdx.x; e < W * W; e += BLK) {
        const unsigned long long v = win[e];
        if (v == 0ull) continue;
        const int gx = wx0 + e % W, gy = wy0 + e / W;   // nonzero cells lie in [1, N-1]^2
        atomicAdd(rho + (size_t)gy * n + gx, v);
    }
}

// ------------------------------------------------------------ deposit kernel --
// One chunk of one bin per block; contributions summed in the bin's LDS
// window and flushed once, spill to global atomics.
template <int TL, int MODE>
__global__ __launch_bounds__(kDepositBlock) void k_deposit_bin(const float* __restrict__ x,
                                                               const float* __restrict__ y,
                                                               const float* __restrict__ m, int n,
                                                               const int64_t* __restrict__ start,
                                                               const int* __restrict__ chunk_start, int chunk,
                                                               float scale, unsigned long long* __restrict__ rho,
                                                               unsigned long long* __restrict__ migrants) {
    constexpr int W = TileGeom<TL>::kWin, E = TileGeom<TL>::kEdge;
    __shared__ unsigned long long win[W * W];
    __shared__ unsigned int s_mig;
    const int tpr = n >> TL;
    int b;
    int64_t s0, s1;
    if (!chunk_of(chunk_start, start, tpr * tpr, blockIdx.x, b, s0, s1)) return;
    const int wx0 = (b % tpr) * E - kHalo, wy0 = (b / tpr) * E - kHalo;
    for (int e = threadIdx.x; e < W * W; e += kDepositBlock) win[e] = 0ull;
    if (threadIdx.x == 0) s_mig = 0;
    __syncthreads();
    const float fhi = (float)(n - 1);
    const float* xb = x + s0;
    const float* yb = y + s0;
    const float* mb = m + s0;
    const int len = (int)(s1 - s0);
    unsigned int mig = 0;
    constexpr int U = 4;
    for (int i0 = threadIdx.x; i0 < len; i0 += U * kDepositBlock) {
        float px[U], py[U], pm[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kDepositBlock;
            px[u] = -1.0f; py[u] = -1.0f; pm[u] = 0.0f;
            if (i < len) { px[u] = xb[i]; py[u] = yb[i]; pm[u] = mb[i]; }
        }
        bool spill[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            spill[u] = in_bounds(px[u], py[u], fhi) && !deposit_local<MODE, W>(win, wx0, wy0, px[u], py[u], pm[u], scale);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (spill[u]) { deposit_global<MODE>(rho, n, px[u], py[u], pm[u], scale); ++mig; }
    }
    if (mig) atomicAdd(&s_mig, mig);
    __syncthreads();
    flush_window<W, kDepositBlock>(win, wx0, wy0, rho, n);
    if (threadIdx.x == 0 && s_mig) atomicAdd(migrants, (unsigned long long)s_mig);
}

// -------------------------------------------------------- fused step kernel --
// Gather + kick-drift of each particle of one chunk (force from the bin's LDS
// force window, or from the lattice when the cell lies outside it), write it
// back, and deposit the new position into the next step's mass lattice.
// COUNT: also histogram the new bins for the next re-sort.
// PIPE: software-pipelined loads (below); 112 instead of 72 VGPRs, so two
// blocks per CU instead of three.
template <int TL> constexpr int fused_block() { return TL <= 5 ? 512 : 1024; }

// Chunk c of one simulation; the single-simulation kernel runs chunk
// blockIdx.x, the batched one (k_push_deposit_batched) maps its blocks over
// the chunks of several simulations.  A chunk of fewer than kSparseChunk
// particles (a sparse tile of a small or spread-out simulation) gathers from
// the force lattice and deposits with global atomics instead of staging the
// two tile windows in LDS: the same arithmetic, bit for bit, without loading
// and flushing the 48 KB of windows (32-cell tiles) for a few dozen particles.
constexpr int kSparseChunk = 256;
template <int TL, int MODE, int WRAP, bool COUNT, bool PIPE>
__device__ __forceinline__ void push_deposit_block(
    const float2* __restrict__ force, float* __restrict__ x, float* __restrict__ y, float* __restrict__ vx,
    float* __restrict__ vy, const float* __restrict__ m, int n, float dt, const int64_t* __restrict__ start,
    const int* __restrict__ chunk_start, float scale, unsigned long long* __restrict__ rho,
    unsigned long long* __restrict__ migrants, int* __restrict__ count, int c) {
    constexpr int W = TileGeom<TL>::kWin, E = TileGeom<TL>::kEdge, BLK = fused_block<TL>();
    __shared__ unsigned long long win[W * W];
    __shared__ float2 fwin[W * W];
    __shared__ unsigned int s_mig;
    extern __shared__ int hist[];
    const int tpr = n >> TL, nin = tpr * tpr;
    int b;
    int64_t s0, s1;
    if (!chunk_of(chunk_start, start, nin, c, b, s0, s1)) return;
    const bool sparse = s1 - s0 < kSparseChunk;
    const int wx0 = (b % tpr) * E - kHalo, wy0 = (b / tpr) * E - kHalo;
    if (!sparse)
        for (int e = threadIdx.x; e < W * W; e += BLK) {
            win[e] = 0ull;
            const int gx = wx0 + e % W, gy = wy0 + e / W;
            fwin[e] = ((unsigned)gx < (unsigned)n && (unsigned)gy < (unsigned)n) ? force[(size_t)gy * n + gx]
                                                                                 : make_float2(0.f, 0.f);
        }
    if (COUNT)
        for (int e = threadIdx.x; e <= nin; e += BLK) hist[e] = 0;
    if (threadIdx.x == 0) s_mig = 0;
    lds_barrier();
    const float fhi = (float)(n - 1);
    // Aligned groups of 4 slots: [q0, q1) covers [s0, s1); slots outside it
    // belong to neighbouring chunks and are neither used nor written.  Reads of
    // the partial groups stay inside the arrays (capacity is padded to 64).
    const int64_t q0 = s0 >> 2, q1 = (s1 + 3) >> 2;
    const int nq = (int)(q1 - q0);
    v4f* X = reinterpret_cast<v4f*>(x) + q0;
    v4f* Y = reinterpret_cast<v4f*>(y) + q0;
    v4f* VX = reinterpret_cast<v4f*>(vx) + q0;
    v4f* VY = reinterpret_cast<v4f*>(vy) + q0;
    const v4f* M = reinterpret_cast<const v4f*>(m) + q0;
    const int lo_k = (int)(s0 - 4 * q0);             // first valid lane of group 0
    const int hi_k = (int)(s1 - 4 * (q1 - 1));       // valid lanes of the last group: [0, hi_k)
    unsigned int mig = 0;
    // Each pass handles U groups per thread.  With PIPE the loads of the next
    // pass are issued before this pass's gathers, pushes and deposits, so each
    // wave keeps a pass of loads in flight while it computes: 4% faster at
    // ~24 particles per cell, where the kernel is bound by HBM traffic, and
    // 6-9% slower at ~240 per cell, where LDS atomic contention needs the
    // third block per CU.
    constexpr int U = 2;
    v4f nx[U], ny[U], nvx[U], nvy[U], nm[U];
    auto load_pass = [&](int j0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * BLK;
            nx[u] = v4f{-1.f, -1.f, -1.f, -1.f};
            ny[u] = nx[u];
            nvx[u] = v4f{0.f, 0.f, 0.f, 0.f};
            nvy[u] = nvx[u];
            nm[u] = nvx[u];
            if (j < nq) {
                nx[u] = __builtin_nontemporal_load(X + j);
                ny[u] = __builtin_nontemporal_load(Y + j);
                nvx[u] = __builtin_nontemporal_load(VX + j);
                nvy[u] = __builtin_nontemporal_load(VY + j);
                nm[u] = __builtin_nontemporal_load(M + j);
            }
        }
    };
    if (PIPE) load_pass(threadIdx.x);
    for (int j0 = threadIdx.x; j0 < nq; j0 += U * BLK) {
        float px[U][4], py[U][4], ux[U][4], uy[U][4], pm[U][4];
        bool ok[U][4];
        v4f cx[U], cy[U], cvx[U], cvy[U], cm[U];
        if (!PIPE) load_pass(j0);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            cx[u] = nx[u];
            cy[u] = ny[u];
            cvx[u] = nvx[u];
            cvy[u] = nvy[u];
            cm[u] = nm[u];
        }
        if (PIPE) load_pass(j0 + U * BLK);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * BLK;
            const v4f a = cx[u], b = cy[u], c = cvx[u], d = cvy[u], e = cm[u];
            px[u][0] = a.x; px[u][1] = a.y; px[u][2] = a.z; px[u][3] = a.w;
            py[u][0] = b.x; py[u][1] = b.y; py[u][2] = b.z; py[u][3] = b.w;
            ux[u][0] = c.x; ux[u][1] = c.y; ux[u][2] = c.z; ux[u][3] = c.w;
            uy[u][0] = d.x; uy[u][1] = d.y; uy[u][2] = d.z; uy[u][3] = d.w;
            pm[u][0] = e.x; pm[u][1] = e.y; pm[u][2] = e.z; pm[u][3] = e.w;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                ok[u][k] = j < nq && (j > 0 || k >= lo_k) && (j < nq - 1 || k < hi_k);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!ok[u][k] || !in_bounds(px[u][k], py[u][k], fhi)) continue;
                const int xi = (int)px[u][k], yi = (int)py[u][k];
                const int lx = xi - wx0, ly = yi - wy0;
                const float fx = frac(px[u][k]), fy = frac(py[u][k]);
                float2 a;
                if (!sparse && (unsigned)lx < (unsigned)(W - 1) && (unsigned)ly < (unsigned)(W - 1))
                    a = gather<MODE>(fwin + ly * W + lx, W, fx, fy);
                else
                    a = gather<MODE>(force + (size_t)yi * n + xi, n, fx, fy);
                kick_drift<WRAP>(a, n, dt, px[u][k], py[u][k], ux[u][k], uy[u][k]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * BLK;
            if (j >= nq) continue;
            if (ok[u][0] && ok[u][3]) {
                __builtin_nontemporal_store(v4f{px[u][0], px[u][1], px[u][2], px[u][3]}, X + j);
                __builtin_nontemporal_store(v4f{py[u][0], py[u][1], py[u][2], py[u][3]}, Y + j);
                __builtin_nontemporal_store(v4f{ux[u][0], ux[u][1], ux[u][2], ux[u][3]}, VX + j);
                __builtin_nontemporal_store(v4f{uy[u][0], uy[u][1], uy[u][2], uy[u][3]}, VY + j);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (!ok[u][k]) continue;
                    float* sx = reinterpret_cast<float*>(X + j) + k;
                    float* sy = reinterpret_cast<float*>(Y + j) + k;
                    float* svx = reinterpret_cast<float*>(VX + j) + k;
                    float* svy = reinterpret_cast<float*>(VY + j) + k;
                    *sx = px[u][k];
                    *sy = py[u][k];
                    *svx = ux[u][k];
                    *svy = uy[u][k];
                }
            }
        }
        bool spill[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                spill[u][k] = false;
                if (!ok[u][k]) continue;
                if (COUNT) wave_count(hist, bin_of(px[u][k], py[u][k], n, TL));
                spill[u][k] = in_bounds(px[u][k], py[u][k], fhi) &&
                              (sparse || !deposit_local<MODE, W>(win, wx0, wy0, px[u][k], py[u][k], pm[u][k], scale));
            }
        }
        // (a sparse chunk deposits everything globally; its migrants are the
        // particles outside the window, as in a dense chunk)
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (spill[u][k]) {
                    deposit_global<MODE>(rho, n, px[u][k], py[u][k], pm[u][k], scale);
                    mig += !sparse || !in_window<W>(wx0, wy0, px[u][k], py[u][k]);
                }
    }
    if (mig) atomicAdd(&s_mig, mig);
    lds_barrier();
    if (!sparse) flush_window<W, BLK>(win, wx0, wy0, rho, n);
    if (COUNT)
        for (int e = threadIdx.x; e <= nin; e += BLK)
            if (hist[e]) atomicAdd(&count[e], hist[e]);
    if (threadIdx.x == 0 && s_mig) atomicAdd(migrants, (unsigned long long)s_mig);
}

template <int TL, int MODE, int WRAP, bool COUNT, bool PIPE>
__global__ __launch_bounds__(fused_block<TL>()) void k_push_deposit(
    const float2* __restrict__ force, float* __restrict__ x, float* __restrict__ y, float* __restrict__ vx,
    float* __restrict__ vy, const float* __restrict__ m, int n, float dt, const int64_t* __restrict__ start,
    const int* __restrict__ chunk_start, float scale, unsigned long long* __restrict__ rho,
    unsigned long long* __restrict__ migrants, int* __restrict__ count) {
    push_deposit_block<TL, MODE, WRAP, COUNT, PIPE>(force, x, y, vx, vy, m, n, dt, start, chunk_start, scale, rho,
                                                    migrants, count, blockIdx.x);
}

// Most blocks of a batched step launch (the blocks stride over the members'
// chunks): about four per resident block slot at two per CU.
constexpr int kBatchedGrid = 2048;

// Simulation of chunk b of a batched launch: the last i with
// chunk_first[i] <= b (chunk_first: exclusive prefix of the members' chunk
// counts, b below their total).
__device__ __forceinline__ int batch_member(const int* __restrict__ chunk_first, int count, int b) {
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (chunk_first[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// The step kernel over several simulations of one lattice size, deposit mode
// and wrap mode (a batched serving round).  The chunks of all members are
// numbered consecutively (chunk_first, filled on the device from the members'
// chunk maps by k_member_chunks) and the blocks stride over them: the grid
// holds at most kBatchedGrid blocks whatever the members' worst-case chunk
// counts, and no block is spent on chunks that do not exist.
// p: the step's parity (which mass lattice is the deposit target).
template <int TL, int MODE, int WRAP, bool COUNT>
__global__ __launch_bounds__(fused_block<TL>()) void k_push_deposit_batched(const TiledDesc* __restrict__ descs,
                                                                           const int* __restrict__ chunk_first,
                                                                           int count, int p) {
    const int total = chunk_first[count];
    for (int b = blockIdx.x; b < total; b += gridDim.x) {
        const int i = batch_member(chunk_first, count, b);
        const TiledDesc& d = descs[i];
        push_deposit_block<TL, MODE, WRAP, COUNT, false>(d.force, d.cur.x, d.cur.y, d.cur.vx, d.cur.vy, d.cur.m, d.n,
                                                        d.dt, d.start, d.chunk_start, d.scale, d.rho_next[p],
                                                        d.migrants, d.next_count, b - chunk_first[i]);
        lds_barrier();                               // the windows are reused by the block's next chunk
    }
}

// --------------------------------------------- fused step kernel with re-sort --
// The re-sort fused into a step: the chunk's particles (read from `src` in the
// order of the current chunk map) are pushed and deposited as above, then
// written -- with mass and load index -- to `dst`, grouped by the bin of
// their position at the start of the step: slot = next_start[bin] + the
// block's base (one global atomic per block, pass and bin, taken by the
// lane of rank 0) + the lane's rank in LDS.  The bins were counted during the
// previous step, so `dst` ends in the order of this step's start positions,
// one step stale, well inside the kHalo margin of the tile windows.
template <int TL, int MODE, int WRAP>
__device__ __forceinline__ void push_deposit_scatter_block(
    const float2* __restrict__ force, const ParticleArrays& src, const ParticleArrays& dst, int n, float dt,
    const int64_t* __restrict__ start, const int* __restrict__ chunk_start, const int64_t* __restrict__ next_start,
    int* __restrict__ next_cursor, float scale, unsigned long long* __restrict__ rho,
    unsigned long long* __restrict__ migrants, int c) {
    constexpr int W = TileGeom<TL>::kWin, E = TileGeom<TL>::kEdge, BLK = fused_block<TL>();
    __shared__ unsigned long long win[W * W];
    __shared__ float2 fwin[W * W];
    __shared__ unsigned int s_mig;
    extern __shared__ int lds[];
    const int tpr = n >> TL, nin = tpr * tpr;
    int* lcnt = lds;             // [nin + 1]
    int* lbase = lds + nin + 1;  // [nin + 1]
    int b;
    int64_t s0, s1;
    if (!chunk_of(chunk_start, start, nin, c, b, s0, s1)) return;
    const bool sparse = s1 - s0 < kSparseChunk;       // gather and deposit through global memory
    const int wx0 = (b % tpr) * E - kHalo, wy0 = (b / tpr) * E - kHalo;
    if (!sparse)
        for (int e = threadIdx.x; e < W * W; e += BLK) {
            win[e] = 0ull;
            const int gx = wx0 + e % W, gy = wy0 + e / W;
            fwin[e] = ((unsigned)gx < (unsigned)n && (unsigned)gy < (unsigned)n) ? force[(size_t)gy * n + gx]
                                                                                 : make_float2(0.f, 0.f);
        }
    for (int e = threadIdx.x; e <= nin; e += BLK) lcnt[e] = 0;
    if (threadIdx.x == 0) s_mig = 0;
    lds_barrier();
    const float fhi = (float)(n - 1);
    const int64_t q0 = s0 >> 2, q1 = (s1 + 3) >> 2;
    const int nq = (int)(q1 - q0);
    const v4f* X = reinterpret_cast<const v4f*>(src.x) + q0;
    const v4f* Y = reinterpret_cast<const v4f*>(src.y) + q0;
    const v4f* VX = reinterpret_cast<const v4f*>(src.vx) + q0;
    const v4f* VY = reinterpret_cast<const v4f*>(src.vy) + q0;
    const v4f* M = reinterpret_cast<const v4f*>(src.m) + q0;
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    const v4u* ID = reinterpret_cast<const v4u*>(src.id) + q0;
    const int lo_k = (int)(s0 - 4 * q0);
    const int hi_k = (int)(s1 - 4 * (q1 - 1));
    unsigned int mig = 0;
    constexpr int U = 2;
    for (int base = 0; base < nq; base += U * BLK) {     // same trip count in every thread (barriers inside)
        float px[U][4], py[U][4], ux[U][4], uy[U][4], pm[U][4];
        unsigned pid[U][4];
        int bn[U][4], rk[U][4];
        bool ok[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = base + u * BLK + threadIdx.x;
            v4f a = {-1.f, -1.f, -1.f, -1.f}, bb = a, c = {0.f, 0.f, 0.f, 0.f}, d = c, e = c;
            v4u id = {0u, 0u, 0u, 0u};
            if (j < nq) {
                a = __builtin_nontemporal_load(X + j);
                bb = __builtin_nontemporal_load(Y + j);
                c = __builtin_nontemporal_load(VX + j);
                d = __builtin_nontemporal_load(VY + j);
                e = __builtin_nontemporal_load(M + j);
                id = __builtin_nontemporal_load(ID + j);
            }
            px[u][0] = a.x; px[u][1] = a.y; px[u][2] = a.z; px[u][3] = a.w;
            py[u][0] = bb.x; py[u][1] = bb.y; py[u][2] = bb.z; py[u][3] = bb.w;
            ux[u][0] = c.x; ux[u][1] = c.y; ux[u][2] = c.z; ux[u][3] = c.w;
            uy[u][0] = d.x; uy[u][1] = d.y; uy[u][2] = d.z; uy[u][3] = d.w;
            pm[u][0] = e.x; pm[u][1] = e.y; pm[u][2] = e.z; pm[u][3] = e.w;
            pid[u][0] = id.x; pid[u][1] = id.y; pid[u][2] = id.z; pid[u][3] = id.w;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ok[u][k] = j < nq && (j > 0 || k >= lo_k) && (j < nq - 1 || k < hi_k);
                bn[u][k] = -1;
                if (ok[u][k]) {
                    bn[u][k] = bin_of(px[u][k], py[u][k], n, TL);    // bin of the start position
                    rk[u][k] = wave_count_add(lcnt, bn[u][k]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!ok[u][k] || !in_bounds(px[u][k], py[u][k], fhi)) continue;
                const int xi = (int)px[u][k], yi = (int)py[u][k];
                const int lx = xi - wx0, ly = yi - wy0;
                const float fx = frac(px[u][k]), fy = frac(py[u][k]);
                float2 a;
                if (!sparse && (unsigned)lx < (unsigned)(W - 1) && (unsigned)ly < (unsigned)(W - 1))
                    a = gather<MODE>(fwin + ly * W + lx, W, fx, fy);
                else
                    a = gather<MODE>(force + (size_t)yi * n + xi, n, fx, fy);
                kick_drift<WRAP>(a, n, dt, px[u][k], py[u][k], ux[u][k], uy[u][k]);
            }
        }
        bool spill[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                spill[u][k] = ok[u][k] && in_bounds(px[u][k], py[u][k], fhi) &&
                              (sparse || !deposit_local<MODE, W>(win, wx0, wy0, px[u][k], py[u][k], pm[u][k], scale));
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (spill[u][k]) {
                    deposit_global<MODE>(rho, n, px[u][k], py[u][k], pm[u][k], scale);
                    mig += !sparse || !in_window<W>(wx0, wy0, px[u][k], py[u][k]);
                }
        lds_barrier();                                   // ranks of this pass complete
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (bn[u][k] >= 0 && rk[u][k] == 0) lbase[bn[u][k]] = atomicAdd(&next_cursor[bn[u][k]], lcnt[bn[u][k]]);
        lds_barrier();                                   // bases known
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (bn[u][k] < 0) continue;
                const int64_t sl = next_start[bn[u][k]] + lbase[bn[u][k]] + rk[u][k];
                dst.x[sl] = px[u][k];
                dst.y[sl] = py[u][k];
                dst.vx[sl] = ux[u][k];
                dst.vy[sl] = uy[u][k];
                dst.m[sl] = pm[u][k];
                dst.id[sl] = pid[u][k];
            }
        lds_barrier();                                   // counts read: reset for the next pass
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (bn[u][k] >= 0 && rk[u][k] == 0) lcnt[bn[u][k]] = 0;
        lds_barrier();
    }
    if (mig) atomicAdd(&s_mig, mig);
    lds_barrier();
    if (!sparse) flush_window<W, BLK>(win, wx0, wy0, rho, n);
    if (threadIdx.x == 0 && s_mig) atomicAdd(migrants, (unsigned long long)s_mig);
}

template <int TL, int MODE, int WRAP>
__global__ __launch_bounds__(fused_block<TL>()) void k_push_deposit_scatter(
    const float2* __restrict__ force, ParticleArrays src, ParticleArrays dst, int n, float dt,
    const int64_t* __restrict__ start, const int* __restrict__ chunk_start, const int64_t* __restrict__ next_start,
    int* __restrict__ next_cursor, float scale, unsigned long long* __restrict__ rho,
    unsigned long long* __restrict__ migrants) {
    push_deposit_scatter_block<TL, MODE, WRAP>(force, src, dst, n, dt, start, chunk_start, next_start, next_cursor,
                                               scale, rho, migrants, blockIdx.x);
}

// The re-sorting step over several simulations (a batched serving round);
// block mapping as k_push_deposit_batched.
template <int TL, int MODE, int WRAP>
__global__ __launch_bounds__(fused_block<TL>()) void k_push_deposit_scatter_batched(
    const TiledDesc* __restrict__ descs, const int* __restrict__ chunk_first, int count, int p) {
    const int total = chunk_first[count];
    for (int b = blockIdx.x; b < total; b += gridDim.x) {
        const int i = batch_member(chunk_first, count, b);
        const TiledDesc& d = descs[i];
        push_deposit_scatter_block<TL, MODE, WRAP>(d.force, d.cur, d.alt, d.n, d.dt, d.start, d.chunk_start,
                                                   d.next_start, d.next_cursor, d.scale, d.rho_next[p], d.migrants,
                                                   b - chunk_first[i]);
        lds_barrier();                               // LDS tables and windows are reused by the next chunk
    }
}

// chunk_first[i] = chunks of members 0 .. i-1 in their current chunk maps
// (chunk_start[nin] of each), chunk_first[count] = the total: one block.
__global__ __launch_bounds__(kBlock) void k_member_chunks(const TiledDesc* __restrict__ descs, int count,
                                                          int* __restrict__ chunk_first) {
    __shared__ int sc[kBlock];
    int off = 0;
    for (int base = 0; base < count; base += kBlock) {
        const int k = base + threadIdx.x;
        const int v = k < count ? descs[k].chunk_start[descs[k].nbins - 1] : 0;
        sc[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < kBlock; o <<= 1) {       // inclusive scan
            const int t = threadIdx.x >= o ? sc[threadIdx.x - o] : 0;
            __syncthreads();
            sc[threadIdx.x] += t;
            __syncthreads();
        }
        if (k < count) chunk_first[k] = off + sc[threadIdx.x] - v;
        off += sc[kBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) chunk_first[count] = off;
}

// Particles standing in the parked bin (never visited by the step kernels)
// move to the front of the parked range of `dst`; the scan started that
// bin's cursor after them.
__global__ __launch_bounds__(kBlock) void k_copy_parked(ParticleArrays src, ParticleArrays dst,
                                                         const int64_t* __restrict__ old_start,
                                                         const int64_t* __restrict__ new_start, int parked) {
    const int64_t a = old_start[parked], cnt = old_start[parked + 1] - a, d = new_start[parked];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * kBlock) {
        dst.x[d + i] = src.x[a + i];
        dst.y[d + i] = src.y[a + i];
        dst.vx[d + i] = src.vx[a + i];
        dst.vy[d + i] = src.vy[a + i];
        dst.m[d + i] = src.m[a + i];
        dst.id[d + i] = src.id[a + i];
    }
}

// ----------------------------------------------------------- plain push --
// Four consecutive particles per lane: 16-byte nontemporal loads and stores
// of x, y, vx, vy; force from the global lattice.  Used by the single-stage
// entry point (the stepping path uses the fused kernel).
template <int MODE, int WRAP>
__global__ __launch_bounds__(kBlock) void k_push4(const float2* __restrict__ force, v4f* __restrict__ x,
                                                   v4f* __restrict__ y, v4f* __restrict__ vx,
                                                   v4f* __restrict__ vy, int64_t n4, int n, float dt) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= n4) return;
    v4f X = __builtin_nontemporal_load(x + q);
    v4f Y = __builtin_nontemporal_load(y + q);
    v4f VX = __builtin_nontemporal_load(vx + q);
    v4f VY = __builtin_nontemporal_load(vy + q);
    float px[4] = {X.x, X.y, X.z, X.w}, py[4] = {Y.x, Y.y, Y.z, Y.w};
    float ux[4] = {VX.x, VX.y, VX.z, VX.w}, uy[4] = {VY.x, VY.y, VY.z, VY.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) push_one<MODE, WRAP>(force, n, dt, px[k], py[k], ux[k], uy[k]);
    __builtin_nontemporal_store(v4f{px[0], px[1], px[2], px[3]}, x + q);
    __builtin_nontemporal_store(v4f{py[0], py[1], py[2], py[3]}, y + q);
    __builtin_nontemporal_store(v4f{ux[0], ux[1], ux[2], ux[3]}, vx + q);
    __builtin_nontemporal_store(v4f{uy[0], uy[1], uy[2], uy[3]}, vy + q);
}

// Remainder particles [first, np) one per lane.
template <int MODE, int WRAP>
__global__ __launch_bounds__(kBlock) void k_push_tail(const float2* __restrict__ force, float* __restrict__ x,
                                                       float* __restrict__ y, float* __restrict__ vx,
                                                       float* __restrict__ vy, int64_t first, int64_t np, int n,
                                                       float dt) {
    const int64_t p = first + threadIdx.x;
    if (p >= np) return;
    float px = x[p], py = y[p], ux = vx[p], uy = vy[p];
    push_one<MODE, WRAP>(force, n, dt, px, py, ux, uy);
    x[p] = px;
    y[p] = py;
    vx[p] = ux;
    vy[p] = uy;
}

// ------------------------------------------------- batched direct push + deposit --
// Block -> simulation of the batch by binary search over block_start; the
// work of a block is direct_push_group (pm_particle_core.h).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_push_deposit_direct(const PartDesc* __restrict__ descs,
                                                                 const int* __restrict__ block_start, int count) {
    int lo = 0, hi = count;                  // last i with block_start[i] <= blockIdx.x
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (block_start[mid] <= (int)blockIdx.x) lo = mid; else hi = mid;
    }
    direct_push_group<MODE>(descs[lo], (int64_t)(blockIdx.x - block_start[lo]) * kBlock + threadIdx.x);
}

// Standing parked particles of simulation blockIdx.y of a batched round (as
// k_copy_parked: from the parked range of cur to the front of alt's).
__global__ __launch_bounds__(kBlock) void k_copy_parked_batched(const TiledDesc* __restrict__ descs) {
    const TiledDesc& d = descs[blockIdx.y];
    const int parked = d.nbins - 1;
    const int64_t a = d.start[parked], cnt = d.start[parked + 1] - a, o = d.next_start[parked];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * kBlock) {
        d.alt.x[o + i] = d.cur.x[a + i];
        d.alt.y[o + i] = d.cur.y[a + i];
        d.alt.vx[o + i] = d.cur.vx[a + i];
        d.alt.vy[o + i] = d.cur.vy[a + i];
        d.alt.m[o + i] = d.cur.m[a + i];
        d.alt.id[o + i] = d.cur.id[a + i];
    }
}

// Batched launchers: dynamic LDS as the single-simulation launchers (the bin
// histogram of the counting step; the scatter's two bin tables).
template <int TL, int MODE, int WRAP, bool COUNT>
hipError_t push_deposit_batched_k(int n, const TiledDesc* d, const int* chunk_first, int count, int total_blocks,
                                  int p, hipStream_t s) {
    constexpr int W = TileGeom<TL>::kWin, BLK = fused_block<TL>();
    const size_t lds = COUNT ? sizeof(int) * ((n >> TL) * (n >> TL) + 1) : 0;
    if (COUNT && lds + 16 * W * W > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_push_deposit_batched<TL, MODE, WRAP, COUNT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_push_deposit_batched<TL, MODE, WRAP, COUNT>), dim3(std::min(total_blocks, kBatchedGrid)),
                       dim3(BLK), lds, s, d, chunk_first, count, p);
    return hipGetLastError();
}

template <int TL>
hipError_t push_deposit_batched_mode(int n, int mode, int wrap_mode, bool count_bins, const TiledDesc* d,
                                     const int* chunk_first, int count, int total_blocks, int p, hipStream_t s) {
    if (mode == 0 && wrap_mode == 0)
        return count_bins ? push_deposit_batched_k<TL, 0, 0, true>(n, d, chunk_first, count, total_blocks, p, s)
                          : push_deposit_batched_k<TL, 0, 0, false>(n, d, chunk_first, count, total_blocks, p, s);
    if (mode == 0)
        return count_bins ? push_deposit_batched_k<TL, 0, 1, true>(n, d, chunk_first, count, total_blocks, p, s)
                          : push_deposit_batched_k<TL, 0, 1, false>(n, d, chunk_first, count, total_blocks, p, s);
    if (wrap_mode == 0)
        return count_bins ? push_deposit_batched_k<TL, 1, 0, true>(n, d, chunk_first, count, total_blocks, p, s)
                          : push_deposit_batched_k<TL, 1, 0, false>(n, d, chunk_first, count, total_blocks, p, s);
    return count_bins ? push_deposit_batched_k<TL, 1, 1, true>(n, d, chunk_first, count, total_blocks, p, s)
                      : push_deposit_batched_k<TL, 1, 1, false>(n, d, chunk_first, count, total_blocks, p, s);
}

template <int TL, int MODE, int WRAP>
hipError_t scatter_batched_t(int n, const TiledDesc* d, const int* chunk_first, int count, int total_blocks, int p,
                             hipStream_t s) {
    const size_t lds = 2 * sizeof(int) * (size_t)(((n >> TL) * (n >> TL)) + 1);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_push_deposit_scatter_batched<TL, MODE, WRAP>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_push_deposit_scatter_batched<TL, MODE, WRAP>), dim3(std::min(total_blocks, kBatchedGrid)),
                       dim3(fused_block<TL>()), lds, s, d, chunk_first, count, p);
    return hipGetLastError();
}

template <int TL>
hipError_t scatter_batched_mode(int n, int mode, int wrap, const TiledDesc* d, const int* chunk_first, int count,
                                int total_blocks, int p, hipStream_t s) {
    if (mode == 0 && wrap == 0) return scatter_batched_t<TL, 0, 0>(n, d, chunk_first, count, total_blocks, p, s);
    if (mode == 0) return scatter_batched_t<TL, 0, 1>(n, d, chunk_first, count, total_blocks, p, s);
    if (wrap == 0) return scatter_batched_t<TL, 1, 0>(n, d, chunk_first, count, total_blocks, p, s);
    return scatter_batched_t<TL, 1, 1>(n, d, chunk_first, count, total_blocks, p, s);
}

// ------------------------------------------------------------------ launchers --
template <int TL>
hipError_t deposit_t(int n, int mode, const float* x, const float* y, const float* m, BinState bs, int max_chunks,
                     float scale, unsigned long long* r, unsigned long long* migrants, hipStream_t s) {
    if (mode == 0)
        hipLaunchKernelGGL((k_deposit_bin<TL, 0>), dim3(max_chunks), dim3(kDepositBlock), 0, s, x, y, m, n, bs.start,
                           bs.chunk_start, bs.chunk, scale, r, migrants);
    else
        hipLaunchKernelGGL((k_deposit_bin<TL, 1>), dim3(max_chunks), dim3(kDepositBlock), 0, s, x, y, m, n, bs.start,
                           bs.chunk_start, bs.chunk, scale, r, migrants);
    return hipGetLastError();
}

template <int TL, int MODE, int WRAP, bool COUNT, bool PIPE>
hipError_t push_deposit_k(int n, const float2* force, ParticleArrays p, BinState bs, int max_chunks, float dt,
                          float scale, unsigned long long* r, unsigned long long* migrants, hipStream_t s) {
    constexpr int W = TileGeom<TL>::kWin, BLK = fused_block<TL>();
    const size_t lds = COUNT ? sizeof(int) * ((n >> TL) * (n >> TL) + 1) : 0;
    if (COUNT && lds + 16 * W * W > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_push_deposit<TL, MODE, WRAP, COUNT, PIPE>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_push_deposit<TL, MODE, WRAP, COUNT, PIPE>), dim3(max_chunks), dim3(BLK), lds, s, force,
                       p.x, p.y, p.vx, p.vy, p.m, n, dt, bs.start, bs.chunk_start, scale, r, migrants,
                       bs.count);
    return hipGetLastError();
}

template <int TL, int MODE, int WRAP>
hipError_t push_deposit_t(int n, const float2* force, ParticleArrays p, BinState bs, int max_chunks, float dt,
                          float scale, unsigned long long* r, unsigned long long* migrants, bool count_bins,
                          bool pipelined, hipStream_t s) {
    if (count_bins)
        return pipelined ? push_deposit_k<TL, MODE, WRAP, true, true>(n, force, p, bs, max_chunks, dt, scale, r, migrants, s)
                         : push_deposit_k<TL, MODE, WRAP, true, false>(n, force, p, bs, max_chunks, dt, scale, r, migrants, s);
    return pipelined ? push_deposit_k<TL, MODE, WRAP, false, true>(n, force, p, bs, max_chunks, dt, scale, r, migrants, s)
                     : push_deposit_k<TL, MODE, WRAP, false, false>(n, force, p, bs, max_chunks, dt, scale, r, migrants, s);
}

template <int TL>
hipError_t push_deposit_mode(int n, int mode, int wrap_mode, const float2* force, ParticleArrays p, BinState bs,
                             int max_chunks, float dt, float scale, unsigned long long* r,
                             unsigned long long* migrants, bool count_bins, bool pipelined, hipStream_t s) {
    if (mode == 0 && wrap_mode == 0)
        return push_deposit_t<TL, 0, 0>(n, force, p, bs, max_chunks, dt, scale, r, migrants, count_bins, pipelined, s);
    if (mode == 0)
        return push_deposit_t<TL, 0, 1>(n, force, p, bs, max_chunks, dt, scale, r, migrants, count_bins, pipelined, s);
    if (wrap_mode == 0)
        return push_deposit_t<TL, 1, 0>(n, force, p, bs, max_chunks, dt, scale, r, migrants, count_bins, pipelined, s);
    return push_deposit_t<TL, 1, 1>(n, force, p, bs, max_chunks, dt, scale, r, migrants, count_bins, pipelined, s);
}

template <int MODE, int WRAP>
hipError_t push_t(const float2* force, float* x, float* y, float* vx, float* vy, int64_t np, int n, float dt,
                  hipStream_t s) {
    const int64_t n4 = np / 4;
    if (n4 > 0) {
        const int64_t blocks = (n4 + kBlock - 1) / kBlock;
        hipLaunchKernelGGL((k_push4<MODE, WRAP>), dim3((unsigned)blocks), dim3(kBlock), 0, s, force,
                           reinterpret_cast<v4f*>(x), reinterpret_cast<v4f*>(y), reinterpret_cast<v4f*>(vx),
                           reinterpret_cast<v4f*>(vy), n4, n, dt);
    }
    if (np > 4 * n4)
        hipLaunchKernelGGL((k_push_tail<MODE, WRAP>), dim3(1), dim3(kBlock), 0, s, force, x, y, vx, vy, 4 * n4, np, n,
                           dt);
    return hipGetLastError();
}

}  // namespace

int deposit_max_chunks(int n, int tl, int64_t capacity, int chunk) {
    const int64_t nin = (int64_t)(n >> tl) * (n >> tl);
    return (int)((capacity + chunk - 1) / chunk + nin);
}

hipError_t deposit(int n, int mode, const float* x, const float* y, const float* m, BinState bs, int max_chunks,
                   float scale, int64_t* rho, unsigned long long* migrants, hipStream_t s) {
    auto* r = reinterpret_cast<unsigned long long*>(rho);
    if (bs.tl == 5) return deposit_t<5>(n, mode, x, y, m, bs, max_chunks, scale, r, migrants, s);
    return deposit_t<6>(n, mode, x, y, m, bs, max_chunks, scale, r, migrants, s);
}

hipError_t push_deposit(int n, int mode, int wrap_mode, const float2* force, ParticleArrays p, BinState bs,
                        int max_chunks, float dt, float scale, int64_t* rho_next, unsigned long long* migrants,
                        bool count_bins, bool pipelined, hipStream_t s) {
    auto* r = reinterpret_cast<unsigned long long*>(rho_next);
    if (bs.tl == 5)
        return push_deposit_mode<5>(n, mode, wrap_mode, force, p, bs, max_chunks, dt, scale, r, migrants, count_bins,
                                    pipelined, s);
    return push_deposit_mode<6>(n, mode, wrap_mode, force, p, bs, max_chunks, dt, scale, r, migrants, count_bins,
                                pipelined, s);
}

int direct_blocks(int64_t np) { return (int)((np + 4 * kBlock - 1) / (4 * kBlock)); }

hipError_t push_deposit_direct(int mode, const PartDesc* d, const int* block_start, int count, int total_blocks,
                               hipStream_t s) {
    if (total_blocks <= 0 || count <= 0) return hipSuccess;
    if (mode == 0)
        hipLaunchKernelGGL((k_push_deposit_direct<0>), dim3(total_blocks), dim3(kBlock), 0, s, d, block_start, count);
    else
        hipLaunchKernelGGL((k_push_deposit_direct<1>), dim3(total_blocks), dim3(kBlock), 0, s, d, block_start, count);
    return hipGetLastError();
}

template <int TL, int MODE, int WRAP>
hipError_t scatter_t(int n, const float2* force, ParticleArrays src, ParticleArrays dst, BinState cur,
                     BinState next, int max_chunks, float dt, float scale, unsigned long long* r,
                     unsigned long long* migrants, hipStream_t s) {
    const size_t lds = 2 * sizeof(int) * (size_t)(((n >> TL) * (n >> TL)) + 1);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_push_deposit_scatter<TL, MODE, WRAP>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_push_deposit_scatter<TL, MODE, WRAP>), dim3(max_chunks), dim3(fused_block<TL>()), lds, s,
                       force, src, dst, n, dt, cur.start, cur.chunk_start, next.start, next.cursor, scale, r,
                       migrants);
    return hipGetLastError();
}

template <int TL>
hipError_t scatter_mode(int n, int mode, int wrap, const float2* force, ParticleArrays src, ParticleArrays dst,
                        BinState cur, BinState next, int max_chunks, float dt, float scale, unsigned long long* r,
                        unsigned long long* migrants, hipStream_t s) {
    if (mode == 0 && wrap == 0) return scatter_t<TL, 0, 0>(n, force, src, dst, cur, next, max_chunks, dt, scale, r, migrants, s);
    if (mode == 0) return scatter_t<TL, 0, 1>(n, force, src, dst, cur, next, max_chunks, dt, scale, r, migrants, s);
    if (wrap == 0) return scatter_t<TL, 1, 0>(n, force, src, dst, cur, next, max_chunks, dt, scale, r, migrants, s);
    return scatter_t<TL, 1, 1>(n, force, src, dst, cur, next, max_chunks, dt, scale, r, migrants, s);
}

bool scatter_fits(int n, int tl) {
    constexpr size_t kLds = 160 * 1024;
    const size_t bins = (size_t)(n >> tl) * (n >> tl) + 1;
    const size_t w = tl == 5 ? TileGeom<5>::kWin : TileGeom<6>::kWin;
    return 2 * sizeof(int) * bins + 16 * w * w + 64 <= kLds;
}

hipError_t push_deposit_scatter(int n, int mode, int wrap_mode, const float2* force, ParticleArrays src,
                                ParticleArrays dst, BinState cur, BinState next, int max_chunks, float dt,
                                float scale, int64_t* rho_next, unsigned long long* migrants, hipStream_t s) {
    auto* r = reinterpret_cast<unsigned long long*>(rho_next);
    hipError_t e;
    if (cur.tl == 5) e = scatter_mode<5>(n, mode, wrap_mode, force, src, dst, cur, next, max_chunks, dt, scale, r, migrants, s);
    else e = scatter_mode<6>(n, mode, wrap_mode, force, src, dst, cur, next, max_chunks, dt, scale, r, migrants, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_copy_parked, dim3(64), dim3(kBlock), 0, s, src, dst, cur.start, next.start,
                       (n >> cur.tl) * (n >> cur.tl));
    return hipGetLastError();
}

hipError_t push(int n, int mode, const float2* force, float* x, float* y, float* vx, float* vy, int64_t np,
                float dt, int wrap_mode, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    if (mode == 0 && wrap_mode == 0) return push_t<0, 0>(force, x, y, vx, vy, np, n, dt, s);
    if (mode == 0) return push_t<0, 1>(force, x, y, vx, vy, np, n, dt, s);
    if (wrap_mode == 0) return push_t<1, 0>(force, x, y, vx, vy, np, n, dt, s);
    return push_t<1, 1>(force, x, y, vx, vy, np, n, dt, s);
}

hipError_t push_deposit_batched(int n, int tl, int mode, int wrap_mode, bool count_bins, const TiledDesc* d,
                                const int* chunk_first, int count, int total_blocks, int p, hipStream_t s) {
    if (count <= 0 || total_blocks <= 0) return hipSuccess;
    if (tl == 5)
        return push_deposit_batched_mode<5>(n, mode, wrap_mode, count_bins, d, chunk_first, count, total_blocks, p, s);
    return push_deposit_batched_mode<6>(n, mode, wrap_mode, count_bins, d, chunk_first, count, total_blocks, p, s);
}

hipError_t member_chunks_batched(const TiledDesc* d, int count, int* chunk_first, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_member_chunks, dim3(1), dim3(kBlock), 0, s, d, count, chunk_first);
    return hipGetLastError();
}

hipError_t push_deposit_scatter_batched(int n, int tl, int mode, int wrap_mode, const TiledDesc* d,
                                        const int* chunk_first, int count, int total_blocks, int p, hipStream_t s) {
    if (count <= 0 || total_blocks <= 0) return hipSuccess;
    hipError_t e = tl == 5 ? scatter_batched_mode<5>(n, mode, wrap_mode, d, chunk_first, count, total_blocks, p, s)
                           : scatter_batched_mode<6>(n, mode, wrap_mode, d, chunk_first, count, total_blocks, p, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_copy_parked_batched, dim3(64, count), dim3(kBlock), 0, s, d);
    return hipGetLastError();
}

}  // namespace pmk
