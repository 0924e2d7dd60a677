// Spatial ordering of the particle arrays by lattice tile.
//
// Particles are kept physically sorted by the 64x64-cell tile ("bin") that
// holds them, refreshed by a counting sort every few steps.  A deposit block
// then owns a contiguous chunk of one bin and accumulates into an LDS window,
// and the force gather of consecutive particles hits the same few cache lines.
// The order of particles inside a bin is arbitrary (atomic slot claims); it
// does not affect results because the deposit is an integer sum and every
// other kernel is per particle.  id[slot] keeps each particle's load index so
// reads return the caller's order.
//
// Bins: b = (y/64) * (N/64) + x/64 for in-bounds particles (1 <= x,y < N-1,
// the reference's rule, cuh:257, 497), b = nbins for the rest ("parked": never
// deposited, never moved).
#include "pm_internal.h"

namespace pmk {
namespace {

__global__ __launch_bounds__(kBlock) void k_bin_count(const float* __restrict__ x, const float* __restrict__ y,
                                                       int64_t np, int n, int tl, int* __restrict__ count) {
    extern __shared__ int hist[];
    const int tpr = n >> tl, nb = tpr * tpr + 1;
    for (int i = threadIdx.x; i < nb; i += kBlock) hist[i] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < np; p += stride)
        atomicAdd(&hist[bin_of(x[p], y[p], n, tl)], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += kBlock)
        if (hist[i]) atomicAdd(&count[i], hist[i]);
}

// One block: exclusive scans of the bin counts (-> start[0..nb]) and of the
// deposit chunk counts of the in-bounds bins (-> chunk_start[0..nb-1]); a
// bin's chunks split it into equal parts.
// Zeroes the scatter cursors.  add_parked: the counts came from the fused
// push, which never visits the parked range; add that range's standing size
// (from the previous start[]) to the parked bin first.
constexpr int kScanBlock = 1024;
// Chunks of an in-bounds bin: ceil(count / chunk), at most kmax.
__device__ __forceinline__ int chunks_of(int count, int chunk, int kmax) {
    const int k = (int)(((int64_t)count + chunk - 1) / chunk);
    return k < kmax ? k : kmax;
}

// old_start: the previous starts (may alias start); standing_cursor: start
// the parked bin's cursor after the standing particles (fused re-sort, which
// moves them separately).
__device__ __forceinline__ void bin_scan_block(int* __restrict__ count, int nb, int chunk, int kmax,
                                               int add_parked, int standing_cursor, const int64_t* old_start,
                                               int64_t* start, int* __restrict__ chunk_start,
                                               int* __restrict__ cursor) {
    __shared__ int64_t s_cnt[kScanBlock];
    __shared__ int s_chk[kScanBlock];
    __shared__ int s_standing;
    const int t = threadIdx.x;
    if (t == 0) {
        s_standing = add_parked ? (int)(old_start[nb] - old_start[nb - 1]) : 0;
        if (add_parked) count[nb - 1] += s_standing;
    }
    __syncthreads();
    const int per = (nb + kScanBlock - 1) / kScanBlock;
    const int b0 = t * per, b1 = min(nb, b0 + per);
    int64_t c = 0;
    int k = 0;
    for (int b = b0; b < b1; ++b) {
        c += count[b];
        if (b < nb - 1) k += chunks_of(count[b], chunk, kmax);
    }
    s_cnt[t] = c;
    s_chk[t] = k;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {      // Hillis-Steele inclusive scan
        const int64_t a = t >= off ? s_cnt[t - off] : 0;
        const int ak = t >= off ? s_chk[t - off] : 0;
        __syncthreads();
        s_cnt[t] += a;
        s_chk[t] += ak;
        __syncthreads();
    }
    int64_t run = s_cnt[t] - c;
    int runk = s_chk[t] - k;
    for (int b = b0; b < b1; ++b) {
        start[b] = run;
        chunk_start[b] = runk;
        cursor[b] = (standing_cursor && b == nb - 1) ? s_standing : 0;
        run += count[b];
        if (b < nb - 1) runk += chunks_of(count[b], chunk, kmax);
    }
    if (t == kScanBlock - 1) start[nb] = s_cnt[t];
}

__global__ __launch_bounds__(kScanBlock) void k_bin_scan(int* __restrict__ count, int nb, int chunk, int kmax,
                                                          int add_parked, int standing_cursor,
                                                          const int64_t* old_start, int64_t* start,
                                                          int* __restrict__ chunk_start, int* __restrict__ cursor) {
    bin_scan_block(count, nb, chunk, kmax, add_parked, standing_cursor, old_start, start, chunk_start, cursor);
}

// scan_next of simulation blockIdx.x of a batched round.
__global__ __launch_bounds__(kScanBlock) void k_bin_scan_batched(const TiledDesc* __restrict__ descs) {
    const TiledDesc& d = descs[blockIdx.x];
    bin_scan_block(d.next_count, d.nbins, d.chunk, d.kmax, 1, 1, d.start, d.next_start, d.next_chunk_start,
                   d.next_cursor);
}

// next_count = 0 for simulation blockIdx.y of a batched round.
__global__ __launch_bounds__(kBlock) void k_zero_next_counts(const TiledDesc* __restrict__ descs) {
    const TiledDesc& d = descs[blockIdx.y];
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < d.nbins; i += gridDim.x * kBlock) d.next_count[i] = 0;
}

// Counting-sort scatter.  A block of kScatBlock threads takes kScatGroups
// aligned groups of 4 particles per thread (16-byte source reads), ranks each
// particle within its bin in LDS (one atomic per wave for the lanes in the
// wave's first bin: nearly all of them when re-sorting particles that were
// sorted a few steps ago), claims per-bin ranges with one global atomic per
// (block, bin), and moves x, y, vx, vy, m, id to their sorted slots in the
// second buffer set.  Lanes of one wave that share a bin receive consecutive
// ranks, so their writes are contiguous.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
constexpr int kScatBlock = 1024;
constexpr int kScatGroups = 2;
constexpr int kScatPerBlock = kScatBlock * kScatGroups * 4;

__global__ __launch_bounds__(kScatBlock) void k_bin_scatter(ParticleArrays src, ParticleArrays dst, int64_t np, int n,
                                                             int tl, const int64_t* __restrict__ start,
                                                             int* __restrict__ cursor) {
    extern __shared__ int lds[];
    const int tpr = n >> tl, nb = tpr * tpr + 1;
    int* lcnt = lds;
    int* lbase = lds + nb;
    for (int i = threadIdx.x; i < nb; i += kScatBlock) lcnt[i] = 0;
    __syncthreads();
    const int64_t g0 = (int64_t)blockIdx.x * kScatBlock * kScatGroups;
    int bq[kScatGroups][4], rq[kScatGroups][4];
    float xq[kScatGroups][4], yq[kScatGroups][4];
#pragma unroll
    for (int g = 0; g < kScatGroups; ++g) {
        const int64_t q = g0 + g * kScatBlock + threadIdx.x;
        v4f X = {0.f, 0.f, 0.f, 0.f}, Y = X;
        if (4 * q < np) {   // the arrays are padded to 64 slots, so the last group reads in bounds
            X = reinterpret_cast<const v4f*>(src.x)[q];
            Y = reinterpret_cast<const v4f*>(src.y)[q];
        }
        xq[g][0] = X.x; xq[g][1] = X.y; xq[g][2] = X.z; xq[g][3] = X.w;
        yq[g][0] = Y.x; yq[g][1] = Y.y; yq[g][2] = Y.z; yq[g][3] = Y.w;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            bq[g][k] = -1;
            if (4 * q + k < np) {
                bq[g][k] = bin_of(xq[g][k], yq[g][k], n, tl);
                rq[g][k] = wave_count_add(lcnt, bq[g][k]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < kScatGroups; ++g)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (bq[g][k] >= 0 && rq[g][k] == 0) lbase[bq[g][k]] = atomicAdd(&cursor[bq[g][k]], lcnt[bq[g][k]]);
    __syncthreads();
#pragma unroll
    for (int g = 0; g < kScatGroups; ++g) {
        const int64_t q = g0 + g * kScatBlock + threadIdx.x;
        if (4 * q >= np) continue;
        const v4f VX = reinterpret_cast<const v4f*>(src.vx)[q];
        const v4f VY = reinterpret_cast<const v4f*>(src.vy)[q];
        const v4f M = reinterpret_cast<const v4f*>(src.m)[q];
        const v4u I = reinterpret_cast<const v4u*>(src.id)[q];
        const float vxq[4] = {VX.x, VX.y, VX.z, VX.w}, vyq[4] = {VY.x, VY.y, VY.z, VY.w};
        const float mq[4] = {M.x, M.y, M.z, M.w};
        const unsigned iq[4] = {I.x, I.y, I.z, I.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (bq[g][k] < 0) continue;
            const int64_t s = start[bq[g][k]] + lbase[bq[g][k]] + rq[g][k];
            dst.x[s] = xq[g][k];
            dst.y[s] = yq[g][k];
            dst.vx[s] = vxq[k];
            dst.vy[s] = vyq[k];
            dst.m[s] = mq[k];
            dst.id[s] = iq[k];
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_iota(uint32_t* __restrict__ id, int64_t np) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < np; p += stride) id[p] = (uint32_t)p;
}

// out[id[s]] = in[s] for up to four arrays (null pointers skipped).
__global__ __launch_bounds__(kBlock) void k_unpermute(const uint32_t* __restrict__ id, int64_t np,
                                                       const float* __restrict__ a0, const float* __restrict__ a1,
                                                       const float* __restrict__ a2, const float* __restrict__ a3,
                                                       float* __restrict__ o0, float* __restrict__ o1,
                                                       float* __restrict__ o2, float* __restrict__ o3) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < np; s += stride) {
        const uint32_t d = id[s];
        if (o0) o0[d] = a0[s];
        if (o1) o1[d] = a1[s];
        if (o2) o2[d] = a2[s];
        if (o3) o3[d] = a3[s];
    }
}

// Dynamic LDS above 64 KiB (large lattices: 16,385 bins) must be allowed per kernel.
hipError_t allow_lds(const void* fn, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

int grid_for(int64_t np, int per_thread, int cap) {
    int64_t b = (np + (int64_t)kBlock * per_thread - 1) / ((int64_t)kBlock * per_thread);
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

}  // namespace

int bin_count_of(int n, int tl) {
    const int tpr = n >> tl;
    return tpr * tpr + 1;
}

hipError_t sort_particles(int n, ParticleArrays cur, ParticleArrays alt, int64_t np, BinState bs, bool counted,
                          hipStream_t s) {
    const int nb = bin_count_of(n, bs.tl);
    hipError_t e;
    if ((e = allow_lds(reinterpret_cast<const void*>(&k_bin_count), sizeof(int) * nb)) != hipSuccess) return e;
    if ((e = allow_lds(reinterpret_cast<const void*>(&k_bin_scatter), 2 * sizeof(int) * nb)) != hipSuccess) return e;
    if (!counted && (e = fill_zero(bs.count, sizeof(int) * (size_t)nb, s)) != hipSuccess) return e;
    if (!counted && np > 0) {
        hipLaunchKernelGGL(k_bin_count, dim3(grid_for(np, 64, 2048)), dim3(kBlock), sizeof(int) * nb, s, cur.x, cur.y,
                           np, n, bs.tl, bs.count);
    }
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kScanBlock), 0, s, bs.count, nb, bs.chunk,
                       bs.max_chunks_per_bin > 0 ? bs.max_chunks_per_bin : (1 << 30), counted ? 1 : 0, 0,
                       static_cast<const int64_t*>(bs.start), bs.start, bs.chunk_start, bs.cursor);
    if (np > 0) {
        const int64_t blocks = (np + kScatPerBlock - 1) / kScatPerBlock;
        hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)blocks), dim3(kScatBlock), 2 * sizeof(int) * nb, s, cur, alt, np,
                           n, bs.tl, bs.start, bs.cursor);
    }
    return hipGetLastError();
}

hipError_t scan_next(int n, BinState next, const int64_t* old_start, hipStream_t s) {
    const int nb = bin_count_of(n, next.tl);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kScanBlock), 0, s, next.count, nb, next.chunk,
                       next.max_chunks_per_bin > 0 ? next.max_chunks_per_bin : (1 << 30), 1, 1, old_start, next.start,
                       next.chunk_start, next.cursor);
    return hipGetLastError();
}

hipError_t scan_next_batched(const TiledDesc* d, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_scan_batched, dim3(count), dim3(kScanBlock), 0, s, d);
    return hipGetLastError();
}

hipError_t zero_next_counts_batched(const TiledDesc* d, int count, int max_nbins, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int bx = (max_nbins + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_zero_next_counts, dim3(bx < 64 ? bx : 64, count), dim3(kBlock), 0, s, d);
    return hipGetLastError();
}

hipError_t iota_ids(uint32_t* id, int64_t np, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(np, 4, 4096)), dim3(kBlock), 0, s, id, np);
    return hipGetLastError();
}

hipError_t unpermute(const uint32_t* id, int64_t np, const float* a0, const float* a1, const float* a2,
                     const float* a3, float* o0, float* o1, float* o2, float* o3, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpermute, dim3(grid_for(np, 4, 4096)), dim3(kBlock), 0, s, id, np, a0, a1, a2, a3, o0, o1,
                       o2, o3);
    return hipGetLastError();
}

}  // namespace pmk
