// Particle kernels: mass deposit (a) and force gather + kick-drift (d).
//
// Deposit restates k_scatterMassOnAccumulator (CosmosCuda.cuh:446-595) with
// int64 fixed-point sums: each contribution w*m (FP32, separately rounded) is
// quantised to round(w*m*2^shift) and added with a 64-bit integer atomic, so
// the lattice is bit-exact and independent of particle order.
//
// Push restates k_forceMultiSampling (CosmosCuda.cuh:231-312) on (x, v) state:
// v' = v + a dt, x' = x + v' dt (the reference's position Verlet with
// v = (x - xOld)/dt), the same in-bounds rule and the same wrap thresholds.
#include "pm_internal.h"

namespace pmk {
namespace {

// (int)x in [1, N-2] for both axes (cuh:257, 497); NaN fails.
__device__ __forceinline__ bool in_bounds(float x, float y, float hi) {
    return x >= 1.0f && x < hi && y >= 1.0f && y < hi;
}

__device__ __forceinline__ unsigned long long quantise(float p, float scale) {
    return (unsigned long long)(long long)rintf(p * scale);
}

// Deposit of one chunk of one bin per block.  Contributions whose cells lie
// in the bin's LDS window (tile + kHalo margin) are summed there with 64-bit
// LDS atomics and flushed once per block; the rest ("migrants": particles that
// left their tile by more than kHalo cells since the last sort) go straight
// to the lattice with 64-bit global atomics.  All sums are integer, so the
// split never changes the result.
template <int MODE>
__global__ __launch_bounds__(kDepositBlock) void k_deposit_bin(const float* __restrict__ x,
                                                               const float* __restrict__ y,
                                                               const float* __restrict__ m, int n,
                                                               const int64_t* __restrict__ start,
                                                               const int* __restrict__ chunk_start, float scale,
                                                               unsigned long long* __restrict__ rho,
                                                               unsigned long long* __restrict__ migrants) {
    __shared__ unsigned long long win[kWin * kWin];
    __shared__ unsigned int s_mig;
    const int tpr = n >> kTileLog, nin = tpr * tpr;      // in-bounds bins
    const int c = blockIdx.x;
    const int total = chunk_start[nin];
    if (c >= total) return;
    int lo = 0, hi = nin;                                // first j with chunk_start[j] > c, minus one
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (chunk_start[mid] <= c) lo = mid + 1; else hi = mid;
    }
    const int b = lo - 1;
    const int64_t s0 = start[b] + (int64_t)(c - chunk_start[b]) * kDepositChunk;
    const int64_t s1 = min(s0 + kDepositChunk, start[b + 1]);
    const int wx0 = (b % tpr) * kTile - kHalo, wy0 = (b / tpr) * kTile - kHalo;
    for (int e = threadIdx.x; e < kWin * kWin; e += kDepositBlock) win[e] = 0ull;
    if (threadIdx.x == 0) s_mig = 0;
    __syncthreads();
    const float fhi = (float)(n - 1);
    unsigned int mig = 0;
    constexpr int U = 4;
    for (int64_t p0 = s0 + threadIdx.x; p0 < s1; p0 += U * kDepositBlock) {
        float px[U], py[U], pm[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t p = p0 + u * kDepositBlock;
            px[u] = -1.0f; py[u] = -1.0f; pm[u] = 0.0f;
            if (p < s1) { px[u] = x[p]; py[u] = y[p]; pm[u] = m[p]; }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!in_bounds(px[u], py[u], fhi)) continue;
            const int xi = (int)px[u], yi = (int)py[u];
            const int lx = xi - wx0, ly = yi - wy0;
            const bool local = (unsigned)lx < (unsigned)(kWin - 1) && (unsigned)ly < (unsigned)(kWin - 1);
            if (MODE == 0) {
                const unsigned long long q = quantise(pm[u], scale);
                if (local) atomicAdd(&win[ly * kWin + lx], q);
                else { atomicAdd(rho + (size_t)yi * n + xi, q); ++mig; }
            } else {
                const float fx = px[u] - (float)xi, fy = py[u] - (float)yi;
                const float ax = 1.0f - fx, ay = 1.0f - fy;
                const unsigned long long q00 = quantise((ax * ay) * pm[u], scale);
                const unsigned long long q10 = quantise((fx * ay) * pm[u], scale);
                const unsigned long long q01 = quantise((ax * fy) * pm[u], scale);
                const unsigned long long q11 = quantise((fx * fy) * pm[u], scale);
                if (local) {
                    unsigned long long* w = win + ly * kWin + lx;
                    atomicAdd(w, q00);
                    atomicAdd(w + 1, q10);
                    atomicAdd(w + kWin, q01);
                    atomicAdd(w + kWin + 1, q11);
                } else {
                    unsigned long long* g = rho + (size_t)yi * n + xi;
                    atomicAdd(g, q00);
                    atomicAdd(g + 1, q10);
                    atomicAdd(g + n, q01);
                    atomicAdd(g + n + 1, q11);
                    ++mig;
                }
            }
        }
    }
    if (mig) atomicAdd(&s_mig, mig);
    __syncthreads();
    for (int e = threadIdx.x; e < kWin * kWin; e += kDepositBlock) {
        const unsigned long long v = win[e];
        if (v == 0ull) continue;
        const int gx = wx0 + e % kWin, gy = wy0 + e / kWin;   // nonzero cells lie in [1, N-1]^2
        atomicAdd(rho + (size_t)gy * n + gx, v);
    }
    if (threadIdx.x == 0 && s_mig) atomicAdd(migrants, (unsigned long long)s_mig);
}

template <int MODE, int WRAP>
__global__ __launch_bounds__(kBlock) void k_push(const float2* __restrict__ force, float* __restrict__ x,
                                                  float* __restrict__ y, float* __restrict__ vx,
                                                  float* __restrict__ vy, int64_t np, int n, float dt) {
    const float hi = (float)(n - 1);
    const float span = (float)(n - 4);
    const float lo_wrap = 2.0f, hi_wrap = (float)(n - 2);
    const float jump = span / dt;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < np; p += stride) {
        const float px = x[p], py = y[p];
        if (!in_bounds(px, py, hi)) continue;
        const int xi = (int)px, yi = (int)py;
        const size_t c = (size_t)yi * n + xi;
        float ax, ay;
        if (MODE == 0) {
            const float2 f = force[c];
            ax = f.x;
            ay = f.y;
        } else {
            const float fx = px - (float)xi, fy = py - (float)yi;
            const float wx = 1.0f - fx, wy = 1.0f - fy;
            const float w00 = wx * wy, w10 = fx * wy, w01 = wx * fy, w11 = fx * fy;
            const float2 f00 = force[c], f10 = force[c + 1], f01 = force[c + n], f11 = force[c + n + 1];
            ax = ((f00.x * w00 + f10.x * w10) + f01.x * w01) + f11.x * w11;
            ay = ((f00.y * w00 + f10.y * w10) + f01.y * w01) + f11.y * w11;
        }
        float nvx = fmaf(ax, dt, vx[p]);
        float nvy = fmaf(ay, dt, vy[p]);
        float nx = fmaf(nvx, dt, px);
        float ny = fmaf(nvy, dt, py);
        if (nx < lo_wrap) { nx += span; if (WRAP) nvx += jump; }
        if (ny < lo_wrap) { ny += span; if (WRAP) nvy += jump; }
        if (nx > hi_wrap) { nx -= span; if (WRAP) nvx -= jump; }
        if (ny > hi_wrap) { ny -= span; if (WRAP) nvy -= jump; }
        x[p] = nx;
        y[p] = ny;
        vx[p] = nvx;
        vy[p] = nvy;
    }
}

int particle_grid(int64_t np) {
    const int64_t blocks = (np + kBlock - 1) / kBlock;
    const int64_t cap = 256 * 64;   // 256 CUs x 64 blocks in flight is plenty for a grid-stride loop
    return (int)(blocks < 1 ? 1 : (blocks > cap ? cap : blocks));
}

}  // namespace

int deposit_max_chunks(int n, int64_t capacity) {
    const int64_t nin = (int64_t)(n >> kTileLog) * (n >> kTileLog);
    return (int)((capacity + kDepositChunk - 1) / kDepositChunk + nin);
}

hipError_t deposit(int n, int mode, const float* x, const float* y, const float* m, BinState bs, int max_chunks,
                   float scale, int64_t* rho, unsigned long long* migrants, hipStream_t s) {
    auto* r = reinterpret_cast<unsigned long long*>(rho);
    const dim3 grid(max_chunks);
    if (mode == 0)
        hipLaunchKernelGGL((k_deposit_bin<0>), grid, dim3(kDepositBlock), 0, s, x, y, m, n, bs.start,
                           bs.chunk_start, scale, r, migrants);
    else
        hipLaunchKernelGGL((k_deposit_bin<1>), grid, dim3(kDepositBlock), 0, s, x, y, m, n, bs.start,
                           bs.chunk_start, scale, r, migrants);
    return hipGetLastError();
}

hipError_t push(int n, int mode, const float2* force, float* x, float* y, float* vx, float* vy, int64_t np,
                float dt, int wrap_mode, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    const dim3 grid(particle_grid(np));
    if (mode == 0 && wrap_mode == 0)
        hipLaunchKernelGGL((k_push<0, 0>), grid, dim3(kBlock), 0, s, force, x, y, vx, vy, np, n, dt);
    else if (mode == 0)
        hipLaunchKernelGGL((k_push<0, 1>), grid, dim3(kBlock), 0, s, force, x, y, vx, vy, np, n, dt);
    else if (wrap_mode == 0)
        hipLaunchKernelGGL((k_push<1, 0>), grid, dim3(kBlock), 0, s, force, x, y, vx, vy, np, n, dt);
    else
        hipLaunchKernelGGL((k_push<1, 1>), grid, dim3(kBlock), 0, s, force, x, y, vx, vy, np, n, dt);
    return hipGetLastError();
}

}  // namespace pmk
