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
                                                               const int* __restrict__ chunk_start, int chunk,
                                                               float scale,
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
    const int64_t s0 = start[b] + (int64_t)(c - chunk_start[b]) * chunk;
    const int64_t s1 = min(s0 + chunk, start[b + 1]);
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

// One particle: gather the force at (x, y) and kick-drift it.  Operation
// order is fixed (no contraction) so the CPU mirror reproduces it bit for bit.
template <int MODE, int WRAP>
__device__ __forceinline__ void push_one(const float2* __restrict__ force, int n, float dt, float& px, float& py,
                                         float& vx, float& vy) {
    if (!in_bounds(px, py, (float)(n - 1))) return;
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
    const float span = (float)(n - 4);
    float nvx = fmaf(ax, dt, vx);
    float nvy = fmaf(ay, dt, vy);
    float nx = fmaf(nvx, dt, px);
    float ny = fmaf(nvy, dt, py);
    if (nx < 2.0f) { nx += span; if (WRAP) nvx += span / dt; }
    if (ny < 2.0f) { ny += span; if (WRAP) nvy += span / dt; }
    if (nx > (float)(n - 2)) { nx -= span; if (WRAP) nvx -= span / dt; }
    if (ny > (float)(n - 2)) { ny -= span; if (WRAP) nvy -= span / dt; }
    px = nx;
    py = ny;
    vx = nvx;
    vy = nvy;
}

// Four consecutive particles per lane: 16-byte nontemporal loads and stores
// of x, y, vx, vy (the streams are touched once per step; keeping them out of
// the caches leaves L2 to the force lattice the gathers reuse).
typedef float v4f __attribute__((ext_vector_type(4)));

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
    X = v4f{px[0], px[1], px[2], px[3]};
    Y = v4f{py[0], py[1], py[2], py[3]};
    VX = v4f{ux[0], ux[1], ux[2], ux[3]};
    VY = v4f{uy[0], uy[1], uy[2], uy[3]};
    __builtin_nontemporal_store(X, x + q);
    __builtin_nontemporal_store(Y, y + q);
    __builtin_nontemporal_store(VX, vx + q);
    __builtin_nontemporal_store(VY, vy + q);
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

// Gather + kick-drift of one particle with the force taken from the bin's
// LDS window when its cell lies inside it, else from the lattice.  Same
// arithmetic as push_one, so results are identical whichever path serves it.
template <int MODE, int WRAP>
__device__ __forceinline__ void push_one_win(const float2* __restrict__ fwin, int wx0, int wy0,
                                             const float2* __restrict__ force, int n, float dt, float& px,
                                             float& py, float& vx, float& vy) {
    if (!in_bounds(px, py, (float)(n - 1))) return;
    const int xi = (int)px, yi = (int)py;
    const int lx = xi - wx0, ly = yi - wy0;
    const bool local = (unsigned)lx < (unsigned)(kWin - 1) && (unsigned)ly < (unsigned)(kWin - 1);
    const float2* f = local ? fwin + ly * kWin + lx : force + (size_t)yi * n + xi;
    const int row = local ? kWin : n;
    float ax, ay;
    if (MODE == 0) {
        const float2 f0 = f[0];
        ax = f0.x;
        ay = f0.y;
    } else {
        const float fx = px - (float)xi, fy = py - (float)yi;
        const float wx = 1.0f - fx, wy = 1.0f - fy;
        const float w00 = wx * wy, w10 = fx * wy, w01 = wx * fy, w11 = fx * fy;
        const float2 f00 = f[0], f10 = f[1], f01 = f[row], f11 = f[row + 1];
        ax = ((f00.x * w00 + f10.x * w10) + f01.x * w01) + f11.x * w11;
        ay = ((f00.y * w00 + f10.y * w10) + f01.y * w01) + f11.y * w11;
    }
    const float span = (float)(n - 4);
    float nvx = fmaf(ax, dt, vx);
    float nvy = fmaf(ay, dt, vy);
    float nx = fmaf(nvx, dt, px);
    float ny = fmaf(nvy, dt, py);
    if (nx < 2.0f) { nx += span; if (WRAP) nvx += span / dt; }
    if (ny < 2.0f) { ny += span; if (WRAP) nvy += span / dt; }
    if (nx > (float)(n - 2)) { nx -= span; if (WRAP) nvx -= span / dt; }
    if (ny > (float)(n - 2)) { ny -= span; if (WRAP) nvy -= span / dt; }
    px = nx;
    py = ny;
    vx = nvx;
    vy = nvy;
}

// Fused push + next deposit: one block per deposit chunk of one bin.  The
// bin's force window (tile + kHalo margin, float2) is staged in LDS so the
// CIC gathers are LDS reads; each particle is kick-drifted, written back, and
// its new position deposited (int64, LDS window with global-atomic spill)
// into the next step's mass lattice.  Optionally the new bins are
// histogrammed for the next re-sort.  Both windows take 99.9 KB of LDS: one
// 1024-thread block per CU.
constexpr int kFusedBlock = 1024;

template <int MODE, int WRAP, bool COUNT>
__global__ __launch_bounds__(kFusedBlock) void k_push_deposit(const float2* __restrict__ force,
                                                              float* __restrict__ x, float* __restrict__ y,
                                                              float* __restrict__ vx, float* __restrict__ vy,
                                                              const float* __restrict__ m, int n, float dt,
                                                              const int64_t* __restrict__ start,
                                                              const int* __restrict__ chunk_start, int chunk,
                                                              float scale, unsigned long long* __restrict__ rho,
                                                              unsigned long long* __restrict__ migrants,
                                                              int* __restrict__ count) {
    __shared__ unsigned long long win[kWin * kWin];
    __shared__ float2 fwin[kWin * kWin];
    __shared__ unsigned int s_mig;
    extern __shared__ int hist[];
    const int tpr = n >> kTileLog, nin = tpr * tpr;
    const int c = blockIdx.x;
    const int total = chunk_start[nin];
    if (c >= total) return;
    int lo = 0, hi = nin;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (chunk_start[mid] <= c) lo = mid + 1; else hi = mid;
    }
    const int b = lo - 1;
    const int64_t s0 = start[b] + (int64_t)(c - chunk_start[b]) * chunk;
    const int64_t s1 = min(s0 + chunk, start[b + 1]);
    const int wx0 = (b % tpr) * kTile - kHalo, wy0 = (b / tpr) * kTile - kHalo;
    for (int e = threadIdx.x; e < kWin * kWin; e += kFusedBlock) {
        win[e] = 0ull;
        const int gx = wx0 + e % kWin, gy = wy0 + e / kWin;
        fwin[e] = ((unsigned)gx < (unsigned)n && (unsigned)gy < (unsigned)n) ? force[(size_t)gy * n + gx]
                                                                             : make_float2(0.f, 0.f);
    }
    if (COUNT)
        for (int e = threadIdx.x; e <= nin; e += kFusedBlock) hist[e] = 0;
    if (threadIdx.x == 0) s_mig = 0;
    __syncthreads();
    const float fhi = (float)(n - 1);
    unsigned int mig = 0;
    constexpr int U = 4;
    for (int64_t p0 = s0 + threadIdx.x; p0 < s1; p0 += U * kFusedBlock) {
        float px[U], py[U], ux[U], uy[U], pm[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t p = p0 + u * kFusedBlock;
            px[u] = -1.0f; py[u] = -1.0f; ux[u] = 0.0f; uy[u] = 0.0f; pm[u] = 0.0f;
            if (p < s1) {
                px[u] = __builtin_nontemporal_load(x + p);
                py[u] = __builtin_nontemporal_load(y + p);
                ux[u] = __builtin_nontemporal_load(vx + p);
                uy[u] = __builtin_nontemporal_load(vy + p);
                pm[u] = __builtin_nontemporal_load(m + p);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            push_one_win<MODE, WRAP>(fwin, wx0, wy0, force, n, dt, px[u], py[u], ux[u], uy[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t p = p0 + u * kFusedBlock;
            if (p >= s1) continue;
            __builtin_nontemporal_store(px[u], x + p);
            __builtin_nontemporal_store(py[u], y + p);
            __builtin_nontemporal_store(ux[u], vx + p);
            __builtin_nontemporal_store(uy[u], vy + p);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t p = p0 + u * kFusedBlock;
            if (p >= s1) continue;
            if (COUNT) atomicAdd(&hist[bin_of(px[u], py[u], n, tpr)], 1);
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
    for (int e = threadIdx.x; e < kWin * kWin; e += kFusedBlock) {
        const unsigned long long v = win[e];
        if (v == 0ull) continue;
        const int gx = wx0 + e % kWin, gy = wy0 + e / kWin;
        atomicAdd(rho + (size_t)gy * n + gx, v);
    }
    if (COUNT)
        for (int e = threadIdx.x; e <= nin; e += kFusedBlock)
            if (hist[e]) atomicAdd(&count[e], hist[e]);
    if (threadIdx.x == 0 && s_mig) atomicAdd(migrants, (unsigned long long)s_mig);
}

template <int MODE, int WRAP>
hipError_t push_deposit_t(int n, const float2* force, ParticleArrays p, BinState bs, int max_chunks, float dt,
                          float scale, int64_t* rho_next, unsigned long long* migrants, bool count_bins,
                          hipStream_t s) {
    auto* r = reinterpret_cast<unsigned long long*>(rho_next);
    const int nb = (n >> kTileLog) * (n >> kTileLog) + 1;
    if (count_bins) {
        const size_t lds = sizeof(int) * nb;
        if (lds + 2 * sizeof(unsigned long long) * kWin * kWin > 65536) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_push_deposit<MODE, WRAP, true>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL((k_push_deposit<MODE, WRAP, true>), dim3(max_chunks), dim3(kFusedBlock), lds, s, force,
                           p.x, p.y, p.vx, p.vy, p.m, n, dt, bs.start, bs.chunk_start, bs.chunk, scale, r, migrants,
                           bs.count);
    } else {
        hipLaunchKernelGGL((k_push_deposit<MODE, WRAP, false>), dim3(max_chunks), dim3(kFusedBlock), 0, s, force,
                           p.x, p.y, p.vx, p.vy, p.m, n, dt, bs.start, bs.chunk_start, bs.chunk, scale, r, migrants,
                           bs.count);
    }
    return hipGetLastError();
}

}  // namespace

int deposit_max_chunks(int n, int64_t capacity, int chunk) {
    const int64_t nin = (int64_t)(n >> kTileLog) * (n >> kTileLog);
    return (int)((capacity + chunk - 1) / chunk + nin);
}

hipError_t deposit(int n, int mode, const float* x, const float* y, const float* m, BinState bs, int max_chunks,
                   float scale, int64_t* rho, unsigned long long* migrants, hipStream_t s) {
    auto* r = reinterpret_cast<unsigned long long*>(rho);
    const dim3 grid(max_chunks);
    if (mode == 0)
        hipLaunchKernelGGL((k_deposit_bin<0>), grid, dim3(kDepositBlock), 0, s, x, y, m, n, bs.start,
                           bs.chunk_start, bs.chunk, scale, r, migrants);
    else
        hipLaunchKernelGGL((k_deposit_bin<1>), grid, dim3(kDepositBlock), 0, s, x, y, m, n, bs.start,
                           bs.chunk_start, bs.chunk, scale, r, migrants);
    return hipGetLastError();
}

hipError_t push_deposit(int n, int mode, int wrap_mode, const float2* force, ParticleArrays p, BinState bs,
                        int max_chunks, float dt, float scale, int64_t* rho_next, unsigned long long* migrants,
                        bool count_bins, hipStream_t s) {
    if (mode == 0 && wrap_mode == 0)
        return push_deposit_t<0, 0>(n, force, p, bs, max_chunks, dt, scale, rho_next, migrants, count_bins, s);
    if (mode == 0)
        return push_deposit_t<0, 1>(n, force, p, bs, max_chunks, dt, scale, rho_next, migrants, count_bins, s);
    if (wrap_mode == 0)
        return push_deposit_t<1, 0>(n, force, p, bs, max_chunks, dt, scale, rho_next, migrants, count_bins, s);
    return push_deposit_t<1, 1>(n, force, p, bs, max_chunks, dt, scale, rho_next, migrants, count_bins, s);
}

hipError_t push(int n, int mode, const float2* force, float* x, float* y, float* vx, float* vy, int64_t np,
                float dt, int wrap_mode, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    if (mode == 0 && wrap_mode == 0) return push_t<0, 0>(force, x, y, vx, vy, np, n, dt, s);
    if (mode == 0) return push_t<0, 1>(force, x, y, vx, vy, np, n, dt, s);
    if (wrap_mode == 0) return push_t<1, 0>(force, x, y, vx, vy, np, n, dt, s);
    return push_t<1, 1>(force, x, y, vx, vy, np, n, dt, s);
}

}  // namespace pmk
