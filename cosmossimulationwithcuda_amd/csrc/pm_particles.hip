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

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_deposit_global(const float* __restrict__ x, const float* __restrict__ y,
                                                            const float* __restrict__ m, int64_t np, int n,
                                                            float scale, unsigned long long* __restrict__ rho) {
    const float hi = (float)(n - 1);
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < np; p += stride) {
        const float px = x[p], py = y[p];
        if (!in_bounds(px, py, hi)) continue;
        const int xi = (int)px, yi = (int)py;
        const size_t c = (size_t)yi * n + xi;
        const float mm = m[p];
        if (MODE == 0) {
            atomicAdd(rho + c, quantise(mm, scale));
        } else {
            const float fx = px - (float)xi, fy = py - (float)yi;
            const float ax = 1.0f - fx, ay = 1.0f - fy;
            atomicAdd(rho + c, quantise((ax * ay) * mm, scale));
            atomicAdd(rho + c + 1, quantise((fx * ay) * mm, scale));
            atomicAdd(rho + c + n, quantise((ax * fy) * mm, scale));
            atomicAdd(rho + c + n + 1, quantise((fx * fy) * mm, scale));
        }
    }
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

hipError_t deposit(int n, int mode, const float* x, const float* y, const float* m, int64_t np, float scale,
                   int64_t* rho, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    auto* r = reinterpret_cast<unsigned long long*>(rho);
    const dim3 grid(particle_grid(np));
    if (mode == 0)
        hipLaunchKernelGGL((k_deposit_global<0>), grid, dim3(kBlock), 0, s, x, y, m, np, n, scale, r);
    else
        hipLaunchKernelGGL((k_deposit_global<1>), grid, dim3(kBlock), 0, s, x, y, m, np, n, scale, r);
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
