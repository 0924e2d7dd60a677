// Device helpers of the particle kernels, shared by pm_particles.hip and the
// persistent round kernel (pm_round.hip): in-bounds rule, fixed-point
// quantisation, CIC weights, force gather, kick-drift, global deposit, and
// the batched direct push + deposit of four particles.
#ifndef COSMOS_PM_PARTICLE_CORE_H
#define COSMOS_PM_PARTICLE_CORE_H

#include "pm_internal.h"

namespace pmk {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// (int)x in [1, N-2] for both axes (cuh:257, 497); NaN fails.
__device__ __forceinline__ bool in_bounds(float x, float y, float hi) {
    return x >= 1.0f && x < hi && y >= 1.0f && y < hi;
}

// round(a * scale) as an unsigned 64-bit integer for a >= 0 and a * scale <
// 2^64 (the mass shift keeps every contribution below 2^62).  v = rint(a s) is
// split at 2^32: hi = floor(v 2^-32) and lo = v - hi 2^32 are exact floats (lo
// keeps a subset of v's significand bits; v has a fraction only below 2^24,
// where hi = 0), so two 32-bit conversions give the integer the generic
// float -> int64 sequence gives, in 7 instructions instead of 12.
__device__ __forceinline__ unsigned long long quantise_mag(float a, float scale) {
    const float v = rintf(a * scale);
    const float hi = floorf(v * 0x1p-32f);
    const float lo = fmaf(hi, -0x1p32f, v);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// Fractional offset x - (int)x for x >= 1 (in-bounds positions): exact either
// way, one instruction.
__device__ __forceinline__ float frac(float x) { return __builtin_amdgcn_fractf(x); }

// The four CIC contributions round(w * m * scale) of one particle, w the
// reference weights (cuh:498-505).  Rounding is symmetric, so a negative mass
// gives the negated contributions of |m|; the negation runs only in waves that
// hold such a particle.
struct Cic4 {
    unsigned long long c00, c10, c01, c11;
};
__device__ __forceinline__ Cic4 cic_contrib(float fx, float fy, float pm, float scale) {
    const float ax = 1.0f - fx, ay = 1.0f - fy, am = fabsf(pm);
    Cic4 c{quantise_mag((ax * ay) * am, scale), quantise_mag((fx * ay) * am, scale),
           quantise_mag((ax * fy) * am, scale), quantise_mag((fx * fy) * am, scale)};
    if (__builtin_expect(__any(pm < 0.0f), 0) && pm < 0.0f) {
        c.c00 = 0ull - c.c00;
        c.c10 = 0ull - c.c10;
        c.c01 = 0ull - c.c01;
        c.c11 = 0ull - c.c11;
    }
    return c;
}
__device__ __forceinline__ unsigned long long ngp_contrib(float pm, float scale) {
    const unsigned long long c = quantise_mag(fabsf(pm), scale);
    return pm < 0.0f ? 0ull - c : c;
}

// CIC (MODE 1) or NGP (MODE 0) force at fractional offsets (fx, fy) in the
// cell whose entry is f[0]; row = stride of f between y and y+1.  The
// reference weights (cuh:262-279) in a fixed order:
//   a = ((F00 w00 + F10 w10) + F01 w01) + F11 w11.
template <int MODE>
__device__ __forceinline__ float2 gather(const float2* f, int row, float fx, float fy) {
    if (MODE == 0) return f[0];
    const float wx = 1.0f - fx, wy = 1.0f - fy;
    const float w00 = wx * wy, w10 = fx * wy, w01 = wx * fy, w11 = fx * fy;
    const float2 f00 = f[0], f10 = f[1], f01 = f[row], f11 = f[row + 1];
    return make_float2(((f00.x * w00 + f10.x * w10) + f01.x * w01) + f11.x * w11,
                       ((f00.y * w00 + f10.y * w10) + f01.y * w01) + f11.y * w11);
}

// v' = v + a dt; x' = x + v' dt; wrap once per side (cuh:286-303).  WRAP=1
// adds the reference's implied velocity jump (its xOld is not shifted).
template <int WRAP>
__device__ __forceinline__ void kick_drift(float2 a, int n, float dt, float& px, float& py, float& vx, float& vy) {
    const float span = (float)(n - 4);
    float nvx = fmaf(a.x, dt, vx);
    float nvy = fmaf(a.y, dt, vy);
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

// One particle of the plain push: force from the global lattice.
template <int MODE, int WRAP>
__device__ __forceinline__ void push_one(const float2* __restrict__ force, int n, float dt, float& px, float& py,
                                         float& vx, float& vy) {
    if (!in_bounds(px, py, (float)(n - 1))) return;
    const int xi = (int)px, yi = (int)py;
    const float2 a = gather<MODE>(force + (size_t)yi * n + xi, n, frac(px), frac(py));
    kick_drift<WRAP>(a, n, dt, px, py, vx, vy);
}

// Deposit of one in-bounds position: contributions round(w*m*scale) into the
// LDS window when the cells (lx, ly) .. (lx+1, ly+1) lie inside it (returns
// true), else nothing (returns false) and the caller spills it to the lattice
// with deposit_global.  The two paths are kept in separate loops so the
// compiler cannot merge an LDS and a global atomic into one flat atomic.
// Whether a deposit at (px, py) lands inside the W x W tile window at
// (wx0, wy0): the test deposit_local makes before it deposits.
template <int W>
__device__ __forceinline__ bool in_window(int wx0, int wy0, float px, float py) {
    return (unsigned)((int)px - wx0) < (unsigned)(W - 1) && (unsigned)((int)py - wy0) < (unsigned)(W - 1);
}

template <int MODE, int W>
__device__ __forceinline__ bool deposit_local(unsigned long long* win, int wx0, int wy0, float px, float py,
                                              float pm, float scale) {
    const int xi = (int)px, yi = (int)py;
    const int lx = xi - wx0, ly = yi - wy0;
    if (!((unsigned)lx < (unsigned)(W - 1) && (unsigned)ly < (unsigned)(W - 1))) return false;
    unsigned long long* w = win + ly * W + lx;
    if (MODE == 0) {
        atomicAdd(w, ngp_contrib(pm, scale));
        return true;
    }
    const Cic4 c = cic_contrib(frac(px), frac(py), pm, scale);
    atomicAdd(w, c.c00);
    atomicAdd(w + 1, c.c10);
    atomicAdd(w + W, c.c01);
    atomicAdd(w + W + 1, c.c11);
    return true;
}

template <int MODE>
__device__ __forceinline__ void deposit_global(unsigned long long* __restrict__ rho, int n, float px, float py,
                                               float pm, float scale) {
    const int xi = (int)px, yi = (int)py;
    unsigned long long* g = rho + (size_t)yi * n + xi;
    if (MODE == 0) {
        atomicAdd(g, ngp_contrib(pm, scale));
        return;
    }
    const Cic4 c = cic_contrib(frac(px), frac(py), pm, scale);
    atomicAdd(g, c.c00);
    atomicAdd(g + 1, c.c10);
    atomicAdd(g + n, c.c01);
    atomicAdd(g + n + 1, c.c11);
}


// Four consecutive particles of one simulation of a batch (group q of its
// slots): force gathered from the lattice, deposit with 64-bit global atomics
// into the next lattice.  Same arithmetic as the tiled kernels.
template <int MODE>
__device__ __forceinline__ void direct_push_group(const PartDesc& d, int64_t q) {
    if (4 * q >= d.np) return;
    const float fhi = (float)(d.n - 1);
    // capacity is padded to 64 slots, so the last group reads in bounds
    v4f X = reinterpret_cast<const v4f*>(d.x)[q], Y = reinterpret_cast<const v4f*>(d.y)[q];
    v4f VX = reinterpret_cast<const v4f*>(d.vx)[q], VY = reinterpret_cast<const v4f*>(d.vy)[q];
    const v4f M = reinterpret_cast<const v4f*>(d.m)[q];
    float px[4] = {X.x, X.y, X.z, X.w}, py[4] = {Y.x, Y.y, Y.z, Y.w};
    float ux[4] = {VX.x, VX.y, VX.z, VX.w}, uy[4] = {VY.x, VY.y, VY.z, VY.w};
    const float pm[4] = {M.x, M.y, M.z, M.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (4 * q + k >= d.np) break;
        if (d.wrap_mode) push_one<MODE, 1>(d.force, d.n, d.dt, px[k], py[k], ux[k], uy[k]);
        else push_one<MODE, 0>(d.force, d.n, d.dt, px[k], py[k], ux[k], uy[k]);
    }
    if (4 * q + 3 < d.np) {
        reinterpret_cast<v4f*>(d.x)[q] = v4f{px[0], px[1], px[2], px[3]};
        reinterpret_cast<v4f*>(d.y)[q] = v4f{py[0], py[1], py[2], py[3]};
        reinterpret_cast<v4f*>(d.vx)[q] = v4f{ux[0], ux[1], ux[2], ux[3]};
        reinterpret_cast<v4f*>(d.vy)[q] = v4f{uy[0], uy[1], uy[2], uy[3]};
    } else {
        for (int k = 0; k < 4 && 4 * q + k < d.np; ++k) {
            d.x[4 * q + k] = px[k];
            d.y[4 * q + k] = py[k];
            d.vx[4 * q + k] = ux[k];
            d.vy[4 * q + k] = uy[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (4 * q + k >= d.np) break;
        if (in_bounds(px[k], py[k], fhi)) deposit_global<MODE>(d.rho_next, d.n, px[k], py[k], pm[k], d.scale);
    }
}

}  // namespace
}  // namespace pmk

#endif  // COSMOS_PM_PARTICLE_CORE_H
