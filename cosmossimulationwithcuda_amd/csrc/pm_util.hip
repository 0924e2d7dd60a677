// Utility kernels: lattice clearing (replaces k_clearAccumulator,
// CosmosCuda.cuh:315-331) and FP64 diagnostic reductions (mass, momentum,
// kinetic and potential energy, max |v|, non-finite count), which have no
// counterpart in the reference.
#include "pm_internal.h"

namespace pmk {
namespace {

// A fill kernel rather than hipMemsetAsync so the step stays capturable in a
// graph without memset nodes.
__global__ __launch_bounds__(kBlock) void k_fill_zero(uint4* __restrict__ p, size_t n16, unsigned char* tail,
                                                       size_t ntail) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

// blockIdx.y selects the buffer; every buffer is a multiple of 16 bytes.
__global__ __launch_bounds__(kBlock) void k_fill_zero_batched(const FillDesc* __restrict__ d) {
    const FillDesc f = d[blockIdx.y];
    uint4* p = reinterpret_cast<uint4*>(f.p);
    const size_t n16 = f.bytes / 16, stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

constexpr int kDiagBlocks = 1024;

__global__ __launch_bounds__(kBlock) void k_diag_particles(const float* __restrict__ x, const float* __restrict__ y,
                                                            const float* __restrict__ vx,
                                                            const float* __restrict__ vy,
                                                            const float* __restrict__ m, int64_t np, int n,
                                                            double* __restrict__ partial) {
    double s[7] = {0, 0, 0, 0, 0, 0, 0};   // mass, px, py, ke, max|v|, nonfinite, in-bounds
    const float hi = (float)(n - 1);
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < np; p += stride) {
        const double mm = m[p], ux = vx[p], uy = vy[p];
        const float px = x[p], py = y[p];
        const bool fin = isfinite(px) && isfinite(py) && isfinite(vx[p]) && isfinite(vy[p]) && isfinite(m[p]);
        if (!fin) { s[5] += 1.0; continue; }
        s[0] += mm;
        s[1] += mm * ux;
        s[2] += mm * uy;
        s[3] += 0.5 * mm * (ux * ux + uy * uy);
        s[4] = fmax(s[4], sqrt(ux * ux + uy * uy));
        if (px >= 1.0f && px < hi && py >= 1.0f && py < hi) s[6] += 1.0;
    }
    __shared__ double red[kBlock / 64][7];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const double v = q == 4 ? wave_max(s[q]) : wave_sum(s[q]);
        if (lane == 0) red[w][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        const int q = threadIdx.x;
        double v = red[0][q];
        for (int k = 1; k < kBlock / 64; ++k) v = q == 4 ? fmax(v, red[k][q]) : v + red[k][q];
        partial[(size_t)blockIdx.x * 8 + q] = v;
    }
    if (threadIdx.x == 7) partial[(size_t)blockIdx.x * 8 + 7] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_diag_lattice(const int64_t* __restrict__ rho, double inv_scale,
                                                          const float* __restrict__ phi, size_t cells,
                                                          double* __restrict__ partial) {
    double s = 0.0;
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < cells; c += stride)
        s += (double)rho[c] * inv_scale * (double)phi[c];
    s = wave_sum(s);
    __shared__ double red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) v += red[k];
        partial[blockIdx.x] = v;
    }
}

struct FieldDesc4 {
    FieldDesc d[4];
};

__global__ void k_set_field_desc(FieldDesc* __restrict__ dst, FieldDesc4 v) {
    if (threadIdx.x < 4) dst[threadIdx.x] = v.d[threadIdx.x];
}

}  // namespace

// The descriptors travel as kernel arguments (copied at launch), so the
// caller's array may go out of scope at once and nothing waits for the copy.
hipError_t set_field_desc(FieldDesc* dst, const FieldDesc* src4, hipStream_t s) {
    FieldDesc4 v;
    for (int i = 0; i < 4; ++i) v.d[i] = src4[i];
    hipLaunchKernelGGL(k_set_field_desc, dim3(1), dim3(64), 0, s, dst, v);
    return hipGetLastError();
}

hipError_t fill_zero(void* p, size_t bytes, hipStream_t s) {
    const size_t n16 = bytes / 16, ntail = bytes - n16 * 16;
    size_t blocks = (n16 + kBlock - 1) / kBlock;
    if (blocks < 1) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(kBlock), 0, s, reinterpret_cast<uint4*>(p), n16,
                       reinterpret_cast<unsigned char*>(p) + n16 * 16, ntail);
    return hipGetLastError();
}

hipError_t fill_zero_batched(const FillDesc* d, int count, size_t max_bytes, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    size_t blocks = (max_bytes / 16 + kBlock - 1) / kBlock;
    if (blocks < 1) blocks = 1;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_fill_zero_batched, dim3((unsigned)blocks, count), dim3(kBlock), 0, s, d);
    return hipGetLastError();
}

int diag_blocks(int64_t np) {
    const int64_t b = (np + kBlock - 1) / kBlock;
    return (int)(b < 1 ? 1 : (b > kDiagBlocks ? kDiagBlocks : b));
}

hipError_t diag_particles(int n, const float* x, const float* y, const float* vx, const float* vy, const float* m,
                          int64_t np, double* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_diag_particles, dim3(diag_blocks(np)), dim3(kBlock), 0, s, x, y, vx, vy, m, np, n, partial);
    return hipGetLastError();
}

int diag_lattice_blocks(int n) {
    const size_t cells = (size_t)n * n;
    const size_t b = (cells + kBlock - 1) / kBlock;
    return (int)(b > kDiagBlocks ? kDiagBlocks : b);
}

hipError_t diag_lattice(int n, const int64_t* rho, double inv_scale, const float* phi, double* partial,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_diag_lattice, dim3(diag_lattice_blocks(n)), dim3(kBlock), 0, s, rho, inv_scale, phi,
                       (size_t)n * n, partial);
    return hipGetLastError();
}

}  // namespace pmk
