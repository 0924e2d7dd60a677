// Utility kernels: lattice clearing (replaces k_clearAccumulator,
// CosmosCuda.cuh:315-331) and FP64 diagnostic reductions (mass, momentum,
// kinetic and potential energy, max |v|, non-finite count), which have no
// counterpart in the reference.
#include <cmath>
#include <mutex>

#include "pm_sim.h"

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

// ---- diagnostics without allocation -------------------------------------------
// pm_diagnostics_noalloc: the reductions of pm_diagnostics over caller-owned
// device arrays, for callers that cannot allocate device memory.  The block
// partial sums live in `scratch`, a fixed array of this code object (each
// device holds its own copy until the process ends); host code reaches it
// only through the per-device address hipGetSymbolAddress returns.
namespace {

constexpr int kPartOff = 0;                                  // particle partials: 8 doubles per block
constexpr int kLatOff = 8 * pmk::kDiagBlocks;                // lattice partials: 1 double per block
constexpr int kScratchDoubles = kLatOff + pmk::kDiagBlocks;  // 9216 doubles = 72 KiB
static_assert(kScratchDoubles * sizeof(double) == PM_DIAG_NOALLOC_SCRATCH_BYTES, "scratch size");

static __device__ __align__(256) double scratch[kScratchDoubles];

// Final reduction in the order pm_diagnostics uses on the host: thread q < 7
// walks column q of the particle partials block by block (q = 4 is a max),
// thread 7 sums the lattice partials; the same inputs in the same order give
// the same bits.
__global__ void k_diag_finish(const double* __restrict__ part, int pb, const double* __restrict__ lat, int lb,
                              double* __restrict__ out) {
    const int q = threadIdx.x;
    if (q < 7) {
        double v = 0.0;
        for (int b = 0; b < pb; ++b) v = q == 4 ? fmax(v, part[b * 8 + 4]) : v + part[b * 8 + q];
        // out: mass, px, py, ke, pe, max|v|, nonfinite, in-bounds
        out[q < 4 ? q : q + 1] = v;
    } else if (q == 7) {
        double pe = 0.0;
        for (int b = 0; b < lb; ++b) pe += lat[b];
        out[4] = 0.5 * pe;
    }
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// looked up once per device.
hipError_t scratch_on(int device, double** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<double*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

}  // namespace

extern "C" pm_status pm_diagnostics_noalloc(int32_t device, void* stream, int32_t n_lattice, const float* x,
                                            const float* y, const float* vx, const float* vy, const float* m,
                                            int64_t np, const int64_t* rho, int32_t mass_shift, const float* phi,
                                            double* out) {
    if (!out) return pmi::fail(nullptr, PM_EINVAL, "null output buffer");
    if (np < 0) return pmi::fail(nullptr, PM_EINVAL, "np < 0");
    if (np > 0 && (!x || !y || !vx || !vy || !m)) return pmi::fail(nullptr, PM_EINVAL, "null particle array");
    if (n_lattice < 64 || n_lattice > 8192 || (n_lattice & (n_lattice - 1)))
        return pmi::fail(nullptr, PM_EINVAL, "n_lattice must be a power of two in [64, 8192]");
    if ((rho == nullptr) != (phi == nullptr)) return pmi::fail(nullptr, PM_EINVAL, "rho and phi: both or neither");
    if (rho && (mass_shift < 0 || mass_shift > 62)) return pmi::fail(nullptr, PM_EINVAL, "mass_shift outside [0, 62]");
    // the partial-sum counts are capped by construction; the scratch holds the caps
    const int pb = pmk::diag_blocks(np > 0 ? np : 1), lb = rho ? pmk::diag_lattice_blocks(n_lattice) : 0;
    if (pb > pmk::kDiagBlocks || lb > pmk::kDiagBlocks)
        return pmi::fail(nullptr, PM_EINVAL, "reduction exceeds the fixed scratch array");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= kMaxDevices)
        return pmi::fail(nullptr, PM_ENODEV, "no such device");
    int prev = -1;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess) e = hipSetDevice(device);
    double* sp = nullptr;
    if (e == hipSuccess) e = scratch_on(device, &sp);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (e == hipSuccess && np > 0) e = pmk::diag_particles(n_lattice, x, y, vx, vy, m, np, sp + kPartOff, s);
    if (e == hipSuccess && rho)
        e = pmk::diag_lattice(n_lattice, rho, std::ldexp(1.0, -mass_shift), phi, sp + kLatOff, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_diag_finish, dim3(1), dim3(64), 0, s, sp + kPartOff, np > 0 ? pb : 0, sp + kLatOff, lb,
                           out);
        e = hipGetLastError();
    }
    if (prev >= 0) (void)hipSetDevice(prev);
    return e == hipSuccess ? PM_OK : pmi::hip_fail(nullptr, e, "pm_diagnostics_noalloc");
}
