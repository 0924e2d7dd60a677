// FP64 setup of the spectral transfer function, once per simulation.
//
// Restates the reference's setup (CosmosCuda.cuh: k_generateFilterLattice
// 89-119, k_generateMassAssignmentKernel 374-415, k_applyPhaseShift 416-445,
// forward FFTs 1018-1027 and 1271-1284) in FP64 on the GPU:
//   Ghat = FFT2(G),  G = 1/r on the wrap-centred lattice, G(0) = green_origin
//   D    = closed form of the phase-shifted mass-kernel spectrum without its
//          (-1)^(kx+ky) factor: NGP 1, CIC cos(pi kx/N) cos(pi ky/N), k signed
//   T    = Ghat * D / (D^2 + eps) / N^2   stored as FP32 over the half spectrum
//          T[kx * N + ky], kx in [0, N/2]
// The (-1)^(kx+ky) factor of the reference's kernel spectrum and its N/2 output
// shift (k_shiftLattice, cuh:333-355) cancel, so neither appears here.
#include "pm_internal.h"

namespace pmk {
namespace {

template <int N>
__global__ __launch_bounds__(kBlock) void k_green64(double2* __restrict__ g, double green_origin) {
    const size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= (size_t)N * N) return;
    const int i = (int)(c & (N - 1)), j = (int)(c / N);
    const double dx = i - (i <= N / 2 ? 0 : N);
    const double dy = j - (j <= N / 2 ? 0 : N);
    const double r = sqrt(dx * dx + dy * dy);
    g[c] = make_double2(r > 0.5 ? 1.0 / r : green_origin, 0.0);
}

// One length-N forward FFT per block along a line of the complex lattice:
// element e of line L sits at data[L * line_stride + e * elem_stride].
// In-place radix-2 Stockham in LDS, twiddles from sincospi in FP64.
template <int N>
__global__ __launch_bounds__(kBlock) void k_fft64_line(double2* __restrict__ data, int64_t line_stride,
                                                        int64_t elem_stride) {
    __shared__ double2 buf[N];
    constexpr int Q = N / 2, PER = (Q + kBlock - 1) / kBlock;
    const int t = threadIdx.x;
    double2* line = data + (int64_t)blockIdx.x * line_stride;
    for (int e = t; e < N; e += kBlock) buf[e] = line[e * elem_stride];
    __syncthreads();
    for (int ns = 1; ns < N; ns *= 2) {
        double2 v0[PER], v1[PER];
        int ob[PER];
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int b = t + p * kBlock;
            if (b < Q) {
                const int k = b & (ns - 1);
                const double2 a0 = buf[b];
                double2 a1 = buf[b + Q];
                double sn, cs;
                sincospi(-(double)k / (double)ns, &sn, &cs);
                a1 = make_double2(a1.x * cs - a1.y * sn, a1.x * sn + a1.y * cs);
                v0[p] = make_double2(a0.x + a1.x, a0.y + a1.y);
                v1[p] = make_double2(a0.x - a1.x, a0.y - a1.y);
                ob[p] = (b - k) * 2 + k;
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int b = t + p * kBlock;
            if (b < Q) {
                buf[ob[p]] = v0[p];
                buf[ob[p] + ns] = v1[p];
            }
        }
        __syncthreads();
    }
    for (int e = t; e < N; e += kBlock) line[e * elem_stride] = buf[e];
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_transfer(const double2* __restrict__ ghat, int mode, double eps,
                                                      float* __restrict__ T) {
    constexpr int NK = N / 2 + 1;
    const size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= (size_t)NK * N) return;
    const int k = (int)(e / N), ky = (int)(e - (size_t)k * N);
    const double g = ghat[(size_t)ky * N + k].x;        // Ghat is real (G real and even)
    double d = 1.0;
    if (mode == 1) {
        const int ksy = ky <= N / 2 ? ky : ky - N;
        d = cospi((double)k / N) * cospi((double)ksy / N);
    }
    const double inv_nn = 1.0 / ((double)N * (double)N);
    T[e] = (float)(g * d / (d * d + eps) * inv_nn);
}

template <int N>
hipError_t build_t(int mode, double green_origin, double eps, double* work64, float* T, hipStream_t s) {
    auto* g = reinterpret_cast<double2*>(work64);
    const size_t cells = (size_t)N * N;
    hipLaunchKernelGGL((k_green64<N>), dim3((unsigned)((cells + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, g,
                       green_origin);
    hipLaunchKernelGGL((k_fft64_line<N>), dim3(N), dim3(kBlock), 0, s, g, (int64_t)N, (int64_t)1);   // rows
    hipLaunchKernelGGL((k_fft64_line<N>), dim3(N), dim3(kBlock), 0, s, g, (int64_t)1, (int64_t)N);   // columns
    const size_t ne = (size_t)(N / 2 + 1) * N;
    hipLaunchKernelGGL((k_transfer<N>), dim3((unsigned)((ne + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, g, mode,
                       eps, T);
    return hipGetLastError();
}

}  // namespace

hipError_t build_transfer(int n, int mode, double green_origin, double eps, double* work64, float* T,
                          hipStream_t s) {
    switch (n) {
        case 64: return build_t<64>(mode, green_origin, eps, work64, T, s);
        case 128: return build_t<128>(mode, green_origin, eps, work64, T, s);
        case 256: return build_t<256>(mode, green_origin, eps, work64, T, s);
        case 512: return build_t<512>(mode, green_origin, eps, work64, T, s);
        case 1024: return build_t<1024>(mode, green_origin, eps, work64, T, s);
        case 2048: return build_t<2048>(mode, green_origin, eps, work64, T, s);
        case 4096: return build_t<4096>(mode, green_origin, eps, work64, T, s);
        case 8192: return build_t<8192>(mode, green_origin, eps, work64, T, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pmk
