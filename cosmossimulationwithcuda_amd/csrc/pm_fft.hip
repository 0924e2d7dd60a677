// FP32 field solve: deposited mass -> potential and force lattices.
//
// Replaces the reference's per-step spectral pipeline (CosmosCuda.cuh:
// k_copyAccumulatorIntoLattice 596-612, cufftExecC2C forward 1265-1270,
// k_multElementwiseLatticeFilter 149-173, k_divElementwiseLatticeFilter
// 120-148, cufftExecC2C inverse 1289-1295, k_shiftLattice 333-355,
// k_calcGradientLattice 174-229) with three LDS-resident passes over a
// real-to-complex half spectrum:
//
//   row_r2c   int64 mass rows -> FP32, two rows per complex FFT, split into two
//             half spectra, stored transposed: H[kx * N + y]
//   col_conv  per kx: forward column FFT, multiply by the real transfer function
//             T (Green's function x deconvolution x 1/N^2, built in FP64), then
//             inverse column FFTs of Phi-hat and of i*d(ky)*Phi-hat (y derivative)
//   row_c2r   per row pair: inverse row FFT of Phi (two rows per complex FFT)
//             and of Fx + i*Fy per row, with Fx-hat = i*d(kx)*Phi-hat
//
// The reference multiplies by conj(D)/(|D|^2+eps) where D carries a (-1)^(kx+ky)
// factor (mass kernel centred at N/2) and then shifts the result by N/2 on
// both axes; the two cancel exactly, so T holds the remaining real, even factor
// and no shift is applied.  d(k) = G*(8 sin(2 pi k/N) - sin(4 pi k/N))/6 is the
// Fourier symbol of the reference's 4th-order stencil; it reproduces that
// stencil exactly on every cell whose 5-point neighbourhood lies inside the
// lattice.  edge_force then recomputes the two outermost rows/columns with the
// reference's clamped stencil on the FP32 potential.  stencil_force applies that
// stencil to every cell (force_mode = stencil).
//
// FFT: in-place Stockham (radix 4, one radix-2 stage when log2 N is odd) on
// LDS, every thread of the block working across all transforms of the block;
// twiddles exp(-2 pi i m/N) come from a table built in FP64.
#include "pm_internal.h"

namespace pmk {
namespace {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Radix-2 first stage (Ns = 1, no twiddles) over NFFT transforms of length N.
template <int N, int NFFT>
__device__ __forceinline__ void stage2_first(float2* buf) {
    constexpr int Q = N / 2, NB = NFFT * Q, PER = (NB + kBlock - 1) / kBlock;
    const int t = threadIdx.x;
    float2 v0[PER], v1[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int bb = t + p * kBlock;
        if (NB % kBlock == 0 || bb < NB) {
            const int f = bb / Q, b = bb - f * Q;
            const float2 a0 = buf[f * N + b], a1 = buf[f * N + b + Q];
            v0[p] = cadd(a0, a1);
            v1[p] = csub(a0, a1);
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int bb = t + p * kBlock;
        if (NB % kBlock == 0 || bb < NB) {
            const int f = bb / Q, b = bb - f * Q;
            buf[f * N + 2 * b] = v0[p];
            buf[f * N + 2 * b + 1] = v1[p];
        }
    }
    __syncthreads();
}

// One radix-4 Stockham stage with sub-transform length NS.
template <int N, int NFFT, int NS, bool INV>
__device__ __forceinline__ void stage4(float2* buf, const float2* __restrict__ tw) {
    constexpr int Q = N / 4, NB = NFFT * Q, PER = (NB + kBlock - 1) / kBlock;
    constexpr int STEP = N / (4 * NS);
    const int t = threadIdx.x;
    float2 v[PER][4];
    int ob[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int bb = t + p * kBlock;
        if (NB % kBlock == 0 || bb < NB) {
            const int f = bb / Q, b = bb - f * Q;
            const float2* base = buf + f * N;
            float2 a0 = base[b], a1 = base[b + Q], a2 = base[b + 2 * Q], a3 = base[b + 3 * Q];
            const int k = b & (NS - 1);
            if (NS > 1) {
                const int e = k * STEP;
                float2 w1 = tw[e], w2 = tw[2 * e], w3 = tw[3 * e];
                if (INV) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
                a1 = cmul(a1, w1);
                a2 = cmul(a2, w2);
                a3 = cmul(a3, w3);
            }
            const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
            // forward: X1 = t1 - i t3, X3 = t1 + i t3; inverse swaps them
            const float2 mi_t3 = make_float2(t3.y, -t3.x);   // -i * t3
            v[p][0] = cadd(t0, t2);
            v[p][2] = csub(t0, t2);
            v[p][1] = INV ? csub(t1, mi_t3) : cadd(t1, mi_t3);
            v[p][3] = INV ? cadd(t1, mi_t3) : csub(t1, mi_t3);
            ob[p] = f * N + (b - k) * 4 + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int bb = t + p * kBlock;
        if (NB % kBlock == 0 || bb < NB) {
            float2* o = buf + ob[p];
            o[0] = v[p][0];
            o[NS] = v[p][1];
            o[2 * NS] = v[p][2];
            o[3 * NS] = v[p][3];
        }
    }
    __syncthreads();
}

template <int N, int NFFT, int NS, bool INV>
__device__ __forceinline__ void stages4_from(float2* buf, const float2* __restrict__ tw) {
    if constexpr (NS < N) {
        stage4<N, NFFT, NS, INV>(buf, tw);
        stages4_from<N, NFFT, NS * 4, INV>(buf, tw);
    }
}

// NFFT in-place transforms of length N at buf + f*N.  Caller synchronises
// before (data in LDS) and may read results right after (ends with a barrier).
template <int N, int NFFT, bool INV>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* __restrict__ tw) {
    if constexpr (ilog2c(N) & 1) {
        stage2_first<N, NFFT>(buf);
        stages4_from<N, NFFT, 2, INV>(buf, tw);
    } else {
        stages4_from<N, NFFT, 1, INV>(buf, tw);
    }
}

// ------------------------------------------------------------------ row R2C --
template <int N> constexpr int row_pairs() { return N >= 8192 ? 1 : 2; }

template <int N, int RP>
__global__ __launch_bounds__(kBlock) void k_row_r2c(const int64_t* __restrict__ rho, float inv_scale,
                                                     const float2* __restrict__ tw, float2* __restrict__ H) {
    __shared__ float2 buf[RP * N];
    const int t = threadIdx.x;
    const int j0 = blockIdx.x * 2 * RP;
    for (int e = t; e < RP * N; e += kBlock) {
        const int f = e / N, i = e - f * N;
        const size_t ra = (size_t)(j0 + 2 * f) * N + i;
        buf[e] = make_float2((float)rho[ra] * inv_scale, (float)rho[ra + N] * inv_scale);
    }
    __syncthreads();
    lds_fft<N, RP, false>(buf, tw);
    constexpr int NK = N / 2 + 1;
    for (int e = t; e < NK * RP; e += kBlock) {
        const int k = e / RP, f = e - k * RP;
        const float2 z = buf[f * N + k];
        const float2 zc = buf[f * N + ((N - k) & (N - 1))];
        // Xa = (Z + conj(Zc))/2,  Xb = (Z - conj(Zc))/(2i)
        const float2 xa = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
        const float2 xb = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
        float4* dst = reinterpret_cast<float4*>(H + (size_t)k * N + j0 + 2 * f);
        *dst = make_float4(xa.x, xa.y, xb.x, xb.y);
    }
}

// --------------------------------------------------------------- column pass --
template <int N> constexpr int col_batch() { return N >= 1024 ? 1 : 1024 / N; }

template <int N, int CB, bool FORCE>
__global__ __launch_bounds__(kBlock) void k_col_conv(float2* __restrict__ H, float2* __restrict__ Hfy,
                                                      const float* __restrict__ T, const float* __restrict__ dk,
                                                      const float2* __restrict__ tw) {
    __shared__ float2 buf[(FORCE ? 2 : 1) * CB * N];
    constexpr int NK = N / 2 + 1;
    const int t = threadIdx.x;
    const int k0 = blockIdx.x * CB;
    for (int e = t; e < CB * N; e += kBlock) {
        const int f = e / N, ky = e - f * N, k = k0 + f;
        buf[e] = k < NK ? H[(size_t)k * N + ky] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    lds_fft<N, CB, false>(buf, tw);
    for (int e = t; e < CB * N; e += kBlock) {
        const int f = e / N, ky = e - f * N, k = k0 + f;
        const float tv = k < NK ? T[(size_t)k * N + ky] : 0.f;
        const float2 y = make_float2(buf[e].x * tv, buf[e].y * tv);
        buf[e] = y;
        if (FORCE) {
            const float d = dk[ky];
            buf[CB * N + e] = make_float2(-y.y * d, y.x * d);   // i * d * y
        }
    }
    __syncthreads();
    lds_fft<N, (FORCE ? 2 : 1) * CB, true>(buf, tw);
    for (int e = t; e < CB * N; e += kBlock) {
        const int f = e / N, j = e - f * N, k = k0 + f;
        if (k < NK) {
            H[(size_t)k * N + j] = buf[e];
            if (FORCE) Hfy[(size_t)k * N + j] = buf[CB * N + e];
        }
    }
}

// ------------------------------------------------------------------ row C2R --
template <int N, int RP, bool FORCE>
__global__ __launch_bounds__(kBlock) void k_row_c2r(const float2* __restrict__ H, const float2* __restrict__ Hfy,
                                                     const float* __restrict__ dk, const float2* __restrict__ tw,
                                                     float* __restrict__ phi, float2* __restrict__ force) {
    __shared__ float2 buf[(FORCE ? 2 : 1) * RP * N];
    constexpr int NK = N / 2 + 1;
    const int t = threadIdx.x;
    const int j0 = blockIdx.x * 2 * RP;
    // potential: rows (j0+2f, j0+2f+1) packed as one complex row
    for (int e = t; e < NK * RP; e += kBlock) {
        const int k = e / RP, f = e - k * RP;
        const float4 v = *reinterpret_cast<const float4*>(H + (size_t)k * N + j0 + 2 * f);
        float2 xa = make_float2(v.x, v.y), xb = make_float2(v.z, v.w);
        if (k == 0 || k == N / 2) { xa.y = 0.f; xb.y = 0.f; }
        buf[f * N + k] = make_float2(xa.x - xb.y, xa.y + xb.x);
        if (k > 0 && k < N / 2) buf[f * N + N - k] = make_float2(xa.x + xb.y, xb.x - xa.y);
    }
    __syncthreads();
    lds_fft<N, RP, true>(buf, tw);
    for (int e = t; e < RP * N; e += kBlock) {
        const int f = e / N, i = e - f * N;
        const size_t ra = (size_t)(j0 + 2 * f) * N + i;
        const float2 z = buf[e];
        phi[ra] = z.x;
        phi[ra + N] = z.y;
    }
    if (!FORCE) return;
    __syncthreads();
    // force: one row per complex FFT, Z = Fx + i Fy
    constexpr int R2 = 2 * RP;
    for (int e = t; e < NK * R2; e += kBlock) {
        const int k = e / R2, q = e - k * R2;
        const size_t src = (size_t)k * N + j0 + q;
        const float2 p = H[src];
        float2 fy = Hfy[src];
        const float d = dk[k];
        const float2 fx = make_float2(-p.y * d, p.x * d);     // i * d(kx) * Phi-hat
        if (k == 0 || k == N / 2) fy.y = 0.f;
        buf[q * N + k] = make_float2(fx.x - fy.y, fx.y + fy.x);
        if (k > 0 && k < N / 2) buf[q * N + N - k] = make_float2(fx.x + fy.y, fy.x - fx.y);
    }
    __syncthreads();
    lds_fft<N, R2, true>(buf, tw);
    for (int e = t; e < R2 * N; e += kBlock) {
        const int q = e / N, i = e - q * N;
        force[(size_t)(j0 + q) * N + i] = buf[e];
    }
}

// ------------------------------------------------------------------ stencils --
// Reference rule (cuh:174-229): F = ((-P[+2] + 8 P[+1]) - 8 P[-1]) + P[-2], /12, *G;
// neighbours outside [0,N) take the centre value.
__device__ __forceinline__ float stencil5(float rr, float r, float l, float ll) {
    return ((-rr + 8.0f * r) - 8.0f * l) + ll;
}

template <int N>
__device__ __forceinline__ float grad_x(const float* __restrict__ phi, int i, int j) {
    const size_t c = (size_t)j * N + i;
    const float p0 = phi[c];
    const float l = i >= 1 ? phi[c - 1] : p0, r = i + 1 < N ? phi[c + 1] : p0;
    const float ll = i >= 2 ? phi[c - 2] : p0, rr = i + 2 < N ? phi[c + 2] : p0;
    return stencil5(rr, r, l, ll);
}

template <int N>
__device__ __forceinline__ float grad_y(const float* __restrict__ phi, int i, int j) {
    const size_t c = (size_t)j * N + i;
    const float p0 = phi[c];
    const float t = j >= 1 ? phi[c - N] : p0, b = j + 1 < N ? phi[c + N] : p0;
    const float tt = j >= 2 ? phi[c - 2 * N] : p0, bb = j + 2 < N ? phi[c + 2 * N] : p0;
    return stencil5(bb, b, t, tt);
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_edge_force(const float* __restrict__ phi, float g,
                                                        float2* __restrict__ force) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 8 * N) return;
    if (e < 4 * N) {                      // columns 0,1,N-2,N-1: Fx
        const int c = e / N, j = e - c * N;
        const int i = c < 2 ? c : N - 4 + c;
        force[(size_t)j * N + i].x = (grad_x<N>(phi, i, j) / 12.0f) * g;
    } else {                              // rows 0,1,N-2,N-1: Fy
        const int e2 = e - 4 * N;
        const int r = e2 / N, i = e2 - r * N;
        const int j = r < 2 ? r : N - 4 + r;
        force[(size_t)j * N + i].y = (grad_y<N>(phi, i, j) / 12.0f) * g;
    }
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_stencil(const float* __restrict__ phi, float g,
                                                     float2* __restrict__ force) {
    const size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= (size_t)N * N) return;
    const int i = (int)(c & (N - 1)), j = (int)(c / N);
    force[c] = make_float2((grad_x<N>(phi, i, j) / 12.0f) * g, (grad_y<N>(phi, i, j) / 12.0f) * g);
}

template <int N>
hipError_t row_r2c_t(const int64_t* rho, float inv_scale, const float2* tw, float2* H, hipStream_t s) {
    constexpr int RP = row_pairs<N>();
    hipLaunchKernelGGL((k_row_r2c<N, RP>), dim3(N / (2 * RP)), dim3(kBlock), 0, s, rho, inv_scale, tw, H);
    return hipGetLastError();
}

template <int N>
hipError_t col_conv_t(float2* H, float2* Hfy, const float* T, const float* dk, const float2* tw, bool force,
                      hipStream_t s) {
    constexpr int CB = col_batch<N>();
    const dim3 grid((N / 2 + 1 + CB - 1) / CB);
    if (force)
        hipLaunchKernelGGL((k_col_conv<N, CB, true>), grid, dim3(kBlock), 0, s, H, Hfy, T, dk, tw);
    else
        hipLaunchKernelGGL((k_col_conv<N, CB, false>), grid, dim3(kBlock), 0, s, H, Hfy, T, dk, tw);
    return hipGetLastError();
}

template <int N>
hipError_t row_c2r_t(const float2* H, const float2* Hfy, const float* dk, const float2* tw, float* phi,
                     float2* force, bool with_force, hipStream_t s) {
    constexpr int RP = row_pairs<N>();
    const dim3 grid(N / (2 * RP));
    if (with_force)
        hipLaunchKernelGGL((k_row_c2r<N, RP, true>), grid, dim3(kBlock), 0, s, H, Hfy, dk, tw, phi, force);
    else
        hipLaunchKernelGGL((k_row_c2r<N, RP, false>), grid, dim3(kBlock), 0, s, H, Hfy, dk, tw, phi, force);
    return hipGetLastError();
}

template <int N>
hipError_t edge_t(const float* phi, float g, float2* force, hipStream_t s) {
    hipLaunchKernelGGL((k_edge_force<N>), dim3((8 * N + kBlock - 1) / kBlock), dim3(kBlock), 0, s, phi, g, force);
    return hipGetLastError();
}

template <int N>
hipError_t stencil_t(const float* phi, float g, float2* force, hipStream_t s) {
    const size_t cells = (size_t)N * N;
    hipLaunchKernelGGL((k_stencil<N>), dim3((unsigned)((cells + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, phi, g,
                       force);
    return hipGetLastError();
}

}  // namespace

#define PM_DISPATCH_N(n, CALL)                 \
    switch (n) {                               \
        case 64: { constexpr int NN = 64; return CALL; }     \
        case 128: { constexpr int NN = 128; return CALL; }   \
        case 256: { constexpr int NN = 256; return CALL; }   \
        case 512: { constexpr int NN = 512; return CALL; }   \
        case 1024: { constexpr int NN = 1024; return CALL; } \
        case 2048: { constexpr int NN = 2048; return CALL; } \
        case 4096: { constexpr int NN = 4096; return CALL; } \
        case 8192: { constexpr int NN = 8192; return CALL; } \
        default: return hipErrorInvalidValue;  \
    }

hipError_t row_r2c(int n, const int64_t* rho, float inv_scale, const float2* tw, float2* H, hipStream_t s) {
    PM_DISPATCH_N(n, (row_r2c_t<NN>(rho, inv_scale, tw, H, s)));
}

hipError_t col_conv(int n, float2* H, float2* Hfy, const float* T, const float* dk, const float2* tw,
                    bool with_force, hipStream_t s) {
    PM_DISPATCH_N(n, (col_conv_t<NN>(H, Hfy, T, dk, tw, with_force, s)));
}

hipError_t row_c2r(int n, const float2* H, const float2* Hfy, const float* dk, const float2* tw, float* phi,
                   float2* force, bool with_force, hipStream_t s) {
    PM_DISPATCH_N(n, (row_c2r_t<NN>(H, Hfy, dk, tw, phi, force, with_force, s)));
}

hipError_t edge_force(int n, const float* phi, float g, float2* force, hipStream_t s) {
    PM_DISPATCH_N(n, (edge_t<NN>(phi, g, force, s)));
}

hipError_t stencil_force(int n, const float* phi, float g, float2* force, hipStream_t s) {
    PM_DISPATCH_N(n, (stencil_t<NN>(phi, g, force, s)));
}

}  // namespace pmk
