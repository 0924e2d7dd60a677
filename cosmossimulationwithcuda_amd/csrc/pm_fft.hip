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
// FFT: in-place Stockham on LDS, every thread of the block working across all
// transforms of the block; stages of radix 4 or 8 (plan choice) after one
// leading radix-2/4 stage for the leftover factor; twiddles exp(-2 pi i m/N)
// from a table built in FP64.
//
// Every pass reads its arrays from a FieldDesc; blockIdx.y selects the
// instance, so one launch serves one simulation or a batch of simulations of
// the same lattice size.
#include "pm_internal.h"

namespace pmk {
namespace {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// -i * a (forward quarter turn); +i * a is its negation
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

template <bool INV>
__device__ __forceinline__ void dft2(float2* v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2* v) {
    const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]), t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    const float2 r = mul_mi(t3);      // forward: X1 = t1 - i t3, X3 = t1 + i t3
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = INV ? csub(t1, r) : cadd(t1, r);
    v[3] = INV ? cadd(t1, r) : csub(t1, r);
}

// 8-point DFT as a radix-2 split into two 4-point DFTs:
// b_k = a_k + a_{k+4}, c_k = (a_k - a_{k+4}) W8^k, X_{2m} = DFT4(b)_m, X_{2m+1} = DFT4(c)_m.
template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
    constexpr float h = 0.70710678118654752440f;
    float2 b[4], c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        b[k] = cadd(v[k], v[k + 4]);
        c[k] = csub(v[k], v[k + 4]);
    }
    // W8^1 = (1 - i)/sqrt2 (forward), W8^2 = -i, W8^3 = (-1 - i)/sqrt2; inverse conjugates
    if (!INV) {
        c[1] = make_float2(h * (c[1].x + c[1].y), h * (c[1].y - c[1].x));
        c[2] = mul_mi(c[2]);
        c[3] = make_float2(h * (c[3].y - c[3].x), -h * (c[3].x + c[3].y));
    } else {
        c[1] = make_float2(h * (c[1].x - c[1].y), h * (c[1].y + c[1].x));
        c[2] = make_float2(-c[2].y, c[2].x);
        c[3] = make_float2(-h * (c[3].x + c[3].y), h * (c[3].x - c[3].y));
    }
    dft4<INV>(b);
    dft4<INV>(c);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        v[2 * m] = b[m];
        v[2 * m + 1] = c[m];
    }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else dft8<INV>(v);
}

// One Stockham stage of radix R with sub-transform length NS over NFFT
// in-place transforms of length N at buf + f*N.  Reads, barrier, writes, barrier.
template <int N, int NFFT, int R, int NS, bool INV>
__device__ __forceinline__ void stage(float2* buf, const float2* __restrict__ tw) {
    constexpr int Q = N / R, NB = NFFT * Q, PER = (NB + kBlock - 1) / kBlock;
    constexpr int STEP = N / (R * NS);
    const int t = threadIdx.x;
    float2 v[PER][R];
    int ob[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int bb = t + p * kBlock;
        if (NB % kBlock == 0 || bb < NB) {
            const int f = bb / Q, b = bb - f * Q;
            const float2* base = buf + f * N;
#pragma unroll
            for (int r = 0; r < R; ++r) v[p][r] = base[b + r * Q];
            const int k = b & (NS - 1);
            if (NS > 1) {
                const int e = k * STEP;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    float2 w = tw[r * e];
                    if (INV) w.y = -w.y;
                    v[p][r] = cmul(v[p][r], w);
                }
            }
            dft<R, INV>(v[p]);
            ob[p] = f * N + (b - k) * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int bb = t + p * kBlock;
        if (NB % kBlock == 0 || bb < NB) {
            float2* o = buf + ob[p];
#pragma unroll
            for (int r = 0; r < R; ++r) o[r * NS] = v[p][r];
        }
    }
    __syncthreads();
}

template <int N, int NFFT, int R, int NS, bool INV>
__device__ __forceinline__ void stages_from(float2* buf, const float2* __restrict__ tw) {
    if constexpr (NS < N) {
        stage<N, NFFT, R, NS, INV>(buf, tw);
        stages_from<N, NFFT, R, NS * R, INV>(buf, tw);
    }
}

// NFFT in-place transforms of length N at buf + f*N, radix RADIX (4 or 8)
// after a leading stage for the leftover factor.  Caller synchronises before;
// ends with a barrier.
template <int N, int NFFT, int RADIX, bool INV>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* __restrict__ tw) {
    constexpr int LR = ilog2c(RADIX), L = ilog2c(N), REM = L % LR;
    if constexpr (REM == 0) {
        stages_from<N, NFFT, RADIX, 1, INV>(buf, tw);
    } else {
        constexpr int R0 = 1 << REM;
        stage<N, NFFT, R0, 1, INV>(buf, tw);
        stages_from<N, NFFT, RADIX, R0, INV>(buf, tw);
    }
}

// ------------------------------------------------------------------ row R2C --
template <int N, int RP, int RADIX>
__global__ __launch_bounds__(kBlock) void k_row_r2c(const FieldDesc* __restrict__ descs,
                                                     const float2* __restrict__ tw) {
    __shared__ float2 buf[RP * N];
    const FieldDesc d = descs[blockIdx.y];
    const int t = threadIdx.x;
    const int j0 = blockIdx.x * 2 * RP;
    for (int e = t; e < RP * N; e += kBlock) {
        const int f = e / N, i = e - f * N;
        const size_t ra = (size_t)(j0 + 2 * f) * N + i;
        buf[e] = make_float2((float)d.rho[ra] * d.inv_scale, (float)d.rho[ra + N] * d.inv_scale);
    }
    __syncthreads();
    lds_fft<N, RP, RADIX, false>(buf, tw);
    constexpr int NK = N / 2 + 1;
    for (int e = t; e < NK * RP; e += kBlock) {
        const int k = e / RP, f = e - k * RP;
        const float2 z = buf[f * N + k];
        const float2 zc = buf[f * N + ((N - k) & (N - 1))];
        // Xa = (Z + conj(Zc))/2,  Xb = (Z - conj(Zc))/(2i)
        const float2 xa = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
        const float2 xb = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
        float4* dst = reinterpret_cast<float4*>(d.H + (size_t)k * N + j0 + 2 * f);
        *dst = make_float4(xa.x, xa.y, xb.x, xb.y);
    }
}

// --------------------------------------------------------------- column pass --
template <int N, int CB, int RADIX, bool FORCE>
__global__ __launch_bounds__(kBlock) void k_col_conv(const FieldDesc* __restrict__ descs,
                                                      const float2* __restrict__ tw) {
    __shared__ float2 buf[(FORCE ? 2 : 1) * CB * N];
    const FieldDesc d = descs[blockIdx.y];
    constexpr int NK = N / 2 + 1;
    const int t = threadIdx.x;
    const int k0 = blockIdx.x * CB;
    for (int e = t; e < CB * N; e += kBlock) {
        const int f = e / N, ky = e - f * N, k = k0 + f;
        buf[e] = k < NK ? d.H[(size_t)k * N + ky] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    lds_fft<N, CB, RADIX, false>(buf, tw);
    for (int e = t; e < CB * N; e += kBlock) {
        const int f = e / N, ky = e - f * N, k = k0 + f;
        const float tv = k < NK ? d.T[(size_t)k * N + ky] : 0.f;
        const float2 y = make_float2(buf[e].x * tv, buf[e].y * tv);
        buf[e] = y;
        if (FORCE) {
            const float dv = d.dk[ky];
            buf[CB * N + e] = make_float2(-y.y * dv, y.x * dv);   // i * d * y
        }
    }
    __syncthreads();
    lds_fft<N, (FORCE ? 2 : 1) * CB, RADIX, true>(buf, tw);
    for (int e = t; e < CB * N; e += kBlock) {
        const int f = e / N, j = e - f * N, k = k0 + f;
        if (k < NK) {
            d.H[(size_t)k * N + j] = buf[e];
            if (FORCE) d.Hfy[(size_t)k * N + j] = buf[CB * N + e];
        }
    }
}

// ------------------------------------------------------------------ row C2R --
template <int N, int RP, int RADIX, bool FORCE>
__global__ __launch_bounds__(kBlock) void k_row_c2r(const FieldDesc* __restrict__ descs,
                                                     const float2* __restrict__ tw) {
    __shared__ float2 buf[(FORCE ? 2 : 1) * RP * N];
    const FieldDesc d = descs[blockIdx.y];
    constexpr int NK = N / 2 + 1;
    const int t = threadIdx.x;
    const int j0 = blockIdx.x * 2 * RP;
    // potential: rows (j0+2f, j0+2f+1) packed as one complex row
    for (int e = t; e < NK * RP; e += kBlock) {
        const int k = e / RP, f = e - k * RP;
        const float4 v = *reinterpret_cast<const float4*>(d.H + (size_t)k * N + j0 + 2 * f);
        float2 xa = make_float2(v.x, v.y), xb = make_float2(v.z, v.w);
        if (k == 0 || k == N / 2) { xa.y = 0.f; xb.y = 0.f; }
        buf[f * N + k] = make_float2(xa.x - xb.y, xa.y + xb.x);
        if (k > 0 && k < N / 2) buf[f * N + N - k] = make_float2(xa.x + xb.y, xb.x - xa.y);
    }
    __syncthreads();
    lds_fft<N, RP, RADIX, true>(buf, tw);
    for (int e = t; e < RP * N; e += kBlock) {
        const int f = e / N, i = e - f * N;
        const size_t ra = (size_t)(j0 + 2 * f) * N + i;
        const float2 z = buf[e];
        d.phi[ra] = z.x;
        d.phi[ra + N] = z.y;
    }
    if (!FORCE) return;
    __syncthreads();
    // force: one row per complex FFT, Z = Fx + i Fy
    constexpr int R2 = 2 * RP;
    for (int e = t; e < NK * R2; e += kBlock) {
        const int k = e / R2, q = e - k * R2;
        const size_t src = (size_t)k * N + j0 + q;
        const float2 p = d.H[src];
        float2 fy = d.Hfy[src];
        const float dv = d.dk[k];
        const float2 fx = make_float2(-p.y * dv, p.x * dv);     // i * d(kx) * Phi-hat
        if (k == 0 || k == N / 2) fy.y = 0.f;
        buf[q * N + k] = make_float2(fx.x - fy.y, fx.y + fy.x);
        if (k > 0 && k < N / 2) buf[q * N + N - k] = make_float2(fx.x + fy.y, fy.x - fx.y);
    }
    __syncthreads();
    lds_fft<N, R2, RADIX, true>(buf, tw);
    for (int e = t; e < R2 * N; e += kBlock) {
        const int q = e / N, i = e - q * N;
        d.force[(size_t)(j0 + q) * N + i] = buf[e];
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
__global__ __launch_bounds__(kBlock) void k_edge_force(const FieldDesc* __restrict__ descs) {
    const FieldDesc d = descs[blockIdx.y];
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 8 * N) return;
    if (e < 4 * N) {                      // columns 0,1,N-2,N-1: Fx
        const int c = e / N, j = e - c * N;
        const int i = c < 2 ? c : N - 4 + c;
        d.force[(size_t)j * N + i].x = (grad_x<N>(d.phi, i, j) / 12.0f) * d.gravity;
    } else {                              // rows 0,1,N-2,N-1: Fy
        const int e2 = e - 4 * N;
        const int r = e2 / N, i = e2 - r * N;
        const int j = r < 2 ? r : N - 4 + r;
        d.force[(size_t)j * N + i].y = (grad_y<N>(d.phi, i, j) / 12.0f) * d.gravity;
    }
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_stencil(const FieldDesc* __restrict__ descs) {
    const FieldDesc d = descs[blockIdx.y];
    const size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= (size_t)N * N) return;
    const int i = (int)(c & (N - 1)), j = (int)(c / N);
    d.force[c] = make_float2((grad_x<N>(d.phi, i, j) / 12.0f) * d.gravity,
                             (grad_y<N>(d.phi, i, j) / 12.0f) * d.gravity);
}

// ------------------------------------------------------------------ launchers --
template <int N, int RP, int RADIX>
hipError_t r2c_t(const FieldDesc* d, int count, const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_row_r2c<N, RP, RADIX>), dim3(N / (2 * RP), count), dim3(kBlock), 0, s, d, tw);
    return hipGetLastError();
}

template <int N, int CB, int RADIX>
hipError_t col_t(const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s) {
    const dim3 grid((N / 2 + 1 + CB - 1) / CB, count);
    if (force) hipLaunchKernelGGL((k_col_conv<N, CB, RADIX, true>), grid, dim3(kBlock), 0, s, d, tw);
    else hipLaunchKernelGGL((k_col_conv<N, CB, RADIX, false>), grid, dim3(kBlock), 0, s, d, tw);
    return hipGetLastError();
}

template <int N, int RP, int RADIX>
hipError_t c2r_t(const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s) {
    const dim3 grid(N / (2 * RP), count);
    if (force) hipLaunchKernelGGL((k_row_c2r<N, RP, RADIX, true>), grid, dim3(kBlock), 0, s, d, tw);
    else hipLaunchKernelGGL((k_row_c2r<N, RP, RADIX, false>), grid, dim3(kBlock), 0, s, d, tw);
    return hipGetLastError();
}

// Row batches RP in {1,2,4} and column batches CB in {1,2,4,8} are compiled
// where the LDS footprint (2*RP*N*8 or 2*CB*N*8 bytes with the force pass)
// fits 160 KB; plan_valid() states the same rule for the host.
template <int N, int RP>
constexpr bool rp_ok() { return 2 * RP * N * 8 <= 160 * 1024 && 2 * RP <= N; }
template <int N, int CB>
constexpr bool cb_ok() { return 2 * CB * N * 8 <= 160 * 1024 && CB <= N / 2 + 1; }

template <int N>
hipError_t r2c_n(const FieldDesc* d, int count, const float2* tw, FftPlan p, hipStream_t s) {
#define PM_R2C(RPV, RV) \
    if constexpr (rp_ok<N, RPV>()) if (p.rp == RPV && p.radix == RV) return r2c_t<N, RPV, RV>(d, count, tw, s);
    PM_R2C(1, 4) PM_R2C(2, 4) PM_R2C(4, 4) PM_R2C(1, 8) PM_R2C(2, 8) PM_R2C(4, 8)
#undef PM_R2C
    return hipErrorInvalidValue;
}

template <int N>
hipError_t col_n(const FieldDesc* d, int count, const float2* tw, FftPlan p, bool force, hipStream_t s) {
#define PM_COL(CBV, RV) \
    if constexpr (cb_ok<N, CBV>()) if (p.cb == CBV && p.radix == RV) return col_t<N, CBV, RV>(d, count, tw, force, s);
    PM_COL(1, 4) PM_COL(2, 4) PM_COL(4, 4) PM_COL(8, 4) PM_COL(1, 8) PM_COL(2, 8) PM_COL(4, 8) PM_COL(8, 8)
#undef PM_COL
    return hipErrorInvalidValue;
}

template <int N>
hipError_t c2r_n(const FieldDesc* d, int count, const float2* tw, FftPlan p, bool force, hipStream_t s) {
#define PM_C2R(RPV, RV) \
    if constexpr (rp_ok<N, RPV>()) if (p.rp == RPV && p.radix == RV) return c2r_t<N, RPV, RV>(d, count, tw, force, s);
    PM_C2R(1, 4) PM_C2R(2, 4) PM_C2R(4, 4) PM_C2R(1, 8) PM_C2R(2, 8) PM_C2R(4, 8)
#undef PM_C2R
    return hipErrorInvalidValue;
}

template <int N>
hipError_t edge_n(const FieldDesc* d, int count, hipStream_t s) {
    hipLaunchKernelGGL((k_edge_force<N>), dim3((8 * N + kBlock - 1) / kBlock, count), dim3(kBlock), 0, s, d);
    return hipGetLastError();
}

template <int N>
hipError_t stencil_n(const FieldDesc* d, int count, hipStream_t s) {
    const size_t cells = (size_t)N * N;
    hipLaunchKernelGGL((k_stencil<N>), dim3((unsigned)((cells + kBlock - 1) / kBlock), count), dim3(kBlock), 0, s, d);
    return hipGetLastError();
}

}  // namespace

#define PM_DISPATCH_N(n, CALL)                               \
    switch (n) {                                             \
        case 64: { constexpr int NN = 64; return CALL; }     \
        case 128: { constexpr int NN = 128; return CALL; }   \
        case 256: { constexpr int NN = 256; return CALL; }   \
        case 512: { constexpr int NN = 512; return CALL; }   \
        case 1024: { constexpr int NN = 1024; return CALL; } \
        case 2048: { constexpr int NN = 2048; return CALL; } \
        case 4096: { constexpr int NN = 4096; return CALL; } \
        case 8192: { constexpr int NN = 8192; return CALL; } \
        default: return hipErrorInvalidValue;                \
    }

bool plan_valid(int n, FftPlan p) {
    if (p.radix != 4 && p.radix != 8) return false;
    if (p.rp != 1 && p.rp != 2 && p.rp != 4) return false;
    if (p.cb != 1 && p.cb != 2 && p.cb != 4 && p.cb != 8) return false;
    const long lds_r = 2L * p.rp * n * 8, lds_c = 2L * p.cb * n * 8;
    return lds_r <= 160 * 1024 && lds_c <= 160 * 1024 && 2 * p.rp <= n && p.cb <= n / 2 + 1;
}

FftPlan default_plan(int n) {
    FftPlan p;
    p.radix = 4;
    p.rp = n >= 8192 ? 1 : 2;
    p.cb = n >= 1024 ? 1 : (n >= 512 ? 2 : 4);
    return p;
}

hipError_t row_r2c(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, hipStream_t s) {
    PM_DISPATCH_N(n, (r2c_n<NN>(d, count, tw, p, s)));
}

hipError_t col_conv(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                    hipStream_t s) {
    PM_DISPATCH_N(n, (col_n<NN>(d, count, tw, p, with_force, s)));
}

hipError_t row_c2r(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                   hipStream_t s) {
    PM_DISPATCH_N(n, (c2r_n<NN>(d, count, tw, p, with_force, s)));
}

hipError_t edge_force(int n, const FieldDesc* d, int count, hipStream_t s) {
    PM_DISPATCH_N(n, (edge_n<NN>(d, count, s)));
}

hipError_t stencil_force(int n, const FieldDesc* d, int count, hipStream_t s) {
    PM_DISPATCH_N(n, (stencil_n<NN>(d, count, s)));
}

}  // namespace pmk
