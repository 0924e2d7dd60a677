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
// reference's clamped stencil on the FP32 potential (the prefetching C2R pass
// of pm_fft_pipe.hip does this itself).  stencil_force applies that stencil to
// every cell (force_mode = stencil).  Inside pm_step the row R2C pass also
// zeroes the other int64 lattice, the next deposit target.
//
// FFT: in-place Stockham on LDS, every thread of the block working across all
// transforms of the block; stages of radix 4 or 8 (plan choice) after one
// leading radix-2/4 stage for the leftover factor; twiddles exp(-2 pi i m/N)
// from a quarter-wave table built in FP64 and staged in LDS (stages whose
// twiddle indices are far apart read per-stage copies of the same values);
// LDS rows are padded against bank conflicts.
//
// Every pass reads its arrays from a FieldDesc; blockIdx.y selects the
// instance, so one launch serves one simulation or a batch of simulations of
// the same lattice size.
#include "pm_fft_core.h"

namespace pmk {
namespace {

// ------------------------------------------------------------------ row R2C --
template <int N, int RP, int RADIX>
__global__ __launch_bounds__((fft_threads<N, RADIX>())) void k_row_r2c(const FieldDesc* __restrict__ descs,
                                                     const float2* __restrict__ tw) {
    constexpr int BT = fft_threads<N, RADIX>();
    __shared__ float2 buf[padded(RP * N)];
    __shared__ float2 twl[tw_lds_entries(N, RADIX)];
    const GField d = gfield(descs[blockIdx.y]);
    const int t = threadIdx.x;
    const int j0 = row_group() * 2 * RP;
    load_twiddles<N, RADIX, BT>(twl, tw);
    for (int e = t; e < RP * N; e += BT) {
        const int f = e / N, i = e - f * N;
        const size_t ra = (size_t)(j0 + 2 * f) * N + i;
        buf[pad(e)] = make_float2((float)d.rho[ra] * d.inv_scale, (float)d.rho[ra + N] * d.inv_scale);
        if (d.rho_zero) {
            d.rho_zero[ra] = 0;
            d.rho_zero[ra + N] = 0;
        }
    }
    lds_barrier();
    lds_fft<N, RP, RADIX, false>(buf, twl);
    constexpr int NK = N / 2 + 1;
    for (int e = t; e < NK * RP; e += BT) {
        const int k = e / RP, f = e - k * RP;
        const float2 z = buf[pad(f * N + k)];
        const float2 zc = buf[pad(f * N + ((N - k) & (N - 1)))];
        // Xa = (Z + conj(Zc))/2,  Xb = (Z - conj(Zc))/(2i)
        const float2 xa = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
        const float2 xb = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
        *(gptr<cplx2>)(d.H + (size_t)k * N + j0 + 2 * f) = cplx2{xa.x, xa.y, xb.x, xb.y};
    }
}

// --------------------------------------------------------------- column pass --
// Blocks cover kx = blockIdx.x * CB + f, f < CB, for kx in [0, N/2).  The
// columns kx = 0 and kx = N/2 are real (both rows of a row-pair FFT are real
// there), so slot kx = 0 carries both as one complex column z = H0 + i*HN2:
// its spectrum Z splits into A = (Z + conj Z(-k))/2 and B = (Z - conj Z(-k))/2i,
// and W = A*T0 + i*B*TN2 transforms back to the two real results (real and
// imaginary parts).  N/2 blocks of CB = 1 then fill the device evenly.
template <int N, int CB, int RADIX, bool FORCE>
__global__ __launch_bounds__((fft_threads<N, RADIX>())) void k_col_conv(const FieldDesc* __restrict__ descs,
                                                      const float2* __restrict__ tw) {
    constexpr int BT = fft_threads<N, RADIX>();
    __shared__ float2 buf[padded((FORCE ? 2 : 1) * CB * N)];
    __shared__ float2 twl[tw_lds_entries(N, RADIX)];
    const GField d = gfield(descs[blockIdx.y]);
    constexpr int NY = N / 2;                 // the Nyquist column's kx
    const int t = threadIdx.x;
    const int k0 = blockIdx.x * CB;
    load_twiddles<N, RADIX, BT>(twl, tw);
    for (int e = t; e < CB * N; e += BT) {
        const int f = e / N, ky = e - f * N, k = k0 + f;
        buf[pad(e)] = k == 0 ? make_float2(d.H[ky].x, d.H[(size_t)NY * N + ky].x) : f2(d.H[(size_t)k * N + ky]);
    }
    lds_barrier();
    lds_fft<N, CB, RADIX, false>(buf, twl);
    for (int e = (k0 == 0 ? N : 0) + t; e < CB * N; e += BT) {
        const int f = e / N, ky = e - f * N, k = k0 + f;
        const float tv = d.T[(size_t)k * N + ky];
        const float2 x = buf[pad(e)];
        const float2 y = make_float2(x.x * tv, x.y * tv);
        buf[pad(e)] = y;
        if (FORCE) {
            const float dv = d.dk[ky];
            buf[pad(CB * N + e)] = make_float2(-y.y * dv, y.x * dv);   // i * d * y
        }
    }
    if (k0 == 0) {                            // packed kx = 0 / N/2 slot, by conjugate pairs
        for (int a = t; a <= N / 2; a += BT) {
            const int b = (N - a) & (N - 1);
            const float2 za = buf[pad(a)], zb = buf[pad(b)];
            const float2 A = make_float2(0.5f * (za.x + zb.x), 0.5f * (za.y - zb.y));
            const float2 B = make_float2(0.5f * (za.y + zb.y), -0.5f * (za.x - zb.x));
            const float t0a = d.T[a], t0b = d.T[b];
            const float tna = d.T[(size_t)NY * N + a], tnb = d.T[(size_t)NY * N + b];
            // W(a) = A T0 + i B TN;  W(b) = conj(A) T0 + i conj(B) TN
            const float2 wa = make_float2(A.x * t0a - B.y * tna, A.y * t0a + B.x * tna);
            const float2 wb = make_float2(A.x * t0b + B.y * tnb, -A.y * t0b + B.x * tnb);
            buf[pad(a)] = wa;
            buf[pad(b)] = wb;
            if (FORCE) {
                const float da = d.dk[a], db = d.dk[b];
                buf[pad(CB * N + a)] = make_float2(-wa.y * da, wa.x * da);
                buf[pad(CB * N + b)] = make_float2(-wb.y * db, wb.x * db);
            }
        }
    }
    lds_barrier();
    lds_fft<N, (FORCE ? 2 : 1) * CB, RADIX, true>(buf, twl);
    for (int e = t; e < CB * N; e += BT) {
        const int f = e / N, j = e - f * N, k = k0 + f;
        const float2 w = buf[pad(e)];
        const float2 v = FORCE ? buf[pad(CB * N + e)] : make_float2(0.f, 0.f);
        if (k == 0) {
            d.H[j] = cplx{w.x, 0.f};
            d.H[(size_t)NY * N + j] = cplx{w.y, 0.f};
            if (FORCE) {
                d.Hfy[j] = cplx{v.x, 0.f};
                d.Hfy[(size_t)NY * N + j] = cplx{v.y, 0.f};
            }
        } else {
            d.H[(size_t)k * N + j] = nv(w);
            if (FORCE) d.Hfy[(size_t)k * N + j] = nv(v);
        }
    }
}

// ------------------------------------------------------------------ row C2R --
template <int N, int RP, int RADIX, bool FORCE>
__global__ __launch_bounds__((fft_threads<N, RADIX>())) void k_row_c2r(const FieldDesc* __restrict__ descs,
                                                     const float2* __restrict__ tw) {
    constexpr int BT = fft_threads<N, RADIX>();
    // MERGE: the potential row pairs and the force rows go through one
    // transform batch (one read of the spectra, half the stage barriers).
    // Measured faster up to N = 1024; above, the larger LDS footprint halves
    // the blocks in flight and costs more than it saves.
    constexpr bool MERGE = FORCE && N <= 1024 && padded(3 * RP * N) * 8 + 8 * tw_lds_entries(N, RADIX) <= 160 * 1024;
    __shared__ float2 buf[padded((MERGE ? 3 : (FORCE ? 2 : 1)) * RP * N)];
    __shared__ float2 twl[tw_lds_entries(N, RADIX)];
    const GField d = gfield(descs[blockIdx.y]);
    constexpr int NK = N / 2 + 1;
    const int t = threadIdx.x;
    const int j0 = row_group() * 2 * RP;
    load_twiddles<N, RADIX, BT>(twl, tw);
    if constexpr (MERGE) {
        for (int e = t; e < NK * RP; e += BT) {
            const int k = e / RP, f = e - k * RP;
            const size_t src = (size_t)k * N + j0 + 2 * f;
            const cplx2 v = *(gptr<const cplx2>)(d.H + src);
            const cplx2 w = *(gptr<const cplx2>)(d.Hfy + src);
            const bool edge = k == 0 || k == N / 2;
            float2 xa = make_float2(v.x, v.y), xb = make_float2(v.z, v.w);
            if (edge) { xa.y = 0.f; xb.y = 0.f; }
            buf[pad(f * N + k)] = make_float2(xa.x - xb.y, xa.y + xb.x);
            if (!edge) buf[pad(f * N + N - k)] = make_float2(xa.x + xb.y, xb.x - xa.y);
            const float dv = d.dk[k];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float2 p = h ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
                float2 fy = h ? make_float2(w.z, w.w) : make_float2(w.x, w.y);
                const float2 fx = make_float2(-p.y * dv, p.x * dv);     // i * d(kx) * Phi-hat
                if (edge) fy.y = 0.f;
                const int row = RP + 2 * f + h;
                buf[pad(row * N + k)] = make_float2(fx.x - fy.y, fx.y + fy.x);
                if (!edge) buf[pad(row * N + N - k)] = make_float2(fx.x + fy.y, fy.x - fx.y);
            }
        }
        lds_barrier();
        lds_fft<N, 3 * RP, RADIX, true>(buf, twl);
        for (int e = t; e < RP * N; e += BT) {
            const int f = e / N, i = e - f * N;
            const size_t ra = (size_t)(j0 + 2 * f) * N + i;
            const float2 z = buf[pad(e)];
            d.phi[ra] = z.x;
            d.phi[ra + N] = z.y;
        }
        for (int e = t; e < 2 * RP * N; e += BT) {
            const int q = e / N, i = e - q * N;
            d.force[(size_t)(j0 + q) * N + i] = nv(buf[pad(RP * N + e)]);
        }
        return;
    }
    // potential: rows (j0+2f, j0+2f+1) packed as one complex row
    for (int e = t; e < NK * RP; e += BT) {
        const int k = e / RP, f = e - k * RP;
        const cplx2 v = *(gptr<const cplx2>)(d.H + (size_t)k * N + j0 + 2 * f);
        float2 xa = make_float2(v.x, v.y), xb = make_float2(v.z, v.w);
        if (k == 0 || k == N / 2) { xa.y = 0.f; xb.y = 0.f; }
        buf[pad(f * N + k)] = make_float2(xa.x - xb.y, xa.y + xb.x);
        if (k > 0 && k < N / 2) buf[pad(f * N + N - k)] = make_float2(xa.x + xb.y, xb.x - xa.y);
    }
    lds_barrier();
    lds_fft<N, RP, RADIX, true>(buf, twl);
    for (int e = t; e < RP * N; e += BT) {
        const int f = e / N, i = e - f * N;
        const size_t ra = (size_t)(j0 + 2 * f) * N + i;
        const float2 z = buf[pad(e)];
        d.phi[ra] = z.x;
        d.phi[ra + N] = z.y;
    }
    if (!FORCE) return;
    lds_barrier();
    // force: one row per complex FFT, Z = Fx + i Fy
    constexpr int R2 = 2 * RP;
    for (int e = t; e < NK * R2; e += BT) {
        const int k = e / R2, q = e - k * R2;
        const size_t src = (size_t)k * N + j0 + q;
        const float2 p = f2(d.H[src]);
        float2 fy = f2(d.Hfy[src]);
        const float dv = d.dk[k];
        const float2 fx = make_float2(-p.y * dv, p.x * dv);     // i * d(kx) * Phi-hat
        if (k == 0 || k == N / 2) fy.y = 0.f;
        buf[pad(q * N + k)] = make_float2(fx.x - fy.y, fx.y + fy.x);
        if (k > 0 && k < N / 2) buf[pad(q * N + N - k)] = make_float2(fx.x + fy.y, fy.x - fx.y);
    }
    lds_barrier();
    lds_fft<N, R2, RADIX, true>(buf, twl);
    for (int e = t; e < R2 * N; e += BT) {
        const int q = e / N, i = e - q * N;
        d.force[(size_t)(j0 + q) * N + i] = nv(buf[pad(e)]);
    }
}

// ------------------------------------------------------------------ stencils --
// Reference rule (cuh:174-229): F = ((-P[+2] + 8 P[+1]) - 8 P[-1]) + P[-2], /12, *G;
// neighbours outside [0,N) take the centre value (stencil5 in pm_fft_core.h).

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
    hipLaunchKernelGGL((k_row_r2c<N, RP, RADIX>), dim3(N / (2 * RP), count), dim3(fft_threads<N, RADIX>()), 0, s, d, tw);
    return hipGetLastError();
}

template <int N, int CB, int RADIX>
hipError_t col_t(const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s) {
    const dim3 grid(N / 2 / CB, count);
    if (force) hipLaunchKernelGGL((k_col_conv<N, CB, RADIX, true>), grid, dim3(fft_threads<N, RADIX>()), 0, s, d, tw);
    else hipLaunchKernelGGL((k_col_conv<N, CB, RADIX, false>), grid, dim3(fft_threads<N, RADIX>()), 0, s, d, tw);
    return hipGetLastError();
}

template <int N, int RP, int RADIX>
hipError_t c2r_t(const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s) {
    const dim3 grid(N / (2 * RP), count);
    if (force) hipLaunchKernelGGL((k_row_c2r<N, RP, RADIX, true>), grid, dim3(fft_threads<N, RADIX>()), 0, s, d, tw);
    else hipLaunchKernelGGL((k_row_c2r<N, RP, RADIX, false>), grid, dim3(fft_threads<N, RADIX>()), 0, s, d, tw);
    return hipGetLastError();
}

// Row batches RP in {1,2,4} and column batches CB in {1,2,4,8} are compiled
// where the LDS footprint (2*RP*N*8 or 2*CB*N*8 bytes with the force pass,
// plus the twiddle tables, 8*tw_lds_entries(N, R) bytes) fits 160 KB;
// plan_valid() states the same rule for the host (with the 1/32 LDS padding).
template <int N, int RP, int R>
constexpr bool rp_ok() { return padded(2 * RP * N) * 8 + 8 * tw_lds_entries(N, R) <= 160 * 1024 && 2 * RP <= N; }
// (k_row_c2r merges its two transform batches when three rows per pair fit.)
template <int N, int CB, int R>
constexpr bool cb_ok() { return padded(2 * CB * N) * 8 + 8 * tw_lds_entries(N, R) <= 160 * 1024 && CB <= N / 2; }

template <int N>
hipError_t r2c_n(const FieldDesc* d, int count, const float2* tw, FftPlan p, hipStream_t s) {
#define PM_R2C(RPV, RV) \
    if constexpr (rp_ok<N, RPV, RV>()) if (p.rp == RPV && p.radix == RV) return r2c_t<N, RPV, RV>(d, count, tw, s);
    PM_R2C(1, 4) PM_R2C(2, 4) PM_R2C(4, 4) PM_R2C(1, 8) PM_R2C(2, 8) PM_R2C(4, 8)
    if constexpr (N >= 256) if (p.rp == 1 && p.radix == 16) return r2c_t<N, 1, 16>(d, count, tw, s);
    if constexpr (N >= 256) if (p.rp == 1 && p.radix == 17) return r2c_t<N, 1, 17>(d, count, tw, s);
#undef PM_R2C
    return hipErrorInvalidValue;
}

template <int N>
hipError_t col_n(const FieldDesc* d, int count, const float2* tw, FftPlan p, bool force, hipStream_t s) {
#define PM_COL(CBV, RV) \
    if constexpr (cb_ok<N, CBV, RV>()) if (p.cb == CBV && p.radix == RV) return col_t<N, CBV, RV>(d, count, tw, force, s);
    PM_COL(1, 4) PM_COL(2, 4) PM_COL(4, 4) PM_COL(8, 4) PM_COL(1, 8) PM_COL(2, 8) PM_COL(4, 8) PM_COL(8, 8)
    if constexpr (N >= 256) if (p.cb == 1 && p.radix == 16) return col_t<N, 1, 16>(d, count, tw, force, s);
    if constexpr (N >= 256) if (p.cb == 1 && p.radix == 17) return col_t<N, 1, 17>(d, count, tw, force, s);
#undef PM_COL
    return hipErrorInvalidValue;
}

template <int N>
hipError_t c2r_n(const FieldDesc* d, int count, const float2* tw, FftPlan p, bool force, hipStream_t s) {
#define PM_C2R(RPV, RV) \
    if constexpr (rp_ok<N, RPV, RV>()) if (p.rp == RPV && p.radix == RV) return c2r_t<N, RPV, RV>(d, count, tw, force, s);
    PM_C2R(1, 4) PM_C2R(2, 4) PM_C2R(4, 4) PM_C2R(1, 8) PM_C2R(2, 8) PM_C2R(4, 8)
    if constexpr (N >= 256) if (p.rp == 1 && p.radix == 16) return c2r_t<N, 1, 16>(d, count, tw, force, s);
    if constexpr (N >= 256) if (p.rp == 1 && p.radix == 17) return c2r_t<N, 1, 17>(d, count, tw, force, s);
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
    if (p.radix == kWaveRadix) return p.rp == 1 && p.cb == 1 && p.it == 1 && wave_plan_compiled(n);
    if (p.it != 1) return p.rp == 1 && p.cb == 1 && pipe_plan_compiled(n, p.radix, p.it);
    if (p.radix == 16 || p.radix == 17) return p.rp == 1 && p.cb == 1 && n >= 256 && n <= 8192;
    if (p.radix != 4 && p.radix != 8) return false;
    if (p.rp != 1 && p.rp != 2 && p.rp != 4) return false;
    if (p.cb != 1 && p.cb != 2 && p.cb != 4 && p.cb != 8) return false;
    const long tw = 8L * tw_lds_entries(n, p.radix);
    const long lds_r = padded(2 * p.rp * n) * 8L + tw, lds_c = padded(2 * p.cb * n) * 8L + tw;
    return lds_r <= 160 * 1024 && lds_c <= 160 * 1024 && 2 * p.rp <= n && p.cb <= n / 2;
}

// Measured on MI355X over n = 256..4096: one row pair / one column per block
// (the most blocks in flight) beats larger batches at every size; radix 8 from
// 2048 up.  The C2R pass with two row pairs per block in sequence, the next
// one prefetched and the edge bands done in the pass (the R2C and column
// passes one per block), takes the field solve from 64.7 to 58.6 us at 1024,
// 107 to 97 us at 2048 and 0.36 to 0.34 ms at 4096 (with 512-thread blocks
// there), measured with two per block in all three passes, which the R2C and
// column passes have since left; four per block are slower.  At 256 and 512
// it serves batches of simulations (one edge launch fewer per step).
FftPlan default_plan(int n) {
    FftPlan p;
    p.radix = n >= 2048 ? 8 : 4;
    p.rp = 1;
    p.cb = 1;
    p.it = pipe_plan_compiled(n, p.radix, 2) ? 2 : 1;
    return p;
}

hipError_t row_r2c(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, hipStream_t s) {
    if (p.radix == kWaveRadix) return wave_pass(0, n, d, count, tw, false, s);
    if (p.it != 1) return row_r2c_pipe(n, d, count, tw, p, s);
    PM_DISPATCH_N(n, (r2c_n<NN>(d, count, tw, p, s)));
}

hipError_t col_conv(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                    hipStream_t s) {
    if (p.radix == kWaveRadix) return wave_pass(1, n, d, count, tw, with_force, s);
    if (p.it != 1) return col_conv_pipe(n, d, count, tw, p, with_force, s);
    PM_DISPATCH_N(n, (col_n<NN>(d, count, tw, p, with_force, s)));
}

hipError_t row_c2r(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                   hipStream_t s) {
    if (p.radix == kWaveRadix) return wave_pass(2, n, d, count, tw, with_force, s);
    if (p.it != 1 && p.radix < 16) return row_c2r_pipe(n, d, count, tw, p, with_force, s);
    PM_DISPATCH_N(n, (c2r_n<NN>(d, count, tw, p, with_force, s)));
}

hipError_t edge_force(int n, const FieldDesc* d, int count, hipStream_t s) {
    PM_DISPATCH_N(n, (edge_n<NN>(d, count, s)));
}

hipError_t stencil_force(int n, const FieldDesc* d, int count, hipStream_t s) {
    PM_DISPATCH_N(n, (stencil_n<NN>(d, count, s)));
}

}  // namespace pmk
