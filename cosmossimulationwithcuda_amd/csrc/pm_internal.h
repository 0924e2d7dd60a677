// Internal launcher interface between the C-ABI layer (pm_api.cpp) and the
// HIP kernels (pm_*.hip).  Not installed; the public interface is include/pm.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmk {

constexpr int kBlock = 256;   // threads per block for every streaming kernel (4 wave64s)

// Lattice sizes with compiled kernels (powers of two).
inline bool lattice_supported(int n) {
    return n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096 ||
           n == 8192;
}

// ----- deposit (pm_deposit.hip) ------------------------------------------------
// rho must be zero on entry; adds int64 fixed-point contributions round(w*m*scale).
hipError_t deposit(int n, int mode, const float* x, const float* y, const float* m, int64_t np,
                   float scale, int64_t* rho, hipStream_t s);

// ----- FFT field solve (pm_fft.hip) -------------------------------------------
// Half spectrum layout: H[k * N + j], k in [0, N/2] (x frequency), j in [0, N)
// (y row before the column pass, y frequency after it).
hipError_t row_r2c(int n, const int64_t* rho, float inv_scale, const float2* tw, float2* H,
                   hipStream_t s);
hipError_t col_conv(int n, float2* H, float2* Hfy, const float* T, const float* dk,
                    const float2* tw, bool with_force, hipStream_t s);
hipError_t row_c2r(int n, const float2* H, const float2* Hfy, const float* dk, const float2* tw,
                   float* phi, float2* force, bool with_force, hipStream_t s);
hipError_t edge_force(int n, const float* phi, float g, float2* force, hipStream_t s);
hipError_t stencil_force(int n, const float* phi, float g, float2* force, hipStream_t s);

// ----- FP64 setup spectra (pm_spectra.hip) --------------------------------------
// T[k*N + ky] = Ghat(k,ky) * c/(c^2+eps) / N^2 as float, k in [0,N/2].
// work64 must hold 2 * N * N doubles (complex lattice).
hipError_t build_transfer(int n, int mode, double green_origin, double eps, double* work64,
                          float* T, hipStream_t s);

// ----- push (pm_push.hip) -----------------------------------------------------------
hipError_t push(int n, int mode, const float2* force, float* x, float* y, float* vx, float* vy,
                int64_t np, float dt, int wrap_mode, hipStream_t s);

// ----- utilities (pm_util.hip) ------------------------------------------------------
hipError_t fill_zero(void* p, size_t bytes, hipStream_t s);
// Per-block FP64 partial sums: out[b*8 + {0..7}] = {mass, px, py, ke, max|v|, nonfinite, n_inb, 0}
int diag_blocks(int64_t np);
hipError_t diag_particles(int n, const float* x, const float* y, const float* vx, const float* vy,
                          const float* m, int64_t np, double* partial, hipStream_t s);
// out[b] = partial sums of rho*phi (mass units)
int diag_lattice_blocks(int n);
hipError_t diag_lattice(int n, const int64_t* rho, double inv_scale, const float* phi,
                        double* partial, hipStream_t s);

}  // namespace pmk
