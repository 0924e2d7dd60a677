// The three FFT passes with IT row pairs (row passes) or IT columns (column
// pass) per block, processed in sequence: the global inputs of the next row
// pair / column are loaded into registers before the current one's LDS
// transform, so its HBM reads overlap the transform instead of every block
// loading, transforming and storing in lockstep.  Same arithmetic per
// transform as the one-per-block passes of pm_fft.hip (identical results);
// the C2R pass stages the potential row pair and both force rows of a pair in
// one batch (three transforms), which frees the prefetch registers before the
// transform starts.  Only the C2R pass runs IT per block (launch_pass).
// Plan field `it` (rp = cb = 1); compiled for the default
// radix at N = 1024 (radix 4), 2048 and 4096 (radix 8), IT in {2, 4}, and for
// the split-BF16 matrix-core radix codes 16 and 17 at N = 1024..4096, IT = 2.
#include "pm_fft_pipe_core.h"

namespace pmk {
namespace {

template <int N, int RADIX, int IT>
__global__ __launch_bounds__((fft_threads<N, RADIX>())) void k_row_r2c_it(const FieldDesc* __restrict__ descs,
                                                        const float2* __restrict__ tw) {
    __shared__ float2 buf[padded(N)];
    __shared__ float2 twl[tw_lds_entries(N, RADIX)];
    r2c_pipe_body<N, RADIX, IT>(gfield(descs[blockIdx.y]), tw, buf, twl, row_group() * IT);
}

template <int N, int RADIX, bool FORCE, int IT>
__global__ __launch_bounds__((fft_threads<N, RADIX>())) void k_col_conv_it(const FieldDesc* __restrict__ descs,
                                                         const float2* __restrict__ tw) {
    __shared__ float2 buf[padded((FORCE ? 2 : 1) * N)];
    __shared__ float2 twl[tw_lds_entries(N, RADIX)];
    col_pipe_body<N, RADIX, FORCE, IT>(gfield(descs[blockIdx.y]), tw, buf, twl, blockIdx.x * IT);
}

// Threads of the C2R pass's blocks.  With the force rows a block holds three
// transforms in LDS (57 KB at 2048^2), so at most two blocks share a CU, and
// at 256 threads the prefetch registers (206 VGPRs) allow two waves per SIMD.
// 512 threads with the registers capped at 128 (no spills) keep four waves
// per SIMD: 2048^2 C2R 38.8-39.4 -> 34.4-34.7 us; at 1024^2 and 512^2 (256
// and 128 blocks, at most one per CU) twice the waves per block: 17.9-18.2
// -> 16.2-16.3 us and 12.1 -> 10.0 us (HIP events, field report).  256^2 (64
// blocks) stays at 256 threads: N must be a multiple of the block size.  At
// 4096^2 (114 KB, one block per CU) 1024 threads spill 42 registers: 140 ->
// 186 us.
template <int N, int RADIX>
constexpr int c2r_threads() {
    return RADIX < 16 && N >= 512 && N <= 2048 ? 512 : fft_threads<N, RADIX>();
}
template <int N, int RADIX, bool FORCE>
constexpr int c2r_min_waves() { return FORCE && RADIX < 16 && N == 2048 ? 4 : 1; }

template <int N, int RADIX, bool FORCE, int IT>
__global__ __launch_bounds__((c2r_threads<N, RADIX>()), (c2r_min_waves<N, RADIX, FORCE>())) void k_row_c2r_it(
    const FieldDesc* __restrict__ descs, const float2* __restrict__ tw) {
    constexpr int ROWS = FORCE ? 3 : 1;
    __shared__ float2 buf[padded(ROWS * N)];
    __shared__ float2 twl[tw_lds_entries(N, RADIX)];
    c2r_pipe_body<N, RADIX, FORCE, IT, c2r_threads<N, RADIX>()>(gfield(descs[blockIdx.y]), tw, buf, twl,
                                                                row_group() * IT);
}

template <int N, int RADIX, int IT>
hipError_t launch_r2c(const FieldDesc* d, int count, const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_row_r2c_it<N, RADIX, IT>), dim3(N / 2 / IT, count), dim3(fft_threads<N, RADIX>()), 0, s, d, tw);
    return hipGetLastError();
}
template <int N, int RADIX, int IT>
hipError_t launch_col(const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s) {
    const dim3 grid(N / 2 / IT, count);
    if (force) hipLaunchKernelGGL((k_col_conv_it<N, RADIX, true, IT>), grid, dim3(fft_threads<N, RADIX>()), 0, s, d, tw);
    else hipLaunchKernelGGL((k_col_conv_it<N, RADIX, false, IT>), grid, dim3(fft_threads<N, RADIX>()), 0, s, d, tw);
    return hipGetLastError();
}
template <int N, int RADIX, int IT>
hipError_t launch_c2r(const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s) {
    const dim3 grid(N / 2 / IT, count);
    if (force) hipLaunchKernelGGL((k_row_c2r_it<N, RADIX, true, IT>), grid, dim3(c2r_threads<N, RADIX>()), 0, s, d, tw);
    else hipLaunchKernelGGL((k_row_c2r_it<N, RADIX, false, IT>), grid, dim3(c2r_threads<N, RADIX>()), 0, s, d, tw);
    return hipGetLastError();
}

// pass 0 = R2C, 1 = column, 2 = C2R.  The R2C and column passes run one row
// pair / column per block (twice the blocks in flight; measured faster than
// two in sequence with the prefetch, alone and batched.  FP32: 2048^2 R2C
// 19.5 -> 18.6 us, column 27.0 -> 24.8 us, per instance of 16 batched 17.2
// -> 14.8 and 18.3 -> 13.8 us; 4096^2 92 -> 83 and 91 -> 74 us.  Split-BF16
// at 4096^2: three-term column 255 -> 183 us, R2C 112 -> 93 us); the C2R
// pass keeps IT row pairs per block, which its edge bands need (the last
// pair reads the potential rows of the pair before it).
template <int N, int RADIX, int IT>
hipError_t launch_pass(int pass, const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s) {
    static_assert(padded(3 * N) * 8 + 8 * tw_lds_entries(N, RADIX) <= 160 * 1024, "C2R batch must fit LDS");
    constexpr int IT_RC = 1;
    if (pass == 0) return launch_r2c<N, RADIX, IT_RC>(d, count, tw, s);
    if (pass == 1) return launch_col<N, RADIX, IT_RC>(d, count, tw, force, s);
    return launch_c2r<N, RADIX, IT>(d, count, tw, force, s);
}

hipError_t pipe_pass(int pass, int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool force,
                     hipStream_t s) {
#define PM_PIPE(NV, RV, ITV) \
    if (n == NV && p.radix == RV && p.it == ITV) return launch_pass<NV, RV, ITV>(pass, d, count, tw, force, s);
    PM_PIPE(256, 4, 2) PM_PIPE(512, 4, 2) PM_PIPE(1024, 4, 2) PM_PIPE(1024, 4, 4) PM_PIPE(2048, 8, 2)
    PM_PIPE(2048, 8, 4) PM_PIPE(4096, 8, 2) PM_PIPE(4096, 8, 4)
    PM_PIPE(1024, 16, 2) PM_PIPE(2048, 16, 2) PM_PIPE(4096, 16, 2)
    PM_PIPE(1024, 17, 2) PM_PIPE(2048, 17, 2) PM_PIPE(4096, 17, 2)
#undef PM_PIPE
    return hipErrorInvalidValue;
}

}  // namespace

bool pipe_plan_compiled(int n, int radix, int it) {
    if ((n == 256 || n == 512) && radix == 4) return it == 2;
    if (radix == 16 || radix == 17) return it == 2 && (n == 1024 || n == 2048 || n == 4096);
    if (it != 2 && it != 4) return false;
    return (n == 1024 && radix == 4) || ((n == 2048 || n == 4096) && radix == 8);
}

hipError_t row_r2c_pipe(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, hipStream_t s) {
    return pipe_pass(0, n, d, count, tw, p, false, s);
}

hipError_t col_conv_pipe(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                         hipStream_t s) {
    return pipe_pass(1, n, d, count, tw, p, with_force, s);
}

hipError_t row_c2r_pipe(int n, const FieldDesc* d, int count, const float2* tw, FftPlan p, bool with_force,
                        hipStream_t s) {
    return pipe_pass(2, n, d, count, tw, p, with_force, s);
}

}  // namespace pmk
