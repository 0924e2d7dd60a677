// Density frames: the step either side of the path that turns the mass
// lattice into the image the reference's UI consumes (render stage,
// CosmosCuda.cuh:613-828, scheduled at 1249-1256):
//   mass per cell -> BLUR_R x BLUR_R filter (weights 1/(r+1) for r < (BLUR_R-1)/2,
//   else 0, zero outside the lattice; cuh:954-966, 771-825) -> min / max of
//   the frame (cuh:631-748) -> running extremes since the first frame
//   (cuh:749-761) -> 1.8 * ((v - min) / (max - min))^0.25 (cuh:762-776).
// Two HBM passes: the filter with a fused min/max reduction, then the scale.
#include "pm_internal.h"

namespace pmk {
namespace {

constexpr int kMaxBlur = 15;

// Order-preserving map of a float to an unsigned key (for integer atomics).
__device__ __forceinline__ unsigned fkey(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// keys[0] = running min key, keys[1] = running max key of this frame;
// smooth[0], smooth[1]: extremes over frames.  The reference starts its max at
// FLT_MIN and the running pair at (1e35, -1) (cuh:627-630, 749-752).
__global__ void k_frame_init(unsigned* keys, float* smooth, int reset) {
    keys[0] = fkey(3.402823466e38f);
    keys[1] = fkey(1.175494351e-38f);
    if (reset) {
        smooth[0] = 1e35f;
        smooth[1] = -1.0f;
    }
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_frame_blur_minmax(const int64_t* __restrict__ rho, int n, float inv_scale,
                                                               float* __restrict__ out, unsigned* keys) {
    constexpr int H = (R - 1) / 2;
    __shared__ float w[R * R];
    for (int e = threadIdx.x; e < R * R; e += kBlock) {
        const int ix = e % R - H, iy = e / R - H;
        const double r = sqrt((double)(ix * ix + iy * iy));
        w[e] = r < H ? 1.0f / ((float)r + 1.0f) : 0.0f;
    }
    __syncthreads();
    float lo = 3.402823466e38f, hi = 1.175494351e-38f;
    const size_t cells = (size_t)n * n, stride = (size_t)gridDim.x * kBlock;
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < cells; c += stride) {
        const int i = (int)(c % n), j = (int)(c / n);
        float acc[2] = {0.f, 0.f};
        for (int iy = -H; iy <= H; ++iy)
            for (int ix = -H; ix <= H; ++ix) {
                const int x = i + ix, y = j + iy;
                const float v = (x >= 0 && x < n && y >= 0 && y < n) ? (float)rho[(size_t)y * n + x] * inv_scale : 0.f;
                acc[(ix + H) & 1] = fmaf(v, w[(ix + H) + (iy + H) * R], acc[(ix + H) & 1]);
            }
        const float v = acc[0] + acc[1];
        out[c] = v;
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o, 64));
        hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&keys[0], fkey(lo));
        atomicMax(&keys[1], fkey(hi));
    }
}

__global__ __launch_bounds__(kBlock) void k_frame_scale(float* __restrict__ v, size_t cells, const unsigned* keys,
                                                         float* smooth) {
    // every block derives the same running extremes; block 0 stores them
    const float s_lo = fminf(smooth[0], funkey(keys[0])), s_hi = fmaxf(smooth[1], funkey(keys[1]));
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        smooth[0] = s_lo;
        smooth[1] = s_hi;
    }
    const float diff = s_hi - s_lo;
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < cells; c += stride)
        v[c] = 1.8f * powf((v[c] - s_lo) / diff, 0.25f);
}

int frame_blocks(int n) {
    const size_t b = ((size_t)n * n + kBlock - 1) / kBlock;
    return (int)(b > 2048 ? 2048 : b);
}

}  // namespace

hipError_t frame(int n, const int64_t* rho, float inv_scale, int blur_r, bool reset, float* out, unsigned* keys,
                 float* smooth, hipStream_t s) {
    if (blur_r < 1 || blur_r > kMaxBlur || (blur_r & 1) == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_frame_init, dim3(1), dim3(1), 0, s, keys, smooth, reset ? 1 : 0);
    const dim3 grid(frame_blocks(n)), block(kBlock);
    switch (blur_r) {
#define PM_FRAME_R(R) \
        case R: hipLaunchKernelGGL((k_frame_blur_minmax<R>), grid, block, 0, s, rho, n, inv_scale, out, keys); break;
        PM_FRAME_R(1) PM_FRAME_R(3) PM_FRAME_R(5) PM_FRAME_R(7) PM_FRAME_R(9) PM_FRAME_R(11) PM_FRAME_R(13)
        PM_FRAME_R(15)
#undef PM_FRAME_R
        default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_frame_scale, grid, block, 0, s, out, (size_t)n * n, keys, smooth);
    return hipGetLastError();
}

}  // namespace pmk
