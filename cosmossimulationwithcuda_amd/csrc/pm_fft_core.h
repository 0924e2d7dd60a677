// Shared device code of the FFT passes: complex helpers, radix-2/4/8 DFTs,
// the padded LDS layout, the quarter-wave twiddle table, Stockham stages and
// the split-BF16 matrix-core radix-16 stages.  Included by pm_fft.hip (one
// transform per block) and pm_fft_pipe.hip (several per block, prefetched);
// everything is in an anonymous namespace, so each file holds its own copy.
#ifndef COSMOS_PM_FFT_CORE_H
#define COSMOS_PM_FFT_CORE_H

#include "pm_internal.h"

namespace pmk {
namespace {

// Global-memory view of a FieldDesc.  Its pointers are loaded from memory,
// so the compiler cannot infer their address space and would use flat
// instructions, which count against the LDS wait counter as well: every LDS
// wait (and so every stage barrier) would also wait for outstanding global
// loads and stores.  Casting to the global address space gives global_load /
// global_store, which only the vector-memory counter tracks.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> glob(T* p) { return (gptr<T>)p; }

// (Complex values are typed as native 2- and 4-float vectors: HIP's float2
// and float4 classes cannot be assigned through address-space-qualified
// pointers.)
typedef float cplx __attribute__((ext_vector_type(2)));
typedef float cplx2 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float2 f2(cplx v) { return make_float2(v.x, v.y); }
__device__ __forceinline__ cplx nv(float2 v) { return cplx{v.x, v.y}; }

struct GField {
    gptr<const int64_t> rho;
    gptr<int64_t> rho_zero;
    gptr<cplx> H, Hfy;
    gptr<const float> T, dk;
    gptr<float> phi;
    gptr<cplx> force;
    float inv_scale, gravity;
};
__device__ __forceinline__ GField gfield(const FieldDesc& d) {
    return GField{glob(d.rho), glob(d.rho_zero), (gptr<cplx>)d.H, (gptr<cplx>)d.Hfy, glob(d.T), glob(d.dk), glob(d.phi),
                  (gptr<cplx>)d.force, d.inv_scale, d.gravity};
}



constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

// Per-stage twiddle tables.  A radix-R Stockham stage with sub-transform
// length NS multiplies by w^(r*k*STEP), w = exp(-2 pi i/N), k = b mod NS,
// STEP = N/(R*NS).  Read from the quarter table, lanes with different k hit
// indices r*STEP apart, which for STEP >= 32 fall in one LDS bank (4- to
// 16-way conflicts at N = 2048).  Stages with STEP > 1 read their own table
// T[(r-1)*NS + k] instead (consecutive k across lanes: conflict-free), laid
// out after the quarter table and filled from it, so the values are the same
// FP32 numbers.  Tables above 1024 entries are left out (LDS budget at 8192).
// Measured on the field solve: 4096^2 326 -> 308 us, 2048^2 83 -> 82 us; at
// 256 and 512 the extra table loads cost more than the conflicts (+3-5%), so
// the tables start at N = 1024.
constexpr int first_radix_ns(int n, int r) {
    return ilog2c(n) % ilog2c(r) ? 1 << (ilog2c(n) % ilog2c(r)) : 1;
}
constexpr bool stage_table(int n, int r, int ns) {
    return n >= 1024 && r < 16 && ns > 1 && n / (r * ns) > 1 && (r - 1) * ns <= 1024;
}
constexpr int stage_table_off(int n, int r, int ns) {
    int off = n / 4;
    if (r >= 16) return off;
    for (int s = first_radix_ns(n, r); s < ns; s *= r)
        if (stage_table(n, r, s)) off += (r - 1) * s;
    return off;
}
// float2 entries of the LDS twiddle array of an n-point, radix-r transform
constexpr int tw_lds_entries(int n, int r) { return stage_table_off(n, r, n); }

// The reference's 4th-order stencil (cuh:174-229) on the FP32 potential:
// ((-P[+2] + 8 P[+1]) - 8 P[-1]) + P[-2]; neighbours outside [0, N) take the
// centre value; the force is this / 12 * G.
__device__ __forceinline__ float stencil5(float rr, float r, float l, float ll) {
    return ((-rr + 8.0f * r) - 8.0f * l) + ll;
}

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

// LDS layout: element i of a block's buffers is stored at pad(i) = i + i/32
// (one spare slot per 32 elements), which spreads the strided Stockham writes
// of the early stages over the LDS banks; strides that are multiples of 32
// elements stay compile-time offsets.
__device__ __forceinline__ int pad(int i) { return i + (i >> 5); }
constexpr int padded(int n) { return n + n / 32; }

// Twiddles exp(-2 pi i m / N) for m in [0, N) from the quarter table
// q[r] = exp(-2 pi i r / N), r < N/4, held in LDS: the quadrant factor
// (-i)^(m / (N/4)) is a swap and sign change, so the values are exactly the
// table's.
template <int N, bool INV>
__device__ __forceinline__ float2 twiddle(const float2* q, int m) {
    constexpr int Q4 = N / 4, LQ = ilog2c(Q4);
    const float2 w = q[m & (Q4 - 1)];
    const unsigned qd = (unsigned)(m >> LQ) & 3u;
    // quadrants 1 and 3 swap the components; signs: x flips for 2, 3; y for 1, 2
    const bool sw = qd & 1u;
    const unsigned sx = (qd >> 1) << 31, sy = ((qd ^ (qd >> 1)) & 1u) << 31;
    const float x = sw ? w.y : w.x, y = sw ? w.x : w.y;
    return make_float2(__uint_as_float(__float_as_uint(x) ^ sx),
                       __uint_as_float(__float_as_uint(y) ^ sy ^ (INV ? 0x80000000u : 0u)));
}

// Row-group index of a row-pass block.  Workgroups are dispatched round-robin
// over the 8 XCDs; the row passes touch the transposed half spectrum in short
// segments (2*RP rows per kx), so consecutive row groups are given to
// consecutive workgroups of the same XCD, whose L2 then serves the rest of
// each 128-byte line to the neighbouring groups.
__device__ __forceinline__ int row_group() {
    const int nb = gridDim.x, b = blockIdx.x;
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// Threads per block of the FFT passes: 512 for N >= 4096 with the FP32
// stages (twice the waves per CU at the same LDS footprint: 4096^2 field
// solve 0.36-0.41 -> 0.34 ms), else 256 (the R2C and column passes are no
// faster with 512 at 2048; the matrix-core radix-16 stages are written for
// 256).  The C2R pass picks its own (c2r_threads, pm_fft_pipe.hip).
template <int N, int RADIX>
constexpr int fft_threads() { return N >= 4096 && RADIX < 16 ? 512 : kBlock; }

template <int N, int R, int NS, int BT>
__device__ __forceinline__ void load_stage_tables(float2* q, const float2* __restrict__ tw) {
    if constexpr (NS < N) {
        if constexpr (stage_table(N, R, NS)) {
            constexpr int OFF = stage_table_off(N, R, NS), STEP = N / (R * NS);
            for (int e = threadIdx.x; e < (R - 1) * NS; e += BT) {
                const int r = e / NS + 1, k = e % NS;
                q[OFF + e] = twiddle<N, false>(tw, r * k * STEP);
            }
        }
        load_stage_tables<N, R, NS * R, BT>(q, tw);
    }
}

// Loads the quarter twiddle table and the per-stage tables of a radix-R
// transform into LDS (tw_lds_entries(N, R) entries; callers synchronise after).
template <int N, int R, int BT = kBlock>
__device__ __forceinline__ void load_twiddles(float2* q, const float2* __restrict__ tw) {
    for (int e = threadIdx.x; e < N / 4; e += BT) q[e] = tw[e];
    if constexpr (R < 16) load_stage_tables<N, R, first_radix_ns(N, R), BT>(q, tw);
}

// One Stockham stage of radix R with sub-transform length NS over NFFT
// in-place transforms of length N at buf + f*N (padded).  Reads, barrier,
// writes, barrier.
template <int N, int NFFT, int R, int NS, bool INV, int BT = kBlock>
__device__ __forceinline__ void stage(float2* buf, const float2* tw) {
    constexpr int Q = N / R, NB = NFFT * Q, PER = (NB + BT - 1) / BT;
    constexpr int STEP = N / (R * NS);
    const int t = threadIdx.x;
    float2 v[PER][R];
    int ob[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int bb = t + p * BT;
        if (NB % BT == 0 || bb < NB) {
            const int f = bb / Q, b = bb - f * Q;
            if constexpr (Q % 32 == 0) {
                const float2* src = buf + pad(f * N + b);
#pragma unroll
                for (int r = 0; r < R; ++r) v[p][r] = src[r * padded(Q)];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) v[p][r] = buf[pad(f * N + b + r * Q)];
            }
            const int k = b & (NS - 1);
            if constexpr (stage_table(N, R, NS)) {
                const float2* st = tw + stage_table_off(N, R, NS) + k;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 w = st[(r - 1) * NS];
                    v[p][r] = cmul(v[p][r], make_float2(w.x, INV ? -w.y : w.y));
                }
            } else if (NS > 1) {
                const int e = k * STEP;
#pragma unroll
                for (int r = 1; r < R; ++r) v[p][r] = cmul(v[p][r], twiddle<N, INV>(tw, r * e));
            }
            dft<R, INV>(v[p]);
            ob[p] = f * N + (b - k) * R + k;
        }
    }
    lds_barrier();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int bb = t + p * BT;
        if (NB % BT == 0 || bb < NB) {
            if constexpr (NS % 32 == 0) {
                float2* dst = buf + pad(ob[p]);
#pragma unroll
                for (int r = 0; r < R; ++r) dst[r * padded(NS)] = v[p][r];
            } else if constexpr (32 % (NS * R) == 0) {   // the R outputs share one 32-element row
                float2* dst = buf + pad(ob[p]);
#pragma unroll
                for (int r = 0; r < R; ++r) dst[r * NS] = v[p][r];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) buf[pad(ob[p] + r * NS)] = v[p][r];
            }
        }
    }
    lds_barrier();
}

template <int N, int NFFT, int R, int NS, bool INV, int BT = kBlock>
__device__ __forceinline__ void stages_from(float2* buf, const float2* tw) {
    if constexpr (NS < N) {
        stage<N, NFFT, R, NS, INV, BT>(buf, tw);
        stages_from<N, NFFT, R, NS * R, INV, BT>(buf, tw);
    }
}

// NFFT in-place transforms of length N at buf + f*N (padded), radix RADIX
// (4 or 8) after a leading stage for the leftover factor; tw is the LDS quarter
// table.  Caller synchronises before; ends with a barrier.
// ---- split-BF16 matrix-core radix-16 stages (precision variants) -------------
// A radix-16 stage over a group of 16 butterflies is one complex 16x16 matrix
// product X = F * x (F[m][r] = exp(-+2 pi i m r / 16)), done as the real
// product [Xr; Xi] = [[Fr, -Fi], [Fi, Fr]] [xr; xi] on the matrix cores with
// v_mfma_f32_16x16x16_bf16, FP32 accumulation.  Every FP32 operand is split
// into bf16 parts:
//   S3 = false (bf16x3): hi + lo, products hi*hi + hi*lo + lo*hi; the lo*lo
//     term (~2^-16 relative) is dropped and each operand keeps ~17 bits;
//   S3 = true (bf16x6): hi + mid + lo (~25 bits), products hi*hi, hi*mid,
//     mid*hi, hi*lo, lo*hi, mid*mid; the dropped terms are ~2^-24 relative.
// Small products are accumulated first.  Operand layout (16x16x16): lane l
// holds A[l%16][4(l/16)+j] and B[4(l/16)+j][l%16] for j < 4, and receives
// D[4(l/16)+i][l%16] for i < 4.
typedef short v4bf __attribute__((ext_vector_type(4)));
typedef float v4acc __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rne(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(short b) { return __uint_as_float((unsigned)(unsigned short)b << 16); }
__device__ __forceinline__ void split_bf16(float x, short& hi, short& lo) {
    hi = bf16_rne(x);
    lo = bf16_rne(x - bf16_val(hi));
}
__device__ __forceinline__ void split3_bf16(float x, short& hi, short& mid, short& lo) {
    hi = bf16_rne(x);
    const float r1 = x - bf16_val(hi);
    mid = bf16_rne(r1);
    lo = bf16_rne(r1 - bf16_val(mid));
}

// cos(2 pi j / 16), j < 16
__constant__ float kCos16[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                                 0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                                 -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
                                 0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};

// DFT-16 operands: [h][kh] = output half (re, im) x input half (re, im).
struct Dft16Ops {
    v4bf hi[2][2], mid[2][2], lo[2][2];
};

template <bool INV, bool S3>
__device__ __forceinline__ void dft16_ops(Dft16Ops& A) {
    const int l = threadIdx.x & 63, m = l & 15, g = l >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int idx = (m * (4 * g + j)) & 15;
        const float c = kCos16[idx], sn = kCos16[(idx + 12) & 15];   // sin(2 pi idx/16)
        const float fr = c, fi = INV ? sn : -sn;
        const float v[2][2] = {{fr, -fi}, {fi, fr}};
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int kh = 0; kh < 2; ++kh) {
                short a, b, c3 = 0;
                if (S3) split3_bf16(v[h][kh], a, b, c3);
                else split_bf16(v[h][kh], a, b);
                A.hi[h][kh][j] = a;
                if (S3) {
                    A.mid[h][kh][j] = b;
                    A.lo[h][kh][j] = c3;
                } else {
                    A.lo[h][kh][j] = b;
                }
            }
    }
}

// One radix-16 Stockham stage (same indexing as stage<>), butterflies taken
// 16 at a time per wavefront.  Results are held in registers until every
// wavefront has read its inputs (in place, like stage<>).
template <int N, int NFFT, int NS, bool INV, bool S3>
__device__ __forceinline__ void stage16(float2* buf, const float2* tw, const Dft16Ops& A) {
    constexpr int R = 16, Q = N / R, NB = NFFT * Q, G = NB / 16, NW = kBlock / 64, GPW = (G + NW - 1) / NW;
    constexpr int STEP = N / (R * NS);
    static_assert(NB % 16 == 0, "radix-16 matrix stage needs a multiple of 16 butterflies");
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, c = l & 15, g = l >> 4;
    v4acc dre[GPW], dim[GPW];
    int ob[GPW];
#pragma unroll
    for (int q = 0; q < GPW; ++q) {
        const int grp = w + q * NW;
        if (G % NW == 0 || grp < G) {              // uniform per wavefront
            const int bb = grp * 16 + c;
            const int f = bb / Q, b = bb - f * Q, k = b & (NS - 1);
            v4bf bhi[2], bmid[2], blo[2];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = 4 * g + j;
                float2 x = buf[pad(f * N + b + r * Q)];
                if (NS > 1) x = cmul(x, twiddle<N, INV>(tw, r * k * STEP));
                short h0, m0 = 0, l0, h1, m1 = 0, l1;
                if (S3) {
                    split3_bf16(x.x, h0, m0, l0);
                    split3_bf16(x.y, h1, m1, l1);
                    bmid[0][j] = m0;
                    bmid[1][j] = m1;
                } else {
                    split_bf16(x.x, h0, l0);
                    split_bf16(x.y, h1, l1);
                }
                bhi[0][j] = h0; blo[0][j] = l0;
                bhi[1][j] = h1; blo[1][j] = l1;
            }
            v4acc acc[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                acc[h] = v4acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kh = 0; kh < 2; ++kh) {
                    acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A.lo[h][kh], bhi[kh], acc[h], 0, 0, 0);
                    acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A.hi[h][kh], blo[kh], acc[h], 0, 0, 0);
                    if (S3) acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A.mid[h][kh], bmid[kh], acc[h], 0, 0, 0);
                }
                if (S3) {
#pragma unroll
                    for (int kh = 0; kh < 2; ++kh) {
                        acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A.mid[h][kh], bhi[kh], acc[h], 0, 0, 0);
                        acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A.hi[h][kh], bmid[kh], acc[h], 0, 0, 0);
                    }
                }
#pragma unroll
                for (int kh = 0; kh < 2; ++kh)
                    acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A.hi[h][kh], bhi[kh], acc[h], 0, 0, 0);
            }
            dre[q] = acc[0];
            dim[q] = acc[1];
            ob[q] = f * N + (b - k) * R + k;
        }
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < GPW; ++q) {
        const int grp = w + q * NW;
        if (G % NW == 0 || grp < G) {
#pragma unroll
            for (int i = 0; i < 4; ++i) buf[pad(ob[q] + (4 * g + i) * NS)] = make_float2(dre[q][i], dim[q][i]);
        }
    }
    lds_barrier();
}

template <int N, int NFFT, int NS, bool INV, bool S3>
__device__ __forceinline__ void stages16_from(float2* buf, const float2* tw, const Dft16Ops& A) {
    if constexpr (NS < N) {
        stage16<N, NFFT, NS, INV, S3>(buf, tw, A);
        stages16_from<N, NFFT, NS * 16, INV, S3>(buf, tw, A);
    }
}

// NFFT transforms (as lds_fft) with radix-16 matrix-core stages after one
// leading FP32 stage of radix 2, 4 or 8 for the leftover factor.
template <int N, int NFFT, bool INV, bool S3>
__device__ __forceinline__ void lds_fft_mfma(float2* buf, const float2* tw) {
    constexpr int L = ilog2c(N), REM = L % 4;
    Dft16Ops A;
    dft16_ops<INV, S3>(A);
    if constexpr (REM == 0) {
        stages16_from<N, NFFT, 1, INV, S3>(buf, tw, A);
    } else {
        constexpr int R0 = 1 << REM;
        stage<N, NFFT, R0, 1, INV>(buf, tw);
        stages16_from<N, NFFT, R0, INV, S3>(buf, tw, A);
    }
}

// Plan radix codes: 4 or 8 = FP32 Stockham stages; kRadixBf16x3 (16) and
// kRadixBf16x6 (17) = the split-BF16 matrix-core radix-16 variants above.
constexpr int kRadixBf16x3 = 16, kRadixBf16x6 = 17;
template <int N, int NFFT, int RADIX, bool INV, int BT = fft_threads<N, RADIX>()>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* tw) {
    constexpr int LR = ilog2c(RADIX), L = ilog2c(N), REM = L % LR;
    if constexpr (RADIX == kRadixBf16x3 || RADIX == kRadixBf16x6) {
        static_assert(BT == kBlock, "the matrix-core stages run 256-thread blocks");
        lds_fft_mfma<N, NFFT, INV, RADIX == kRadixBf16x6>(buf, tw);
    } else if constexpr (REM == 0) {
        stages_from<N, NFFT, RADIX, 1, INV, BT>(buf, tw);
    } else {
        constexpr int R0 = 1 << REM;
        stage<N, NFFT, R0, 1, INV, BT>(buf, tw);
        stages_from<N, NFFT, RADIX, R0, INV, BT>(buf, tw);
    }
}

}  // namespace
}  // namespace pmk

#endif  // COSMOS_PM_FFT_CORE_H
