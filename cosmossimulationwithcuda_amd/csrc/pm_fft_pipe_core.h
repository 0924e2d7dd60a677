// The three FFT passes with IT row pairs (row passes) or IT columns (column
// pass) handled in sequence by one block, the next one's global inputs loaded
// into registers before the current one's LDS transform; as device functions
// over explicit LDS buffers, so the pass kernels (pm_fft_pipe.hip) and the
// persistent round kernel (pm_round.hip) run the same code.
#ifndef COSMOS_PM_FFT_PIPE_CORE_H
#define COSMOS_PM_FFT_PIPE_CORE_H

#include "pm_fft_core.h"

namespace pmk {
namespace {

// ------------------------------------------------------------------ row R2C --
// Row pairs g0 .. g0 + IT - 1 of one simulation; buf: padded(N) float2 and
// twl: tw_lds_entries(N, RADIX) float2 of LDS.
template <int N, int RADIX, int IT>
__device__ __forceinline__ void r2c_pipe_body(const GField& d, const float2* __restrict__ tw, float2* buf,
                                              float2* twl, int g0) {
    constexpr int BT = fft_threads<N, RADIX>();
    constexpr int PER = N / BT, NK = N / 2 + 1;
    const int t = threadIdx.x;
    int64_t ra[PER], rb[PER];
    auto fetch = [&](int g) {
        const gptr<const int64_t> r = d.rho + (size_t)(2 * g) * N + t;
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            ra[p] = r[p * BT];
            rb[p] = r[N + p * BT];
        }
    };
    load_twiddles<N, RADIX, BT>(twl, tw);
    fetch(g0);
    for (int it = 0; it < IT; ++it) {
        const int j0 = 2 * (g0 + it);
#pragma unroll
        for (int p = 0; p < PER; ++p)
            buf[pad(t + p * BT)] = make_float2((float)ra[p] * d.inv_scale, (float)rb[p] * d.inv_scale);
        if (d.rho_zero) {
            const gptr<int64_t> z = d.rho_zero + (size_t)j0 * N + t;
#pragma unroll
            for (int p = 0; p < PER; ++p) {
                z[p * BT] = 0;
                z[N + p * BT] = 0;
            }
        }
        lds_barrier();
        if (it + 1 < IT) fetch(g0 + it + 1);    // in flight during the transform
        lds_fft<N, 1, RADIX, false>(buf, twl);
        for (int k = t; k < NK; k += BT) {
            const float2 z = buf[pad(k)];
            const float2 zc = buf[pad((N - k) & (N - 1))];
            // Xa = (Z + conj(Zc))/2,  Xb = (Z - conj(Zc))/(2i)
            const float2 xa = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
            const float2 xb = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
            *(gptr<cplx2>)(d.H + (size_t)k * N + j0) = cplx2{xa.x, xa.y, xb.x, xb.y};
        }
        lds_barrier();                         // buf is rewritten by the next row pair
    }
}

// --------------------------------------------------------------- column pass --
// Columns kx = kb + it, it < IT; kx = 0 carries the packed kx = 0 / N/2
// pair exactly as the one-column pass does.
// Columns kb .. kb + IT - 1; buf: padded((FORCE ? 2 : 1) N) float2 of LDS.
template <int N, int RADIX, bool FORCE, int IT>
__device__ __forceinline__ void col_pipe_body(const GField& d, const float2* __restrict__ tw, float2* buf,
                                              float2* twl, int kb) {
    constexpr int BT = fft_threads<N, RADIX>();
    constexpr int NY = N / 2, PER = N / BT;
    const int t = threadIdx.x;
    float2 h[PER];
    float tv[PER];
    auto fetch = [&](int k) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int ky = t + p * BT;
            h[p] = k == 0 ? make_float2(d.H[ky].x, d.H[(size_t)NY * N + ky].x) : f2(d.H[(size_t)k * N + ky]);
            tv[p] = d.T[(size_t)k * N + ky];
        }
    };
    load_twiddles<N, RADIX, BT>(twl, tw);
    fetch(kb);
    for (int it = 0; it < IT; ++it) {
        const int k = kb + it;
        float tc[PER];
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            buf[pad(t + p * BT)] = h[p];
            tc[p] = tv[p];
        }
        lds_barrier();
        if (it + 1 < IT) fetch(k + 1);           // in flight during the transforms
        lds_fft<N, 1, RADIX, false>(buf, twl);
        if (k != 0) {
#pragma unroll
            for (int p = 0; p < PER; ++p) {
                const int e = t + p * BT;
                const float2 x = buf[pad(e)];
                const float2 y = make_float2(x.x * tc[p], x.y * tc[p]);
                buf[pad(e)] = y;
                if (FORCE) {
                    const float dv = d.dk[e];
                    buf[pad(N + e)] = make_float2(-y.y * dv, y.x * dv);   // i * d * y
                }
            }
        } else {                                 // packed kx = 0 / N/2 slot, by conjugate pairs
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
                    buf[pad(N + a)] = make_float2(-wa.y * da, wa.x * da);
                    buf[pad(N + b)] = make_float2(-wb.y * db, wb.x * db);
                }
            }
        }
        lds_barrier();
        lds_fft<N, FORCE ? 2 : 1, RADIX, true>(buf, twl);
        for (int j = t; j < N; j += BT) {
            const float2 w = buf[pad(j)];
            const float2 v = FORCE ? buf[pad(N + j)] : make_float2(0.f, 0.f);
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
        lds_barrier();                         // buf is rewritten by the next column
    }
}

// ------------------------------------------------------------------ row C2R --
// Per row pair: the potential rows (two per complex transform) and, with
// FORCE, the force rows (Z = Fx + i Fy per row, Fx-hat = i d(kx) Phi-hat) as
// one batch of 1 or 3 transforms.
// Row pairs g0 .. g0 + IT - 1 (the edge bands need IT >= 2 with g0 a
// multiple of IT); buf: padded((FORCE ? 3 : 1) N) float2 of LDS.
// BT: threads of the block (the pass kernel runs c2r_threads, the round
// kernel fft_threads).
template <int N, int RADIX, bool FORCE, int IT, int BT = fft_threads<N, RADIX>()>
__device__ __forceinline__ void c2r_pipe_body(const GField& d, const float2* __restrict__ tw, float2* buf,
                                              float2* twl, int g0) {
    static_assert(N % BT == 0, "the edge bands read back potential stores made by the same thread");
    constexpr int ROWS = FORCE ? 3 : 1, NK = N / 2 + 1, PK = (NK + BT - 1) / BT;
    const int t = threadIdx.x;
    cplx2 hv[PK], hw[PK];
    auto fetch = [&](int g) {
#pragma unroll
        for (int q = 0; q < PK; ++q) {
            const int k = t + q * BT;
            if (k < NK) {
                const size_t src = (size_t)k * N + 2 * g;
                hv[q] = *(gptr<const cplx2>)(d.H + src);
                if (FORCE) hw[q] = *(gptr<const cplx2>)(d.Hfy + src);
            }
        }
    };
    load_twiddles<N, RADIX, BT>(twl, tw);
    fetch(g0);
    for (int it = 0; it < IT; ++it) {
        const int j0 = 2 * (g0 + it);
#pragma unroll
        for (int q = 0; q < PK; ++q) {
            const int k = t + q * BT;
            if (k >= NK) continue;
            const cplx2 v = hv[q];
            const bool edge = k == 0 || k == N / 2;
            float2 xa = make_float2(v.x, v.y), xb = make_float2(v.z, v.w);
            if (edge) { xa.y = 0.f; xb.y = 0.f; }
            buf[pad(k)] = make_float2(xa.x - xb.y, xa.y + xb.x);
            if (!edge) buf[pad(N - k)] = make_float2(xa.x + xb.y, xb.x - xa.y);
            if (FORCE) {
                const cplx2 w = hw[q];
                const float dv = d.dk[k];
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const float2 p = r ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
                    float2 fy = r ? make_float2(w.z, w.w) : make_float2(w.x, w.y);
                    const float2 fx = make_float2(-p.y * dv, p.x * dv);     // i * d(kx) * Phi-hat
                    if (edge) fy.y = 0.f;
                    const int row = 1 + r;
                    buf[pad(row * N + k)] = make_float2(fx.x - fy.y, fx.y + fy.x);
                    if (!edge) buf[pad(row * N + N - k)] = make_float2(fx.x + fy.y, fy.x - fx.y);
                }
            }
        }
        lds_barrier();
        if (it + 1 < IT) fetch(g0 + it + 1);    // in flight during the transforms
        lds_fft<N, ROWS, RADIX, true, BT>(buf, twl);
        for (int i = t; i < N; i += BT) {
            const size_t ra = (size_t)j0 * N + i;
            const float2 z = buf[pad(i)];
            d.phi[ra] = z.x;
            d.phi[ra + N] = z.y;
        }
        if (FORCE) {
            // The reference's clamped stencil on the edge bands (what the
            // separate edge pass does after the one-pair-per-block passes):
            // Fx on columns 0, 1, N-2, N-1 from the row's potential in LDS; Fy
            // on rows N-2, N-1 with rows N-4, N-3 from the previous pair (this
            // block's own stores, read back by the thread that wrote them);
            // Fy on rows 0, 1 once rows 2, 3 are here (below).
            auto P = [&](int q, int i) { const float2 z = buf[pad(i)]; return q ? z.y : z.x; };
            for (int e = t; e < 2 * N; e += BT) {
                const int q = e / N, i = e - q * N, j = j0 + q;
                cplx v = nv(buf[pad(N + e)]);
                if (i < 2 || i >= N - 2) {
                    const float p0 = P(q, i);
                    const float l = i >= 1 ? P(q, i - 1) : p0, r = i + 1 < N ? P(q, i + 1) : p0;
                    const float ll = i >= 2 ? P(q, i - 2) : p0, rr = i + 2 < N ? P(q, i + 2) : p0;
                    v.x = (stencil5(rr, r, l, ll) / 12.0f) * d.gravity;
                }
                if (j >= N - 2) {
                    const float p0 = P(q, i);
                    const float up = d.phi[(size_t)(N - 3) * N + i], up2 = d.phi[(size_t)(N - 4) * N + i];
                    const float t1 = q ? P(0, i) : up, t2 = q ? up : up2;     // P[j-1], P[j-2]
                    const float b1 = q ? p0 : P(1, i);                         // P[j+1] (clamped)
                    v.y = (stencil5(p0, b1, t1, t2) / 12.0f) * d.gravity;
                }
                d.force[(size_t)j * N + i] = v;
            }
            if (j0 == 2)                         // rows 0 and 1: P[2], P[3] are this pair
                for (int i = t; i < N; i += BT) {
                    const float p0 = d.phi[i], p1 = d.phi[(size_t)N + i], p2 = P(0, i), p3 = P(1, i);
                    ((gptr<float>)(d.force + i))[1] = (stencil5(p2, p1, p0, p0) / 12.0f) * d.gravity;
                    ((gptr<float>)(d.force + (size_t)N + i))[1] = (stencil5(p3, p2, p0, p1) / 12.0f) * d.gravity;
                }
        }
        lds_barrier();                         // buf is rewritten by the next row pair
    }
}

}  // namespace
}  // namespace pmk

#endif  // COSMOS_PM_FFT_PIPE_CORE_H
