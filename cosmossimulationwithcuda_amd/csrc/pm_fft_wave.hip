// Wave-local FFT passes (plan radix kWaveRadix): every 64-lane wavefront
// transforms one whole row pair or column of N = 64 V points on its own,
// with no workgroup barrier after the twiddle table is staged.
//
// A transform of n = l + 64 j (lane l holds points j = 0..V-1, so every
// global access is one contiguous 64-lane stretch) runs as
//   1. an in-lane DFT of length V over j          -> Y[l][k1]
//   2. the twiddles W_N^(l k1)
//   3. one transpose through the wave's LDS buffer: with B = 64 / V lanes per
//      k1, lane (k1, q) takes l = B l' + q for l' < V
//   4. an in-lane DFT of length V over l'          -> E_q[k2']
//   5. the twiddles W_64^(q k2')
//   6. a radix-2 DIF over the B lanes of each k1 (lane exchanges, no LDS),
//      which leaves X[k1 + V k2' + V^2 m] in lane (k1, q), register k2',
//      with m the bit reversal of q.
// The inverse runs the conjugate transpose of these steps in reverse order,
// so it starts from that distribution and ends with n = l + 64 j again: the
// column pass goes forward, multiplies, and back without reordering, and the
// C2R pass gathers its spectrum straight into the forward distribution.
// Compared with the block-wide Stockham passes (pm_fft.hip, pm_fft_pipe.hip),
// which take two workgroup barriers and an LDS round trip per radix stage,
// a transform here makes one LDS round trip, and each wavefront loads,
// transforms and stores at its own pace, so the waves of a pass overlap
// their memory and compute phases instead of moving in lockstep.
//
// Same data layout and arithmetic per element as the block passes (half
// spectrum H[kx * N + y], packed kx = 0 / N/2 column, transfer function T,
// stencil symbol dk); the FP32 rounding of the transforms differs.  The edge
// bands of the force lattice are left to edge_force, as with the
// one-pair-per-block passes.  N = 256 .. 2048.
#include <type_traits>

#include "pm_fft_core.h"

namespace pmk {
namespace {

constexpr int kWaves = 4;                 // wavefronts (independent jobs) per block

// ---- compile-time helpers -------------------------------------------------------
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

constexpr int bitrev_c(int x, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((x >> b) & 1) << (bits - 1 - b);
    return r;
}

// sin and cos by their Taylor series on |x| <= pi (double, 20 terms: the
// truncation error is below 1e-17), for the constant twiddles of the in-lane
// DFTs; rounded once to float.
constexpr double taylor_sin(double x) {
    double t = x, s = x;
    for (int k = 1; k < 20; ++k) {
        t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        s += t;
    }
    return s;
}
constexpr double taylor_cos(double x) {
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 20; ++k) {
        t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        s += t;
    }
    return s;
}

// exp(-2 pi i M / NN) with M reduced to (-NN/2, NN/2]
template <int NN, int M>
struct CTw {
    static constexpr int m0 = ((M % NN) + NN) % NN;
    static constexpr int m = 2 * m0 > NN ? m0 - NN : m0;
    static constexpr double a = -6.283185307179586476925286766559 * (double)m / (double)NN;
    static constexpr float c = (float)taylor_cos(a), s = (float)taylor_sin(a);
};

__device__ __forceinline__ cplx cmulc(cplx a, float2 w) {
    return cplx{fmaf(a.x, w.x, -a.y * w.y), fmaf(a.x, w.y, a.y * w.x)};
}

// a * W_NN^M (forward) or a * conj(W_NN^M) (INV); exact for the quarter turns
template <int NN, int M, bool INV>
__device__ __forceinline__ cplx twc(cplx a) {
    constexpr int m0 = CTw<NN, M>::m0;
    if constexpr (m0 == 0) {
        return a;
    } else if constexpr (2 * m0 == NN) {
        return -a;
    } else if constexpr (4 * m0 == NN) {            // forward -i
        return INV ? cplx{-a.y, a.x} : cplx{a.y, -a.x};
    } else if constexpr (4 * m0 == 3 * NN) {        // forward +i
        return INV ? cplx{a.y, -a.x} : cplx{-a.y, a.x};
    } else {
        constexpr float c = CTw<NN, M>::c, s = INV ? -CTw<NN, M>::s : CTw<NN, M>::s;
        return cplx{fmaf(a.x, c, -a.y * s), fmaf(a.x, s, a.y * c)};
    }
}

// In-lane DFT of length V, natural order in and out: radix-2 decimation in
// frequency with constant twiddles, then the bit reversal as a renaming.
template <int V, int H, bool INV>
__device__ __forceinline__ void dif_stages(cplx (&x)[V]) {
    if constexpr (H >= 1) {
        sfor<0, V / (2 * H)>([&](auto g) {
            sfor<0, H>([&](auto u) {
                constexpr int i0 = decltype(g)::value * 2 * H + decltype(u)::value, i1 = i0 + H;
                const cplx a = x[i0], b = x[i1];
                x[i0] = a + b;
                x[i1] = twc<2 * H, decltype(u)::value, INV>(a - b);
            });
        });
        dif_stages<V, H / 2, INV>(x);
    }
}

template <int V, bool INV>
__device__ __forceinline__ void reg_dft(cplx (&x)[V]) {
    if constexpr (V > 1) {
        dif_stages<V, V / 2, INV>(x);
        constexpr int LV = ilog2c(V);
        cplx y[V];
        sfor<0, V>([&](auto r) { y[bitrev_c(decltype(r)::value, LV)] = x[decltype(r)::value]; });
        sfor<0, V>([&](auto r) { x[decltype(r)::value] = y[decltype(r)::value]; });
    }
}

// Orders one wavefront's LDS accesses (LDS executes a wave's instructions in
// order; the fences keep the compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Keeps the scheduler from hoisting one phase's loads into the previous
// phase (which would hold them in registers across a whole DFT).
__device__ __forceinline__ void phase_fence() { __builtin_amdgcn_sched_barrier(0); }

__device__ __forceinline__ cplx shfl_xor_c(cplx v, int mask) {
    return cplx{__shfl_xor(v.x, mask, 64), __shfl_xor(v.y, mask, 64)};
}

// ---- one wavefront's transform of N = 64 V points ------------------------------------
template <int N>
struct WaveFft {
    static constexpr int V = N / 64, B = 64 / V, LB = ilog2c(B);
    static constexpr int S = 64 + B;                       // transpose row stride (conflict-free both ways)
    static constexpr int BUF = V * S > padded(N) ? V * S : padded(N);   // float2 per wave
    static_assert(V >= 4 && V <= 64 && 64 % V == 0, "N = 256 .. 4096");

    // natural index of register k2 of lane `lane` after fwd()
    __device__ static __forceinline__ int out_index(int lane, int k2) {
        const int k1 = lane >> LB, q = lane & (B - 1);
        const int m = LB ? (int)(__brev((unsigned)q) >> (32 - LB)) : 0;
        return k1 + V * k2 + V * V * m;
    }

    // forward DFT: x[j] = point lane + 64 j  ->  x[k2] = X[out_index(lane, k2)]
    __device__ static __forceinline__ void fwd(cplx (&x)[V], float2* wb, const float2* qtw, int lane) {
        reg_dft<V, false>(x);
        phase_fence();
        sfor<1, V>([&](auto k1) {
            constexpr int K1 = decltype(k1)::value;
            x[K1] = cmulc(x[K1], twiddle<N, false>(qtw, (lane * K1) & (N - 1)));
        });
        phase_fence();
        const int k1r = lane >> LB, q = lane & (B - 1);
        sfor<0, V>([&](auto k1) { wb[decltype(k1)::value * S + lane] = f2(x[decltype(k1)::value]); });
        wave_lds_sync();
        sfor<0, V>([&](auto lp) { x[decltype(lp)::value] = nv(wb[k1r * S + B * decltype(lp)::value + q]); });
        wave_lds_sync();
        reg_dft<V, false>(x);
        phase_fence();
        if constexpr (B > 1) {
            sfor<1, V>([&](auto k2) {
                constexpr int K2 = decltype(k2)::value;
                x[K2] = cmulc(x[K2], twiddle<N, false>(qtw, (q * K2 * (N / 64)) & (N - 1)));
            });
            phase_fence();
            sfor<0, LB>([&](auto lv) {                     // h = B/2, B/4, ..., 1
                constexpr int h = B >> (1 + decltype(lv)::value);
                const int t = q & (h - 1);
                const bool up = (q & h) != 0;
                const float2 w = twiddle<N, false>(qtw, t * (N / (2 * h)));
                sfor<0, V>([&](auto j) {
                    constexpr int J = decltype(j)::value;
                    const cplx a = x[J], p = shfl_xor_c(a, h);
                    x[J] = up ? cmulc(p - a, w) : a + p;
                });
            });
        }
    }

    // inverse DFT (unnormalised): the conjugate transpose of fwd()
    __device__ static __forceinline__ void inv(cplx (&x)[V], float2* wb, const float2* qtw, int lane) {
        const int k1r = lane >> LB, q = lane & (B - 1);
        if constexpr (B > 1) {
            sfor<0, LB>([&](auto lv) {                     // h = 1, 2, ..., B/2
                constexpr int h = 1 << decltype(lv)::value;
                const int t = q & (h - 1);
                const bool up = (q & h) != 0;
                const float2 w = twiddle<N, true>(qtw, t * (N / (2 * h)));
                sfor<0, V>([&](auto j) {
                    constexpr int J = decltype(j)::value;
                    const cplx a = x[J], p = shfl_xor_c(a, h);
                    x[J] = up ? p - cmulc(a, w) : a + cmulc(p, w);
                });
            });
            phase_fence();
            sfor<1, V>([&](auto k2) {
                constexpr int K2 = decltype(k2)::value;
                x[K2] = cmulc(x[K2], twiddle<N, true>(qtw, (q * K2 * (N / 64)) & (N - 1)));
            });
            phase_fence();
        }
        reg_dft<V, true>(x);
        sfor<0, V>([&](auto lp) { wb[k1r * S + B * decltype(lp)::value + q] = f2(x[decltype(lp)::value]); });
        wave_lds_sync();
        sfor<0, V>([&](auto k1) { x[decltype(k1)::value] = nv(wb[decltype(k1)::value * S + lane]); });
        wave_lds_sync();
        phase_fence();
        sfor<1, V>([&](auto k1) {
            constexpr int K1 = decltype(k1)::value;
            x[K1] = cmulc(x[K1], twiddle<N, true>(qtw, (lane * K1) & (N - 1)));
        });
        phase_fence();
        reg_dft<V, true>(x);
    }
};

template <int N>
__device__ __forceinline__ void stage_quarter_table(float2* qtw, const float2* __restrict__ tw) {
    for (int e = threadIdx.x; e < N / 4; e += kWaves * 64) qtw[e] = tw[e];
    lds_barrier();
}

// ------------------------------------------------------------------ row R2C --
template <int N>
__global__ __launch_bounds__(kWaves * 64) void k_row_r2c_w(const FieldDesc* __restrict__ descs,
                                                        const float2* __restrict__ tw) {
    using W = WaveFft<N>;
    constexpr int V = W::V;
    __shared__ float2 qtw[N / 4];
    __shared__ float2 wbs[kWaves][W::BUF];
    const GField d = gfield(descs[blockIdx.y]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    stage_quarter_table<N>(qtw, tw);
    float2* wb = wbs[wv];
    const int y0 = 2 * (row_group() * kWaves + wv);
    cplx x[V];
    {
        const gptr<const int64_t> r = d.rho + (size_t)y0 * N + lane;
        sfor<0, V>([&](auto j) {
            constexpr int J = decltype(j)::value;
            x[J] = cplx{(float)r[64 * J] * d.inv_scale, (float)r[N + 64 * J] * d.inv_scale};
        });
    }
    if (d.rho_zero) {                          // the next deposit target: 2 rows, 16 bytes per lane and store
        typedef unsigned int u4v __attribute__((ext_vector_type(4)));
        const gptr<u4v> z = (gptr<u4v>)(d.rho_zero + (size_t)y0 * N) + lane;
        sfor<0, V>([&](auto j) { z[64 * decltype(j)::value] = u4v{0u, 0u, 0u, 0u}; });
    }
    W::fwd(x, wb, qtw, lane);
    sfor<0, V>([&](auto k2) { wb[pad(W::out_index(lane, decltype(k2)::value))] = f2(x[decltype(k2)::value]); });
    wave_lds_sync();
    constexpr int NK = N / 2 + 1;
    for (int k = lane; k < NK; k += 64) {
        const float2 z = wb[pad(k)];
        const float2 zc = wb[pad((N - k) & (N - 1))];
        // Xa = (Z + conj(Zc))/2,  Xb = (Z - conj(Zc))/(2i)
        const float2 xa = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
        const float2 xb = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
        *(gptr<cplx2>)(d.H + (size_t)k * N + y0) = cplx2{xa.x, xa.y, xb.x, xb.y};
    }
}

// --------------------------------------------------------------- column pass --
// Wave w of block b takes column slot kx = 4 b + w; slot 0 carries the real
// columns kx = 0 and N/2 as one complex column (split by conjugate pairs
// after the forward transform, as in the block passes).
template <int N, bool FORCE>
__global__ __launch_bounds__(kWaves * 64) void k_col_conv_w(const FieldDesc* __restrict__ descs,
                                                         const float2* __restrict__ tw) {
    using W = WaveFft<N>;
    constexpr int V = W::V, NY = N / 2;
    __shared__ float2 qtw[N / 4];
    __shared__ float2 wbs[kWaves][W::BUF];
    const GField d = gfield(descs[blockIdx.y]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    stage_quarter_table<N>(qtw, tw);
    float2* wb = wbs[wv];
    const int kx = blockIdx.x * kWaves + wv;
    cplx x[V], f[V];
    if (kx != 0) {
        const gptr<const cplx> src = d.H + (size_t)kx * N + lane;
        sfor<0, V>([&](auto j) { x[decltype(j)::value] = src[64 * decltype(j)::value]; });
    } else {
        sfor<0, V>([&](auto j) {
            const int ky = lane + 64 * decltype(j)::value;
            x[decltype(j)::value] = cplx{d.H[ky].x, d.H[(size_t)NY * N + ky].x};
        });
    }
    W::fwd(x, wb, qtw, lane);
    if (kx != 0) {
        sfor<0, V>([&](auto k2) {
            constexpr int K2 = decltype(k2)::value;
            const int ky = W::out_index(lane, K2);
            const float tv = d.T[(size_t)kx * N + ky];
            const cplx y = cplx{x[K2].x * tv, x[K2].y * tv};
            x[K2] = y;
            if (FORCE) {
                const float dv = d.dk[ky];
                f[K2] = cplx{-y.y * dv, y.x * dv};          // i * d * y
            }
        });
    } else {
        // packed kx = 0 / N/2 slot: conjugate pairs (a, N - a) in natural order
        sfor<0, V>([&](auto k2) { wb[pad(W::out_index(lane, decltype(k2)::value))] = f2(x[decltype(k2)::value]); });
        wave_lds_sync();
        constexpr int NP = NY / 64 + 1;                     // a = lane + 64 t <= N/2
        float2 wa[NP], wbv[NP], fa[NP], fb[NP];
        sfor<0, NP>([&](auto tt) {
            constexpr int T = decltype(tt)::value;
            const int a = lane + 64 * T;
            if (a <= NY) {
                const int b = (N - a) & (N - 1);
                const float2 za = wb[pad(a)], zb = wb[pad(b)];
                const float2 A = make_float2(0.5f * (za.x + zb.x), 0.5f * (za.y - zb.y));
                const float2 Bv = make_float2(0.5f * (za.y + zb.y), -0.5f * (za.x - zb.x));
                const float t0a = d.T[a], t0b = d.T[b];
                const float tna = d.T[(size_t)NY * N + a], tnb = d.T[(size_t)NY * N + b];
                // W(a) = A T0 + i B TN;  W(b) = conj(A) T0 + i conj(B) TN
                wa[T] = make_float2(A.x * t0a - Bv.y * tna, A.y * t0a + Bv.x * tna);
                wbv[T] = make_float2(A.x * t0b + Bv.y * tnb, -A.y * t0b + Bv.x * tnb);
                if (FORCE) {
                    const float da = d.dk[a], db = d.dk[b];
                    fa[T] = make_float2(-wa[T].y * da, wa[T].x * da);
                    fb[T] = make_float2(-wbv[T].y * db, wbv[T].x * db);
                }
            }
        });
        wave_lds_sync();
        sfor<0, NP>([&](auto tt) {
            constexpr int T = decltype(tt)::value;
            const int a = lane + 64 * T;
            if (a <= NY) {
                wb[pad(a)] = wa[T];
                wb[pad((N - a) & (N - 1))] = wbv[T];
            }
        });
        wave_lds_sync();
        sfor<0, V>([&](auto k2) { x[decltype(k2)::value] = nv(wb[pad(W::out_index(lane, decltype(k2)::value))]); });
        if (FORCE) {
            wave_lds_sync();
            sfor<0, NP>([&](auto tt) {
                constexpr int T = decltype(tt)::value;
                const int a = lane + 64 * T;
                if (a <= NY) {
                    wb[pad(a)] = fa[T];
                    wb[pad((N - a) & (N - 1))] = fb[T];
                }
            });
            wave_lds_sync();
            sfor<0, V>([&](auto k2) { f[decltype(k2)::value] = nv(wb[pad(W::out_index(lane, decltype(k2)::value))]); });
        }
        wave_lds_sync();
    }
    W::inv(x, wb, qtw, lane);
    if (kx != 0) {
        const gptr<cplx> dst = d.H + (size_t)kx * N + lane;
        sfor<0, V>([&](auto j) { dst[64 * decltype(j)::value] = x[decltype(j)::value]; });
    } else {
        sfor<0, V>([&](auto j) {
            const int ky = lane + 64 * decltype(j)::value;
            d.H[ky] = cplx{x[decltype(j)::value].x, 0.f};
            d.H[(size_t)NY * N + ky] = cplx{x[decltype(j)::value].y, 0.f};
        });
    }
    if (FORCE) {
        W::inv(f, wb, qtw, lane);
        if (kx != 0) {
            const gptr<cplx> dst = d.Hfy + (size_t)kx * N + lane;
            sfor<0, V>([&](auto j) { dst[64 * decltype(j)::value] = f[decltype(j)::value]; });
        } else {
            sfor<0, V>([&](auto j) {
                const int ky = lane + 64 * decltype(j)::value;
                d.Hfy[ky] = cplx{f[decltype(j)::value].x, 0.f};
                d.Hfy[(size_t)NY * N + ky] = cplx{f[decltype(j)::value].y, 0.f};
            });
        }
    }
}

// ------------------------------------------------------------------ row C2R --
// The spectrum of a row pair is gathered straight into the forward
// distribution: register k2 of a lane needs Z[k], k = out_index(lane, k2),
// built from H[kk] with kk = min(k, N - k) (Hermitian extension).  With
// FORCE the two force rows (Z = Fx + i Fy per row, Fx-hat = i d(kx) Phi-hat)
// follow, both at once up to V = 32 registers, one after the other above.
template <int N, bool FORCE>
__global__ __launch_bounds__(kWaves * 64) void k_row_c2r_w(const FieldDesc* __restrict__ descs,
                                                        const float2* __restrict__ tw) {
    using W = WaveFft<N>;
    constexpr int V = W::V, NY = N / 2;
    __shared__ float2 qtw[N / 4];
    __shared__ float2 wbs[kWaves][W::BUF];
    const GField d = gfield(descs[blockIdx.y]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    stage_quarter_table<N>(qtw, tw);
    float2* wb = wbs[wv];
    const int y0 = 2 * (row_group() * kWaves + wv);
    cplx x[V];
    sfor<0, V>([&](auto k2) {
        const int k = W::out_index(lane, decltype(k2)::value);
        const int kk = k <= NY ? k : N - k;
        const cplx2 v = *(gptr<const cplx2>)(d.H + (size_t)kk * N + y0);
        const bool edge = kk == 0 || kk == NY;
        const float2 xa = make_float2(v.x, edge ? 0.f : v.y), xb = make_float2(v.z, edge ? 0.f : v.w);
        x[decltype(k2)::value] = k <= NY ? cplx{xa.x - xb.y, xa.y + xb.x} : cplx{xa.x + xb.y, xb.x - xa.y};
    });
    W::inv(x, wb, qtw, lane);
    {
        const gptr<float> p0 = d.phi + (size_t)y0 * N + lane;
        sfor<0, V>([&](auto j) {
            constexpr int J = decltype(j)::value;
            p0[64 * J] = x[J].x;
            p0[N + 64 * J] = x[J].y;
        });
    }
    if constexpr (FORCE) {
        // force row h of the pair from register values of H and Hfy
        auto force_in = [&](int k2, const cplx2& v, const cplx2& wv2, int h) {
            const int k = W::out_index(lane, k2);
            const int kk = k <= NY ? k : N - k;
            const bool edge = kk == 0 || kk == NY;
            const float dv = d.dk[kk];
            const float2 p = h ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
            float2 fy = h ? make_float2(wv2.z, wv2.w) : make_float2(wv2.x, wv2.y);
            const float2 fx = make_float2(-p.y * dv, p.x * dv);       // i * d(kx) * Phi-hat
            if (edge) fy.y = 0.f;
            return k <= NY ? cplx{fx.x - fy.y, fx.y + fy.x} : cplx{fx.x + fy.y, fy.x - fx.y};
        };
        auto spectra = [&](int k2, cplx2& v, cplx2& w2) {
            const int k = W::out_index(lane, k2);
            const int kk = k <= NY ? k : N - k;
            v = *(gptr<const cplx2>)(d.H + (size_t)kk * N + y0);
            w2 = *(gptr<const cplx2>)(d.Hfy + (size_t)kk * N + y0);
        };
        auto store_row = [&](const cplx (&z)[V], int row) {
            const gptr<cplx> dst = d.force + (size_t)row * N + lane;
            sfor<0, V>([&](auto j) { dst[64 * decltype(j)::value] = z[decltype(j)::value]; });
        };
        if constexpr (V <= 32) {
            cplx f1[V];
            sfor<0, V>([&](auto k2) {
                constexpr int K2 = decltype(k2)::value;
                cplx2 v, w2;
                spectra(K2, v, w2);
                x[K2] = force_in(K2, v, w2, 0);
                f1[K2] = force_in(K2, v, w2, 1);
            });
            W::inv(x, wb, qtw, lane);
            store_row(x, y0);
            W::inv(f1, wb, qtw, lane);
            store_row(f1, y0 + 1);
        } else {
#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                sfor<0, V>([&](auto k2) {
                    constexpr int K2 = decltype(k2)::value;
                    cplx2 v, w2;
                    spectra(K2, v, w2);
                    x[K2] = force_in(K2, v, w2, h);
                });
                W::inv(x, wb, qtw, lane);
                store_row(x, y0 + h);
            }
        }
    }
}

template <int N>
hipError_t wave_launch(int pass, const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s) {
    const dim3 grid(N / 2 / kWaves, count), block(kWaves * 64);
    if (pass == 0) {
        hipLaunchKernelGGL((k_row_r2c_w<N>), grid, block, 0, s, d, tw);
    } else if (pass == 1) {
        if (force) hipLaunchKernelGGL((k_col_conv_w<N, true>), grid, block, 0, s, d, tw);
        else hipLaunchKernelGGL((k_col_conv_w<N, false>), grid, block, 0, s, d, tw);
    } else {
        if (force) hipLaunchKernelGGL((k_row_c2r_w<N, true>), grid, block, 0, s, d, tw);
        else hipLaunchKernelGGL((k_row_c2r_w<N, false>), grid, block, 0, s, d, tw);
    }
    return hipGetLastError();
}

}  // namespace

bool wave_plan_compiled(int n) { return n == 256 || n == 512 || n == 1024 || n == 2048; }

hipError_t wave_pass(int pass, int n, const FieldDesc* d, int count, const float2* tw, bool force, hipStream_t s) {
    switch (n) {
        case 256: return wave_launch<256>(pass, d, count, tw, force, s);
        case 512: return wave_launch<512>(pass, d, count, tw, force, s);
        case 1024: return wave_launch<1024>(pass, d, count, tw, force, s);
        case 2048: return wave_launch<2048>(pass, d, count, tw, force, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pmk
