/*
 * pm_cpu32.c -- the per-step particle-mesh algorithm in FP32 on the host's
 * cores, timed by bench.py as the CPU baseline (its cpu_baseline leg) and
 * checked by tests/test_oracle.py.  Test and measurement infrastructure
 * only: the product library never links or calls it.
 *
 * One step follows the reference order (CosmosCuda.cuh, "cuh" below,
 * worker loop cuh:1336-1341):
 *   deposit   int64 fixed point, the oracle's rule (orc_deposit, cuh:446-595),
 *             into per-thread lattices that are allocated once and zeroed
 *             while they are summed (exact integer reduction);
 *   spectrum  L = complex(rho * 2^-shift, 0); L = FFT(L); L *= T;
 *             L = IFFT(L) (cuh:596-612, 1266, 1286-1287, 1291), FP32
 *             throughout (the reference's single-precision cuFFT pipeline),
 *             with T the FP64 composite of orc_transfer rounded to FP32;
 *   shift     Phi[(i+N/2)%N, (j+N/2)%N] = Re L[i,j] / N^2 (cuh:333-355);
 *   gradient  the reference's 4th-order stencil in FP32 (orc_gradient32,
 *             cuh:174-229);
 *   push      CIC/NGP gather + kick-drift in FP32 (orc_push32, cuh:231-312).
 * The FFT is an iterative radix-2 transform over split real/imaginary rows
 * with per-stage contiguous twiddle tables (from FP64), rows in parallel; the
 * column pass transposes in 32x32 blocks, transforms rows and transposes
 * back.  Every buffer is allocated by cpu32_create.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_gradient32(int n, float g, const float* phi, float* fxy);
void orc_push32(int n, int mode, float dt, int wrap_mode, int64_t np, float* x, float* y,
                float* vx, float* vy, const float* fxy);

typedef struct cpu32 {
    int n, mode, wrap_mode, shift, nth, lg;
    float dt, g;
    int64_t* priv;     /* nth lattices of N*N int64 (zero between steps) */
    float *re, *im;    /* N*N complex lattice, split */
    float *tre, *tim;  /* N*N transposition buffers */
    float *Tr, *Ti;    /* transfer function, FP32 */
    float* phi;        /* N*N potential */
    float* fxy;        /* 2*N*N force */
    float *wr, *wi;    /* per-stage twiddles: stage of length 2h at offset h - 1 */
    int* rev;          /* bit reversal of 0..N-1 */
} cpu32;

static int ilog2i(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }

/* In-place radix-2 FFT of one row (split), sign -1 forward, +1 inverse. */
static void fft_row(const cpu32* h, float* restrict xr, float* restrict xi, int sign)
{
    const int n = h->n;
    for (int i = 0; i < n; ++i) {
        const int r = h->rev[i];
        if (r > i) {
            float t = xr[i]; xr[i] = xr[r]; xr[r] = t;
            t = xi[i]; xi[i] = xi[r]; xi[r] = t;
        }
    }
    for (int half = 1; half < n; half <<= 1) {
        const float* wr = h->wr + half - 1;
        const float* wi = h->wi + half - 1;
        for (int s = 0; s < n; s += 2 * half) {
            float* ar = xr + s; float* ai = xi + s;
            float* br = xr + s + half; float* bi = xi + s + half;
            for (int k = 0; k < half; ++k) {
                const float c = wr[k], d = sign < 0 ? wi[k] : -wi[k];
                const float tr = br[k] * c - bi[k] * d;
                const float ti = br[k] * d + bi[k] * c;
                br[k] = ar[k] - tr; bi[k] = ai[k] - ti;
                ar[k] += tr; ai[k] += ti;
            }
        }
    }
}

static void transpose(int n, const float* restrict a, float* restrict b)
{
    #pragma omp parallel for schedule(static) collapse(2)
    for (int jb = 0; jb < n; jb += 32)
        for (int ib = 0; ib < n; ib += 32)
            for (int j = jb; j < jb + 32 && j < n; ++j)
                for (int i = ib; i < ib + 32 && i < n; ++i) b[(size_t)i * n + j] = a[(size_t)j * n + i];
}

static void fft2d(cpu32* h, int sign)
{
    const int n = h->n;
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) fft_row(h, h->re + (size_t)j * n, h->im + (size_t)j * n, sign);
    transpose(n, h->re, h->tre);
    transpose(n, h->im, h->tim);
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) fft_row(h, h->tre + (size_t)j * n, h->tim + (size_t)j * n, sign);
    transpose(n, h->tre, h->re);
    transpose(n, h->tim, h->im);
}

cpu32* cpu32_create(int n, int mode, float dt, float g, int wrap_mode, int shift, const double* Tre,
                    const double* Tim)
{
    if (n < 8 || (n & (n - 1))) return NULL;
    cpu32* h = (cpu32*)calloc(1, sizeof(cpu32));
    if (!h) return NULL;
    const size_t nn = (size_t)n * n;
    h->n = n; h->mode = mode; h->dt = dt; h->g = g; h->wrap_mode = wrap_mode; h->shift = shift;
    h->lg = ilog2i(n);
    h->nth = 1;
#ifdef _OPENMP
    h->nth = omp_get_max_threads();
#endif
    h->priv = (int64_t*)calloc((size_t)h->nth * nn, sizeof(int64_t));
    h->re = (float*)malloc(sizeof(float) * nn);
    h->im = (float*)malloc(sizeof(float) * nn);
    h->tre = (float*)malloc(sizeof(float) * nn);
    h->tim = (float*)malloc(sizeof(float) * nn);
    h->Tr = (float*)malloc(sizeof(float) * nn);
    h->Ti = (float*)malloc(sizeof(float) * nn);
    h->phi = (float*)malloc(sizeof(float) * nn);
    h->fxy = (float*)malloc(sizeof(float) * 2 * nn);
    h->wr = (float*)malloc(sizeof(float) * n);
    h->wi = (float*)malloc(sizeof(float) * n);
    h->rev = (int*)malloc(sizeof(int) * n);
    if (!h->priv || !h->re || !h->im || !h->tre || !h->tim || !h->Tr || !h->Ti || !h->phi || !h->fxy || !h->wr ||
        !h->wi || !h->rev) {
        free(h->priv); free(h->re); free(h->im); free(h->tre); free(h->tim); free(h->Tr); free(h->Ti);
        free(h->phi); free(h->fxy); free(h->wr); free(h->wi); free(h->rev); free(h);
        return NULL;
    }
    for (size_t c = 0; c < nn; ++c) { h->Tr[c] = (float)Tre[c]; h->Ti[c] = (float)Tim[c]; }
    const double pi = 3.14159265358979323846;
    for (int half = 1; half < n; half <<= 1)
        for (int k = 0; k < half; ++k) {
            h->wr[half - 1 + k] = (float)cos(-pi * k / half);
            h->wi[half - 1 + k] = (float)sin(-pi * k / half);
        }
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < h->lg; ++b) r |= ((i >> b) & 1) << (h->lg - 1 - b);
        h->rev[i] = r;
    }
    return h;
}

void cpu32_destroy(cpu32* h)
{
    if (!h) return;
    free(h->priv); free(h->re); free(h->im); free(h->tre); free(h->tim); free(h->Tr); free(h->Ti);
    free(h->phi); free(h->fxy); free(h->wr); free(h->wi); free(h->rev); free(h);
}

/* Potential of the last step (N*N floats, index i + j*N). */
const float* cpu32_potential(const cpu32* h) { return h->phi; }

static void deposit(cpu32* h, int64_t np, const float* x, const float* y, const float* m)
{
    const int n = h->n, mode = h->mode;
    const size_t nn = (size_t)n * n;
    const float scale = ldexpf(1.0f, h->shift), hi = (float)(n - 1);
    #pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t* acc = h->priv + (size_t)t * nn;
        #pragma omp for schedule(static)
        for (int64_t p = 0; p < np; ++p) {
            const float px = x[p], py = y[p];
            if (!(px >= 1.0f && px < hi && py >= 1.0f && py < hi)) continue;
            const int xi = (int)px, yi = (int)py;
            const size_t c = (size_t)xi + (size_t)yi * n;
            const float mm = m[p];
            if (mode == 0) { acc[c] += (int64_t)rintf(mm * scale); continue; }
            const float fx = px - (float)xi, fy = py - (float)yi, ax = 1.0f - fx, ay = 1.0f - fy;
            acc[c] += (int64_t)rintf((ax * ay) * mm * scale);
            acc[c + 1] += (int64_t)rintf((fx * ay) * mm * scale);
            acc[c + n] += (int64_t)rintf((ax * fy) * mm * scale);
            acc[c + n + 1] += (int64_t)rintf((fx * fy) * mm * scale);
        }
    }
    /* exact integer reduction straight into the complex lattice, zeroing the
     * per-thread lattices for the next step */
    const float inv = ldexpf(1.0f, -h->shift);
    const int nth = h->nth;
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)nn; ++c) {
        int64_t s = 0;
        for (int t = 0; t < nth; ++t) {
            int64_t* q = h->priv + (size_t)t * nn + c;
            s += *q;
            *q = 0;
        }
        h->re[c] = (float)s * inv;
        h->im[c] = 0.0f;
    }
}

/* Advances the FP32 state (x, y, vx, vy) by `steps` steps in place. */
void cpu32_step(cpu32* h, int64_t np, float* x, float* y, float* vx, float* vy, const float* m, int steps)
{
    const int n = h->n;
    const size_t nn = (size_t)n * n;
    for (int s = 0; s < steps; ++s) {
        deposit(h, np, x, y, m);
        fft2d(h, -1);
        #pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < (int64_t)nn; ++c) {
            const float a = h->re[c], b = h->im[c];
            h->re[c] = a * h->Tr[c] - b * h->Ti[c];
            h->im[c] = a * h->Ti[c] + b * h->Tr[c];
        }
        fft2d(h, +1);
        const float inv = 1.0f / ((float)n * (float)n);
        #pragma omp parallel for schedule(static)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const int sx = (i + n / 2) % n, sy = (j + n / 2) % n;
                h->phi[sx + (size_t)sy * n] = h->re[i + (size_t)j * n] * inv;
            }
        orc_gradient32(n, h->g, h->phi, h->fxy);
        orc_push32(n, h->mode, h->dt, h->wrap_mode, np, x, y, vx, vy, h->fxy);
    }
}
