/*
 * pm_oracle.c -- CPU restatement of the particle-mesh (PM) gravity step of
 * CosmosSimulationWithCuda, used ONLY as test infrastructure: the pytest
 * suite, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it as
 * the checker.  Nothing in the product library links or calls it.
 *
 * Every function names the reference lines it restates (CosmosCuda.cuh is
 * abbreviated "cuh").  Conventions follow the reference:
 *   lattice index idx = i + j*N  (i = x, fastest;  cuh:102-103, 195-196)
 *   forward DFT  X[k] = sum_n x[n] exp(-2 pi i k n / N), inverse unnormalised
 *   (cuFFT conventions, cuh:1266-1294).
 *
 * Pinning (what checks each function, so a plausible slip fails a test):
 *   orc_fft1d / orc_fft2d   brute-force DFT at N<=32, numpy.fft FP64 fixtures,
 *                           Parseval, delta -> constant, round trip.
 *   orc_green               closed form 1/r at integer offsets, origin value.
 *   orc_transfer            closed forms: NGP  T = Ghat*(-1)^(kx+ky)/(1+eps),
 *                           CIC  T = Ghat*s*c/(c^2+eps) with c = cos(pi kx/N)cos(pi ky/N);
 *                           Ghat imaginary part == 0 (G real-even).
 *   orc_deposit             brute force on tiny inputs (numpy fixture), exact
 *                           mass sum for NGP, CIC partition of unity,
 *                           permutation invariance, bounds rule.
 *   orc_potential           numpy FP64 convolution fixture (independent FFT),
 *                           direct O(N^4) periodic convolution at N=8/16,
 *                           single particle: peak lands at its own cell.
 *   orc_gradient*           exact on cubic polynomials in the interior,
 *                           edge clamp rule checked cell by cell.
 *   orc_push32 / orc_push64 linear force field exactness, rest start closed
 *                           form x1 = x0 + a dt^2, wrap thresholds.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

static int ilog2(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* ------------------------------------------------------------------ FFT -- */
/* Radix-2 decimation-in-time, in place, FP64, sign = -1 forward, +1 inverse
 * (unnormalised).  Stands in for cufftExecC2C (cuh:1266-1294); cuFFT itself is
 * closed source, so the restatement follows the DFT definition. */
int orc_fft1d(int n, double* re, double* im, int sign)
{
    if (!is_pow2(n)) return -1;
    int lg = ilog2(n);
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
        if (r > i) {
            double t = re[i]; re[i] = re[r]; re[r] = t;
            t = im[i]; im[i] = im[r]; im[r] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            /* twiddle from the exact angle, no recurrence: keeps FP64 error ~eps*log n */
            double ang = sign * 2.0 * ORC_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (int s = 0; s < n; s += len) {
                int a = s + k, b = a + half;
                double tr = re[b] * wr - im[b] * wi;
                double ti = re[b] * wi + im[b] * wr;
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] += tr; im[a] += ti;
            }
        }
    }
    return 0;
}

/* 2-D transform of an N x N complex lattice: rows (x) then columns (y),
 * the order README.md:28-32 describes. */
int orc_fft2d(int n, double* re, double* im, int sign)
{
    if (!is_pow2(n)) return -1;
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) orc_fft1d(n, re + (size_t)j * n, im + (size_t)j * n, sign);
    #pragma omp parallel
    {
        double* cr = (double*)malloc(sizeof(double) * n);
        double* ci = (double*)malloc(sizeof(double) * n);
        #pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < n; ++j) { cr[j] = re[i + (size_t)j * n]; ci[j] = im[i + (size_t)j * n]; }
            orc_fft1d(n, cr, ci, sign);
            for (int j = 0; j < n; ++j) { re[i + (size_t)j * n] = cr[j]; im[i + (size_t)j * n] = ci[j]; }
        }
        free(cr); free(ci);
    }
    return 0;
}

/* ------------------------------------------------------- setup spectra -- */
/* Periodic 1/r lattice, cuh:89-119: dx = i - (i <= N/2 ? 0 : N), same for y,
 * r = sqrt(dx^2+dy^2); G = 1/r where r > 0.5, else the origin value (3.0 in
 * the reference, cuh:113-116; a runtime parameter here). */
void orc_green(int n, double g0, double* G)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double dx = i - (i <= n / 2 ? 0 : n);
            double dy = j - (j <= n / 2 ? 0 : n);
            double r = sqrt(dx * dx + dy * dy);
            G[i + (size_t)j * n] = r > 0.5 ? 1.0 / r : g0;
        }
}

/* Spatial mass-assignment kernel, cuh:374-415: NGP -> 1 at (N/2,N/2);
 * CIC -> 0.25 at (N/2+a, N/2+b), a,b in {0,1}. */
void orc_mass_kernel(int n, int mode, double* D)
{
    memset(D, 0, sizeof(double) * (size_t)n * n);
    int c = n / 2;
    if (mode == 0) { D[c + (size_t)c * n] = 1.0; return; }
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) D[(c + a) + (size_t)(c + b) * n] = 0.25;
}

/* Full transfer function of the reference's spectral step, in FP64:
 *   Ghat = FFT(G)                      cuh:1018, 1273-1276
 *   Dhat = FFT(D)                      cuh:1020, 1280-1283
 *   CIC: Dhat *= exp(i*phi), phi = -2 pi (kx*fx + ky*fy)/N, fx=fy=-0.5,
 *        kx signed (k <= N/2 ? k : k-N)                     cuh:416-445, 1027
 *   T = Ghat * conj(Dhat) / (|Dhat|^2 + eps)                 cuh:120-173
 * The per-step product applies mult then div (cuh:1286-1287); T is their
 * composite.  Output: complex T over the full N x N lattice. */
void orc_transfer(int n, int mode, double g0, double eps, double* Tre, double* Tim)
{
    size_t nn = (size_t)n * n;
    double* gr = (double*)malloc(sizeof(double) * nn);
    double* gi = (double*)calloc(nn, sizeof(double));
    double* dr = (double*)malloc(sizeof(double) * nn);
    double* di = (double*)calloc(nn, sizeof(double));
    orc_green(n, g0, gr);
    orc_mass_kernel(n, mode, dr);
    orc_fft2d(n, gr, gi, -1);
    orc_fft2d(n, dr, di, -1);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            size_t idx = i + (size_t)j * n;
            double xr = dr[idx], xi = di[idx];
            if (mode == 1) {
                int kx = i <= n / 2 ? i : i - n;
                int ky = j <= n / 2 ? j : j - n;
                double ph = -2.0 * ORC_PI * ((double)kx * -0.5 / n + (double)ky * -0.5 / n);
                double c = cos(ph), s = sin(ph);
                double yr = xr * c - xi * s, yi = xr * s + xi * c;
                xr = yr; xi = yi;
            }
            /* L*Ghat (complex), then L*conj(D)/(|D|^2+eps); composite factor: */
            double den = xr * xr + xi * xi + eps;
            double qr = xr / den, qi = -xi / den;            /* conj(D)/(|D|^2+eps) */
            Tre[idx] = gr[idx] * qr - gi[idx] * qi;
            Tim[idx] = gr[idx] * qi + gi[idx] * qr;
        }
    free(gr); free(gi); free(dr); free(di);
}

/* --------------------------------------------------------------- deposit -- */
/* In-bounds rule shared by deposit and push, cuh:257, 497:
 * (int)x in [1, N-2] and (int)y in [1, N-2].  For finite x this equals
 * 1 <= x < N-1; NaN fails both comparisons, as it must. */
static inline int in_bounds(float x, float y, int n)
{
    return x >= 1.0f && x < (float)(n - 1) && y >= 1.0f && y < (float)(n - 1);
}

/* Fixed-point quantisation of one contribution p = w*m (FP32 product):
 * q = round-to-nearest-even(p * 2^shift).  p*2^shift is exact in FP32. */
static inline int64_t quantise(float p, float scale)
{
    return (int64_t)rintf(p * scale);
}

/* Mass deposit, cuh:446-595 (bounds cuh:497, weights cuh:498-505, targets
 * cuh:508-511), restated with int64 fixed-point sums instead of float
 * atomics.  NGP: +m at (xi,yi).  CIC: fx = x - xi, ax = 1 - fx (same for y);
 * w00 = ax*ay -> (xi,yi), w10 = fx*ay -> (xi+1,yi), w01 = ax*fy -> (xi,yi+1),
 * w11 = fx*fy -> (xi+1,yi+1); contribution = w*m in FP32, each product
 * rounded separately.  The reference skips the Np mod 4 tail (cuh:239,454);
 * this restatement, like the product, deposits every particle.
 * rho must hold N*N int64 and is overwritten.  Returns 0. */
int orc_deposit(int n, int mode, int shift, int64_t np, const float* x, const float* y,
                const float* m, int64_t* rho)
{
    size_t nn = (size_t)n * n;
    const float scale = ldexpf(1.0f, shift);
    memset(rho, 0, sizeof(int64_t) * nn);
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    int64_t* priv = (int64_t*)calloc((size_t)nth * nn, sizeof(int64_t));
    #pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t* acc = priv + (size_t)t * nn;
        #pragma omp for schedule(static)
        for (int64_t p = 0; p < np; ++p) {
            float px = x[p], py = y[p];
            if (!in_bounds(px, py, n)) continue;
            int xi = (int)px, yi = (int)py;
            size_t c = (size_t)xi + (size_t)yi * n;
            if (mode == 0) { acc[c] += quantise(m[p], scale); continue; }
            float fx = px - (float)xi, fy = py - (float)yi;
            float ax = 1.0f - fx, ay = 1.0f - fy;
            float w00 = ax * ay, w10 = fx * ay, w01 = ax * fy, w11 = fx * fy;
            float mm = m[p];
            acc[c] += quantise(w00 * mm, scale);
            acc[c + 1] += quantise(w10 * mm, scale);
            acc[c + n] += quantise(w01 * mm, scale);
            acc[c + n + 1] += quantise(w11 * mm, scale);
        }
    }
    /* integer reduction: exact, so the thread count cannot change the bits */
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)nn; ++c) {
        int64_t s = 0;
        for (int t = 0; t < nth; ++t) s += priv[(size_t)t * nn + c];
        rho[c] = s;
    }
    free(priv);
    return 0;
}

/* FP64 deposit for the all-double trajectory oracle (same rule, double weights). */
int orc_deposit64(int n, int mode, int64_t np, const double* x, const double* y,
                  const double* m, double* rho)
{
    size_t nn = (size_t)n * n;
    memset(rho, 0, sizeof(double) * nn);
    for (int64_t p = 0; p < np; ++p) {
        double px = x[p], py = y[p];
        if (!(px >= 1.0 && px < n - 1.0 && py >= 1.0 && py < n - 1.0)) continue;
        int xi = (int)px, yi = (int)py;
        size_t c = (size_t)xi + (size_t)yi * n;
        if (mode == 0) { rho[c] += m[p]; continue; }
        double fx = px - xi, fy = py - yi, ax = 1.0 - fx, ay = 1.0 - fy;
        rho[c] += ax * ay * m[p];
        rho[c + 1] += fx * ay * m[p];
        rho[c + n] += ax * fy * m[p];
        rho[c + n + 1] += fx * fy * m[p];
    }
    return 0;
}

/* ------------------------------------------------------------- potential -- */
/* Per-step spectral pipeline in FP64, following the reference literally:
 *   L = complex(rho, 0)                          cuh:596-612, 1245
 *   L = FFT(L); L *= T  (mult + div composite)   cuh:1266, 1286-1287
 *   L = IFFT(L)  (unnormalised)                  cuh:1291
 *   Phi[(i+N/2)%N, (j+N/2)%N] = Re L[i,j] / N^2  cuh:333-355
 * rho_d is the mass lattice in mass units (double). */
void orc_potential_d(int n, const double* Tre, const double* Tim, const double* rho_d, double* phi)
{
    size_t nn = (size_t)n * n;
    double* lr = (double*)malloc(sizeof(double) * nn);
    double* li = (double*)calloc(nn, sizeof(double));
    memcpy(lr, rho_d, sizeof(double) * nn);
    orc_fft2d(n, lr, li, -1);
    for (size_t c = 0; c < nn; ++c) {
        double a = lr[c], b = li[c];
        lr[c] = a * Tre[c] - b * Tim[c];
        li[c] = a * Tim[c] + b * Tre[c];
    }
    orc_fft2d(n, lr, li, +1);
    double inv = 1.0 / ((double)n * (double)n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            int sx = (i + n / 2) % n, sy = (j + n / 2) % n;
            phi[sx + (size_t)sy * n] = lr[i + (size_t)j * n] * inv;
        }
    free(lr); free(li);
}

/* Same, from the int64 fixed-point lattice: rho = q * 2^-shift. */
void orc_potential(int n, const double* Tre, const double* Tim, int shift, const int64_t* rho, double* phi)
{
    size_t nn = (size_t)n * n;
    double* rd = (double*)malloc(sizeof(double) * nn);
    double s = ldexp(1.0, -shift);
    for (size_t c = 0; c < nn; ++c) rd[c] = (double)rho[c] * s;
    orc_potential_d(n, Tre, Tim, rd, phi);
    free(rd);
}

/* -------------------------------------------------------------- gradient -- */
/* 4th-order central difference, cuh:174-229:
 *   Fx = (-P[i+2] + 8 P[i+1] - 8 P[i-1] + P[i-2]) / 12 * G,  Fy likewise in j;
 * any neighbour outside [0,N) is replaced by the centre value (cuh:186-220). */
void orc_gradient64(int n, double g, const double* phi, double* fx, double* fy)
{
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            size_t c = i + (size_t)j * n;
            double p0 = phi[c];
            double l = i - 1 >= 0 ? phi[c - 1] : p0, r = i + 1 < n ? phi[c + 1] : p0;
            double ll = i - 2 >= 0 ? phi[c - 2] : p0, rr = i + 2 < n ? phi[c + 2] : p0;
            double t = j - 1 >= 0 ? phi[c - n] : p0, b = j + 1 < n ? phi[c + n] : p0;
            double tt = j - 2 >= 0 ? phi[c - 2 * (size_t)n] : p0, bb = j + 2 < n ? phi[c + 2 * (size_t)n] : p0;
            fx[c] = (-rr + 8.0 * r - 8.0 * l + ll) / 12.0 * g;
            fy[c] = (-bb + 8.0 * b - 8.0 * t + tt) / 12.0 * g;
        }
}

/* FP32 mirror of the product's stencil kernel (same operation order, no
 * contraction):  s = ((-rr + 8*r) - 8*l) + ll;  F = (s / 12) * G.
 * fxy is interleaved (Fx, Fy) per cell.  Used to check the product's stencil
 * bit for bit on an identical FP32 potential. */
void orc_gradient32(int n, float g, const float* phi, float* fxy)
{
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            size_t c = i + (size_t)j * n;
            float p0 = phi[c];
            float l = i - 1 >= 0 ? phi[c - 1] : p0, r = i + 1 < n ? phi[c + 1] : p0;
            float ll = i - 2 >= 0 ? phi[c - 2] : p0, rr = i + 2 < n ? phi[c + 2] : p0;
            float t = j - 1 >= 0 ? phi[c - n] : p0, b = j + 1 < n ? phi[c + n] : p0;
            float tt = j - 2 >= 0 ? phi[c - 2 * (size_t)n] : p0, bb = j + 2 < n ? phi[c + 2 * (size_t)n] : p0;
            float sx = ((-rr + 8.0f * r) - 8.0f * l) + ll;
            float sy = ((-bb + 8.0f * b) - 8.0f * t) + tt;
            fxy[2 * c] = (sx / 12.0f) * g;
            fxy[2 * c + 1] = (sy / 12.0f) * g;
        }
}

/* ------------------------------------------------------------------ push -- */
/* Force gather + integration, cuh:231-312 (bounds cuh:257, CIC cuh:262-279,
 * update cuh:286-287, wrap cuh:288-303), with the state (x, v) in place of
 * the reference's (x, xOld):  v' = v + a dt;  x' = x + v' dt  -- the same
 * algebra as x' = 2x - xOld + a dt^2 when v = (x - xOld)/dt.
 * FP32 mirror of the product kernel, operation for operation:
 *   CIC a = ((F00*w00 + F10*w10) + F01*w01) + F11*w11   (no contraction)
 *   v' = fmaf(a, dt, v);  x' = fmaf(v', dt, x)
 *   wrap: x' < 2 -> x' += N-4;  x' > N-2 -> x' -= N-4 (one test per side, y alike)
 *   wrap_mode 1 (reference quirk, xOld not shifted, cuh:288-303): the implied
 *   velocity jumps by the same amount, so v' +/-= (N-4)/dt.
 * Out-of-bounds particles are left untouched (cuh:257). */
void orc_push32(int n, int mode, float dt, int wrap_mode, int64_t np, float* x, float* y,
                float* vx, float* vy, const float* fxy)
{
    const float span = (float)(n - 4);
    const float jump = span / dt;
    #pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        float px = x[p], py = y[p];
        if (!in_bounds(px, py, n)) continue;
        int xi = (int)px, yi = (int)py;
        size_t c = (size_t)xi + (size_t)yi * n;
        float ax, ay;
        if (mode == 0) {
            ax = fxy[2 * c]; ay = fxy[2 * c + 1];
        } else {
            float fx = px - (float)xi, fy = py - (float)yi;
            float wx = 1.0f - fx, wy = 1.0f - fy;
            float w00 = wx * wy, w10 = fx * wy, w01 = wx * fy, w11 = fx * fy;
            const float* f00 = fxy + 2 * c;
            const float* f10 = fxy + 2 * (c + 1);
            const float* f01 = fxy + 2 * (c + n);
            const float* f11 = fxy + 2 * (c + n + 1);
            ax = ((f00[0] * w00 + f10[0] * w10) + f01[0] * w01) + f11[0] * w11;
            ay = ((f00[1] * w00 + f10[1] * w10) + f01[1] * w01) + f11[1] * w11;
        }
        float nvx = fmaf(ax, dt, vx[p]);
        float nvy = fmaf(ay, dt, vy[p]);
        float nx = fmaf(nvx, dt, px);
        float ny = fmaf(nvy, dt, py);
        if (nx < 2.0f) { nx += span; if (wrap_mode == 1) nvx += jump; }
        if (ny < 2.0f) { ny += span; if (wrap_mode == 1) nvy += jump; }
        if (nx > (float)(n - 2)) { nx -= span; if (wrap_mode == 1) nvx -= jump; }
        if (ny > (float)(n - 2)) { ny -= span; if (wrap_mode == 1) nvy -= jump; }
        x[p] = nx; y[p] = ny; vx[p] = nvx; vy[p] = nvy;
    }
}

/* All-double counterpart (trajectory oracle). */
void orc_push64(int n, int mode, double dt, int wrap_mode, int64_t np, double* x, double* y,
                double* vx, double* vy, const double* fx, const double* fy)
{
    const double span = n - 4.0;
    #pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        double px = x[p], py = y[p];
        if (!(px >= 1.0 && px < n - 1.0 && py >= 1.0 && py < n - 1.0)) continue;
        int xi = (int)px, yi = (int)py;
        size_t c = (size_t)xi + (size_t)yi * n;
        double ax, ay;
        if (mode == 0) { ax = fx[c]; ay = fy[c]; }
        else {
            double ex = px - xi, ey = py - yi, wx = 1.0 - ex, wy = 1.0 - ey;
            ax = fx[c] * wx * wy + fx[c + 1] * ex * wy + fx[c + n] * wx * ey + fx[c + n + 1] * ex * ey;
            ay = fy[c] * wx * wy + fy[c + 1] * ex * wy + fy[c + n] * wx * ey + fy[c + n + 1] * ex * ey;
        }
        double nvx = vx[p] + ax * dt, nvy = vy[p] + ay * dt;
        double nx = px + nvx * dt, ny = py + nvy * dt;
        if (nx < 2.0) { nx += span; if (wrap_mode == 1) nvx += span / dt; }
        if (ny < 2.0) { ny += span; if (wrap_mode == 1) nvy += span / dt; }
        if (nx > n - 2.0) { nx -= span; if (wrap_mode == 1) nvx -= span / dt; }
        if (ny > n - 2.0) { ny -= span; if (wrap_mode == 1) nvy -= span / dt; }
        x[p] = nx; y[p] = ny; vx[p] = nvx; vy[p] = nvy;
    }
}

/* ----------------------------------------------------------------- steps -- */
/* One step of the reference order (cuh:1336-1341: scatter -> FFT -> mult/div
 * -> IFFT -> shift -> gradient -> gather/update) on FP32 particle state:
 * int64 deposit (bit-exact rule), FP64 field, FP64 force rounded to FP32,
 * FP32 push.  Scratch: rho N*N int64, phi/fx/fy N*N double, fxy 2*N*N float. */
void orc_step_mixed(int n, int mode, float dt, float g, int wrap_mode, int shift,
                    const double* Tre, const double* Tim, int64_t np,
                    float* x, float* y, float* vx, float* vy, const float* m,
                    int64_t* rho, double* phi, double* fx, double* fy, float* fxy)
{
    size_t nn = (size_t)n * n;
    orc_deposit(n, mode, shift, np, x, y, m, rho);
    orc_potential(n, Tre, Tim, shift, rho, phi);
    orc_gradient64(n, (double)g, phi, fx, fy);
    for (size_t c = 0; c < nn; ++c) { fxy[2 * c] = (float)fx[c]; fxy[2 * c + 1] = (float)fy[c]; }
    orc_push32(n, mode, dt, wrap_mode, np, x, y, vx, vy, fxy);
}

/* One all-double step: the FP64 trajectory oracle. */
void orc_step64(int n, int mode, double dt, double g, int wrap_mode,
                const double* Tre, const double* Tim, int64_t np,
                double* x, double* y, double* vx, double* vy, const double* m,
                double* rho, double* phi, double* fx, double* fy)
{
    orc_deposit64(n, mode, np, x, y, m, rho);
    orc_potential_d(n, Tre, Tim, rho, phi);
    orc_gradient64(n, g, phi, fx, fy);
    orc_push64(n, mode, dt, wrap_mode, np, x, y, vx, vy, fx, fy);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int t)
{
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}
