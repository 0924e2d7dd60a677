// C-ABI layer of libpm.so (declared in include/pm.h): handle lifecycle,
// uploads/read-back, step orchestration on one HIP stream, diagnostics and
// per-kernel event timing.  Replaces the device side of the reference's
// struct Universe (CosmosCuda.cuh:830-1428): ctor device setup 969-1035,
// step methods 1209-1300, worker-loop body 1336-1341, dtor 1396-1427.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <climits>
#include <cstring>
#include <utility>
#include <new>
#include <string>
#include <vector>

#include "pm_sim.h"

namespace {

thread_local std::string g_last_error;   // errors of calls made without a valid handle

}  // namespace

namespace pmi {

pm_status fail(pm_sim* h, pm_status st, const std::string& msg) {
    if (h) h->err = msg;
    g_last_error = msg;
    return st;
}

pm_status hip_fail(pm_sim* h, hipError_t e, const char* what) {
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    return fail(h, e == hipErrorOutOfMemory ? PM_ENOMEM : PM_EDEVICE, msg);
}


bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

pm_status validate(const pm_config* c, std::string* why) {
    if (!c) { *why = "null config"; return PM_EINVAL; }
    if (!pmk::lattice_supported(c->n_lattice)) { *why = "n_lattice must be a power of two in [64, 8192]"; return PM_EINVAL; }
    if (c->mode != PM_NGP && c->mode != PM_CIC) { *why = "mode must be PM_NGP or PM_CIC"; return PM_EINVAL; }
    if (c->n_particles < 0 || c->n_particles > (int64_t)INT32_MAX) { *why = "n_particles must be in [0, 2^31-1]"; return PM_EINVAL; }
    if (c->resort_interval < 0) { *why = "resort_interval must be >= 0"; return PM_EINVAL; }
    if (!std::isfinite(c->dt) || !std::isfinite(c->gravity)) { *why = "dt and gravity must be finite"; return PM_EINVAL; }
    if (!(c->deconv_eps >= 0.0) || !std::isfinite(c->green_origin)) { *why = "bad deconv_eps or green_origin"; return PM_EINVAL; }
    if (c->wrap_mode != PM_WRAP_KEEP_VELOCITY && c->wrap_mode != PM_WRAP_REFERENCE) { *why = "bad wrap_mode"; return PM_EINVAL; }
    if (c->force_mode != PM_FORCE_SPECTRAL && c->force_mode != PM_FORCE_STENCIL) { *why = "bad force_mode"; return PM_EINVAL; }
    if (c->precision != PM_FFT_FP32 && c->precision != PM_FFT_BF16X3 && c->precision != PM_FFT_BF16X6) {
        *why = "precision: unknown";
        return PM_EINVAL;
    }
    if (c->precision != PM_FFT_FP32 && (c->n_lattice < 256 || c->n_lattice > 8192)) {
        *why = "the split-BF16 precisions need 256 <= n_lattice <= 8192";
        return PM_EINVAL;
    }
    if (c->mass_shift < 0 || c->mass_shift > 62) { *why = "mass_shift must be in [0, 62]"; return PM_EINVAL; }
    if (c->wrap_mode == PM_WRAP_REFERENCE && c->dt == 0.0f) { *why = "wrap_mode reference needs dt != 0"; return PM_EINVAL; }
    return PM_OK;
}

int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

// FP32: Stockham radix 4/8 (default plan for the size); BF16X3: radix-16
// matrix-core stages, one row pair / column per block.
// Re-sort every 9 steps at load (7 for dense runs, more than 64 particles per
// cell); the drift rule (cycle_boundary) shortens the interval as the
// particles speed up.  The counting and re-sorting steps of a cycle grow with
// the drift since the last sort (particles outside their tile windows,
// scattered re-sort writes).  Two colliding disks, 1e8 particles, 2048^2,
// CIC, 11-cell window margin, ms per step over steps 5-25 / 25-61 / 61-200:
// 7 with the rule at 2000 per million 0.871-0.878 / 0.849-0.855 / 0.862; 9 at
// 2000 0.852-0.856 / 0.838 / 0.893 (the rule does not fire: few particles
// leave the wider windows, the re-sort steps grow instead); 9 at 200 0.851-
// 0.854 / 0.840-0.843 / 0.862-0.865; fixed 8, 9, 10 late 0.879 / 0.893 /
// 0.928.  NGP disks 1e8 over steps 5-25: 0.830 at 9 against 0.844 at 7; the
// 2e7 single disk the same either way; 1e9 NGP (dense) over steps 5-15 11.40
// at 9 against 9.93 at 7.
constexpr int kResortIntervalDefault = 9;
constexpr int kResortIntervalDense = 7;
// Large dense CIC runs (at least 3e7 particles, more than 64 per cell on
// average) re-sort every 2 steps with the separate sort: there the fused
// re-sort step and the window spills of drifted particles (four contended
// global atomics each) cost more than a sort.  Measured on two colliding
// disks of 1e9 particles on 2048^2 over steps 5-30 and 30-55: CIC 16.3 and
// 22.1 ms per step at the default, 14.4 and 14.0 at 2; NGP 10.0 and 10.9 at
// the default, 14.3 and 13.8 at 2.  Chosen at each load unless an interval
// was set.
int auto_resort_interval(int mode, int64_t np, int n) {
    const bool dense = np > (int64_t)pmk::kPipelineMaxDensity * n * n;
    if (!dense) return kResortIntervalDefault;
    return mode == PM_CIC && np >= pmk::kSmallRunParticles ? 2 : kResortIntervalDense;
}

pmk::FftPlan plan_for(const pm_config& c) {
    if (c.precision != PM_FFT_FP32) {
        pmk::FftPlan p;
        p.rp = 1; p.cb = 1; p.radix = c.precision == PM_FFT_BF16X6 ? 17 : 16;
        p.it = pmk::pipe_plan_compiled(c.n_lattice, p.radix, 2) ? 2 : 1;
        return p;
    }
    return pmk::default_plan(c.n_lattice);
}

struct Layout {
    int64_t particles, bins, rho, fdesc, H, T, dk, tw, phi, force, diag, total, setup_work;
};

Layout layout_of(const pm_config* c) {
    const int64_t n = c->n_lattice, nk = n / 2 + 1, cap = std::max<int64_t>(c->n_particles, 1);
    Layout L{};
    L.particles = align256(4 * cap);
    const int64_t nb = pmk::bin_count_of((int)n, pmk::tile_log_for((int)n));
    L.bins = align256(4 * (nb + 1)) * 3 + align256(8 * (nb + 2)) + 256;
    L.rho = align256(8 * n * n);
    L.H = align256(8 * nk * n);
    L.T = align256(4 * nk * n);
    L.dk = align256(4 * n);
    L.tw = align256(8 * n);
    L.phi = align256(4 * n * n);
    L.force = align256(8 * n * n);
    const int64_t dblocks = std::max(pmk::diag_blocks(cap), pmk::diag_lattice_blocks((int)n));
    L.diag = align256(8 * 8 * dblocks);
    L.fdesc = align256(4 * sizeof(pmk::FieldDesc));
    L.total = 12 * L.particles + 2 * L.bins + 2 * L.rho + L.fdesc + 2 * L.H + L.T + L.dk + L.tw + L.phi + L.force +
              L.diag;
    L.setup_work = 16 * n * n;
    return L;
}

pm_status dalloc(pm_sim* h, void** p, int64_t bytes) {
    hipError_t e = hipMalloc(p, (size_t)bytes);
    if (e != hipSuccess) return hip_fail(h, e, "hipMalloc");
    h->allocs.push_back(*p);
    h->bytes += bytes;
    return PM_OK;
}

void release(pm_sim* h) {
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->own_stream && h->own_stream != h->stream) (void)hipStreamSynchronize(h->own_stream);
    for (auto& r : h->recs) { (void)hipEventDestroy(r.start); (void)hipEventDestroy(r.end); }
    for (auto e : h->pool) (void)hipEventDestroy(e);
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    if (h->mig_event) (void)hipEventDestroy(h->mig_event);
    if (h->mig_host) (void)hipHostFree(h->mig_host);
    if (h->tables_host) (void)hipHostFree(h->tables_host);
    if (h->tables_event) (void)hipEventDestroy(h->tables_event);
    h->tables_event = nullptr;
    h->mig_event = nullptr;
    h->mig_host = nullptr;
    h->tables_host = nullptr;
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    h->own_stream = nullptr;
    h->stream = nullptr;
}

// ---- per-kernel event timing ------------------------------------------------
hipEvent_t pool_event(pm_sim* h) {
    if (!h->pool.empty()) { hipEvent_t e = h->pool.back(); h->pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct KernelTimer {
    pm_sim* h; int kernel; hipEvent_t s = nullptr;
    KernelTimer(pm_sim* hh, int k) : h(hh), kernel(k) {
        if (h->profiling && ((h->prof_mask >> k) & 1) && (s = pool_event(h))) (void)hipEventRecord(s, h->stream);
    }
    ~KernelTimer() {
        if (!s) return;
        hipEvent_t e = pool_event(h);
        if (!e) { h->pool.push_back(s); return; }
        (void)hipEventRecord(e, h->stream);
        h->recs.push_back({kernel, s, e});
    }
};

pm_status resolve_profile(pm_sim* h) {
    if (h->recs.empty()) return PM_OK;
    PM_HIP(h, hipStreamSynchronize(h->stream));
    for (auto& r : h->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.start, r.end) == hipSuccess) {
            h->prof_ms[r.kernel] += ms;
            h->prof_n[r.kernel] += 1;
        }
        h->pool.push_back(r.start);
        h->pool.push_back(r.end);
    }
    h->recs.clear();
    return PM_OK;
}

// Device descriptors of the field passes, one per mass-lattice parity; [2 + p]
// also has the row R2C pass zero the other lattice (the next deposit target).
pm_status upload_field_desc(pm_sim* h) {
    pmk::FieldDesc d[4];
    for (int p = 0; p < 2; ++p) {
        d[p].rho = h->rho_buf[p];
        d[p].H = h->H;
        d[p].Hfy = h->Hfy;
        d[p].T = h->T;
        d[p].dk = h->dk;
        d[p].phi = h->phi;
        d[p].force = h->force;
        d[p].inv_scale = std::ldexp(1.0f, -h->shift);
        d[p].gravity = h->cfg.gravity;
        d[2 + p] = d[p];
        d[2 + p].rho_zero = h->rho_buf[p ^ 1];
    }
    PM_HIP(h, pmk::set_field_desc(h->fdesc, d, h->stream));   // no host wait
    return PM_OK;
}

// ---- stages -----------------------------------------------------------------
// A re-sort cycle begins (every push so far ran in the old order).  The cycle
// before the one that just ended has all its migrant reads in (they lag at
// most two pushes and a cycle is at least three), so its peak rate steers
// the automatic interval: as the particles speed up (collapsing disks), the
// step kernels of a long cycle's last steps find more and more of them
// outside their tile windows (global atomics) and the re-sorting step
// scatters more of them across tiles.  The interval only shortens (a rule
// that also lengthened it below a lower rate oscillated and was slower).
// Intervals never change results.
void cycle_boundary(pm_sim* h) {
    if (h->resort_auto && h->resort_permille > 0 && h->drift_ppm > 0 && h->resort_base >= 4) {
        const double r = h->prev_peak * 1e6 / (double)std::max<int64_t>(h->np, 1);
        if (r > (double)h->drift_ppm && h->resort_interval > 3) {
            h->resort_interval -= 1;
            h->interval_changes += 1;
        }
    }
    h->prev_peak = h->cyc_peak;
    h->cyc_peak = 0.0;
    h->prev_begin = h->cyc_begin;
    h->cyc_begin = h->pushes;
}

pm_status sort_now(pm_sim* h, bool counted) {
    h->scatter_pending = false;
    h->resort_request = false;
    KernelTimer kt(h, PM_K_SORT);
    PM_HIP(h, pmk::sort_particles(h->cfg.n_lattice, h->cur, h->alt, h->np, h->bins, counted, h->stream));
    std::swap(h->cur, h->alt);
    h->steps_since_sort = 0;
    h->sorts += 1;
    cycle_boundary(h);
    return PM_OK;
}

pm_status stage_deposit(pm_sim* h) {
    const int n = h->cfg.n_lattice;
    if (h->steps_since_sort >= h->resort_interval) {
        pm_status st = sort_now(h, false);
        if (st != PM_OK) return st;
    }
    {
        KernelTimer kt(h, PM_K_CLEAR);
        PM_HIP(h, pmk::fill_zero(h->rho_buf[h->rho_cur], (size_t)8 * n * n, h->stream));
    }
    {
        KernelTimer kt(h, PM_K_DEPOSIT);
        PM_HIP(h, pmk::deposit(n, h->cfg.mode, h->cur.x, h->cur.y, h->cur.m, h->bins, h->max_chunks,
                               std::ldexp(1.0f, h->shift), h->rho_buf[h->rho_cur], h->migrants, h->stream));
    }
    h->rho_valid = true;
    return PM_OK;
}

// clear_next: the row R2C pass also zeroes the other mass lattice, which the
// push that follows deposits into (pm_step; the push then skips its clear).
pm_status stage_field(pm_sim* h, bool clear_next) {
    const int n = h->cfg.n_lattice;
    const bool spectral = h->cfg.force_mode == PM_FORCE_SPECTRAL;
    const pmk::FieldDesc* d = h->fdesc + h->rho_cur + (clear_next ? 2 : 0);
    {
        KernelTimer kt(h, PM_K_ROW_R2C);
        PM_HIP(h, pmk::row_r2c(n, d, 1, h->tw, h->plan, h->stream));
        h->field_parity = h->rho_cur;
    }
    {
        KernelTimer kt(h, PM_K_COL_CONV);
        PM_HIP(h, pmk::col_conv(n, d, 1, h->tw, h->plan, spectral, h->stream));
    }
    {
        KernelTimer kt(h, PM_K_ROW_C2R);
        PM_HIP(h, pmk::row_c2r(n, d, 1, h->tw, h->plan, spectral, h->stream));
    }
    if (spectral) {
        if (!pmk::c2r_does_edges(h->plan)) {
            KernelTimer kt(h, PM_K_EDGE);
            PM_HIP(h, pmk::edge_force(n, d, 1, h->stream));
        }
    } else {
        KernelTimer kt(h, PM_K_STENCIL);
        PM_HIP(h, pmk::stencil_force(n, d, 1, h->stream));
    }
    return PM_OK;
}

pm_status stage_push(pm_sim* h) {
    h->scatter_pending = false;          // positions change: the counted bins no longer apply
    KernelTimer kt(h, PM_K_PUSH);
    PM_HIP(h, pmk::push(h->cfg.n_lattice, h->cfg.mode, h->force, h->cur.x, h->cur.y, h->cur.vx, h->cur.vy, h->np,
                        h->cfg.dt, h->cfg.wrap_mode, h->stream));
    h->steps_since_sort += 1;
    h->rho_valid = false;
    return PM_OK;
}

// Step tail with the next deposit fused into the push; the mass lattice of
// the new positions is left in rho_buf[rho_cur] for the next field solve.
bool capturing(const pm_sim* h) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(h->stream, &st) != hipSuccess || st != hipStreamCaptureStatusNone;
}

// Reads the migrant count copied after an earlier push if it has landed (no
// wait) and, when the per-step rate is above the threshold, re-sorts: with
// the fused re-sort by starting its cycle at this push (count now, write the
// new order in the next step), else with the separate sort right away.
// For large simulations the host waits once the copy is two pushes old, so
// the decision is never staler than that (a step there is far longer than
// the launch latency, so the device queue does not drain).
pm_status adaptive_resort(pm_sim* h, bool fused) {
    if (!h->mig_inflight) return PM_OK;
    if (h->np >= (1 << 20) && h->pushes - h->pushes_at_copy >= 2) {
        PM_HIP(h, hipEventSynchronize(h->mig_event));
    } else if (hipEventQuery(h->mig_event) != hipSuccess) {
        (void)hipGetLastError();      // "not ready" is not an error for the launches that follow
        return PM_OK;
    }
    h->mig_inflight = false;
    const unsigned long long v = *h->mig_host;
    const int64_t steps = h->pushes_at_copy - h->pushes_at_read;
    const double per_step = steps > 0 ? (double)(v - h->mig_last) / (double)steps : 0.0;
    h->mig_last = v;
    h->pushes_at_read = h->pushes_at_copy;
    if (h->pushes_at_copy > h->cyc_begin) h->cyc_peak = std::max(h->cyc_peak, per_step);
    else if (h->pushes_at_copy > h->prev_begin) h->prev_peak = std::max(h->prev_peak, per_step);
    if (h->steps_since_sort > 0 && per_step * 1000.0 > (double)h->resort_permille * (double)h->np) {
        h->adaptive_sorts += 1;
        if (fused) {
            h->resort_request = true;
            return PM_OK;
        }
        return sort_now(h, false);
    }
    return PM_OK;
}

// next_cleared: the field stage just zeroed the deposit target.
pm_status stage_push_deposit(pm_sim* h, bool next_cleared) {
    const int n = h->cfg.n_lattice;
    const bool adaptive = h->resort_permille > 0 && h->mig_host && !capturing(h);
    const bool fused = h->fused_sort && h->resort_interval >= 3 && pmk::scatter_fits(n, h->bins.tl);
    if (adaptive && !(h->scatter_pending && fused)) {
        pm_status st = adaptive_resort(h, fused);
        if (st != PM_OK) return st;
    }
    // Fused re-sort (interval >= 3): the step where steps_since_sort + 2 reaches
    // the interval (or an adaptive request is pending) counts the new bins
    // into the second bin set; the next step writes the particles into the
    // other buffers in that order, so the cycle is `interval` steps with no
    // separate sort pass.  Otherwise the count step is followed by the
    // separate counting sort.
    if (h->scatter_pending && fused) {
        if (!next_cleared) {
            KernelTimer kt(h, PM_K_CLEAR);
            PM_HIP(h, pmk::fill_zero(h->rho_buf[h->rho_cur ^ 1], (size_t)8 * n * n, h->stream));
        }
        {
            KernelTimer kt(h, PM_K_PUSH_DEPOSIT_SORT);
            PM_HIP(h, pmk::push_deposit_scatter(n, h->cfg.mode, h->cfg.wrap_mode, h->force, h->cur, h->alt, h->bins,
                                                next_bins(h), h->max_chunks, h->cfg.dt, std::ldexp(1.0f, h->shift),
                                                h->rho_buf[h->rho_cur ^ 1], h->migrants, h->stream));
        }
        std::swap(h->cur, h->alt);
        std::swap(h->bins.count, h->nx_count);
        std::swap(h->bins.start, h->nx_start);
        std::swap(h->bins.chunk_start, h->nx_chunk_start);
        std::swap(h->bins.cursor, h->nx_cursor);
        h->scatter_pending = false;
        h->rho_cur ^= 1;
        h->rho_valid = true;
        h->steps_since_sort = 0;
        h->sorts += 1;
        h->pushes += 1;
        cycle_boundary(h);
        return PM_OK;
    }
    h->scatter_pending = false;
    const bool due = (fused && h->resort_request) || h->steps_since_sort + (fused ? 2 : 1) >= h->resort_interval;
    h->resort_request = false;
    int* count = fused ? h->nx_count : h->bins.count;
    if (!next_cleared || due) {
        KernelTimer kt(h, PM_K_CLEAR);
        if (!next_cleared) PM_HIP(h, pmk::fill_zero(h->rho_buf[h->rho_cur ^ 1], (size_t)8 * n * n, h->stream));
        if (due) PM_HIP(h, pmk::fill_zero(count, sizeof(int) * (size_t)h->nbins, h->stream));
    }
    {
        KernelTimer kt(h, due ? PM_K_PUSH_DEPOSIT_COUNT : PM_K_PUSH_DEPOSIT);
        pmk::BinState bs = h->bins;
        bs.count = count;
        const bool pipelined = h->np < (int64_t)pmk::kPipelineMaxDensity * n * n;
        PM_HIP(h, pmk::push_deposit(n, h->cfg.mode, h->cfg.wrap_mode, h->force, h->cur, bs, h->max_chunks,
                                    h->cfg.dt, std::ldexp(1.0f, h->shift), h->rho_buf[h->rho_cur ^ 1], h->migrants,
                                    due, pipelined, h->stream));
    }
    h->rho_cur ^= 1;
    h->rho_valid = true;
    h->steps_since_sort += 1;
    h->pushes += 1;
    if (due && fused) {
        KernelTimer kt(h, PM_K_SORT);
        PM_HIP(h, pmk::scan_next(n, next_bins(h), h->bins.start, h->stream));
        h->scatter_pending = true;
    }
    if (adaptive && !h->mig_inflight) {
        PM_HIP(h, hipMemcpyAsync(h->mig_host, h->migrants, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                 h->stream));
        PM_HIP(h, hipEventRecord(h->mig_event, h->stream));
        h->mig_inflight = true;
        h->pushes_at_copy = h->pushes;
    }
    if (due && !fused) return sort_now(h, true);
    return PM_OK;
}

// Fixed-point shift from sum|m| (eight interleaved partial sums, added in a
// fixed order: deterministic, and ~4x faster than one dependent chain).
int auto_shift(const float* m, int64_t n) {
    double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = 0;
    for (; i + 8 <= n; i += 8)
        for (int k = 0; k < 8; ++k) p[k] += std::fabs((double)m[i + k]);
    for (; i < n; ++i) p[i & 7] += std::fabs((double)m[i]);
    const double s = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
    if (!(s > 0.0) || !std::isfinite(s)) return 40;
    // cell sums stay below sum|m| * 2^shift + 2n <= 2^61 + 2^35 < 2^63
    int sh = (int)std::floor(61.0 - std::log2(s));
    return std::max(0, std::min(62, sh));
}

pm_status upload(pm_sim* h, const float* x, const float* y, const float* vx, const float* vy, const float* m,
                 int64_t n) {
    const size_t b = (size_t)n * 4;
    if (b) {
        PM_HIP(h, hipMemcpyAsync(h->cur.x, x, b, hipMemcpyHostToDevice, h->stream));
        PM_HIP(h, hipMemcpyAsync(h->cur.y, y, b, hipMemcpyHostToDevice, h->stream));
        PM_HIP(h, hipMemcpyAsync(h->cur.vx, vx, b, hipMemcpyHostToDevice, h->stream));
        PM_HIP(h, hipMemcpyAsync(h->cur.vy, vy, b, hipMemcpyHostToDevice, h->stream));
        PM_HIP(h, hipMemcpyAsync(h->cur.m, m, b, hipMemcpyHostToDevice, h->stream));
    }
    h->np = n;
    h->shift = h->cfg.mass_shift > 0 ? h->cfg.mass_shift : auto_shift(m, n);
    if (h->resort_auto) {
        h->resort_interval = auto_resort_interval(h->cfg.mode, n, h->cfg.n_lattice);
        h->resort_base = h->resort_interval;
    }
    h->cyc_begin = h->prev_begin = h->pushes;
    h->cyc_peak = h->prev_peak = 0.0;
    const int nl = h->cfg.n_lattice;
    const bool fused_dense = n > (int64_t)pmk::kPipelineMaxDensity * nl * nl && h->fused_sort &&
                             h->resort_interval >= 3 && pmk::scatter_fits(nl, h->bins.tl);
    if (h->chunk_auto && pmk::auto_chunk(n, nl, fused_dense) != h->bins.chunk) {
        h->bins.chunk = pmk::auto_chunk(n, nl, fused_dense);
        h->max_chunks = pmk::deposit_max_chunks(h->cfg.n_lattice, h->bins.tl, h->cap, h->bins.chunk);
    }
    h->rho_valid = false;
    h->field_parity = -1;
    // inv_scale follows the shift; uploaded before the sort kernels, so its
    // blocking copy waits for the particle uploads only
    pm_status st = upload_field_desc(h);
    if (st != PM_OK) return st;
    PM_HIP(h, pmk::iota_ids(h->cur.id, n, h->stream));
    return sort_now(h, false);
}

// Stencil symbol d(k) = G (8 sin t - sin 2t)/6, t = 2 pi k/N (exactly 0 at
// k = 0 and k = N/2), and twiddles exp(-2 pi i k/N), both evaluated in FP64.
pm_status upload_tables(pm_sim* h) {
    const int n = h->cfg.n_lattice;
    // staged in a pinned buffer of the handle, so the copies need no host
    // wait; the buffer is rewritten only after the previous copies ran
    if (!h->tables_host) {
        PM_HIP(h, hipHostMalloc(&h->tables_host, 12 * (size_t)n, hipHostMallocDefault));
        PM_HIP(h, hipEventCreateWithFlags(&h->tables_event, hipEventDisableTiming));
    } else {
        PM_HIP(h, hipEventSynchronize(h->tables_event));
    }
    float2* tw = static_cast<float2*>(h->tables_host);
    float* dk = reinterpret_cast<float*>(tw + n);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < n; ++k) {
        const double a = 2.0 * pi * k / n;
        tw[k] = make_float2((float)std::cos(a), (float)-std::sin(a));
        dk[k] = (k == 0 || k == n / 2)
                    ? 0.0f
                    : (float)((double)h->cfg.gravity * (8.0 * std::sin(a) - std::sin(2.0 * a)) / 6.0);
    }
    PM_HIP(h, hipMemcpyAsync(h->tw, tw, 8 * (size_t)n, hipMemcpyHostToDevice, h->stream));
    PM_HIP(h, hipMemcpyAsync(h->dk, dk, 4 * (size_t)n, hipMemcpyHostToDevice, h->stream));
    PM_HIP(h, hipEventRecord(h->tables_event, h->stream));
    return PM_OK;
}

// The FP64 setup work (2 N^2 doubles) lives in the two mass lattices, which
// are contiguous and cleared before every deposit: no allocation, so no
// device-wide synchronisation of a free.  With T_src (a transfer function
// built for the same lattice, mode, G(0) and eps) the spectra are copied.
pm_status build_spectra(pm_sim* h, const float* T_src) {
    const int n = h->cfg.n_lattice;
    hipError_t e;
    if (T_src) {                              // a device copy: stream order suffices, no host wait
        e = hipMemcpyAsync(h->T, T_src, (size_t)4 * (n / 2 + 1) * n, hipMemcpyDeviceToDevice, h->stream);
        h->rho_valid = false;
        return e != hipSuccess ? hip_fail(h, e, "transfer copy") : PM_OK;
    }
    double* work = reinterpret_cast<double*>(h->rho_buf[0]);
    e = pmk::build_transfer(n, h->cfg.mode, h->cfg.green_origin, h->cfg.deconv_eps, work, h->T, h->stream);
    // built spectra may be read on other streams next (a server caches them)
    hipError_t e2 = hipStreamSynchronize(h->stream);
    h->rho_valid = false;
    if (e != hipSuccess) return hip_fail(h, e, "build_transfer");
    if (e2 != hipSuccess) return hip_fail(h, e2, "build_transfer sync");
    return PM_OK;
}

pm_status reconfigure(pm_sim* h, const pm_config* cfg, const float* T_src) {
    std::string why;
    if (validate(cfg, &why) != PM_OK) return fail(h, PM_EINVAL, why);
    if (cfg->n_lattice != h->cfg.n_lattice || cfg->n_particles > h->cap || cfg->device != h->cfg.device)
        return fail(h, PM_EINVAL, "reconfigure: lattice, device or capacity differ");
    const bool spectra = cfg->mode != h->cfg.mode || cfg->green_origin != h->cfg.green_origin ||
                         cfg->deconv_eps != h->cfg.deconv_eps;
    const bool tables = cfg->gravity != h->cfg.gravity;
    if (cfg->precision != h->cfg.precision) h->plan = plan_for(*cfg);
    const int64_t cap = h->cap;
    h->cfg = *cfg;
    h->cfg.n_particles = cap;
    h->resort_interval = cfg->resort_interval > 0 ? cfg->resort_interval : kResortIntervalDefault;
    h->resort_auto = cfg->resort_interval <= 0;
    pm_status st = PM_OK;
    if (spectra && (st = build_spectra(h, T_src)) != PM_OK) return st;
    if (tables && (st = upload_tables(h)) != PM_OK) return st;
    h->np = 0;
    h->rho_valid = false;
    h->field_parity = -1;
    return upload_field_desc(h);
}

pmk::BinState next_bins(const pm_sim* h) {
    pmk::BinState b = h->bins;
    b.count = h->nx_count;
    b.start = h->nx_start;
    b.chunk_start = h->nx_chunk_start;
    b.cursor = h->nx_cursor;
    return b;
}

HostState save_state(const pm_sim* h) {
    return HostState{h->cur,          h->alt,         h->bins,           next_bins(h),        h->rho_cur,
                     h->field_parity, h->steps_since_sort, h->rho_valid, h->scatter_pending, h->sorts};
}

void restore_state(pm_sim* h, const HostState& s) {
    h->cur = s.cur;
    h->alt = s.alt;
    h->bins = s.bins;
    h->nx_count = s.next.count;
    h->nx_start = s.next.start;
    h->nx_chunk_start = s.next.chunk_start;
    h->nx_cursor = s.next.cursor;
    h->rho_cur = s.rho_cur;
    h->field_parity = s.field_parity;
    h->steps_since_sort = s.steps_since_sort;
    h->rho_valid = s.rho_valid;
    h->scatter_pending = s.scatter_pending;
    h->sorts = s.sorts;
}

}  // namespace pmi

using namespace pmi;

extern "C" {

void pm_config_default(pm_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->n_lattice = 2048;       // Constants::N (cuh:38)
    c->mode = PM_CIC;
    c->n_particles = 0;
    c->dt = 50.0f;             // Constants::dt (cuh:44)
    c->gravity = 1.0f;         // Constants::gravityMultiplier (cuh:46)
    c->green_origin = 3.0;     // cuh:113-116
    c->deconv_eps = 1e-3;      // cuh:139
    c->wrap_mode = PM_WRAP_KEEP_VELOCITY;
    c->force_mode = PM_FORCE_SPECTRAL;
    c->precision = PM_FFT_FP32;
    c->mass_shift = 0;
    c->device = 0;
}

int64_t pm_instance_bytes(const pm_config* cfg) {
    std::string why;
    if (validate(cfg, &why) != PM_OK) return -1;
    return layout_of(cfg).total;
}

int pm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"

namespace pmi {

pm_status create(const pm_config* cfg, pm_sim** out, const float* T_src, hipStream_t work_stream) {
    if (!out) return fail(nullptr, PM_EINVAL, "null output handle");
    *out = nullptr;
    std::string why;
    if (validate(cfg, &why) != PM_OK) return fail(nullptr, PM_EINVAL, why);
    int ndev = pm_device_count();
    if (ndev <= 0) return fail(nullptr, PM_ENODEV, "no HIP device visible");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, PM_EINVAL, "device ordinal out of range");
    if (hipSetDevice(cfg->device) != hipSuccess) return fail(nullptr, PM_EDEVICE, "hipSetDevice failed");

    pm_sim* h = new (std::nothrow) pm_sim();
    if (!h) return fail(nullptr, PM_ENOMEM, "host allocation failed");
    h->cfg = *cfg;
    h->cap = cfg->n_particles;
    const Layout L = layout_of(cfg);
    const int n = cfg->n_lattice, nk = n / 2 + 1;
    pm_status st = PM_OK;
    auto bail = [&](pm_status s) { release(h); std::string e = h->err; delete h; g_last_error = e; return s; };
    if (hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking) != hipSuccess) {
        h->own_stream = nullptr;
        return bail(fail(h, PM_EDEVICE, "hipStreamCreate failed"));
    }
    h->stream = work_stream ? work_stream : h->own_stream;
    // every array in one device allocation (one hipMalloc instead of two dozen;
    // the Layout sizes are multiples of 256 bytes)
    void* slab = nullptr;
    if ((st = dalloc(h, &slab, L.total)) != PM_OK) return bail(st);
    char* next = static_cast<char*>(slab);
    auto carve = [&](int64_t bytes) { char* q = next; next += bytes; return static_cast<void*>(q); };
    void* p;
    float** parts[10] = {&h->cur.x, &h->cur.y, &h->cur.vx, &h->cur.vy, &h->cur.m,
                         &h->alt.x, &h->alt.y, &h->alt.vx, &h->alt.vy, &h->alt.m};
    for (auto q : parts) *q = static_cast<float*>(carve(L.particles));
    h->cur.id = static_cast<uint32_t*>(carve(L.particles));
    h->alt.id = static_cast<uint32_t*>(carve(L.particles));
    h->bins.chunk = pmk::kDepositChunkDefault;
    h->bins.tl = pmk::tile_log_for(n);
    h->nbins = pmk::bin_count_of(n, h->bins.tl);
    h->nbins_alloc = h->nbins;
    h->max_chunks = pmk::deposit_max_chunks(n, h->bins.tl, h->cap, h->bins.chunk);
    h->resort_interval = cfg->resort_interval > 0 ? cfg->resort_interval : kResortIntervalDefault;
    h->resort_auto = cfg->resort_interval <= 0;
    p = carve(L.bins);
    {
        char* bp = static_cast<char*>(p);
        const int64_t ib = align256(4 * (int64_t)(h->nbins + 1));
        h->bins.count = reinterpret_cast<int*>(bp);
        h->bins.chunk_start = reinterpret_cast<int*>(bp + ib);
        h->bins.cursor = reinterpret_cast<int*>(bp + 2 * ib);
        h->bins.start = reinterpret_cast<int64_t*>(bp + 3 * ib);
        h->migrants = reinterpret_cast<unsigned long long*>(bp + 3 * ib + align256(8 * (int64_t)(h->nbins + 2)));
    }
    p = carve(L.bins);
    {
        char* bp = static_cast<char*>(p);
        const int64_t ib = align256(4 * (int64_t)(h->nbins + 1));
        h->nx_count = reinterpret_cast<int*>(bp);
        h->nx_chunk_start = reinterpret_cast<int*>(bp + ib);
        h->nx_cursor = reinterpret_cast<int*>(bp + 2 * ib);
        h->nx_start = reinterpret_cast<int64_t*>(bp + 3 * ib);
        if (pmk::fill_zero(p, (size_t)L.bins, h->stream) != hipSuccess) return bail(fail(h, PM_EDEVICE, "clear"));
    }
    // both mass lattices in one block: 2 * 8 N^2 bytes, also the FP64 setup work area
    h->rho_buf[0] = static_cast<int64_t*>(carve(2 * L.rho));
    h->rho_buf[1] = h->rho_buf[0] + L.rho / 8;
    h->fdesc = static_cast<pmk::FieldDesc*>(carve(L.fdesc));
    h->plan = plan_for(*cfg);
    h->H = static_cast<float2*>(carve(L.H));
    h->Hfy = static_cast<float2*>(carve(L.H));
    h->T = static_cast<float*>(carve(L.T));
    h->dk = static_cast<float*>(carve(L.dk));
    h->tw = static_cast<float2*>(carve(L.tw));
    h->phi = static_cast<float*>(carve(L.phi));
    h->force = static_cast<float2*>(carve(L.force));
    h->diag = static_cast<double*>(carve(L.diag));
    if (next - static_cast<char*>(slab) != L.total) return bail(fail(h, PM_EDEVICE, "layout mismatch"));

    if (hipHostMalloc(reinterpret_cast<void**>(&h->mig_host), sizeof(unsigned long long), hipHostMallocDefault) !=
            hipSuccess ||
        hipEventCreateWithFlags(&h->mig_event, hipEventDisableTiming) != hipSuccess)
        return bail(fail(h, PM_ENOMEM, "pinned migrant counter"));
    *h->mig_host = 0;
    if ((st = upload_tables(h)) != PM_OK) return bail(st);
    if ((st = build_spectra(h, T_src)) != PM_OK) return bail(st);
    if ((st = upload_field_desc(h)) != PM_OK) return bail(st);
    // deterministic initial lattices (force/potential read before any step are
    // zero).  When the set-up ran on a caller's stream (a server's creation
    // stream), the handle's own stream is ordered after it by an event: no
    // host wait, so creating a simulation does not queue behind other work.
    if (pmk::fill_zero(h->phi, (size_t)4 * n * n, h->stream) != hipSuccess ||
        pmk::fill_zero(h->force, (size_t)8 * n * n, h->stream) != hipSuccess ||
        pmk::fill_zero(h->rho_buf[0], (size_t)8 * n * n, h->stream) != hipSuccess ||
        pmk::fill_zero(h->H, (size_t)8 * nk * n, h->stream) != hipSuccess ||
        pmk::fill_zero(h->bins.count, (size_t)L.bins, h->stream) != hipSuccess)
        return bail(fail(h, PM_EDEVICE, "lattice clear failed"));
    if (h->stream != h->own_stream &&
        (hipEventRecord(h->tables_event, h->stream) != hipSuccess ||
         hipStreamWaitEvent(h->own_stream, h->tables_event, 0) != hipSuccess))
        return bail(fail(h, PM_EDEVICE, "set-up ordering failed"));
    *out = h;
    return PM_OK;
}

}  // namespace pmi

extern "C" {

pm_status pm_create(const pm_config* cfg, pm_handle* out) { return pmi::create(cfg, out, nullptr, nullptr); }

pm_status pm_destroy(pm_handle h) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    (void)hipSetDevice(h->cfg.device);
    release(h);
    delete h;
    return PM_OK;
}

pm_status pm_load(pm_handle h, const float* x, const float* y, const float* vx, const float* vy, const float* m,
                  int64_t n) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    if (n < 0 || n > h->cap) return fail(h, PM_EINVAL, "particle count exceeds capacity");
    if (n > 0 && (!x || !y || !vx || !vy || !m)) return fail(h, PM_EINVAL, "null particle array");
    PM_HIP(h, hipSetDevice(h->cfg.device));
    return upload(h, x, y, vx, vy, m, n);
}

pm_status pm_load_verlet(pm_handle h, const float* x, const float* y, const float* x_old, const float* y_old,
                         const float* m, int64_t n) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    if (n < 0 || n > h->cap) return fail(h, PM_EINVAL, "particle count exceeds capacity");
    if (n > 0 && (!x || !y || !x_old || !y_old || !m)) return fail(h, PM_EINVAL, "null particle array");
    if (h->cfg.dt == 0.0f) return fail(h, PM_EINVAL, "Verlet state needs dt != 0");
    std::vector<float> vx((size_t)n), vy((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        vx[i] = (x[i] - x_old[i]) / h->cfg.dt;
        vy[i] = (y[i] - y_old[i]) / h->cfg.dt;
    }
    PM_HIP(h, hipSetDevice(h->cfg.device));
    return upload(h, x, y, vx.data(), vy.data(), m, n);
}

pm_status pm_stage(pm_handle h, int32_t stage) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    PM_HIP(h, hipSetDevice(h->cfg.device));
    switch (stage) {
        case PM_STAGE_DEPOSIT: return stage_deposit(h);
        case PM_STAGE_FIELD: return stage_field(h);
        case PM_STAGE_PUSH: return stage_push(h);
        default: return fail(h, PM_EINVAL, "unknown stage");
    }
}

pm_status pm_step(pm_handle h, int32_t nsteps) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    if (nsteps < 0) return fail(h, PM_EINVAL, "nsteps < 0");
    PM_HIP(h, hipSetDevice(h->cfg.device));
    for (int32_t s = 0; s < nsteps; ++s) {
        pm_status st;
        if (!h->rho_valid && (st = stage_deposit(h)) != PM_OK) return st;
        if ((st = stage_field(h, true)) != PM_OK) return st;
        if ((st = stage_push_deposit(h, true)) != PM_OK) return st;
    }
    return PM_OK;
}

pm_status pm_sync(pm_handle h) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    PM_HIP(h, hipSetDevice(h->cfg.device));
    PM_HIP(h, hipStreamSynchronize(h->stream));
    return PM_OK;
}

pm_status pm_read(pm_handle h, float* x, float* y, float* vx, float* vy) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    PM_HIP(h, hipSetDevice(h->cfg.device));
    const pm_status st = pmi::read_async(h, x, y, vx, vy);
    if (st != PM_OK) return st;
    PM_HIP(h, hipStreamSynchronize(h->stream));
    return PM_OK;
}

}  // extern "C"

pm_status pmi::read_async(pm_sim* h, float* x, float* y, float* vx, float* vy) {
    const size_t b = (size_t)h->np * 4;
    if (b) {
        // slots hold particles in bin order; scatter them back to load order in the scratch set
        PM_HIP(h, pmk::unpermute(h->cur.id, h->np, h->cur.x, h->cur.y, h->cur.vx, h->cur.vy, x ? h->alt.x : nullptr,
                                 y ? h->alt.y : nullptr, vx ? h->alt.vx : nullptr, vy ? h->alt.vy : nullptr,
                                 h->stream));
        if (x) PM_HIP(h, hipMemcpyAsync(x, h->alt.x, b, hipMemcpyDeviceToHost, h->stream));
        if (y) PM_HIP(h, hipMemcpyAsync(y, h->alt.y, b, hipMemcpyDeviceToHost, h->stream));
        if (vx) PM_HIP(h, hipMemcpyAsync(vx, h->alt.vx, b, hipMemcpyDeviceToHost, h->stream));
        if (vy) PM_HIP(h, hipMemcpyAsync(vy, h->alt.vy, b, hipMemcpyDeviceToHost, h->stream));
    }
    return PM_OK;
}

extern "C" {

pm_status pm_read_verlet(pm_handle h, float* x, float* y, float* x_old, float* y_old) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    std::vector<float> vx((size_t)h->np), vy((size_t)h->np), px((size_t)h->np), py((size_t)h->np);
    pm_status st = pm_read(h, px.data(), py.data(), vx.data(), vy.data());
    if (st != PM_OK) return st;
    for (int64_t i = 0; i < h->np; ++i) {
        if (x) x[i] = px[i];
        if (y) y[i] = py[i];
        if (x_old) x_old[i] = px[i] - vx[i] * h->cfg.dt;
        if (y_old) y_old[i] = py[i] - vy[i] * h->cfg.dt;
    }
    return PM_OK;
}

pm_status pm_read_lattice(pm_handle h, int32_t which, void* dst, size_t bytes) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    if (!dst) return fail(h, PM_EINVAL, "null destination");
    const size_t n = (size_t)h->cfg.n_lattice, nk = n / 2 + 1;
    const void* src = nullptr;
    size_t need = 0;
    switch (which) {
        case PM_LATTICE_MASS: src = h->rho_buf[h->rho_cur]; need = 8 * n * n; break;
        case PM_LATTICE_POTENTIAL: src = h->phi; need = 4 * n * n; break;
        case PM_LATTICE_FORCE: src = h->force; need = 8 * n * n; break;
        case PM_LATTICE_TRANSFER: src = h->T; need = 4 * nk * n; break;
        default: return fail(h, PM_EINVAL, "unknown lattice");
    }
    if (bytes < need) return fail(h, PM_EINVAL, "destination too small");
    PM_HIP(h, hipSetDevice(h->cfg.device));
    PM_HIP(h, hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, h->stream));
    PM_HIP(h, hipStreamSynchronize(h->stream));
    return PM_OK;
}

pm_status pm_frame(pm_handle h, int32_t blur_r, int32_t reset_extremes, float* frame) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    if (!frame) return fail(h, PM_EINVAL, "null frame buffer");
    if (blur_r < 1 || blur_r > 15 || (blur_r & 1) == 0) return fail(h, PM_EINVAL, "blur_r must be odd in [1, 15]");
    PM_HIP(h, hipSetDevice(h->cfg.device));
    const int n = h->cfg.n_lattice;
    const size_t cells = (size_t)n * n;
    if (!h->frame) {
        void* p = nullptr;
        pm_status st = dalloc(h, &p, align256(4 * (int64_t)cells) + 256);
        if (st != PM_OK) return st;
        h->frame = static_cast<float*>(p);
    }
    if (!h->rho_valid) {
        pm_status st = stage_deposit(h);
        if (st != PM_OK) return st;
    }
    auto* tail = reinterpret_cast<char*>(h->frame) + align256(4 * (int64_t)cells);
    PM_HIP(h, pmk::frame(n, h->rho_buf[h->rho_cur], std::ldexp(1.0f, -h->shift), blur_r,
                         reset_extremes != 0 || !h->frame_started, h->frame, reinterpret_cast<unsigned*>(tail),
                         reinterpret_cast<float*>(tail + 16), h->stream));
    h->frame_started = true;
    PM_HIP(h, hipMemcpyAsync(frame, h->frame, 4 * cells, hipMemcpyDeviceToHost, h->stream));
    PM_HIP(h, hipStreamSynchronize(h->stream));
    return PM_OK;
}

pm_status pm_diagnostics(pm_handle h, double out[8]) {
    if (!h || !out) return fail(h, PM_EINVAL, "null argument");
    PM_HIP(h, hipSetDevice(h->cfg.device));
    const int n = h->cfg.n_lattice;
    const int pb = pmk::diag_blocks(std::max<int64_t>(h->np, 1));
    std::vector<double> part((size_t)pb * 8, 0.0);
    if (h->np > 0) {
        PM_HIP(h, pmk::diag_particles(n, h->cur.x, h->cur.y, h->cur.vx, h->cur.vy, h->cur.m, h->np, h->diag,
                                      h->stream));
        PM_HIP(h, hipMemcpyAsync(part.data(), h->diag, part.size() * 8, hipMemcpyDeviceToHost, h->stream));
    }
    PM_HIP(h, hipStreamSynchronize(h->stream));
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < (h->np > 0 ? pb : 0); ++b) {
        for (int q = 0; q < 7; ++q) {
            if (q == 4) s[4] = std::max(s[4], part[b * 8 + 4]);
            else s[q] += part[b * 8 + q];
        }
    }
    const int lb = pmk::diag_lattice_blocks(n);
    std::vector<double> lp((size_t)lb, 0.0);
    double pe = 0.0;
    if (h->field_parity >= 0) {
        PM_HIP(h, pmk::diag_lattice(n, h->rho_buf[h->field_parity], std::ldexp(1.0, -h->shift), h->phi, h->diag,
                                    h->stream));
        PM_HIP(h, hipMemcpyAsync(lp.data(), h->diag, lp.size() * 8, hipMemcpyDeviceToHost, h->stream));
        PM_HIP(h, hipStreamSynchronize(h->stream));
        for (double v : lp) pe += v;
    }
    out[0] = s[0]; out[1] = s[1]; out[2] = s[2]; out[3] = s[3];
    out[4] = 0.5 * pe; out[5] = s[4]; out[6] = s[5]; out[7] = s[6];
    return PM_OK;
}

pm_status pm_set_param(pm_handle h, const char* name, int64_t value) {
    if (!h || !name) return fail(h, PM_EINVAL, "null argument");
    const std::string k(name);
    if (k == "resort_interval") {
        if (value < 1 || value > (1 << 20)) return fail(h, PM_EINVAL, "resort_interval must be in [1, 2^20]");
        h->resort_interval = (int)value;
        h->resort_auto = false;
        return PM_OK;
    }
    if (k == "fused_sort") {
        if (value != 0 && value != 1) return fail(h, PM_EINVAL, "fused_sort must be 0 or 1");
        h->fused_sort = (int)value;
        PM_HIP(h, hipSetDevice(h->cfg.device));
        return sort_now(h, false);
    }
    if (k == "bin_max_chunks") {
        if (value < 0 || value > (1 << 20)) return fail(h, PM_EINVAL, "bin_max_chunks must be in [0, 2^20] (0 = no cap)");
        h->bins.max_chunks_per_bin = (int)value;
        PM_HIP(h, hipSetDevice(h->cfg.device));
        return sort_now(h, false);
    }
    if (k == "profile_mask") {
        h->prof_mask = value;
        return PM_OK;
    }
    if (k == "resort_migrants_permille") {
        if (value < 0 || value > 1000) return fail(h, PM_EINVAL, "resort_migrants_permille must be in [0, 1000]");
        h->resort_permille = (int)value;
        return PM_OK;
    }
    if (k == "resort_drift_ppm") {
        if (value < 0 || value > 1000000) return fail(h, PM_EINVAL, "resort_drift_ppm must be in [0, 1e6]");
        h->drift_ppm = (int)value;
        return PM_OK;
    }
    if (k == "deposit_chunk") {
        if (value < pmk::kDepositBlock || value > (1 << 24) || (value % pmk::kDepositBlock) != 0)
            return fail(h, PM_EINVAL, "deposit_chunk must be a multiple of 512 in [512, 2^24]");
        // chunk starts depend on the chunk size: re-bin now so every chunk-mapped
        // kernel (deposit, fused push + deposit) sees a consistent map
        h->bins.chunk = (int)value;
        h->chunk_auto = false;
        h->max_chunks = pmk::deposit_max_chunks(h->cfg.n_lattice, h->bins.tl, h->cap, h->bins.chunk);
        PM_HIP(h, hipSetDevice(h->cfg.device));
        return sort_now(h, false);
    }
    if (k == "tile_log") {
        if (value != 5 && value != 6) return fail(h, PM_EINVAL, "tile_log must be 5 or 6");
        if (pmk::bin_count_of(h->cfg.n_lattice, (int)value) > h->nbins_alloc)
            return fail(h, PM_EINVAL, "tile_log too small for the bin arrays of this lattice");
        h->bins.tl = (int)value;
        h->nbins = pmk::bin_count_of(h->cfg.n_lattice, h->bins.tl);
        h->max_chunks = pmk::deposit_max_chunks(h->cfg.n_lattice, h->bins.tl, h->cap, h->bins.chunk);
        PM_HIP(h, hipSetDevice(h->cfg.device));
        return sort_now(h, false);
    }
    if (k == "fft_rp" || k == "fft_cb" || k == "fft_radix" || k == "fft_it") {
        pmk::FftPlan p = h->plan;
        (k == "fft_rp" ? p.rp : k == "fft_cb" ? p.cb : k == "fft_radix" ? p.radix : p.it) = (int)value;
        // the prefetch depth changes no result: fall back to 1 where the new
        // batch or radix has no prefetching variant
        if (k != "fft_it" && p.it != 1 && !pmk::plan_valid(h->cfg.n_lattice, p)) p.it = 1;
        if (!pmk::plan_valid(h->cfg.n_lattice, p)) return fail(h, PM_EINVAL, "FFT plan not available for this lattice");
        if (p.radix != plan_for(h->cfg).radix && ((p.radix >= 16 && p.radix <= 17) || h->cfg.precision != PM_FFT_FP32))
            return fail(h, PM_EINVAL,
                        "fft_radix 16 is the PM_FFT_BF16X3 precision, 17 PM_FFT_BF16X6; 4, 8 or 64 are FP32");
        h->plan = p;
        return PM_OK;
    }
    return fail(h, PM_EINVAL, "unknown parameter: " + k);
}

pm_status pm_get_param(pm_handle h, const char* name, int64_t* value) {
    if (!h || !name || !value) return fail(h, PM_EINVAL, "null argument");
    const std::string k(name);
    if (k == "resort_interval") *value = h->resort_interval;
    else if (k == "resort_migrants_permille") *value = h->resort_permille;
    else if (k == "resort_drift_ppm") *value = h->drift_ppm;
    else if (k == "interval_changes") *value = h->interval_changes;
    else if (k == "bin_max_chunks") *value = h->bins.max_chunks_per_bin;
    else if (k == "fused_sort") *value = h->fused_sort;
    else if (k == "adaptive_sorts") *value = h->adaptive_sorts;
    else if (k == "deposit_chunk") *value = h->bins.chunk;
    else if (k == "tile_log") *value = h->bins.tl;
    else if (k == "fft_rp") *value = h->plan.rp;
    else if (k == "fft_cb") *value = h->plan.cb;
    else if (k == "fft_radix") *value = h->plan.radix;
    else if (k == "fft_it") *value = h->plan.it;
    else if (k == "profile_mask") *value = h->prof_mask;
    else return fail(h, PM_EINVAL, "unknown parameter: " + k);
    return PM_OK;
}

}  // extern "C"

namespace {

// Device time of `reps` calls of fn on the handle's stream (HIP events).
template <typename F>
pm_status time_on_stream(pm_sim* h, int reps, F fn, float* ms) {
    hipEvent_t a = nullptr, b = nullptr;
    PM_HIP(h, hipEventCreate(&a));
    PM_HIP(h, hipEventCreate(&b));
    pm_status st = PM_OK;
    if ((st = fn()) == PM_OK) {                        // warm-up (first-touch, code load)
        (void)hipEventRecord(a, h->stream);
        for (int r = 0; r < reps && st == PM_OK; ++r) st = fn();
        (void)hipEventRecord(b, h->stream);
        if (st == PM_OK && hipEventSynchronize(b) == hipSuccess) (void)hipEventElapsedTime(ms, a, b);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return st;
}

}  // namespace

pm_status pm_autotune(pm_handle h, int32_t flags, int32_t reps) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    if (reps <= 0) reps = 5;
    PM_HIP(h, hipSetDevice(h->cfg.device));
    const bool prof = h->profiling;
    h->profiling = false;
    pm_status st = PM_OK;
    const int n = h->cfg.n_lattice;
    if ((flags & PM_TUNE_FFT) && h->cfg.precision == PM_FFT_FP32) {   // BF16X3 has one plan
        // Field solves only read the mass lattice: repeating them changes no state.
        if (!h->rho_valid && (st = stage_deposit(h)) != PM_OK) { h->profiling = prof; return st; }
        pmk::FftPlan best = h->plan;
        float best_ms = 1e30f;
        for (int radix : {4, 8, pmk::kWaveRadix})
            for (int rp : {1, 2, 4})
                for (int cb : {1, 2, 4, 8})
                for (int it : {1, 2, 4}) {
                    pmk::FftPlan p;
                    p.rp = rp; p.cb = cb; p.radix = radix; p.it = it;
                    if (!pmk::plan_valid(n, p)) continue;
                    h->plan = p;
                    float ms = 0.f;
                    if ((st = time_on_stream(h, reps, [&] { return stage_field(h); }, &ms)) != PM_OK) break;
                    if (ms < best_ms) { best_ms = ms; best = p; }
                }
        h->plan = best;
        if (st != PM_OK) { h->profiling = prof; return st; }
        if ((st = stage_field(h)) != PM_OK) { h->profiling = prof; return st; }
    }
    if ((flags & PM_TUNE_PARTICLES) && h->np > 0) {
        // Particle trials advance the state: snapshot it and restore it after.
        const size_t pb = (size_t)h->np * 4, lb = (size_t)8 * n * n;
        char* snap = nullptr;
        hipError_t e = hipMalloc(&snap, 6 * pb + 2 * lb);
        if (e != hipSuccess) { h->profiling = prof; return hip_fail(h, e, "hipMalloc(autotune snapshot)"); }
        void* src[6] = {h->cur.x, h->cur.y, h->cur.vx, h->cur.vy, h->cur.m, h->cur.id};
        for (int a = 0; a < 6; ++a)
            (void)hipMemcpyAsync(snap + a * pb, src[a], pb, hipMemcpyDeviceToDevice, h->stream);
        (void)hipMemcpyAsync(snap + 6 * pb, h->rho_buf[0], lb, hipMemcpyDeviceToDevice, h->stream);
        (void)hipMemcpyAsync(snap + 6 * pb + lb, h->rho_buf[1], lb, hipMemcpyDeviceToDevice, h->stream);
        const pmk::ParticleArrays cur0 = h->cur, alt0 = h->alt;
        const int rho_cur0 = h->rho_cur, since0 = h->steps_since_sort, fpar0 = h->field_parity;
        const bool valid0 = h->rho_valid;
        const int64_t sorts0 = h->sorts;
        const int chunk0 = h->bins.chunk, tl0 = h->bins.tl, resort0 = h->resort_interval;
        int best_chunk = chunk0, best_tl = tl0, best_resort = resort0;
        float best_rate = 1e30f;
        const int tl_default = pmk::tile_log_for(n);
        for (int tl : {5, 6}) {
            if (tl < tl_default) continue;
            std::vector<int> chunks = {8192, 12288, 16384, 32768, 65536, 131072};
            if (h->np < pmk::kSmallRunParticles) chunks.insert(chunks.begin(), {2048, 4096});
            for (int chunk : chunks) {
                for (int resort : {4, 8, 12}) {
                    h->bins.tl = tl;
                    h->nbins = pmk::bin_count_of(n, tl);
                    h->bins.chunk = chunk;
                    h->max_chunks = pmk::deposit_max_chunks(n, tl, h->cap, chunk);
                    h->resort_interval = resort;
                    if ((st = sort_now(h, false)) != PM_OK) break;
                    if (!h->rho_valid && (st = stage_deposit(h)) != PM_OK) break;
                    float ms = 0.f;
                    st = time_on_stream(h, resort, [&] {
                        pm_status s1 = stage_field(h);
                        return s1 != PM_OK ? s1 : stage_push_deposit(h);
                    }, &ms);
                    if (st != PM_OK) break;
                    const float per_step = ms / (float)resort;
                    if (per_step < best_rate) { best_rate = per_step; best_chunk = chunk; best_tl = tl; best_resort = resort; }
                }
                if (st != PM_OK) break;
            }
            if (st != PM_OK) break;
        }
        // restore the state exactly as it was
        h->cur = cur0;
        h->alt = alt0;
        void* dst[6] = {h->cur.x, h->cur.y, h->cur.vx, h->cur.vy, h->cur.m, h->cur.id};
        for (int a = 0; a < 6; ++a)
            (void)hipMemcpyAsync(dst[a], snap + a * pb, pb, hipMemcpyDeviceToDevice, h->stream);
        (void)hipMemcpyAsync(h->rho_buf[0], snap + 6 * pb, lb, hipMemcpyDeviceToDevice, h->stream);
        (void)hipMemcpyAsync(h->rho_buf[1], snap + 6 * pb + lb, lb, hipMemcpyDeviceToDevice, h->stream);
        (void)hipStreamSynchronize(h->stream);
        (void)hipFree(snap);
        h->rho_cur = rho_cur0;
        h->steps_since_sort = since0;
        h->field_parity = fpar0;
        h->rho_valid = valid0;
        h->sorts = sorts0;
        h->bins.tl = best_tl;
        h->nbins = pmk::bin_count_of(n, best_tl);
        h->bins.chunk = best_chunk;
        h->chunk_auto = false;
        h->max_chunks = pmk::deposit_max_chunks(n, best_tl, h->cap, best_chunk);
        h->resort_interval = best_resort;
        h->resort_auto = false;
        // the bin tables must match the restored particle order and the chosen tiles
        if (st == PM_OK) st = sort_now(h, false);
        // potential and force as the last field solve left them
        if (st == PM_OK && fpar0 >= 0) {
            h->rho_cur = fpar0;
            st = stage_field(h);
            h->rho_cur = rho_cur0;
        }
    }
    h->profiling = prof;
    if (st == PM_OK) PM_HIP(h, hipStreamSynchronize(h->stream));
    return st;
}

extern "C" {


pm_status pm_get_info(pm_handle h, pm_info* info) {
    if (!h || !info) return fail(h, PM_EINVAL, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->n_lattice = h->cfg.n_lattice;
    info->mass_shift = h->shift;
    info->n_particles = h->np;
    info->capacity = h->cap;
    info->device_bytes = h->bytes;
    info->device = h->cfg.device;
    info->sorts = h->sorts;
    unsigned long long mig = 0;
    PM_HIP(h, hipSetDevice(h->cfg.device));
    PM_HIP(h, hipMemcpyAsync(&mig, h->migrants, sizeof(mig), hipMemcpyDeviceToHost, h->stream));
    PM_HIP(h, hipStreamSynchronize(h->stream));
    info->migrants = (int64_t)mig;
    return PM_OK;
}

pm_status pm_profile(pm_handle h, int32_t enable) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    h->profiling = enable != 0;
    return PM_OK;
}

pm_status pm_profile_read(pm_handle h, int32_t kernel, double* total_ms, int64_t* launches) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    if (kernel < 0 || kernel >= PM_K_COUNT) return fail(h, PM_EINVAL, "unknown kernel id");
    PM_HIP(h, hipSetDevice(h->cfg.device));
    pm_status st = resolve_profile(h);
    if (st != PM_OK) return st;
    if (total_ms) *total_ms = h->prof_ms[kernel];
    if (launches) *launches = h->prof_n[kernel];
    return PM_OK;
}

pm_status pm_profile_reset(pm_handle h) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    pm_status st = resolve_profile(h);
    if (st != PM_OK) return st;
    for (int k = 0; k < PM_K_COUNT; ++k) { h->prof_ms[k] = 0.0; h->prof_n[k] = 0; }
    return PM_OK;
}

void* pm_stream(pm_handle h) { return h ? (void*)h->stream : nullptr; }

const char* pm_last_error(pm_handle h) { return h ? h->err.c_str() : g_last_error.c_str(); }

const char* pm_status_string(pm_status s) {
    switch (s) {
        case PM_OK: return "ok";
        case PM_EINVAL: return "invalid argument";
        case PM_ENOMEM: return "out of device memory";
        case PM_EDEVICE: return "device error";
        case PM_EBUSY: return "busy";
        case PM_ECANCELLED: return "cancelled";
        case PM_ENODEV: return "no device";
    }
    return "unknown status";
}

}  // extern "C"

// ---- checkpoints -----------------------------------------------------------------
namespace {

struct CkptHeader {
    char magic[8];
    uint32_t header_bytes, version;
    pm_config cfg;
    int32_t mass_shift, fft_rp, fft_cb, fft_radix;
    int32_t resort_interval, resort_permille, deposit_chunk, tile_log;
    int64_t n_particles, steps;
    uint64_t checksum;
};
constexpr char kCkptMagic[8] = {'P', 'M', 'C', 'K', 'P', 'T', 0, 1};

uint64_t fnv1a(const unsigned char* p, size_t n, uint64_t hsh = 1469598103934665603ull) {
    for (size_t i = 0; i < n; ++i) hsh = (hsh ^ p[i]) * 1099511628211ull;
    return hsh;
}

}  // namespace

extern "C" {

int64_t pm_checkpoint_bytes(pm_handle h) {
    if (!h) return -1;
    return (int64_t)sizeof(CkptHeader) + 20 * h->np;
}

pm_status pm_checkpoint_write(pm_handle h, void* dst, int64_t bytes) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    if (!dst || bytes < pm_checkpoint_bytes(h)) return fail(h, PM_EINVAL, "checkpoint buffer too small");
    const int64_t n = h->np;
    auto* out = static_cast<unsigned char*>(dst);
    float* x = reinterpret_cast<float*>(out + sizeof(CkptHeader));
    float *y = x + n, *vx = y + n, *vy = vx + n, *m = vy + n;
    pm_status st = pm_read(h, x, y, vx, vy);
    if (st != PM_OK) return st;
    if (n) {   // masses in load order: the slots' load indices scatter the sorted masses back
        std::vector<uint32_t> id((size_t)n);
        std::vector<float> ms((size_t)n);
        PM_HIP(h, hipMemcpyAsync(id.data(), h->cur.id, 4 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
        PM_HIP(h, hipMemcpyAsync(ms.data(), h->cur.m, 4 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
        PM_HIP(h, hipStreamSynchronize(h->stream));
        for (int64_t k = 0; k < n; ++k) m[id[(size_t)k]] = ms[(size_t)k];
    }
    CkptHeader hd;
    std::memset(&hd, 0, sizeof(hd));
    std::memcpy(hd.magic, kCkptMagic, 8);
    hd.header_bytes = sizeof(CkptHeader);
    hd.version = 1;
    hd.cfg = h->cfg;
    hd.mass_shift = h->shift;
    hd.fft_rp = h->plan.rp;
    hd.fft_cb = h->plan.cb;
    hd.fft_radix = h->plan.radix;
    hd.resort_interval = h->resort_interval;
    hd.resort_permille = h->resort_permille;
    hd.deposit_chunk = h->bins.chunk;
    hd.tile_log = h->bins.tl;
    hd.n_particles = n;
    hd.steps = h->pushes;
    hd.checksum = fnv1a(out + sizeof(CkptHeader), 20 * (size_t)n);
    std::memcpy(out, &hd, sizeof(hd));
    return PM_OK;
}

pm_status pm_checkpoint_restore(const void* src, int64_t bytes, int32_t device, pm_handle* out) {
    if (!src || !out) return fail(nullptr, PM_EINVAL, "null argument");
    *out = nullptr;
    if (bytes < (int64_t)sizeof(CkptHeader)) return fail(nullptr, PM_EINVAL, "checkpoint truncated");
    CkptHeader hd;
    std::memcpy(&hd, src, sizeof(hd));
    if (std::memcmp(hd.magic, kCkptMagic, 8) != 0 || hd.version != 1 || hd.header_bytes != sizeof(CkptHeader))
        return fail(nullptr, PM_EINVAL, "not a checkpoint of this format");
    // n comes from the file: bound it before any size arithmetic (20 * n must not overflow)
    const int64_t n = hd.n_particles;
    if (n < 0 || n > (bytes - (int64_t)sizeof(CkptHeader)) / 20)
        return fail(nullptr, PM_EINVAL, "checkpoint truncated");
    const auto* in = static_cast<const unsigned char*>(src);
    if (fnv1a(in + sizeof(CkptHeader), 20 * (size_t)n) != hd.checksum)
        return fail(nullptr, PM_EINVAL, "checkpoint checksum mismatch");
    pm_config c = hd.cfg;
    c.device = device;
    c.mass_shift = hd.mass_shift;            // the same fixed point as when saved
    c.n_particles = std::max<int64_t>(c.n_particles, n);
    pm_sim* h = nullptr;
    pm_status st = pm_create(&c, &h);
    if (st != PM_OK) return st;
    h->plan.rp = hd.fft_rp;
    h->plan.cb = hd.fft_cb;
    h->plan.radix = hd.fft_radix;
    // the prefetch depth is not stored (it changes no result bit): take the
    // default's where the stored plan allows it
    h->plan.it = 1;
    if (hd.fft_rp == 1 && hd.fft_cb == 1) {
        pmk::FftPlan q = h->plan;
        q.it = plan_for(c).it;
        if (pmk::plan_valid(c.n_lattice, q)) h->plan = q;
    }
    if (!pmk::plan_valid(c.n_lattice, h->plan)) h->plan = plan_for(c);
    h->resort_interval = hd.resort_interval > 0 ? hd.resort_interval : kResortIntervalDefault;
    h->resort_auto = false;                 // the interval in force when the state was saved
    h->resort_permille = hd.resort_permille;
    if (hd.deposit_chunk > 0) {
        h->bins.chunk = hd.deposit_chunk;
        h->chunk_auto = false;
    }
    h->max_chunks = pmk::deposit_max_chunks(c.n_lattice, h->bins.tl, h->cap, h->bins.chunk);
    if (hd.tile_log != h->bins.tl && pm_set_param(h, "tile_log", hd.tile_log) != PM_OK) {
        pm_destroy(h);
        return fail(nullptr, PM_EINVAL, "checkpoint tile size not available");
    }
    const float* x = reinterpret_cast<const float*>(in + sizeof(CkptHeader));
    st = pm_load(h, x, x + n, x + 2 * n, x + 3 * n, x + 4 * n, n);
    if (st != PM_OK) {
        std::string e = h->err;
        pm_destroy(h);
        return fail(nullptr, st, e);
    }
    h->pushes = hd.steps;
    *out = h;
    return PM_OK;
}

pm_status pm_checkpoint_save(pm_handle h, const char* path) {
    if (!h) return fail(nullptr, PM_EINVAL, "null handle");
    if (!path) return fail(h, PM_EINVAL, "null path");
    std::vector<unsigned char> buf((size_t)pm_checkpoint_bytes(h));
    pm_status st = pm_checkpoint_write(h, buf.data(), (int64_t)buf.size());
    if (st != PM_OK) return st;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(h, PM_EINVAL, std::string("cannot open ") + path);
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok) return fail(h, PM_EINVAL, std::string("write failed: ") + path);
    return PM_OK;
}

pm_status pm_checkpoint_open(const char* path, int32_t device, pm_handle* out) {
    if (!path || !out) return fail(nullptr, PM_EINVAL, "null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(nullptr, PM_EINVAL, std::string("cannot open ") + path);
    std::vector<unsigned char> buf;
    unsigned char tmp[1 << 16];
    size_t got;
    while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
    std::fclose(f);
    return pm_checkpoint_restore(buf.data(), (int64_t)buf.size(), device, out);
}

}  // extern "C"
