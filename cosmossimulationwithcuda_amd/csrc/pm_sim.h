// Internal state of one simulation handle and the host-side helpers shared by
// the C-ABI layer (pm_api.cpp) and the serving engine (pm_serve.cpp).  Not
// installed; the public interface is include/pm.h.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/pm.h"
#include "pm_internal.h"

namespace pmi {

struct ProfRec {
    int kernel;
    hipEvent_t start, end;
};

}  // namespace pmi

struct pm_sim {
    pm_config cfg{};
    hipStream_t stream = nullptr;          // stream the handle's work is issued on
    hipStream_t own_stream = nullptr;      // created with the handle (a server may issue on its own streams)
    int64_t cap = 0, np = 0;
    int shift = 40;
    pmk::ParticleArrays cur{}, alt{};      // sorted state and the counting-sort target / read scratch
    pmk::BinState bins{};
    // second bin set: filled by the counting step of a fused re-sort while
    // `bins` still maps the current order; swapped in after the re-sort step
    int* nx_count = nullptr;
    int64_t* nx_start = nullptr;
    int* nx_chunk_start = nullptr;
    int* nx_cursor = nullptr;
    bool scatter_pending = false;          // the next push writes the re-sorted order
    bool resort_request = false;           // adaptive re-sort: start the fused cycle at the next push
    int fused_sort = 1;                    // re-sort inside the step kernel when it fits
    bool chunk_auto = true;                // deposit chunk follows the particle count (auto_chunk) at load
    int nbins = 0;                         // in-bounds bins + parked bin
    int nbins_alloc = 0;                   // bins the bin arrays were sized for
    int max_chunks = 0;
    int resort_interval = 9;
    bool resort_auto = true;               // interval chosen at load by auto_resort_interval
    int steps_since_sort = 0;
    // adaptive re-sort: the migrant counter is copied to pinned host memory
    // after a push and read without blocking at a later push; when more than
    // resort_permille / 1000 of the particles left their tile windows per
    // step, the particles are re-sorted before the next push
    int resort_permille = 10;
    unsigned long long* mig_host = nullptr;
    void* tables_host = nullptr;           // pinned staging of the twiddle / stencil tables
    hipEvent_t tables_event = nullptr;     // recorded after their last upload
    hipEvent_t mig_event = nullptr;
    bool mig_inflight = false;
    unsigned long long mig_last = 0;
    int64_t pushes = 0, pushes_at_copy = 0, pushes_at_read = 0;
    int64_t adaptive_sorts = 0;
    // drift-following interval (automatic interval, adaptive re-sort on): the
    // peak migrant rate of each re-sort cycle, its reads attributed to cycles
    // by push index, shortens the interval by one (down to 3) when above
    // drift_ppm particles per million per step
    int resort_base = 9;                   // interval chosen at load
    int drift_ppm = 200;
    int64_t cyc_begin = 0, prev_begin = 0;  // pushes done when the current / previous cycle began
    double cyc_peak = 0.0, prev_peak = 0.0;
    int64_t interval_changes = 0;
    int64_t sorts = 0;
    unsigned long long* migrants = nullptr;
    int64_t* rho_buf[2] = {nullptr, nullptr};
    int rho_cur = 0;                       // rho_buf[rho_cur]: mass lattice of the current positions (when rho_valid)
    int field_parity = -1;                 // rho_buf index the last field solve consumed (potential energy)
    pmk::FieldDesc* fdesc = nullptr;       // device descriptors, [p] reads rho_buf[p]
    pmk::FftPlan plan{};
    bool rho_valid = false;
    float2 *H = nullptr, *Hfy = nullptr;
    float* T = nullptr;
    float* dk = nullptr;
    float2* tw = nullptr;
    float* phi = nullptr;
    float2* force = nullptr;
    double* diag = nullptr;
    float* frame = nullptr;                // density frame work area (allocated by the first pm_frame)
    bool frame_started = false;
    int64_t bytes = 0;
    std::vector<void*> allocs;
    std::string err;
    bool profiling = false;
    int64_t prof_mask = -1;                // kernels timed while profiling (bit = PM_K_* id)
    std::vector<pmi::ProfRec> recs;
    std::vector<hipEvent_t> pool;
    double prof_ms[PM_K_COUNT] = {};
    int64_t prof_n[PM_K_COUNT] = {};
};

namespace pmi {

pm_status fail(pm_sim* h, pm_status st, const std::string& msg);
pm_status hip_fail(pm_sim* h, hipError_t e, const char* what);

#define PM_HIP(h, call)                                        \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return pmi::hip_fail((h), e_, #call); \
    } while (0)

pm_status validate(const pm_config* c, std::string* why);
int64_t align256(int64_t b);
pm_status upload_field_desc(pm_sim* h);
// pm_read without the final wait: the copies into the caller's buffers are
// enqueued on h->stream (complete once the stream is synchronised)
pm_status read_async(pm_sim* h, float* x, float* y, float* vx, float* vy);
pm_status sort_now(pm_sim* h, bool counted);
// Fixed-point shift pm_load picks for these masses (cfg.mass_shift 0).
int auto_shift(const float* m, int64_t n);
pm_status stage_deposit(pm_sim* h);
pm_status stage_field(pm_sim* h, bool clear_next = false);
pm_status stage_push(pm_sim* h);
pm_status stage_push_deposit(pm_sim* h, bool next_cleared = false);
// pm_create with an optional ready transfer function (device, [N/2+1][N]
// floats, built for the same lattice, mode, G(0) and eps) copied instead of
// rebuilt, and an optional stream for the setup work (the handle then issues
// on it; its own stream is still created).
pm_status create(const pm_config* cfg, pm_sim** out, const float* T_src, hipStream_t work_stream);
// Re-targets an existing handle (same lattice size, capacity >= its needs) to
// another configuration: rebuilds (or copies from T_src) the transfer function
// and the stencil symbol when the physics parameters differ.
pm_status reconfigure(pm_sim* h, const pm_config* cfg, const float* T_src);

// Host-side state that pm_step advances (for replaying captured step graphs).
struct HostState {
    pmk::ParticleArrays cur, alt;
    pmk::BinState bins, next;
    int rho_cur, field_parity, steps_since_sort;
    bool rho_valid, scatter_pending;
    int64_t sorts;
};
// The bin set the counting step of a fused re-sort fills (parameters of `bins`).
pmk::BinState next_bins(const pm_sim* h);
HostState save_state(const pm_sim* h);
void restore_state(pm_sim* h, const HostState& s);

}  // namespace pmi
