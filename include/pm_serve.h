/*
 * pm_serve.h -- serving engine over the PM gravity engine (libpm.so).
 *
 * No counterpart in the reference, whose host side runs one simulation on a
 * worker thread per GPU and hands frames to one consumer through a mutex
 * ring (cuh:872-878, 1302-1395).  Here independent simulation requests of
 * different lattice sizes and particle counts are submitted at any time; a
 * worker thread admits them in FIFO order while their device memory fits a
 * budget, keeps their state resident, advances all admitted requests
 * together in rounds of `round_steps` steps, and returns each result when its
 * step count is reached.
 *
 * Modes:
 *   PM_SERVE_PLAIN    one request at a time, each stepped on its own to the
 *                     end (the one-at-a-time loop baseline);
 *   PM_SERVE_BATCHED  many requests resident; simulations with lattices up to
 *                     `batch_max_lattice` are advanced by batched launches (one
 *                     launch per FFT pass and lattice size, one per particle
 *                     kernel and group, for all of them): those with at most
 *                     `batch_max_particles` particles by the direct kernel,
 *                     the larger ones (`batch_tiled`) by the tile-sorted step
 *                     kernels with their re-sort cycle aligned to the round;
 *                     the others on their own streams, concurrently; rounds
 *                     replayed from captured HIP graphs when `use_graphs` is
 *                     set.
 *
 * Results do not depend on the mode or on which other requests share a
 * round: every simulation's arithmetic is the same as alone.
 */
#ifndef COSMOS_PM_SERVE_H
#define COSMOS_PM_SERVE_H

#include "pm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_server* pm_server_handle;

enum { PM_SERVE_PLAIN = 0, PM_SERVE_BATCHED = 1 };
enum { PM_REQ_QUEUED = 0, PM_REQ_ACTIVE = 1, PM_REQ_DONE = 2, PM_REQ_FAILED = 3, PM_REQ_UNKNOWN = 4 };

typedef struct pm_server_config {
    int64_t hbm_budget_bytes;   /* device memory the server may hold (active + pooled simulations) */
    int32_t max_active;         /* admitted requests at most (PLAIN: forced to 1) */
    int32_t mode;               /* PM_SERVE_* */
    int32_t round_steps;        /* steps per round (even; default 8) */
    int32_t use_graphs;         /* replay rounds from captured HIP graphs */
    int32_t batch_max_lattice;  /* lattices up to this size (default 2048) ... */
    int32_t device;
    int32_t batch_max_particles; /* ... with at most this many particles (default 131072) run in
                                  * batched launches (BATCHED) */
    int32_t pinned_staging_mib;  /* pinned host memory (MiB, taken at creation; default 1024) that
                                  * holds submitted arrays until their results are collected:
                                  * uploads and read-backs from it are asynchronous DMA.  A request
                                  * that finds no room uses pageable memory.  0 = none. */
    int32_t persistent;          /* BATCHED: run each round's steps of the members with lattices up
                                  * to 1024 (default FFT plans) in one persistent kernel that pulls
                                  * row, column and particle items from a device-side queue.
                                  * Disabled in this release (the kernel did not complete its first
                                  * rounds on the device; cause not found): must be 0, any other
                                  * value is PM_EINVAL. */
    int32_t batch_tiled;         /* BATCHED (default 0): simulations above batch_max_particles with
                                  * lattices up to batch_max_lattice also run in the batched launches,
                                  * with the tile-sorted step kernels over descriptor tables (their
                                  * re-sort interval is the round length; the request must not set its
                                  * own); 0 = on their own streams */
    int32_t reserved[4];
} pm_server_config;

typedef struct pm_server_stats {
    int64_t submitted, completed, failed, active, queued;
    double particle_steps;           /* completed particle-steps */
    double steps;                    /* completed simulation steps */
    double first_submit_s;           /* wall clock of the first submission (s, steady clock) */
    double last_done_s;              /* wall clock of the last completion */
    double step_latency_p50_ms, step_latency_p99_ms;       /* per simulation step, over all steps */
    double request_latency_p50_ms, request_latency_p99_ms; /* submit -> result */
    int64_t bytes_in_use, peak_bytes;
    int64_t rounds, batched_launch_rounds, graph_launches, graph_captures, pool_hits;
    /* worker-thread wall time by phase (s): admission (allocation, upload, first
     * deposit), launching rounds, reading back results, waiting for the device */
    double admit_s, launch_s, readback_s, wait_s;
    /* admission split: creating simulations, and uploading + first deposit */
    double create_s, upload_s;
} pm_server_stats;

void pm_server_config_default(pm_server_config* cfg);

/* Admission planner used by the server (pure host logic): how many requests
 * from the head of a FIFO queue, needing need_bytes[i] each, are admitted
 * with free_bytes available and at most max_new slots.  Strict FIFO: the
 * first request that does not fit stops admission. */
int32_t pm_plan_admission(const int64_t* need_bytes, int32_t n_queued, int64_t free_bytes, int32_t max_new);

pm_status pm_server_create(const pm_server_config* cfg, pm_server_handle* out);

/* Queue a request; the arrays are copied.  cfg->n_particles is ignored (n
 * is used).  PM_ENOMEM when the request alone exceeds the budget. */
pm_status pm_server_submit(pm_server_handle s, const pm_config* cfg, const float* x, const float* y,
                           const float* vx, const float* vy, const float* m, int64_t n, int32_t steps,
                           int64_t* request_id);

pm_status pm_server_status(pm_server_handle s, int64_t request_id, int32_t* state);

/* Block until the request is done (or timeout_s elapses: PM_EBUSY), copy its
 * final state into the buffers (any may be NULL) and forget it. */
pm_status pm_server_wait(pm_server_handle s, int64_t request_id, float* x, float* y, float* vx, float* vy,
                         double timeout_s);

/* Block until every submitted request is done or failed. */
pm_status pm_server_drain(pm_server_handle s);

pm_status pm_server_get_stats(pm_server_handle s, pm_server_stats* out);

/* Message of the last failure seen by pm_server_submit / pm_server_wait (or
 * of the worker); valid until the next call on the server. */
const char* pm_server_last_error(pm_server_handle s);

/* Stop the worker (after finishing admitted work) and free everything. */
pm_status pm_server_destroy(pm_server_handle s);

#ifdef __cplusplus
}
#endif
#endif /* COSMOS_PM_SERVE_H */
