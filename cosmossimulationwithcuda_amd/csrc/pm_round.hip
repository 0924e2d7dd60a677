// Persistent round kernel: R steps of every simulation of a batch in one
// launch, the blocks pulling work items from a device-side queue.
//
// No counterpart in the reference (one simulation per GPU, one kernel per
// stage, CosmosCuda.cuh:1336-1341).  The queue holds, step by step and pass
// by pass, the items of every member: R2C items (two row pairs), column
// items (two columns), C2R items (two row pairs, edge bands included) and
// push items (1024 particles).  A block takes the next item with one atomic
// add, waits until the member's previous pass of that step (or its last
// push of the step before) has completed, runs the item with the same device
// code the pass kernels use (pm_fft_pipe_core.h, pm_particle_core.h), and
// counts it done.  Items are taken in queue order and every item depends
// only on items earlier in the queue, so the oldest unfinished item never
// waits: the queue always drains, whatever the number of blocks.  A wait
// that exceeds the spin limit sets the error flag and every block leaves.
//
// Memory ordering between blocks (possibly on different XCDs), in the
// pattern of a grid barrier: the block barrier after an item completes every
// wave's stores, then one thread issues an agent-scope release fence (the
// XCD's L2 written back) and increments the done counter; the waiting thread
// polls the counter with relaxed agent-scope loads (coherent reads, no cache
// maintenance), issues one agent-scope acquire fence (L2 and L1 invalidated)
// once the count is reached, and the block barrier hands that on to the other
// waves.  An acquire load per poll would invalidate the L2 on every
// iteration of every waiting block: hundreds of blocks polling then keep the
// L2s of all XCDs under continuous invalidation and starve the items they
// wait for.  Polls back off exponentially (64 to 2048 cycles), and the wait
// limit is wall-clock time (the constant-rate counter), not a poll count.
#include "pm_fft_pipe_core.h"
#include "pm_particle_core.h"

namespace pmk {
namespace {

constexpr int kRoundIt = 2;                       // row pairs / columns per FFT item

template <int N>
struct RoundFft {
    static constexpr int R = N >= 2048 ? 8 : 4;   // the default plan's radix (pipe plans, it = 2)
    static constexpr int LDS = (padded(3 * N) + tw_lds_entries(N, R)) * 8;
};

// Items are separate functions (not inlined into the queue loop): the
// kernel's register budget is then the largest item's, not the sum the
// inlined switch needs.
template <int N>
__device__ __noinline__ void fft_item(const RoundInst& I, int phase, int p, int k, float2* lds) {
    using F = RoundFft<N>;
    float2* buf = lds;
    float2* twl = lds + padded(3 * N);
    const GField d = gfield(I.fd[p]);
    if (phase == 0) r2c_pipe_body<N, F::R, kRoundIt>(d, I.tw, buf, twl, k * kRoundIt);
    else if (phase == 1) col_pipe_body<N, F::R, true, kRoundIt>(d, I.tw, buf, twl, k * kRoundIt);
    else c2r_pipe_body<N, F::R, true, kRoundIt>(d, I.tw, buf, twl, k * kRoundIt);
}

template <int MODE>
__device__ __noinline__ void push_item(const PartDesc& d, int64_t q) {
    direct_push_group<MODE>(d, q);
}

__device__ __forceinline__ int load_relaxed(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits (one thread) until cnt >= target, without cache maintenance (the
// caller fences once afterwards); false when `limit` wall-clock ticks pass or
// another block raised the error flag.
__device__ __forceinline__ bool wait_count(const int* cnt, int target, int* err, long long limit) {
    if (load_relaxed(cnt) >= target) return true;
    const long long t0 = wall_clock64();
    for (int k = 0;; ++k) {
        switch (k < 5 ? k : 5) {                       // s_sleep takes an immediate: 64 * 2^k cycles
            case 0: __builtin_amdgcn_s_sleep(1); break;
            case 1: __builtin_amdgcn_s_sleep(2); break;
            case 2: __builtin_amdgcn_s_sleep(4); break;
            case 3: __builtin_amdgcn_s_sleep(8); break;
            case 4: __builtin_amdgcn_s_sleep(16); break;
            default: __builtin_amdgcn_s_sleep(32); break;
        }
        if (load_relaxed(cnt) >= target) return true;
        if ((k & 15) == 15 && (load_relaxed(err) || wall_clock64() - t0 > limit)) {
            atomicExch(err, 1);
            return false;
        }
    }
}

// ctl: [0] queue head, [1] error flag, [2 + 4 i + phase] items of member i's
// pass `phase` done so far (all zero at launch).  MAXN: the largest member
// lattice (the 2048^2 C2R item alone needs ~200 VGPRs; batches up to 1024^2
// get a lighter kernel).
template <int MAXN>
__global__ __launch_bounds__(kBlock) void k_round(const RoundInst* __restrict__ inst,
                                                  const RoundSeg* __restrict__ seg, int nseg, int total, int* ctl,
                                                  long long spin_limit) {
    extern __shared__ float2 lds[];
    __shared__ int s_item, s_seg, s_go;
    int* head = ctl;
    int* err = ctl + 1;
    int* cnt = ctl + 2;
    for (;;) {
        if (threadIdx.x == 0) {
            const int it = atomicAdd(head, 1);
            int sg = -1, go = 0;
            if (it < total && !load_relaxed(err)) {
                int lo = 0, hi = nseg;                 // last segment with start <= it
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (seg[mid].start <= it) lo = mid; else hi = mid;
                }
                sg = lo;
                const RoundSeg S = seg[sg];
                const int* c = cnt + 4 * S.inst;
                const RoundInst& I = inst[S.inst];
                // r2c of step s: the member's C2R and push of step s - 1 are done
                // (H, the force lattice and the lattice it zeroes are free); the
                // other passes: the previous pass of the same step is done
                bool ok = true;
                if (S.phase == 0) {
                    if (S.step > 0)
                        ok = wait_count(c + 2, S.step * I.items[2], err, spin_limit) &&
                             wait_count(c + 3, S.step * I.items[3], err, spin_limit);
                } else {
                    ok = wait_count(c + S.phase - 1, (S.step + 1) * I.items[S.phase - 1], err, spin_limit);
                }
                go = ok ? 1 : 0;
                // acquire: the producers' stores (written back by their
                // release) are visible to this block's loads
                if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            s_item = it;
            s_seg = sg;
            s_go = go;
        }
        __syncthreads();
        if (!s_go) break;
        const int it = s_item;
        const RoundSeg S = seg[s_seg];
        const RoundInst& I = inst[S.inst];
        const int k = it - S.start, p = S.step & 1;
        if (S.phase == 3) {
            const int64_t q = (int64_t)k * kBlock + threadIdx.x;
            if (I.mode == 0) push_item<0>(I.pd[p], q);
            else push_item<1>(I.pd[p], q);
        } else {
            switch (I.n) {
                case 256: fft_item<256>(I, S.phase, p, k, lds); break;
                case 512: fft_item<512>(I, S.phase, p, k, lds); break;
                case 1024: fft_item<1024>(I, S.phase, p, k, lds); break;
                default:
                    if constexpr (MAXN >= 2048) fft_item<2048>(I, S.phase, p, k, lds);
                    break;
            }
        }
        __syncthreads();                               // every wave's stores of the item are complete
        if (threadIdx.x == 0)                          // release: written back, then counted
            __hip_atomic_fetch_add(cnt + 4 * S.inst + S.phase, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

bool round_supported(int n, const FftPlan& p) {
    if (n != 256 && n != 512 && n != 1024 && n != 2048) return false;
    return p.rp == 1 && p.cb == 1 && p.it == kRoundIt && p.radix == (n >= 2048 ? 8 : 4);
}

void round_items(int n, int64_t np, int items[4]) {
    items[0] = items[1] = items[2] = n / 2 / kRoundIt;
    items[3] = direct_blocks(np);
}

size_t round_lds_bytes(int max_n) {
    switch (max_n) {
        case 256: return RoundFft<256>::LDS;
        case 512: return RoundFft<512>::LDS;
        case 1024: return RoundFft<1024>::LDS;
        default: return RoundFft<2048>::LDS;
    }
}

static const void* round_kernel(int max_n) {
    return max_n >= 2048 ? reinterpret_cast<const void*>(&k_round<2048>) : reinterpret_cast<const void*>(&k_round<1024>);
}

long long round_wait_ticks(double seconds) {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        khz = 100000;                                  // the MI300-series constant counter: 100 MHz
    return (long long)(seconds * 1e3 * khz);
}

int round_max_blocks(int max_n, size_t lds_bytes) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipFuncSetAttribute(round_kernel(max_n), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, round_kernel(max_n), kBlock, lds_bytes) != hipSuccess)
        return 0;
    return per_cu * cus;
}

hipError_t round_persistent(const RoundInst* inst, const RoundSeg* seg, int nseg, int total, int* ctl, int ninst,
                            int max_n, size_t lds_bytes, int blocks, long long spin_limit, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    // (the dynamic LDS limit was raised by round_max_blocks, outside any capture)
    const hipError_t e = fill_zero(ctl, round_ctl_bytes(ninst), s);
    if (e != hipSuccess) return e;
    if (max_n >= 2048)
        hipLaunchKernelGGL(k_round<2048>, dim3(blocks), dim3(kBlock), lds_bytes, s, inst, seg, nseg, total, ctl,
                           spin_limit);
    else
        hipLaunchKernelGGL(k_round<1024>, dim3(blocks), dim3(kBlock), lds_bytes, s, inst, seg, nseg, total, ctl,
                           spin_limit);
    return hipGetLastError();
}

}  // namespace pmk
