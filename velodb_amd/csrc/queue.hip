// queue.hip -- the resident command-queue kernel (TaskManager replacement).
//
// Reference: CommandQueueController / persistentKernel / TaskManager
// (include/cuda/commands.hpp:16-201, src/cuda/commands.cpp:11-91,
// src/cuda/persistent_kernel.cu:18-124, src/cuda/task_manager.cu:26-124):
// a ring of Commands in mapped host memory, one cooperative kernel polling
// it, grid-wide barriers between commands, in-order completion.
//
// Here:
//   * every opcode runs resident: vo_submit records the command's program
//     (resident.hpp: the launches its host driver would make, as steps) into
//     a host-mapped arena and publishes a ring entry; the resident kernel
//     copies the program into device memory and runs it step by step, each
//     step's virtual blocks spread over the lane's blocks, with a barrier of
//     the lane's blocks between steps;
//   * the grid is split into lanes (block b serves lane b mod L, so with
//     L | 8 a lane keeps to its own XCDs and their L2), each with its own
//     ring, barrier and leader: commands of different lanes run side by side
//     and finish out of order.  Commands of one lane run in order;
//   * completion is per command: the lane leader stores the command's
//     sequence id in flag[seq mod capacity] (fine-grained host memory) after
//     the results are written back, and bumps its lane's `done`;
//   * every wait on the device is bounded: the kernel leaves when every lane
//     has been idle for the idle period, and a barrier that does not
//     complete in time aborts the kernel, so the grid always drains.  The
//     host relaunches on the next submit or wait after an idle exit.
#include "compact_kernels.hpp"
#include "gather_obl_kernels.hpp"
#include "join_kernels.hpp"
#include "join_pairs_kernels.hpp"
#include "misc_kernels.hpp"
#include "predicate_kernels.hpp"
#include "resident_tables.hpp"
#include "serve_kernels.hpp"
#include "shuffle_kernels.hpp"
#include "sort_kernels.hpp"
#include "sort_multi_kernels.hpp"

#include <chrono>
#include <deque>
#include <thread>

namespace vo {
namespace {

using namespace sortimpl;

constexpr uint32_t RB = kResidentThreads;
constexpr uint32_t kMaxLanes = 8;
constexpr uint32_t kMaxQueueBlocks = 2048;
constexpr uint32_t kFlagStride = 16;   // u32 words per arrival flag: one 64-byte line
constexpr uint32_t kMaxSteps = 4096;   // steps of one command (1 MiB of device program per lane)
constexpr uint64_t kArenaSteps = 32768; // host program arena (8 MiB)

// One ring entry (fine-grained host memory).
struct QEntry {
    uint64_t seq;      // global sequence id
    uint32_t opcode;
    uint32_t n_steps;  // program length (0 for NOP / TERMINATE)
    const Step *prog;  // host-mapped program
    uint32_t act;      // blocks of the lane the program uses (its largest step grid)
    uint32_t n_bar;    // barriers between its steps
    uint64_t pad_[4];
};
static_assert(sizeof(QEntry) == 64, "ring entries are one line");

struct QLaneHost {
    uint64_t head; // entries published on the lane
    uint64_t done; // entries completed on the lane
    uint64_t pad_[6];
};

// Control words shared with the host (fine-grained host memory).
struct QCtrl {
    QLaneHost lane[kMaxLanes];
    uint32_t stop;        // host asks the kernel to leave (after draining published entries)
    uint32_t alive;       // 1 while the kernel runs
    uint32_t error;       // 1 = a barrier timed out
    uint32_t exit_reason; // 1 stop, 2 idle, 3 terminate, 4 abort
};

// Per-lane device state (persists across launches).
struct QLaneDev {
    uint32_t release; // barrier generation released by the lane leader
    uint32_t exit;    // exit code decided by the leader's fetch
    uint32_t epoch;   // barrier generation when the lane last left
    uint32_t n_steps; // program length of the current entry
    uint64_t consumed;
    uint64_t seq;     // sequence id of the current entry
    uint64_t prog;    // host address of its program
    uint32_t opcode;
    uint32_t act;     // blocks of the lane the current entry uses
    uint32_t n_bar;   // barriers between its steps
    uint32_t pad_[3];
};
static_assert(sizeof(QLaneDev) == 64, "lane state is one line");

struct QDev {
    uint32_t abort;
    uint32_t exit_all;   // every lane leaves (idle vote complete, or TERMINATE)
    uint32_t idle_votes; // lanes idle for the idle period
    uint32_t blocks_out; // blocks that have left this launch
    uint32_t reason;     // exit reason of the launch (max over lanes)
    uint32_t pad_[11];
    QLaneDev lane[kMaxLanes];
    uint32_t arrive[kMaxQueueBlocks * kFlagStride];
};

struct QArgs {
    const QEntry *ring; // lanes x cap_lane entries
    QCtrl *ctrl;
    QDev *qd;
    Step *prog;         // lanes x kMaxSteps (device)
    uint64_t *flag;     // cap completion flags (host)
    uint32_t cap;       // flags
    uint32_t cap_lane;  // ring entries per lane
    uint32_t lanes;
    uint32_t pad_;
    uint64_t idle_ticks, bar_ticks;
};

__device__ __forceinline__ uint32_t ld_sys32(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void st_sys(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void st_sys32(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ uint32_t ld_ag(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct NoHook {
    __device__ void operator()() const {}
};

__device__ __forceinline__ bool timed_out(QDev *qd, uint64_t t0, uint64_t limit)
{
    if (ld_ag(&qd->abort)) return true;
    if (wall_clock64() - t0 > limit) {
        __hip_atomic_store(&qd->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
    return false;
}

// Where a block serves: lane = block mod L, rank within the lane, lane size.
struct LanePos {
    uint32_t lane, rank, nbl, L;
};

// Barrier number ++epoch of one lane.  The lane leader (rank 0) waits until
// every other block of the lane has arrived, runs hook() with all its
// threads (the hook sees every block's writes), then releases the lane.
// `pub`: this block wrote data that others read after the barrier (its
// arrival is a release); `sub`: it reads data others wrote before it (an
// acquire follows the release).  A block without either touches no cache
// (flags are read and written by device-coherent atomics, polled relaxed).
// Every wait is bounded: a barrier that does not complete within
// limit_ticks sets `abort` and returns false in every block.
template <class Hook>
__device__ bool lbarrier(QDev *qd, const LanePos &lp, uint32_t &epoch, uint64_t limit_ticks, Hook hook, bool pub,
                         bool sub)
{
    __shared__ int ok;
    QLaneDev *ld = &qd->lane[lp.lane];
    const uint32_t e = ++epoch;
    __syncthreads();
    if (lp.rank != 0) {
        if (threadIdx.x == 0) {
            uint32_t *a = &qd->arrive[blockIdx.x * kFlagStride];
            if (pub)
                __hip_atomic_store(a, e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            else
                __hip_atomic_store(a, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int good = 1;
            const uint64_t t0 = wall_clock64();
            while ((int32_t)(__hip_atomic_load(&ld->release, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - e) < 0) {
                if (timed_out(qd, t0, limit_ticks)) {
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (sub) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            ok = good;
        }
        __syncthreads();
        return ok != 0;
    }
    if (threadIdx.x < 64) {
        int good = 1;
        const uint64_t t0 = wall_clock64();
        for (;;) {
            // up to 8 flags per lane per round, loaded before any is tested:
            // one memory round trip per 512 blocks
            int all = 1;
            for (uint32_t k0 = 1 + threadIdx.x; k0 < lp.nbl; k0 += 512) {
                uint32_t v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t k = k0 + 64u * j;
                    v[j] = k < lp.nbl ? __hip_atomic_load(&qd->arrive[(lp.lane + k * lp.L) * kFlagStride],
                                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                      : e;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) all &= (int32_t)(v[j] - e) >= 0;
            }
            if (__all(all)) break;
            if (timed_out(qd, t0, limit_ticks)) { // uniform: scalar clock, same flag for every lane
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (threadIdx.x == 0) ok = good;
    }
    __syncthreads();
    if (!ok) return false;
    hook();
    __syncthreads();
    // the leader always releases: the hook wrote the lane state and program
    if (threadIdx.x == 0) __hip_atomic_store(&ld->release, e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// ------------------------------------------------------------ step bodies

// A kernel parameter read from the program with every 32-bit word made
// wave-uniform, so pointers and sizes stay in scalar registers.
template <class T>
__device__ __forceinline__ T uread(const char *p)
{
    T v;
    __builtin_memcpy(&v, p, sizeof(T));
    if constexpr (sizeof(T) % 4 == 0) {
        uint32_t w[sizeof(T) / 4];
        __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
        for (uint32_t i = 0; i < sizeof(T) / 4; ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
        __builtin_memcpy(&v, w, sizeof(T));
    } else {
        v = (T)__builtin_amdgcn_readfirstlane((uint32_t)v);
    }
    return v;
}

template <class S, class F, size_t... I>
__device__ __forceinline__ void call_blob_u_i(const char *blob, F &&f, std::index_sequence<I...>)
{
    f(uread<typename S::template T<I>>(blob + S::off(I))...);
}

template <class K, class F>
__device__ __forceinline__ void call_blob_u(const char *blob, F &&f)
{
    using S = KSigOf<K>;
    call_blob_u_i<S>(blob, f, std::make_index_sequence<S::n>{});
}

// A pointer made wave-uniform (function arguments arrive in vector registers).
__device__ __forceinline__ const char *uptr(const char *p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const char *)(((uint64_t)hi << 32) | lo);
}

// One runner per table row: the step's arguments once, then every virtual
// block of the step that falls to this block.  Not inlined, so each body is
// compiled as its own function.
// The resident block's dynamic LDS (every body's `lds`).
extern __shared__ __align__(16) char q_lds[];

#define VO_RUNNER(ID, K, CALL, L)                                                                                  \
    __device__ __noinline__ void run_##ID(const char *blob_v, uint32_t grid_v, uint32_t rank_v, uint32_t nbl_v)     \
    {                                                                                                              \
        char *lds = q_lds;                                                                                         \
        const char *blob = uptr(blob_v);                                                                           \
        const uint32_t grid = __builtin_amdgcn_readfirstlane(grid_v);                                              \
        const uint32_t rank = __builtin_amdgcn_readfirstlane(rank_v);                                              \
        const uint32_t nbl = __builtin_amdgcn_readfirstlane(nbl_v);                                                \
        call_blob_u<decltype(K)>(blob, [&](auto... a) {                                                            \
            const uint32_t nvb = grid;                                                                             \
            (void)nvb;                                                                                             \
            (void)lds;                                                                                             \
            for (uint32_t vb = rank; vb < grid; vb += nbl) {                                                       \
                CALL;                                                                                              \
                __syncthreads();                                                                                   \
            }                                                                                                      \
        });                                                                                                        \
    }
VO_ALL_BODIES(VO_RUNNER)
#undef VO_RUNNER

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Built-in fill and copy: the whole lane strides over the bytes.
__device__ void step_fill(const FillArgs &a, uint32_t value, uint32_t rank, uint32_t nbl)
{
    const uint64_t stride = (uint64_t)nbl * RB;
    const uint64_t t = (uint64_t)rank * RB + threadIdx.x;
    char *d = (char *)a.dst;
    const uint32_t b = value & 0xFF;
    if ((((uintptr_t)d | a.bytes) & 15) == 0) {
        const uint32_t w = b * 0x01010101u;
        const u32x4 v = {w, w, w, w};
        for (uint64_t i = t; i < a.bytes / 16; i += stride) ((u32x4 *)d)[i] = v;
    } else {
        for (uint64_t i = t; i < a.bytes; i += stride) d[i] = (char)b;
    }
}

__device__ void step_copy(const CopyArgs &a, uint32_t rank, uint32_t nbl)
{
    const uint64_t stride = (uint64_t)nbl * RB;
    const uint64_t t = (uint64_t)rank * RB + threadIdx.x;
    char *d = (char *)a.dst;
    const char *s = (const char *)a.src;
    if ((((uintptr_t)d | (uintptr_t)s | a.bytes) & 15) == 0) {
        for (uint64_t i = t; i < a.bytes / 16; i += stride) ((u32x4 *)d)[i] = ((const u32x4 *)s)[i];
    } else if ((((uintptr_t)d | (uintptr_t)s | a.bytes) & 7) == 0) {
        for (uint64_t i = t; i < a.bytes / 8; i += stride) ((uint64_t *)d)[i] = ((const uint64_t *)s)[i];
    } else {
        for (uint64_t i = t; i < a.bytes; i += stride) d[i] = s[i];
    }
}

__device__ void run_step(const Step *st, uint32_t rank, uint32_t nbl)
{
    const uint32_t id = __builtin_amdgcn_readfirstlane(st->id);
    const uint32_t grid = __builtin_amdgcn_readfirstlane(st->grid);
    switch (id) {
    // fill and copy: the first min(grid, nbl) blocks stride over the bytes
    case kStepFill:
        if (rank < grid) step_fill(uread<FillArgs>(st->blob), __builtin_amdgcn_readfirstlane((uint32_t)st->aux), rank,
                                   grid < nbl ? grid : nbl);
        break;
    case kStepCopy:
        if (rank < grid) step_copy(uread<CopyArgs>(st->blob), rank, grid < nbl ? grid : nbl);
        break;
#define VO_CASE(ID, K, CALL, L)                                                                                    \
    case ID: run_##ID(st->blob, grid, rank, nbl); break;
        VO_ALL_BODIES(VO_CASE)
#undef VO_CASE
    default: break; // ids are checked on the host before a program is published
    }
}

// ------------------------------------------------------------ fetch

// The lane leader's work at the barrier that ends one command and starts
// the next: publish completion of the command every block of the lane just
// finished, wait (bounded) for the next ring entry, and copy its program into
// the lane's device buffer (one PCIe round trip per 4 KiB).
struct FetchHook {
    const QArgs *qa;
    uint32_t lane;
    uint64_t consumed; // entries of this lane completed so far
    uint64_t done_seq; // sequence id of the command just finished (0: none)

    __device__ void operator()() const
    {
        __shared__ uint32_t shared_ns, shared_ex;
        QDev *qd = qa->qd;
        QLaneDev *ld = &qd->lane[lane];
        QCtrl *ctrl = qa->ctrl;
        if (threadIdx.x < 64) {
            if (done_seq && threadIdx.x == 0) {
                // the flag's release covers the lane count stored after it
                st_sys(&qa->flag[done_seq % qa->cap], done_seq);
                __hip_atomic_store(&ctrl->lane[lane].done, consumed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            const uint64_t t0 = wall_clock64();
            uint32_t ex = 0;
            bool voted = false;
            for (;;) { // every exit condition reads one address for all lanes: uniform
                // the four words of one round are loaded before any is tested
                const uint32_t abort = ld_ag(&qd->abort);
                const uint64_t head =
                    __hip_atomic_load(&ctrl->lane[lane].head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                const uint32_t exit_all = ld_ag(&qd->exit_all);
                const uint32_t stop = ld_sys32(&ctrl->stop);
                if (abort) {
                    ex = 4;
                    break;
                }
                if (head > consumed) {
                    if (voted && threadIdx.x == 0)
                        __hip_atomic_fetch_sub(&qd->idle_votes, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    voted = false;
                    // the others may have decided to leave meanwhile: leave too
                    // (the host starts the kernel again for this entry)
                    if (exit_all) ex = 2;
                    break;
                }
                if (exit_all) {
                    ex = 2;
                    break;
                }
                if (stop) {
                    ex = 1;
                    break;
                }
                if (!voted && wall_clock64() - t0 > qa->idle_ticks) {
                    voted = true;
                    uint32_t v = 0;
                    if (threadIdx.x == 0)
                        v = __hip_atomic_fetch_add(&qd->idle_votes, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
                    v = __builtin_amdgcn_readfirstlane(v);
                    if (v >= qa->lanes) {
                        if (threadIdx.x == 0)
                            __hip_atomic_store(&qd->exit_all, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ex = 2;
                        break;
                    }
                }
                __builtin_amdgcn_s_sleep(2);
            }
            // the entry (published after its program) and the program: one
            // acquire for the poll that saw the head move
            if (ex == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
            // the entry's first four words, one per lane: one PCIe round trip
            uint64_t wv = 0;
            if (ex == 0 && threadIdx.x < 4) {
                const QEntry *en = &qa->ring[lane * qa->cap_lane + consumed % qa->cap_lane];
                wv = __hip_atomic_load(reinterpret_cast<const uint64_t *>(en) + threadIdx.x, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
            }
            const uint64_t w0 = __shfl(wv, 0), w1 = __shfl(wv, 1), w2 = __shfl(wv, 2), w3 = __shfl(wv, 3);
            if (threadIdx.x == 0) {
                uint32_t ns = 0;
                if (ex == 0) {
                    ns = (uint32_t)(w1 >> 32);
                    ns = ns <= kMaxSteps ? ns : 0;
                    ld->seq = w0;
                    ld->opcode = (uint32_t)w1;
                    ld->prog = w2;
                    ld->act = ns ? (uint32_t)w3 : 0;
                    ld->n_bar = ns ? (uint32_t)(w3 >> 32) : 0;
                }
                ld->n_steps = ns;
                ld->exit = ex;
                shared_ns = ns;
                shared_ex = ex;
            }
        }
        __syncthreads();
        if (shared_ex != 0) return;
        const uint32_t ns = shared_ns;
        if (ns == 0) return;
        const u32x4 *src = (const u32x4 *)ld->prog;
        u32x4 *dst = (u32x4 *)(qa->prog + (uint64_t)lane * kMaxSteps);
        const uint32_t n16 = ns * (sizeof(Step) / 16);
        for (uint32_t i = threadIdx.x; i < n16; i += 4 * RB) {
            u32x4 v[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint32_t j = i + m * RB;
                if (j < n16) v[m] = __builtin_nontemporal_load(src + j);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint32_t j = i + m * RB;
                if (j < n16) dst[j] = v[m];
            }
        }
    }
};

__global__ __launch_bounds__(RB, 2) void k_queue(QArgs qa)
{
    QDev *qd = qa.qd;
    LanePos lp;
    lp.L = qa.lanes;
    lp.lane = blockIdx.x % lp.L;
    lp.rank = blockIdx.x / lp.L;
    lp.nbl = (gridDim.x - lp.lane + lp.L - 1) / lp.L;
    QLaneDev *ld = &qd->lane[lp.lane];
    if (blockIdx.x == 0 && threadIdx.x == 0) st_sys32(&qa.ctrl->alive, 1);
    uint64_t consumed = ld->consumed;
    uint32_t epoch = ld->epoch;
    uint32_t reason = 0;
    uint64_t done_seq = 0;
    const Step *prog = qa.prog + (uint64_t)lp.lane * kMaxSteps;
    for (;;) {
        // the other blocks wait here while the leader polls the ring: allow the idle time
        const FetchHook hook{&qa, lp.lane, consumed, done_seq};
        // the previous command's results were written back by the release
        // below; the lane state is read with device-coherent atomics
        if (!lbarrier(qd, lp, epoch, qa.idle_ticks + qa.bar_ticks, hook, false, false)) {
            reason = 4;
            break;
        }
        const uint32_t ex = ld_ag(&ld->exit);
        if (ex) {
            reason = ex;
            break;
        }
        const uint32_t ns = __builtin_amdgcn_readfirstlane(ld_ag(&ld->n_steps));
        // blocks of rank >= act have no virtual block in any step: they skip
        // the program and only keep the lane's barrier count
        const bool active = lp.rank < __builtin_amdgcn_readfirstlane(ld_ag(&ld->act));
        // inputs may have been written by the host or a copy engine since
        // this CU last looked, and the program is new: drop cached lines
        if (active && threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
        __syncthreads();
        const uint32_t opcode = __builtin_amdgcn_readfirstlane(ld_ag(&ld->opcode));
        const uint64_t seq = __hip_atomic_load(&ld->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ++consumed;
        done_seq = seq;
        if (opcode == VO_OP_TERMINATE) {
            if (lp.rank == 0 && threadIdx.x == 0) {
                st_sys(&qa.flag[seq % qa.cap], seq);
                st_sys(&qa.ctrl->lane[lp.lane].done, consumed);
                __hip_atomic_store(&qd->exit_all, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            reason = 3;
            break;
        }
        bool ok = true;
        if (active) {
            bool wrote = false; // a virtual block of this block since the last barrier
            for (uint32_t s = 0; s < ns && ok;) {
                wrote |= lp.rank < __builtin_amdgcn_readfirstlane(prog[s].grid);
                run_step(prog + s, lp.rank, lp.nbl);
                const uint32_t sl = __builtin_amdgcn_readfirstlane((uint32_t)prog[s].slots);
                const uint32_t fl = __builtin_amdgcn_readfirstlane((uint32_t)prog[s].aux);
                s += sl ? sl : 1;
                // the next step reads nothing this one wrote: run on into it
                if (s < ns && !(fl & kStepNoBarrier)) {
                    ok = lbarrier(qd, lp, epoch, qa.bar_ticks, NoHook{}, wrote, true);
                    wrote = false;
                }
            }
            // results out of every XCD's L2 before the leader tells the host
            __syncthreads();
            if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        } else {
            const uint32_t nb = __builtin_amdgcn_readfirstlane(ld_ag(&ld->n_bar));
            for (uint32_t k = 0; k < nb && ok; ++k) ok = lbarrier(qd, lp, epoch, qa.bar_ticks, NoHook{}, false, false);
        }
        if (!ok) {
            reason = 4;
            break;
        }
    }
    if (lp.rank == 0 && threadIdx.x == 0) {
        ld->consumed = consumed;
        ld->epoch = epoch;
        __hip_atomic_fetch_max(&qd->reason, reason, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // the last block out reports the launch's end and resets the launch state
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t out = __hip_atomic_fetch_add(&qd->blocks_out, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1;
        if (out == gridDim.x) {
            const uint32_t r = ld_ag(&qd->reason);
            qd->blocks_out = 0;
            qd->exit_all = 0;
            qd->idle_votes = 0;
            qd->reason = 0;
            qd->abort = 0;
            __threadfence();
            if (r == 4) st_sys32(&qa.ctrl->error, 1);
            st_sys32(&qa.ctrl->exit_reason, r);
            st_sys32(&qa.ctrl->alive, 0);
        }
    }
}

// ----------------------------------------------------------------- host

struct QueueState {
    QEntry *ring = nullptr;
    QCtrl *ctrl = nullptr;
    QDev *qd = nullptr;
    Step *dprog = nullptr;
    uint64_t *flag = nullptr;
    Step *arena = nullptr;
    hipStream_t stream = nullptr;
    uint32_t cap = 0, cap_lane = 0, lanes = 1;
    uint64_t submitted = 0;                 // global sequence ids handed out
    uint64_t lane_submitted[kMaxLanes] = {}; // entries published per lane
    uint64_t arena_head = 0;                // steps handed out (absolute)
    struct Live {
        uint64_t seq, start, end;
    };
    std::deque<Live> live;                  // arena regions of commands not yet known complete
    uint32_t grid = 0;
    size_t lds = 0;
    uint64_t idle_ticks = 0, bar_ticks = 0;
    bool terminated = false;
};

constexpr double kBarrierSeconds = 10.0;

volatile QCtrl *vctrl(QueueState *q) { return q->ctrl; }

uint64_t lane_done(QueueState *q, uint32_t l) { return __atomic_load_n(&q->ctrl->lane[l].done, __ATOMIC_ACQUIRE); }

bool seq_done(QueueState *q, uint64_t seq)
{
    return __atomic_load_n(&q->flag[seq % q->cap], __ATOMIC_ACQUIRE) >= seq;
}

bool all_done(QueueState *q)
{
    for (uint32_t l = 0; l < q->lanes; ++l)
        if (lane_done(q, l) < q->lane_submitted[l]) return false;
    return true;
}

void launch_queue_kernel(vo_ctx *ctx, QueueState *q)
{
    vctrl(q)->alive = 0;
    vctrl(q)->stop = 0;
    vctrl(q)->exit_reason = 0;
    QArgs qa{q->ring, q->ctrl, q->qd, q->dprog, q->flag, q->cap, q->cap_lane, q->lanes, 0, q->idle_ticks,
             q->bar_ticks};
    void *args[] = {&qa};
    const hipError_t le =
        hipLaunchCooperativeKernel((const void *)k_queue, dim3(q->grid), dim3(RB), args, (unsigned)q->lds, q->stream);
    if (le != hipSuccess) {
        (void)hipGetLastError();
        std::string m = std::string("queue kernel launch: ") + hipGetErrorString(le);
        // seen when another process has made a cooperative launch on this GPU
        if (le == hipErrorIllegalState)
            m += " (is another process running resident queues on this GPU? cooperative launches are one process "
                 "per GPU on this runtime)";
        throw Error{VO_ERR_HIP, m};
    }
    // wait until the kernel reports that it runs
    const auto t0 = std::chrono::steady_clock::now();
    while (vctrl(q)->alive == 0) {
        hipError_t e = hipStreamQuery(q->stream);
        if (e != hipErrorNotReady && e != hipSuccess)
            throw Error{VO_ERR_HIP, std::string("queue kernel: ") + hipGetErrorString(e)};
        if (e == hipSuccess) break; // it already left (nothing to do): the caller checks again
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(30))
            throw Error{VO_ERR_TIMEOUT, "queue kernel did not start"};
        std::this_thread::yield();
    }
    (void)ctx;
}

void reap_error(QueueState *q)
{
    if (vctrl(q)->error) throw Error{VO_ERR_HIP, "queue kernel aborted (a grid barrier timed out)"};
    hipError_t e = hipStreamQuery(q->stream);
    if (e != hipErrorNotReady && e != hipSuccess)
        throw Error{VO_ERR_HIP, std::string("queue kernel: ") + hipGetErrorString(e)};
}

int fail_q(vo_ctx *ctx, const Error &e)
{
    ctx->last_error = e.msg;
    ctx->last_status = e.status;
    return e.status;
}

// Restart the kernel when it left after an idle period with entries still
// unconsumed (it may have decided to leave just before one was published).
// Caller holds q's submission lock (ctx->mu).
void ensure_running_locked(vo_ctx *ctx, QueueState *q)
{
    reap_error(q);
    if (vctrl(q)->alive) return;
    VO_HIP(hipStreamSynchronize(q->stream)); // the previous launch has drained
    reap_error(q);
    if (q->terminated || vctrl(q)->exit_reason == 3) {
        q->terminated = true;
        if (!all_done(q)) throw Error{VO_ERR_QUEUE, "queue terminated"};
        return;
    }
    if (!all_done(q)) launch_queue_kernel(ctx, q);
}

// Waits (relaunching an idle kernel) until pred() holds.  Caller holds ctx->mu.
template <class P>
void wait_locked(vo_ctx *ctx, QueueState *q, P pred, const char *what)
{
    const auto t0 = std::chrono::steady_clock::now();
    while (!pred()) {
        ensure_running_locked(ctx, q);
        const auto el = std::chrono::steady_clock::now() - t0;
        if (el > std::chrono::milliseconds(ctx->wait_timeout_ms)) throw Error{VO_ERR_TIMEOUT, what};
        if (el > std::chrono::milliseconds(2))
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        else if (el > std::chrono::microseconds(50))
            std::this_thread::yield();
    }
}

// Copies a recorded program into the host arena; returns its address.
const Step *place_program(vo_ctx *ctx, QueueState *q, Recorder &rec, uint64_t seq)
{
    const uint64_t data_steps = (rec.data.size() + sizeof(Step) - 1) / sizeof(Step);
    const uint64_t n = rec.steps.size() + data_steps;
    VO_CHECK(rec.steps.size() <= kMaxSteps, VO_ERR_UNSUPPORTED,
             "command program of " + std::to_string(rec.steps.size()) + " steps exceeds the resident limit");
    VO_CHECK(n <= kArenaSteps / 2, VO_ERR_UNSUPPORTED, "command program exceeds the program arena");
    uint64_t start = q->arena_head;
    if (start % kArenaSteps + n > kArenaSteps) start += kArenaSteps - start % kArenaSteps; // no wrap inside
    const uint64_t end = start + n;
    auto reclaim = [&] {
        while (!q->live.empty() && seq_done(q, q->live.front().seq)) q->live.pop_front();
        const uint64_t oldest = q->live.empty() ? end : q->live.front().start;
        return end - oldest <= kArenaSteps;
    };
    wait_locked(ctx, q, reclaim, "program arena full");
    Step *dst = q->arena + start % kArenaSteps;
    char *data = (char *)(dst + rec.steps.size());
    for (const auto &ref : rec.data_refs) {
        CopyArgs ca;
        std::memcpy(&ca, rec.steps[ref.first].blob, sizeof ca);
        ca.src = data + ref.second;
        std::memcpy(rec.steps[ref.first].blob, &ca, sizeof ca);
    }
    std::memcpy(dst, rec.steps.data(), rec.steps.size() * sizeof(Step));
    if (!rec.data.empty()) std::memcpy(data, rec.data.data(), rec.data.size());
    q->arena_head = end;
    q->live.push_back({seq, start, end});
    return dst;
}

// The blocks a program needs (its largest step grid: blocks of higher rank
// have no virtual block in any step) and the barriers between its steps, the
// same count the resident kernel's step loop takes.
void program_shape(const std::vector<Step> &st, uint32_t &act, uint32_t &n_bar)
{
    act = 0;
    n_bar = 0;
    for (size_t s = 0; s < st.size();) {
        act = std::max(act, st[s].grid);
        const size_t next = s + std::max<uint16_t>(1, st[s].slots);
        if (next < st.size() && !(st[s].aux & kStepNoBarrier)) ++n_bar;
        s = next;
    }
}

// Publishes one entry on `lane`.  Caller holds ctx->mu.
uint64_t publish_locked(vo_ctx *ctx, QueueState *q, uint32_t lane, uint32_t opcode, Recorder *rec)
{
    VO_CHECK(!q->terminated, VO_ERR_QUEUE, "queue terminated");
    const uint64_t seq = q->submitted + 1;
    // the lane's ring slot and the sequence id's completion flag must be free
    wait_locked(ctx, q, [&] { return q->lane_submitted[lane] - lane_done(q, lane) < q->cap_lane; },
                "command ring full");
    if (seq > q->cap) wait_locked(ctx, q, [&] { return seq_done(q, seq - q->cap); }, "completion flags full");
    const Step *prog = nullptr;
    uint32_t ns = 0, act = 0, n_bar = 0;
    if (rec && !rec->steps.empty()) {
        prog = place_program(ctx, q, *rec, seq);
        ns = (uint32_t)rec->steps.size();
        program_shape(rec->steps, act, n_bar);
    }
    QEntry *en = &q->ring[lane * q->cap_lane + q->lane_submitted[lane] % q->cap_lane];
    en->seq = seq;
    en->opcode = opcode;
    en->n_steps = ns;
    en->prog = prog;
    en->act = act;
    en->n_bar = n_bar;
    q->submitted = seq;
    q->lane_submitted[lane] += 1;
    __atomic_store_n(&q->ctrl->lane[lane].head, q->lane_submitted[lane], __ATOMIC_RELEASE);
    // checked after publishing: a kernel that left before it saw the entry
    // is started again
    ensure_running_locked(ctx, q);
    return seq;
}

} // namespace

uint64_t queue_submit_lane(vo_ctx *ctx, vo_command *cmd, uint32_t lane)
{
    QueueState *q = (QueueState *)ctx->queue;
    try {
        VO_CHECK(lane < q->lanes, VO_ERR_INVALID, "lane " + std::to_string(lane) + " of " + std::to_string(q->lanes));
        std::lock_guard<std::mutex> lk(ctx->mu);
        Recorder rec;
        rec.lane = lane;
        if (cmd->opcode != VO_OP_NOP && cmd->opcode != VO_OP_TERMINATE) record_command(ctx, *cmd, rec);
        const uint64_t seq = publish_locked(ctx, q, lane, cmd->opcode, &rec);
        cmd->sequence_id = seq;
        return seq;
    } catch (const Error &e) {
        fail_q(ctx, e);
        return 0;
    }
}

uint64_t queue_submit(vo_ctx *ctx, vo_command *cmd) { return queue_submit_lane(ctx, cmd, 0); }

// Publishes a program recorded by the caller (several operators as one
// entry); caller holds ctx->mu.
uint64_t queue_publish_locked(vo_ctx *ctx, uint32_t lane, Recorder &rec)
{
    QueueState *q = (QueueState *)ctx->queue;
    VO_CHECK(lane < q->lanes, VO_ERR_INVALID, "lane " + std::to_string(lane) + " of " + std::to_string(q->lanes));
    return publish_locked(ctx, q, lane, kOpProgram, &rec);
}

void queue_wait_throw(vo_ctx *ctx, uint64_t seq, uint32_t timeout_ms)
{
    QueueState *q = (QueueState *)ctx->queue;
    if (timeout_ms == 0) timeout_ms = ctx->wait_timeout_ms;
    VO_CHECK(seq >= 1 && seq <= q->submitted, VO_ERR_INVALID, "wait on a sequence id never submitted");
    // Spin, then yield, then sleep: a sleep lasts the timer slack (tens of
    // microseconds) at least, longer than most resident commands.
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t spins = 0; !seq_done(q, seq); ++spins) {
        if (vctrl(q)->alive == 0 || vctrl(q)->error) {
            std::lock_guard<std::mutex> lk(ctx->mu);
            ensure_running_locked(ctx, q);
        }
        if ((spins & 63) != 63) continue;
        const auto el = std::chrono::steady_clock::now() - t0;
        if (el > std::chrono::milliseconds(timeout_ms))
            throw Error{VO_ERR_TIMEOUT, "vo_wait: command " + std::to_string(seq) + " not finished"};
        if (el > std::chrono::milliseconds(2))
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        else if (el > std::chrono::microseconds(50))
            std::this_thread::yield();
    }
}

int queue_wait(vo_ctx *ctx, uint64_t seq, uint32_t timeout_ms)
{
    try {
        queue_wait_throw(ctx, seq, timeout_ms);
        return VO_OK;
    } catch (const Error &e) {
        return fail_q(ctx, e);
    }
}

int queue_poll(vo_ctx *ctx, uint64_t seq)
{
    QueueState *q = (QueueState *)ctx->queue;
    try {
        reap_error(q);
        if (seq_done(q, seq)) return 1;
        // a kernel that left on idle just as this command was published is
        // started again here too, so a caller that only polls makes progress
        if (vctrl(q)->alive == 0) {
            std::lock_guard<std::mutex> lk(ctx->mu);
            ensure_running_locked(ctx, q);
        }
        return seq_done(q, seq) ? 1 : 0;
    } catch (const Error &e) {
        return -fail_q(ctx, e);
    }
}

// Waits for every submitted command; caller holds ctx->mu (used by the
// scratch arena before it is replaced).
void queue_drain_locked(vo_ctx *ctx)
{
    QueueState *q = (QueueState *)ctx->queue;
    wait_locked(ctx, q, [&] { return all_done(q); }, "queue drain");
}

int queue_park(vo_ctx *ctx)
{
    QueueState *q = (QueueState *)ctx->queue;
    try {
        std::lock_guard<std::mutex> lk(ctx->mu);
        reap_error(q);
        // the kernel takes every published entry before it looks at `stop`
        vctrl(q)->stop = 1;
        const auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            const hipError_t e = hipStreamQuery(q->stream);
            if (e == hipSuccess) break;
            if (e != hipErrorNotReady) throw Error{VO_ERR_HIP, std::string("queue kernel: ") + hipGetErrorString(e)};
            if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(ctx->wait_timeout_ms))
                throw Error{VO_ERR_TIMEOUT, "queue kernel did not leave"};
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        vctrl(q)->stop = 0;
        reap_error(q);
        // an idle exit may have raced with the last publish: run what is left
        if (!all_done(q)) {
            launch_queue_kernel(ctx, q);
            vctrl(q)->stop = 1;
            VO_HIP(hipStreamSynchronize(q->stream));
            vctrl(q)->stop = 0;
            reap_error(q);
        }
        VO_CHECK(all_done(q), VO_ERR_QUEUE, "queue kernel left with commands outstanding");
        return VO_OK;
    } catch (const Error &e) {
        return fail_q(ctx, e);
    }
}

int queue_sync(vo_ctx *ctx)
{
    try {
        std::lock_guard<std::mutex> lk(ctx->mu);
        queue_drain_locked(ctx);
        return VO_OK;
    } catch (const Error &e) {
        return fail_q(ctx, e);
    }
}

uint32_t queue_lanes(vo_ctx *ctx) { return ctx->queue ? ((QueueState *)ctx->queue)->lanes : 0; }

// The sequence id the command being recorded will receive (ids are handed
// out in order under ctx->mu).
uint64_t queue_next_seq(vo_ctx *ctx) { return ((QueueState *)ctx->queue)->submitted + 1; }

// Waits for one submitted command; caller holds ctx->mu.
void queue_wait_seq_locked(vo_ctx *ctx, uint64_t seq)
{
    QueueState *q = (QueueState *)ctx->queue;
    VO_CHECK(seq >= 1 && seq <= q->submitted, VO_ERR_INVALID, "wait on a sequence id never submitted");
    wait_locked(ctx, q, [&] { return seq_done(q, seq); }, "command not finished");
}

} // namespace vo

using namespace vo;

extern "C" {

int vo_queue_start(vo_ctx *ctx)
{
    if (!ctx) return VO_ERR_INVALID;
    if (ctx->queue) return fail_q(ctx, Error{VO_ERR_QUEUE, "queue already running"});
    if (!(ctx->flags & VO_CFG_PERSISTENT)) return VO_OK; // launch mode: nothing to start
    QueueState *q = new QueueState();
    try {
        VO_HIP(hipSetDevice(ctx->device));
        q->lanes = std::max<uint32_t>(1, ctx->queue_lanes);
        VO_CHECK(q->lanes <= kMaxLanes, VO_ERR_INVALID, "queue_lanes above 8");
        q->cap = ctx->queue_capacity;
        VO_CHECK(q->cap >= q->lanes, VO_ERR_INVALID, "queue_capacity below the lane count");
        q->cap_lane = q->cap / q->lanes;
        VO_HIP(hipHostMalloc((void **)&q->ring, sizeof(QEntry) * q->lanes * q->cap_lane,
                             hipHostMallocCoherent | hipHostMallocMapped));
        VO_HIP(hipHostMalloc((void **)&q->ctrl, sizeof(QCtrl), hipHostMallocCoherent | hipHostMallocMapped));
        VO_HIP(hipHostMalloc((void **)&q->flag, sizeof(uint64_t) * q->cap, hipHostMallocCoherent | hipHostMallocMapped));
        VO_HIP(hipHostMalloc((void **)&q->arena, sizeof(Step) * kArenaSteps, hipHostMallocCoherent | hipHostMallocMapped));
        std::memset((void *)q->ctrl, 0, sizeof(QCtrl));
        std::memset((void *)q->ring, 0, sizeof(QEntry) * q->lanes * q->cap_lane);
        std::memset((void *)q->flag, 0, sizeof(uint64_t) * q->cap);
        VO_HIP(hipMalloc((void **)&q->qd, sizeof(QDev)));
        VO_HIP(hipMalloc((void **)&q->dprog, sizeof(Step) * kMaxSteps * q->lanes));
        VO_HIP(hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking));
        VO_HIP(hipMemsetAsync(q->qd, 0, sizeof(QDev), q->stream));
        VO_HIP(hipStreamSynchronize(q->stream));
        q->lds = kResidentLds;
        VO_HIP(hipFuncSetAttribute((const void *)k_queue, hipFuncAttributeMaxDynamicSharedMemorySize, (int)q->lds));
        int per_cu = 0;
        VO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_queue, RB, q->lds));
        VO_CHECK(per_cu >= 1, VO_ERR_UNSUPPORTED, "queue kernel does not fit on a CU");
        q->grid = std::min<uint32_t>((uint32_t)per_cu * (uint32_t)ctx->num_cus, kMaxQueueBlocks);
        VO_CHECK(q->grid >= q->lanes, VO_ERR_INVALID, "fewer resident blocks than lanes");
        int rate_khz = 0;
        VO_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
        const double hz = rate_khz > 0 ? rate_khz * 1e3 : 1e8;
        q->idle_ticks = (uint64_t)(ctx->queue_idle_us * 1e-6 * hz);
        q->bar_ticks = (uint64_t)(kBarrierSeconds * hz);
        VO_HIP(hipStreamSynchronize(ctx->stream)); // earlier launch-mode work first
        launch_queue_kernel(ctx, q);
    } catch (const Error &e) {
        if (q->stream) (void)hipStreamDestroy(q->stream);
        if (q->qd) (void)hipFree(q->qd);
        if (q->dprog) (void)hipFree(q->dprog);
        if (q->ctrl) (void)hipHostFree(q->ctrl);
        if (q->ring) (void)hipHostFree(q->ring);
        if (q->flag) (void)hipHostFree(q->flag);
        if (q->arena) (void)hipHostFree(q->arena);
        delete q;
        return fail_q(ctx, e);
    }
    ctx->queue = q;
    return VO_OK;
}

int vo_queue_stop(vo_ctx *ctx, uint32_t timeout_ms)
{
    if (!ctx) return VO_ERR_INVALID;
    QueueState *q = (QueueState *)ctx->queue;
    if (!q) return VO_OK;
    int rc = VO_OK;
    if (vctrl(q)->alive && !q->terminated) {
        vo_command term{};
        term.opcode = VO_OP_TERMINATE;
        if (queue_submit(ctx, &term) == 0) rc = ctx->last_status;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const auto limit = std::chrono::milliseconds(timeout_ms ? timeout_ms : 5000);
    bool asked_stop = false;
    for (;;) {
        hipError_t e = hipStreamQuery(q->stream);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) {
            rc = fail_q(ctx, Error{VO_ERR_HIP, std::string("queue kernel: ") + hipGetErrorString(e)});
            break;
        }
        if (std::chrono::steady_clock::now() - t0 > limit) {
            if (!asked_stop) {
                // the reference resets the device here (task_manager.cu:110-115); ask the kernel to leave instead
                vctrl(q)->stop = 1;
                asked_stop = true;
                continue;
            }
            if (std::chrono::steady_clock::now() - t0 > limit + std::chrono::seconds(15)) {
                return fail_q(ctx, Error{VO_ERR_TIMEOUT, "queue kernel did not stop"});
            }
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    (void)hipStreamDestroy(q->stream);
    (void)hipFree(q->qd);
    (void)hipFree(q->dprog);
    (void)hipHostFree(q->ctrl);
    (void)hipHostFree(q->ring);
    (void)hipHostFree(q->flag);
    (void)hipHostFree(q->arena);
    delete q;
    ctx->queue = nullptr;
    return rc;
}

int vo_queue_running(const vo_ctx *ctx)
{
    if (!ctx || !ctx->queue) return 0;
    return ((QueueState *)ctx->queue)->ctrl->alive ? 1 : 0;
}

uint64_t vo_submit_lane(vo_ctx *ctx, vo_command *cmd, uint32_t lane)
{
    if (!ctx || !cmd) return 0;
    if (!ctx->queue) {
        if (lane != 0) {
            fail_q(ctx, Error{VO_ERR_INVALID, "lanes need the resident queue (VO_CFG_PERSISTENT)"});
            return 0;
        }
        return vo_submit(ctx, cmd);
    }
    return queue_submit_lane(ctx, cmd, lane);
}

int vo_queue_lanes(const vo_ctx *ctx) { return ctx && ctx->queue ? (int)((QueueState *)ctx->queue)->lanes : 0; }

} // extern "C"
