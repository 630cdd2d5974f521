// queue.hip -- the resident command-queue kernel (TaskManager replacement).
//
// Reference: CommandQueueController / persistentKernel / TaskManager
// (include/cuda/commands.hpp:16-201, src/cuda/commands.cpp:11-91,
// src/cuda/persistent_kernel.cu:18-124, src/cuda/task_manager.cu:26-124):
// a ring of Commands in mapped host memory, one cooperative kernel polling
// it, grid-wide barriers between commands.
//
// Here:
//   * ring and control words live in fine-grained host-coherent memory
//     (hipHostMalloc coherent + mapped); the host writes a command, then
//     publishes `head` with a release store; the device leader reads `head`
//     with a system-scope acquire load and copies the command into device
//     memory for every block;
//   * one cooperative launch of 512-thread blocks, as many as fit on every
//     CU (all co-resident), runs the same device bodies as the launch-mode
//     kernels, separated by a sense-counting grid barrier in device memory;
//   * completion is a per-command count (`done`) published with a system
//     release after the results are written back;
//   * every wait on the device is bounded: the leader exits after an idle
//     period, and a barrier that does not complete in time aborts the
//     kernel, so the grid always drains.  The host relaunches on the next
//     submit after an idle exit.
// Commands available in the resident kernel: NOP, SORT_KP, PERMUTE, GATHER,
// COMPARE_EQ, COMPOSE_POSITION_MAP, ENRICH_POSITIONS, TERMINATE; the others
// run in launch mode.
#include "sort_kernels.hpp"

#include <chrono>
#include <thread>

namespace vo {
namespace {

using namespace sortimpl;

constexpr int QBLOCK = 512;
constexpr int QT_LOG = 13, QR_LOG = 4;
constexpr int kQueueSMax = 4, kQueueHMax = 7; // widths 1-4 and 6-7 are built here
using QTile = dev::TileCT<RecKP, QT_LOG, QR_LOG, false>;

// Control words shared with the host (fine-grained host memory).
struct QCtrl {
    uint64_t head;  // commands published by the host
    uint64_t done;  // commands completed by the device
    uint32_t stop;  // host asks the kernel to exit
    uint32_t alive; // 1 while the kernel runs
    uint32_t error; // device error code (0 = none, 1 = barrier timeout)
    uint32_t exit_reason; // 1 stop, 2 idle, 3 terminate, 4 abort
};

static_assert(sizeof(vo_command) % 8 == 0, "commands are copied in 8-byte words");

// Grid coordination state (device memory).
// Each block announces its arrival at a barrier in its own 64-byte line of
// `arrive` (no two blocks contend for one address); block 0 collects the
// arrivals with one wavefront of parallel loads and releases the grid with
// one store to `release`.
constexpr uint32_t kMaxQueueBlocks = 2048;
constexpr uint32_t kFlagStride = 16; // u32 words per flag: one 64-byte line

struct QDev {
    uint32_t release;
    uint32_t abort;
    uint32_t exit;
    uint32_t epoch; // barrier generation when the kernel last left
    uint64_t consumed;
    uint64_t pad_[5];
    vo_command cmd;
    uint32_t arrive[kMaxQueueBlocks * kFlagStride];
};

__device__ __forceinline__ uint64_t ld_sys(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}

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

struct NoHook {
    __device__ void operator()() const {}
};

__device__ __forceinline__ bool timed_out(QDev *qd, uint64_t t0, uint64_t limit)
{
    if (__hip_atomic_load(&qd->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return true;
    if (wall_clock64() - t0 > limit) {
        __hip_atomic_store(&qd->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
    return false;
}

// Grid barrier number ++epoch.  Block 0 waits until every other block has
// arrived, runs hook() with all its threads (the hook sees every block's
// writes), then releases the grid.  Every wait is bounded: a barrier that
// does not complete within limit_ticks sets `abort` and returns false in
// every block, so the grid drains.
template <class Hook>
__device__ bool qbarrier(QDev *qd, uint32_t nb, uint32_t &epoch, uint64_t limit_ticks, Hook hook)
{
    __shared__ int ok;
    const uint32_t e = ++epoch;
    __syncthreads();
    if (blockIdx.x != 0) {
        if (threadIdx.x == 0) {
            __threadfence();
            __hip_atomic_store(&qd->arrive[blockIdx.x * kFlagStride], e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            int good = 1;
            const uint64_t t0 = wall_clock64();
            while ((int32_t)(__hip_atomic_load(&qd->release, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) - e) < 0) {
                if (timed_out(qd, t0, limit_ticks)) {
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __threadfence();
            ok = good;
        }
        __syncthreads();
        return ok != 0;
    }
    if (threadIdx.x < 64) {
        int good = 1;
        const uint64_t t0 = wall_clock64();
        for (;;) {
            int all = 1;
            for (uint32_t blk = 1 + threadIdx.x; blk < nb; blk += 64)
                all &= (int32_t)(__hip_atomic_load(&qd->arrive[blk * kFlagStride], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT) - e) >= 0;
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
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_store(&qd->release, e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

__device__ __forceinline__ bool qbarrier(QDev *qd, uint32_t nb, uint32_t &epoch, uint64_t limit_ticks)
{
    return qbarrier(qd, nb, epoch, limit_ticks, NoHook{});
}

// ------------------------------------------------------------ op bodies

template <int S>
__device__ __forceinline__ void sort_global_items(const dev::Bufs<RecKP> &b, const dev::IMap &map, int p, int q_lo,
                                                  int p_top, uint64_t n_pad, uint32_t nb)
{
    // 256-thread items of the launch-mode pass; a 512-thread block runs two
    const uint32_t items = (uint32_t)((n_pad >> S) / 256);
    const uint32_t half = threadIdx.x >> 8;
    for (uint32_t w = blockIdx.x; 2 * w < items; w += nb) {
        const uint32_t it = 2 * w + half;
        if (it < items) dev::body_global<RecKP, S, false>(b, map, p, q_lo, p_top, nullptr, it, threadIdx.x & 255);
    }
}

// SORT_KP: the launch-mode pass schedule (plan_passes with register passes
// of at most 4 stages and hybrid passes of 6-7), one grid barrier per pass.
__device__ bool exec_sort_kp(const vo_sort_kp_args &a, char *lds, QDev *qd, uint32_t nb, uint32_t &epoch,
                             uint64_t bar)
{
    dev::Bufs<RecKP> b;
    b.k[0] = a.keys;
    b.p = a.payload;
    const dev::IMap map{nullptr, 0, nullptr};
    int n_log = 0;
    while ((1ull << n_log) < a.n_pad) ++n_log;
    if (n_log < QR_LOG) {
        dev::body_tiny<RecKP, false>(b, n_log, a.n, 1, nullptr, blockIdx.x);
        return qbarrier(qd, nb, epoch, bar);
    }
    if (n_log < QT_LOG) {
        if (blockIdx.x == 0)
            dev::body_local<RecKP, QT_LOG, QR_LOG, QBLOCK, false>(b, map, n_log, n_log, n_log, a.n, 1, nullptr, lds, 0);
        return qbarrier(qd, nb, epoch, bar);
    }
    const uint32_t tiles = (uint32_t)(a.n_pad >> QT_LOG);
    for (uint32_t w = blockIdx.x; w < tiles; w += nb)
        dev::body_local_ct<RecKP, QT_LOG, QR_LOG, QBLOCK, false>(b, map, n_log, a.n, 1, nullptr, lds, w);
    if (!qbarrier(qd, nb, epoch, bar)) return false;
    for (int p = QT_LOG + 1; p <= n_log; ++p) {
        const int count = p - QT_LOG;
        const int groups = plan_groups(count, kQueueSMax, kQueueHMax, 6); // widths built below: 1-4, 6-7
        int q = p - 1;
        for (int gi = 0; gi < groups; ++gi) {
            const int s = plan_width(count, groups, gi);
            const int q_lo = q - s + 1;
            q -= s;
            switch (s) {
            case 1: sort_global_items<1>(b, map, p, q_lo, n_log, a.n_pad, nb); break;
            case 2: sort_global_items<2>(b, map, p, q_lo, n_log, a.n_pad, nb); break;
            case 3: sort_global_items<3>(b, map, p, q_lo, n_log, a.n_pad, nb); break;
            case 4: sort_global_items<4>(b, map, p, q_lo, n_log, a.n_pad, nb); break;
            case 6:
                for (uint32_t w = blockIdx.x; w < tiles; w += nb)
                    dev::body_hybrid<RecKP, QT_LOG, QR_LOG, QBLOCK, 6, false>(b, map, p, q_lo, n_log, nullptr, lds, w);
                break;
            case 7:
                for (uint32_t w = blockIdx.x; w < tiles; w += nb)
                    dev::body_hybrid<RecKP, QT_LOG, QR_LOG, QBLOCK, 7, false>(b, map, p, q_lo, n_log, nullptr, lds, w);
                break;
            default: break; // plan_groups above yields only the widths built here
            }
            if (!qbarrier(qd, nb, epoch, bar)) return false;
        }
        for (uint32_t w = blockIdx.x; w < tiles; w += nb)
            dev::body_merge<RecKP, QT_LOG, QR_LOG, QBLOCK, false>(b, map, p, n_log, nullptr, lds, w);
        if (!qbarrier(qd, nb, epoch, bar)) return false;
    }
    return true;
}

__device__ __forceinline__ uint32_t width_of(uint32_t t)
{
    switch (t) {
    case VO_T_BOOLEAN: case VO_T_TINYINT: return 1;
    case VO_T_SMALLINT: return 2;
    case VO_T_INTEGER: case VO_T_FLOAT: case VO_T_DATE: return 4;
    default: return 8;
    }
}

__device__ __forceinline__ uint64_t ld_w(const void *p, uint64_t i, uint32_t w)
{
    switch (w) {
    case 1: return ((const uint8_t *)p)[i];
    case 2: return ((const uint16_t *)p)[i];
    case 4: return ((const uint32_t *)p)[i];
    default: return ((const uint64_t *)p)[i];
    }
}

__device__ __forceinline__ void st_w(void *p, uint64_t i, uint32_t w, uint64_t v)
{
    switch (w) {
    case 1: ((uint8_t *)p)[i] = (uint8_t)v; break;
    case 2: ((uint16_t *)p)[i] = (uint16_t)v; break;
    case 4: ((uint32_t *)p)[i] = (uint32_t)v; break;
    default: ((uint64_t *)p)[i] = v; break;
    }
}

// Element-wise commands (reference permute.cuh, gather.cuh, compare.cuh,
// oblivious_shuffle.cuh) as grid-stride loops over the resident grid.
__device__ void exec_elementwise(const vo_command &c, uint32_t nb)
{
    const uint64_t stride = (uint64_t)nb * QBLOCK;
    const uint64_t i0 = (uint64_t)blockIdx.x * QBLOCK + threadIdx.x;
    switch (c.opcode) {
    case VO_OP_PERMUTE: {
        const vo_permute_args &a = c.args.permute;
        const uint32_t w = width_of(a.type_id);
        for (uint64_t i = i0; i < a.n; i += stride) st_w(a.out_data, i, w, ld_w(a.in_data, (uint64_t)a.in_indices[i], w));
        break;
    }
    case VO_OP_GATHER: {
        const vo_gather_args &a = c.args.gather;
        const uint32_t w = width_of(a.type_id);
        for (uint64_t i = i0; i < a.n; i += stride)
            if (a.in_mask == nullptr || a.in_mask[i]) st_w(a.out_data, (uint64_t)a.in_indices[i], w, ld_w(a.in_data, i, w));
        break;
    }
    case VO_OP_COMPARE_EQ: {
        const vo_compare_eq_args &a = c.args.compare_eq;
        const uint32_t w = width_of(a.type_id);
        for (uint64_t i = i0; i < a.n; i += stride) {
            bool eq;
            if (a.type_id == VO_T_FLOAT)
                eq = ((const float *)a.left_data)[i] == ((const float *)a.right_data)[i];
            else if (a.type_id == VO_T_DOUBLE)
                eq = ((const double *)a.left_data)[i] == ((const double *)a.right_data)[i];
            else
                eq = ld_w(a.left_data, i, w) == ld_w(a.right_data, i, w);
            a.out_mask[i] = eq ? 1 : 0;
        }
        break;
    }
    case VO_OP_COMPOSE_POSITION_MAP: {
        const vo_compose_args &a = c.args.compose_position_map;
        for (uint64_t i = i0; i < a.n; i += stride) a.new_map[i] = a.old_map[a.sigma[i]];
        break;
    }
    case VO_OP_ENRICH_POSITIONS: {
        const vo_enrich_args &a = c.args.enrich_positions;
        for (uint64_t i = i0; i < a.n; i += stride) a.positions[i] = a.position_map[a.record_ids[i]];
        break;
    }
    default: break;
    }
}

// Block 0's work at the barrier that ends one command and starts the next:
// publish completion of the command every block just finished, then wait
// (bounded by the idle time) for the host to publish the next one and copy it
// into device memory with one wavefront (one PCIe round trip per 64 words).
struct FetchHook {
    const vo_command *ring;
    QCtrl *ctrl;
    QDev *qd;
    uint32_t cap;
    uint64_t consumed; // commands completed so far
    bool publish;      // a command finished since the last fetch
    uint64_t idle_ticks;

    __device__ void operator()() const
    {
        if (threadIdx.x >= 64) return;
        const uint32_t lane = threadIdx.x;
        if (publish && lane == 0) st_sys(&ctrl->done, consumed);
        const uint64_t t0 = wall_clock64();
        uint32_t ex = 0;
        for (;;) {
            if (ld_sys(&ctrl->head) > consumed) break; // same address in every lane: uniform
            if (ld_sys32(&ctrl->stop)) {
                ex = 1;
                break;
            }
            if (wall_clock64() - t0 > idle_ticks) {
                ex = 2;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (ex == 0) {
            const uint64_t *src = reinterpret_cast<const uint64_t *>(&ring[consumed % cap]);
            uint64_t *dst = reinterpret_cast<uint64_t *>(&qd->cmd);
            for (uint32_t k = lane; k < sizeof(vo_command) / 8; k += 64)
                dst[k] = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        if (lane == 0) qd->exit = ex;
    }
};

__global__ __launch_bounds__(QBLOCK) void k_queue(const vo_command *ring, QCtrl *ctrl, QDev *qd, uint32_t cap,
                                                   uint64_t idle_ticks, uint64_t bar_ticks)
{
    extern __shared__ __align__(16) char lds[];
    const uint32_t nb = gridDim.x;
    const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
    if (leader) st_sys32(&ctrl->alive, 1);
    uint64_t consumed = qd->consumed;
    uint32_t epoch = qd->epoch;
    uint32_t reason = 0;
    bool publish = false;
    for (;;) {
        // the other blocks wait here while block 0 polls the ring: allow the idle time
        const FetchHook hook{ring, ctrl, qd, cap, consumed, publish, idle_ticks};
        if (!qbarrier(qd, nb, epoch, idle_ticks + bar_ticks, hook)) {
            reason = 4;
            break;
        }
        const uint32_t ex = __hip_atomic_load(&qd->exit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ex) {
            reason = ex;
            break;
        }
        // inputs may have been written by the host or a copy engine since
        // this CU last looked: drop cached lines (every XCD has blocks here)
        if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
        __syncthreads();
        const vo_command cmd = qd->cmd;
        if (cmd.opcode == VO_OP_TERMINATE) {
            ++consumed;
            if (leader) st_sys(&ctrl->done, consumed);
            reason = 3;
            break;
        }
        bool ok = true;
        if (cmd.opcode == VO_OP_SORT_KP)
            ok = exec_sort_kp(cmd.args.sort_kp, lds, qd, nb, epoch, bar_ticks);
        else
            exec_elementwise(cmd, nb);
        if (!ok) {
            reason = 4;
            break;
        }
        // results out of every XCD's L2 before block 0 tells the host
        __syncthreads();
        if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        ++consumed;
        publish = true;
    }
    if (leader) {
        qd->consumed = consumed;
        qd->epoch = epoch;
        if (reason == 4) st_sys32(&ctrl->error, 1);
        st_sys32(&ctrl->exit_reason, reason);
        st_sys32(&ctrl->alive, 0);
    }
}

// ----------------------------------------------------------------- host

struct QueueState {
    vo_command *ring = nullptr;
    QCtrl *ctrl = nullptr;
    QDev *qd = nullptr;
    hipStream_t stream = nullptr;
    uint32_t cap = 0;
    uint64_t submitted = 0;
    uint32_t grid = 0;
    size_t lds = 0;
    uint64_t idle_ticks = 0, bar_ticks = 0;
};

constexpr double kBarrierSeconds = 10.0;

volatile QCtrl *vctrl(QueueState *q) { return q->ctrl; }

void launch_queue_kernel(vo_ctx *ctx, QueueState *q)
{
    vctrl(q)->alive = 0;
    vctrl(q)->stop = 0;
    vctrl(q)->exit_reason = 0;
    void *args[] = {&q->ring, &q->ctrl, &q->qd, &q->cap, &q->idle_ticks, &q->bar_ticks};
    VO_HIP(hipLaunchCooperativeKernel((const void *)k_queue, dim3(q->grid), dim3(QBLOCK), args, (unsigned)q->lds,
                                      q->stream));
    // wait until the kernel reports that it runs
    const auto t0 = std::chrono::steady_clock::now();
    while (vctrl(q)->alive == 0) {
        hipError_t e = hipStreamQuery(q->stream);
        if (e != hipErrorNotReady && e != hipSuccess) throw Error{VO_ERR_HIP, std::string("queue kernel: ") + hipGetErrorString(e)};
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(30))
            throw Error{VO_ERR_TIMEOUT, "queue kernel did not start"};
        std::this_thread::yield();
    }
    (void)ctx;
}

bool supported(uint32_t op)
{
    switch (op) {
    case VO_OP_NOP: case VO_OP_SORT_KP: case VO_OP_PERMUTE: case VO_OP_GATHER: case VO_OP_COMPARE_EQ:
    case VO_OP_COMPOSE_POSITION_MAP: case VO_OP_ENRICH_POSITIONS: case VO_OP_TERMINATE: return true;
    default: return false;
    }
}

uint32_t elem_bytes(uint32_t t)
{
    switch (t) {
    case VO_T_BOOLEAN: case VO_T_TINYINT: return 1;
    case VO_T_SMALLINT: return 2;
    case VO_T_INTEGER: case VO_T_FLOAT: case VO_T_DATE: return 4;
    case VO_T_BIGINT: case VO_T_DOUBLE: case VO_T_VARCHAR: case VO_T_UBIGINT: return 8;
    default: return 0;
    }
}

void validate(const vo_command &c)
{
    VO_CHECK(supported(c.opcode), VO_ERR_UNSUPPORTED,
             "opcode " + std::to_string(c.opcode) + " runs in launch mode only (not in the resident queue)");
    switch (c.opcode) {
    case VO_OP_PERMUTE: {
        const vo_permute_args &a = c.args.permute;
        VO_CHECK(elem_bytes(a.type_id) && (a.n == 0 || (a.out_data && a.in_data && a.in_indices)), VO_ERR_INVALID,
                 "permute: bad type or buffers");
        break;
    }
    case VO_OP_GATHER: {
        const vo_gather_args &a = c.args.gather;
        VO_CHECK(elem_bytes(a.type_id) && (a.n == 0 || (a.out_data && a.in_data && a.in_indices)), VO_ERR_INVALID,
                 "gather: bad type or buffers");
        break;
    }
    case VO_OP_COMPARE_EQ: {
        const vo_compare_eq_args &a = c.args.compare_eq;
        VO_CHECK(elem_bytes(a.type_id) && (a.n == 0 || (a.left_data && a.right_data && a.out_mask)), VO_ERR_INVALID,
                 "compare_eq: bad type or buffers");
        break;
    }
    case VO_OP_COMPOSE_POSITION_MAP: {
        const vo_compose_args &a = c.args.compose_position_map;
        VO_CHECK(a.n == 0 || (a.old_map && a.sigma && a.new_map), VO_ERR_INVALID, "compose: null buffer");
        break;
    }
    case VO_OP_ENRICH_POSITIONS: {
        const vo_enrich_args &a = c.args.enrich_positions;
        VO_CHECK(a.n == 0 || (a.record_ids && a.position_map && a.positions), VO_ERR_INVALID, "enrich: null buffer");
        break;
    }
    default: break;
    }
    if (c.opcode == VO_OP_SORT_KP) {
        const vo_sort_kp_args &a = c.args.sort_kp;
        VO_CHECK(a.keys && a.payload && is_pow2(a.n_pad) && a.n <= a.n_pad && a.n_pad <= (1ull << 29), VO_ERR_INVALID,
                 "sort_kp: bad buffers or sizes");
    }
}

void reap_error(vo_ctx *ctx, QueueState *q)
{
    if (vctrl(q)->error) throw Error{VO_ERR_HIP, "queue kernel aborted (a grid barrier timed out)"};
    hipError_t e = hipStreamQuery(q->stream);
    if (e != hipErrorNotReady && e != hipSuccess) throw Error{VO_ERR_HIP, std::string("queue kernel: ") + hipGetErrorString(e)};
    (void)ctx;
}

int fail_q(vo_ctx *ctx, const Error &e)
{
    ctx->last_error = e.msg;
    ctx->last_status = e.status;
    return e.status;
}

} // namespace

// Restart the kernel when it left after an idle period with commands still
// unconsumed (it may have decided to leave just before a command was
// published).  Caller holds ctx->mu.
void ensure_running_locked(vo_ctx *ctx, QueueState *q)
{
    reap_error(ctx, q);
    if (vctrl(q)->alive) return;
    VO_HIP(hipStreamSynchronize(q->stream)); // the previous kernel has drained
    if (vctrl(q)->exit_reason == 3) throw Error{VO_ERR_QUEUE, "queue terminated"};
    if (__atomic_load_n(&q->ctrl->done, __ATOMIC_ACQUIRE) < q->submitted)
        launch_queue_kernel(ctx, q);
}

uint64_t queue_submit(vo_ctx *ctx, vo_command *cmd)
{
    QueueState *q = (QueueState *)ctx->queue;
    try {
        validate(*cmd);
        std::lock_guard<std::mutex> lk(ctx->mu);
        const auto t0 = std::chrono::steady_clock::now();
        while (q->submitted - __atomic_load_n(&q->ctrl->done, __ATOMIC_ACQUIRE) >= q->cap) { // ring full
            ensure_running_locked(ctx, q);
            if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(ctx->wait_timeout_ms))
                throw Error{VO_ERR_TIMEOUT, "command ring full"};
            std::this_thread::yield();
        }
        const uint64_t seq = q->submitted + 1;
        cmd->sequence_id = seq;
        q->ring[q->submitted % q->cap] = *cmd;
        __atomic_store_n(&q->ctrl->head, seq, __ATOMIC_RELEASE);
        q->submitted = seq;
        // checked after publishing: a kernel that left before it saw the
        // command is started again
        ensure_running_locked(ctx, q);
        return seq;
    } catch (const Error &e) {
        fail_q(ctx, e);
        return 0;
    }
}

int queue_wait(vo_ctx *ctx, uint64_t seq, uint32_t timeout_ms)
{
    QueueState *q = (QueueState *)ctx->queue;
    try {
        if (timeout_ms == 0) timeout_ms = ctx->wait_timeout_ms;
        VO_CHECK(seq <= q->submitted, VO_ERR_INVALID, "wait on a sequence id never submitted");
        const auto t0 = std::chrono::steady_clock::now();
        int spins = 0;
        while (__atomic_load_n(&q->ctrl->done, __ATOMIC_ACQUIRE) < seq) {
            if (vctrl(q)->alive == 0 || vctrl(q)->error) {
                std::lock_guard<std::mutex> lk(ctx->mu);
                ensure_running_locked(ctx, q);
            }
            if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(timeout_ms))
                throw Error{VO_ERR_TIMEOUT, "vo_wait: command " + std::to_string(seq) + " not finished"};
            if (++spins > 2000) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        return VO_OK;
    } catch (const Error &e) {
        return fail_q(ctx, e);
    }
}

int queue_poll(vo_ctx *ctx, uint64_t seq)
{
    QueueState *q = (QueueState *)ctx->queue;
    try {
        reap_error(ctx, q);
        if (__atomic_load_n(&q->ctrl->done, __ATOMIC_ACQUIRE) >= seq) return 1;
        // a kernel that left on idle just as this command was published is
        // started again here too, so a caller that only polls makes progress
        if (vctrl(q)->alive == 0) {
            std::lock_guard<std::mutex> lk(ctx->mu);
            ensure_running_locked(ctx, q);
        }
        return __atomic_load_n(&q->ctrl->done, __ATOMIC_ACQUIRE) >= seq ? 1 : 0;
    } catch (const Error &e) {
        return -fail_q(ctx, e);
    }
}

bool queue_supports(uint32_t opcode) { return supported(opcode); }

int queue_park(vo_ctx *ctx)
{
    QueueState *q = (QueueState *)ctx->queue;
    try {
        std::lock_guard<std::mutex> lk(ctx->mu);
        reap_error(ctx, q);
        // the kernel takes every published command before it looks at `stop`
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
        reap_error(ctx, q);
        VO_CHECK(__atomic_load_n(&q->ctrl->done, __ATOMIC_ACQUIRE) == q->submitted, VO_ERR_QUEUE,
                 "queue kernel left with commands outstanding");
        return VO_OK;
    } catch (const Error &e) {
        return fail_q(ctx, e);
    }
}

int queue_sync(vo_ctx *ctx)
{
    QueueState *q = (QueueState *)ctx->queue;
    return q->submitted ? queue_wait(ctx, q->submitted, 0) : VO_OK;
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
        q->cap = ctx->queue_capacity;
        VO_HIP(hipHostMalloc((void **)&q->ring, sizeof(vo_command) * q->cap, hipHostMallocCoherent | hipHostMallocMapped));
        VO_HIP(hipHostMalloc((void **)&q->ctrl, sizeof(QCtrl), hipHostMallocCoherent | hipHostMallocMapped));
        std::memset((void *)q->ctrl, 0, sizeof(QCtrl));
        std::memset((void *)q->ring, 0, sizeof(vo_command) * q->cap);
        VO_HIP(hipMalloc((void **)&q->qd, sizeof(QDev)));
        VO_HIP(hipMemset(q->qd, 0, sizeof(QDev)));
        VO_HIP(hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking));
        // LDS: the larger of the tile column buffer and the small-array tile
        q->lds = std::max(QTile::lds_bytes(), dev::lds_bytes<RecKP>(QT_LOG - 1));
        VO_HIP(hipFuncSetAttribute((const void *)k_queue, hipFuncAttributeMaxDynamicSharedMemorySize, (int)q->lds));
        int per_cu = 0;
        VO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_queue, QBLOCK, q->lds));
        VO_CHECK(per_cu >= 1, VO_ERR_UNSUPPORTED, "queue kernel does not fit on a CU");
        q->grid = std::min<uint32_t>((uint32_t)per_cu * (uint32_t)ctx->num_cus, kMaxQueueBlocks);
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
        if (q->ctrl) (void)hipHostFree(q->ctrl);
        if (q->ring) (void)hipHostFree(q->ring);
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
    if (vctrl(q)->alive) {
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
    (void)hipHostFree(q->ctrl);
    (void)hipHostFree(q->ring);
    delete q;
    ctx->queue = nullptr;
    return rc;
}

int vo_queue_running(const vo_ctx *ctx)
{
    if (!ctx || !ctx->queue) return 0;
    return ((QueueState *)ctx->queue)->ctrl->alive ? 1 : 0;
}

} // extern "C"
