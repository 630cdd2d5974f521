// sort.hip -- oblivious bitonic sort kernels and their pass schedule.
//
// Reference: executeBitonicSort (include/cuda/sort.cuh:70-118) runs one
// grid-wide pass per (k, j) stage, S = log2N(log2N+1)/2 passes.  Here the
// same compare-exchange network is grouped into far fewer HBM passes:
//   pass 0       k_sort_local : every tile of 2^T_LOG records fully sorted
//                               (all levels p <= T_LOG) in registers + LDS;
//   per level p > T_LOG:
//                k_sort_global: stages q >= T_LOG, up to S_MAX per pass;
//                k_sort_merge : stages q < T_LOG in one tile pass.
// The schedule is a function of n_pad only.
#include "sort_core.hpp"

#include <algorithm>

namespace vo {
namespace dev {

// ----------------------------------------------------------- tiny arrays

// n_pad <= 8: one thread runs the network directly.
template <class R, bool TRACE>
__global__ void k_sort_tiny(Bufs<R> b, int n_log, uint64_t n_valid, int fill, TraceSlot *slot)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t h = 0;
    uint32_t nops = 0;
    const uint32_t N = 1u << n_log;
    R r[8];
    for (uint32_t i = 0; i < 8; ++i) {
        if (i < N) {
            gload<R, TRACE>(r[i], b, i, h, nops);
            if (fill) pad_select(r[i], i, n_valid);
        }
    }
    for (int p = 1; p <= n_log; ++p)
        for (int q = p - 1; q >= 0; --q)
            for (uint32_t i = 0; i < N; ++i) {
                const uint32_t ixj = i ^ (1u << q);
                if (i < ixj) cmpx(r[i], r[ixj], dir_asc(i, p, n_log));
            }
    for (uint32_t i = 0; i < 8; ++i)
        if (i < N) gstore<R, TRACE>(r[i], b, i, h, nops);
    if constexpr (TRACE) tcommit(slot, 0, h, nops);
}

// ---------------------------------------------------------- tile kernels

// Sorts tiles of 2^t records (t <= T_LOG): levels p = 1 .. p_max with the
// top level p_top.  Loads and stores use the coalesced layout a = t - R_LOG.
template <class R, int T_LOG, int R_LOG, int BLOCK, bool TRACE>
__global__ __launch_bounds__(BLOCK) void k_sort_local(Bufs<R> b, IMap map, int t, int p_max, int p_top,
                                                      uint64_t n_valid, int fill, TraceSlot *slot)
{
    extern __shared__ __align__(16) char lds_raw[];
    const uint32_t T = 1u << t;
    const Lds<R> s = lds_view<R>(lds_raw, T + (T >> 4));
    const uint32_t tau = threadIdx.x;
    const bool act = tau < (T >> R_LOG);
    const uint64_t vbase = (uint64_t)blockIdx.x << t;
    uint64_t h = 0;
    uint32_t nops = 0;
    R r[1 << R_LOG];
    int a = t - R_LOG;
    if (act) {
        const uint32_t base = lay_base<R_LOG>(tau, a);
#pragma unroll
        for (int m = 0; m < (1 << R_LOG); ++m) {
            const uint64_t v = vbase + (base | ((uint32_t)m << a));
            gload<R, TRACE>(r[m], b, phys(map, v), h, nops);
            if (fill) pad_select(r[m], v, n_valid);
        }
    }
    for (int p = 1; p <= p_max; ++p) tile_level<R, R_LOG, TRACE>(r, s, tau, act, a, p, p - 1, p_top, vbase, h, nops);
    const int a_out = t - R_LOG;
    if (a != a_out) {
        relayout<R, R_LOG, TRACE>(r, s, tau, act, a, a_out, h, nops);
        a = a_out;
    }
    if (act) {
        const uint32_t base = lay_base<R_LOG>(tau, a);
#pragma unroll
        for (int m = 0; m < (1 << R_LOG); ++m) {
            const uint64_t v = vbase + (base | ((uint32_t)m << a));
            gstore<R, TRACE>(r[m], b, phys(map, v), h, nops);
        }
    }
    if constexpr (TRACE) tcommit(slot, (uint64_t)blockIdx.x * BLOCK + tau, h, nops);
}

// Stages (p, T_LOG-1 .. 0) of one level p > T_LOG on tiles of 2^T_LOG.
template <class R, int T_LOG, int R_LOG, int BLOCK, bool TRACE>
__global__ __launch_bounds__(BLOCK) void k_sort_merge(Bufs<R> b, IMap map, int p, int p_top, TraceSlot *slot)
{
    extern __shared__ __align__(16) char lds_raw[];
    constexpr uint32_t T = 1u << T_LOG;
    const Lds<R> s = lds_view<R>(lds_raw, T + (T >> 4));
    const uint32_t tau = threadIdx.x;
    const uint64_t vbase = (uint64_t)blockIdx.x << T_LOG;
    uint64_t h = 0;
    uint32_t nops = 0;
    R r[1 << R_LOG];
    int a = T_LOG - R_LOG;
    {
        const uint32_t base = lay_base<R_LOG>(tau, a);
#pragma unroll
        for (int m = 0; m < (1 << R_LOG); ++m)
            gload<R, TRACE>(r[m], b, phys(map, vbase + (base | ((uint32_t)m << a))), h, nops);
    }
    tile_level<R, R_LOG, TRACE>(r, s, tau, true, a, p, T_LOG - 1, p_top, vbase, h, nops);
    if (a != T_LOG - R_LOG) {
        relayout<R, R_LOG, TRACE>(r, s, tau, true, a, T_LOG - R_LOG, h, nops);
        a = T_LOG - R_LOG;
    }
    {
        const uint32_t base = lay_base<R_LOG>(tau, a);
#pragma unroll
        for (int m = 0; m < (1 << R_LOG); ++m)
            gstore<R, TRACE>(r[m], b, phys(map, vbase + (base | ((uint32_t)m << a))), h, nops);
    }
    if constexpr (TRACE) tcommit(slot, (uint64_t)blockIdx.x * BLOCK + tau, h, nops);
}

// ------------------------------------------------------------ global pass

// S fused stages (p, q_lo+S-1 .. q_lo): thread g owns the 2^S records
// base + m * 2^q_lo, base = (g mod 2^q_lo) + (g / 2^q_lo) * 2^(q_lo+S).
// Consecutive lanes touch consecutive records, so every load and store of a
// wavefront is one contiguous 512-byte (keys) / 256-byte (payload) run.
template <class R, int S, bool TRACE>
__global__ __launch_bounds__(256) void k_sort_global(Bufs<R> b, IMap map, int p, int q_lo, int p_top,
                                                     TraceSlot *slot)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint64_t low = g & ((1ull << q_lo) - 1);
    const uint64_t base = low | ((g >> q_lo) << (q_lo + S));
    uint64_t h = 0;
    uint32_t nops = 0;
    R r[1 << S];
#pragma unroll
    for (int m = 0; m < (1 << S); ++m) gload<R, TRACE>(r[m], b, phys(map, base + ((uint64_t)m << q_lo)), h, nops);
    const bool asc = dir_asc(base, p, p_top);
#pragma unroll
    for (int qq = S - 1; qq >= 0; --qq) {
#pragma unroll
        for (int m = 0; m < (1 << S); ++m) {
            if (m & (1 << qq)) continue;
            cmpx(r[m], r[m | (1 << qq)], asc);
        }
    }
#pragma unroll
    for (int m = 0; m < (1 << S); ++m) gstore<R, TRACE>(r[m], b, phys(map, base + ((uint64_t)m << q_lo)), h, nops);
    if constexpr (TRACE) tcommit(slot, g, h, nops);
}

} // namespace dev

// ================================================================== host

namespace {

using dev::Bufs;
using dev::IMap;

struct Pass {
    int kind; // 0 tiny, 1 local, 2 global, 3 merge
    int p, q_lo, s;
};

// Pass schedule for n_log = log2(n_pad) with tiles of 2^t_log and at most
// s_max fused global stages.  Depends on public sizes only.
std::vector<Pass> plan_passes(int n_log, int t_log, int s_max, int r_log)
{
    std::vector<Pass> v;
    if (n_log < r_log) {
        v.push_back({0, n_log, 0, 0});
        return v;
    }
    const int t = std::min(n_log, t_log);
    v.push_back({1, t, 0, t});
    for (int p = t_log + 1; p <= n_log; ++p) {
        int q = p - 1;
        while (q >= t_log) {
            int s = std::min(s_max, q - t_log + 1);
            v.push_back({2, p, q - s + 1, s});
            q -= s;
        }
        v.push_back({3, p, 0, 0});
    }
    return v;
}

template <class R, int T_LOG, int R_LOG, int BLOCK, int S_MAX>
struct SortImpl {
    template <bool TRACE>
    static void global_pass(vo_ctx *ctx, Bufs<R> b, IMap map, int p, int q_lo, int s, int p_top, uint64_t n_virtual)
    {
        const uint64_t threads = n_virtual >> s;
        const uint32_t block = (uint32_t)std::min<uint64_t>(256, threads);
        VO_CHECK(block == 256, VO_ERR_INVALID, "global pass below one block");
        LaunchInfo li{"sort_global", dim3((uint32_t)(threads / 256)), dim3(256), 0,
                      hash_mix(hash_mix(p, q_lo), hash_mix(s, p_top))};
        TraceSlot *slot = trace_slot(ctx);
        switch (s) {
        case 1: launch(ctx, li, dev::k_sort_global<R, 1, TRACE>, b, map, p, q_lo, p_top, slot); break;
        case 2: launch(ctx, li, dev::k_sort_global<R, 2, TRACE>, b, map, p, q_lo, p_top, slot); break;
        case 3:
            if constexpr (S_MAX >= 3) {
                launch(ctx, li, dev::k_sort_global<R, 3, TRACE>, b, map, p, q_lo, p_top, slot);
                break;
            }
        case 4:
            if constexpr (S_MAX >= 4) {
                launch(ctx, li, dev::k_sort_global<R, 4, TRACE>, b, map, p, q_lo, p_top, slot);
                break;
            }
        case 5:
            if constexpr (S_MAX >= 5) {
                launch(ctx, li, dev::k_sort_global<R, 5, TRACE>, b, map, p, q_lo, p_top, slot);
                break;
            }
        default: throw Error{VO_ERR_INVALID, "bad fused stage count"};
        }
    }

    template <bool TRACE>
    static void run_t(vo_ctx *ctx, Bufs<R> b, int n_log, uint64_t n_valid, int fill, IMap map, int p_top,
                      uint64_t n_virtual)
    {
        const int s_max = std::max(1, std::min(S_MAX, ctx->tuning.max_fused_stages));
        for (const Pass &ps : plan_passes(n_log, T_LOG, s_max, R_LOG)) {
            switch (ps.kind) {
            case 0: {
                LaunchInfo li{"sort_tiny", dim3(1), dim3(64), 0, hash_mix(n_log, fill)};
                launch(ctx, li, dev::k_sort_tiny<R, TRACE>, b, n_log, n_valid, fill, trace_slot(ctx));
                break;
            }
            case 1: {
                const int t = ps.p;
                const size_t lds = dev::lds_bytes<R>(t);
                LaunchInfo li{"sort_local", dim3((uint32_t)(n_virtual >> t)), dim3(BLOCK), (uint32_t)lds,
                              hash_mix(hash_mix(t, p_top), fill)};
                auto k = dev::k_sort_local<R, T_LOG, R_LOG, BLOCK, TRACE>;
                VO_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                launch(ctx, li, k, b, map, t, std::min(t, p_top), p_top, n_valid, fill, trace_slot(ctx));
                break;
            }
            case 2: global_pass<TRACE>(ctx, b, map, ps.p, ps.q_lo, ps.s, p_top, n_virtual); break;
            case 3: {
                const size_t lds = dev::lds_bytes<R>(T_LOG);
                LaunchInfo li{"sort_merge", dim3((uint32_t)(n_virtual >> T_LOG)), dim3(BLOCK), (uint32_t)lds,
                              hash_mix(ps.p, p_top)};
                auto k = dev::k_sort_merge<R, T_LOG, R_LOG, BLOCK, TRACE>;
                VO_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                launch(ctx, li, k, b, map, ps.p, p_top, trace_slot(ctx));
                break;
            }
            }
        }
    }

    static void run(vo_ctx *ctx, Bufs<R> b, int n_log, uint64_t n_valid, int fill)
    {
        IMap map{nullptr, 0};
        if (ctx->tracing)
            run_t<true>(ctx, b, n_log, n_valid, fill, map, n_log, 1ull << n_log);
        else
            run_t<false>(ctx, b, n_log, n_valid, fill, map, n_log, 1ull << n_log);
    }

    static int passes(vo_ctx *ctx, int n_log)
    {
        const int s_max = std::max(1, std::min(S_MAX, ctx->tuning.max_fused_stages));
        return (int)plan_passes(n_log, T_LOG, s_max, R_LOG).size();
    }
};

// Record kinds and their tile shapes (LDS = 2^T_LOG * 17/16 * record bytes).
using RecKP = dev::Rec<1, uint32_t, true>;  // 12 B: u64 key + u32 payload
using RecKV = dev::Rec<1, uint64_t, true>;  // 16 B: u64 key + u64 value
using RecM1 = dev::Rec<1, int64_t, false>;  // multi-column, packed key words + index
using RecM2 = dev::Rec<2, int64_t, false>;
using RecM4 = dev::Rec<4, int64_t, false>;
using RecM8 = dev::Rec<8, int64_t, false>;

using SortKP = SortImpl<RecKP, 13, 4, 512, 5>;
using SortKV = SortImpl<RecKV, 12, 4, 256, 4>;
using SortM1 = SortImpl<RecM1, 12, 4, 256, 4>;
using SortM2 = SortImpl<RecM2, 12, 4, 256, 3>;
using SortM4 = SortImpl<RecM4, 11, 3, 256, 2>;
using SortM8 = SortImpl<RecM8, 10, 2, 256, 2>;

// ------------------------------------------------- multi-column (OP_SORT)

// Order-preserving unsigned image of a column value.  DESC columns are
// bit-inverted, so lexicographic ascending order of the packed words is the
// reference's column-by-column order with per-column direction
// (sort.cuh:31-40).  FLOAT/DOUBLE map -0.0 to +0.0 (the reference treats
// them as equal); NaN is not supported.
__device__ __forceinline__ uint64_t load_raw(const void *col, uint32_t type, uint64_t i)
{
    switch (type) {
    case VO_T_BOOLEAN:
    case VO_T_TINYINT: return ((const uint8_t *)col)[i];
    case VO_T_SMALLINT: return ((const uint16_t *)col)[i];
    case VO_T_INTEGER:
    case VO_T_FLOAT:
    case VO_T_DATE: return ((const uint32_t *)col)[i];
    default: return ((const uint64_t *)col)[i];
    }
}

// Order-preserving unsigned image of a column value given its raw bits.
__device__ __forceinline__ uint64_t norm_bits(uint32_t type, uint64_t u)
{
    switch (type) {
    case VO_T_TINYINT: return (u ^ 0x80u) & 0xFFu;
    case VO_T_SMALLINT: return (u ^ 0x8000u) & 0xFFFFu;
    case VO_T_INTEGER: return (u ^ 0x80000000u) & 0xFFFFFFFFu;
    case VO_T_BIGINT: return u ^ 0x8000000000000000ull;
    case VO_T_FLOAT: {
        uint32_t f = (uint32_t)u;
        f = (f == 0x80000000u) ? 0u : f;
        return (f & 0x80000000u) ? (uint32_t)~f : (f | 0x80000000u);
    }
    case VO_T_DOUBLE: {
        u = (u == 0x8000000000000000ull) ? 0ull : u;
        return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    }
    default: return u; // BOOLEAN, VARCHAR, DATE, UBIGINT: already unsigned
    }
}

__device__ __forceinline__ void denorm_store(void *col, uint32_t type, uint64_t i, uint64_t v)
{
    switch (type) {
    case VO_T_BOOLEAN: ((uint8_t *)col)[i] = (uint8_t)v; break;
    case VO_T_TINYINT: ((int8_t *)col)[i] = (int8_t)((uint8_t)v ^ 0x80); break;
    case VO_T_SMALLINT: ((int16_t *)col)[i] = (int16_t)((uint16_t)v ^ 0x8000); break;
    case VO_T_INTEGER: ((int32_t *)col)[i] = (int32_t)((uint32_t)v ^ 0x80000000u); break;
    case VO_T_BIGINT: ((int64_t *)col)[i] = (int64_t)(v ^ 0x8000000000000000ull); break;
    case VO_T_FLOAT: {
        uint32_t u = (uint32_t)v;
        u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
        ((uint32_t *)col)[i] = u;
        break;
    }
    case VO_T_DOUBLE: {
        uint64_t u = v;
        u = (u & 0x8000000000000000ull) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u;
        ((uint64_t *)col)[i] = u;
        break;
    }
    case VO_T_VARCHAR:
    case VO_T_UBIGINT: ((uint64_t *)col)[i] = v; break;
    case VO_T_DATE: ((uint32_t *)col)[i] = (uint32_t)v; break;
    default: break;
    }
}

struct PackSpec {
    const void *cols[VO_MAX_SORT_COLUMNS];
    uint32_t types[VO_MAX_SORT_COLUMNS];
    uint32_t bits[VO_MAX_SORT_COLUMNS];  // width of the column image
    uint32_t word[VO_MAX_SORT_COLUMNS];  // destination word
    uint32_t shift[VO_MAX_SORT_COLUMNS]; // left shift inside the word
    uint32_t desc[VO_MAX_SORT_COLUMNS];
    uint64_t fill[VO_MAX_SORT_COLUMNS];  // raw bits of the pad value
    uint32_t ncols;
    uint32_t nwords;  // words holding columns
    uint32_t nwrite;  // words written (>= nwords; the rest are zero)
};

// Packs the normalised columns into key words (MSB-first per word) and
// writes index = row.  Pad rows take the image of max for ASC columns and
// of lowest for DESC columns, the values sort.cuh:79-100 writes.
__global__ void k_pack_keys(PackSpec ps, uint64_t *w0, uint64_t *w1, uint64_t *w2, uint64_t *w3, uint64_t *w4,
                            uint64_t *w5, uint64_t *w6, uint64_t *w7, int64_t *idx, uint64_t n, uint64_t n_pad)
{
    uint64_t *ws[8] = {w0, w1, w2, w3, w4, w5, w6, w7};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const bool pad = i >= n;
        const uint64_t src = pad ? 0 : i;
        for (uint32_t c = 0; c < ps.ncols; ++c) {
            const uint64_t mask = ps.bits[c] == 64 ? ~0ull : ((1ull << ps.bits[c]) - 1);
            const uint64_t raw = load_raw(ps.cols[c], ps.types[c], src);
            uint64_t v = norm_bits(ps.types[c], pad ? ps.fill[c] : raw);
            v = ps.desc[c] ? (~v & mask) : v;
            acc[ps.word[c]] |= v << ps.shift[c];
        }
        for (uint32_t w = 0; w < ps.nwrite; ++w) ws[w][i] = acc[w];
        idx[i] = (int64_t)i;
    }
}

__global__ void k_unpack_keys(PackSpec ps, const uint64_t *w0, const uint64_t *w1, const uint64_t *w2,
                              const uint64_t *w3, const uint64_t *w4, const uint64_t *w5, const uint64_t *w6,
                              const uint64_t *w7, uint64_t n_pad)
{
    const uint64_t *ws[8] = {w0, w1, w2, w3, w4, w5, w6, w7};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
        for (uint32_t c = 0; c < ps.ncols; ++c) {
            const uint64_t mask = ps.bits[c] == 64 ? ~0ull : ((1ull << ps.bits[c]) - 1);
            uint64_t v = (ws[ps.word[c]][i] >> ps.shift[c]) & mask;
            v = ps.desc[c] ? (~v & mask) : v;
            denorm_store(const_cast<void *>(ps.cols[c]), ps.types[c], i, v);
        }
    }
}

uint32_t type_bits(uint32_t t)
{
    switch (t) {
    case VO_T_BOOLEAN: case VO_T_TINYINT: return 8;
    case VO_T_SMALLINT: return 16;
    case VO_T_INTEGER: case VO_T_FLOAT: case VO_T_DATE: return 32;
    case VO_T_BIGINT: case VO_T_DOUBLE: case VO_T_VARCHAR: case VO_T_UBIGINT: return 64;
    default: return 0;
    }
}

// Raw bits of numeric_limits<T>::max() (ASC) / lowest() (DESC) for the
// column's device type (include/data/type_traits.hpp:10-19).
uint64_t pad_raw(uint32_t t, bool asc)
{
    switch (t) {
    case VO_T_BOOLEAN: return asc ? 0xFFu : 0u;
    case VO_T_TINYINT: return asc ? 0x7Fu : 0x80u;
    case VO_T_SMALLINT: return asc ? 0x7FFFu : 0x8000u;
    case VO_T_INTEGER: return asc ? 0x7FFFFFFFu : 0x80000000u;
    case VO_T_BIGINT: return asc ? 0x7FFFFFFFFFFFFFFFull : 0x8000000000000000ull;
    case VO_T_FLOAT: return asc ? 0x7F7FFFFFu : 0xFF7FFFFFu;
    case VO_T_DOUBLE: return asc ? 0x7FEFFFFFFFFFFFFFull : 0xFFEFFFFFFFFFFFFFull;
    case VO_T_DATE: return asc ? 0xFFFFFFFFu : 0u;
    default: return asc ? ~0ull : 0ull; // VARCHAR ordinal, UBIGINT
    }
}

} // namespace

// ------------------------------------------------------------ entry points

void op_sort_kp(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n, uint64_t n_pad)
{
    VO_CHECK(is_pow2(n_pad) && n <= n_pad, VO_ERR_INVALID, "n_pad must be a power of two >= n");
    VO_CHECK(n_pad <= (1ull << 32), VO_ERR_INVALID, "n_pad above 2^32 (u32 payload pad index)");
    VO_CHECK(keys && payload, VO_ERR_INVALID, "null buffer");
    Bufs<RecKP> b;
    b.k[0] = keys;
    b.p = payload;
    SortKP::run(ctx, b, ilog2(n_pad), n, 1);
}

void sort_kp_full(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n_pad)
{
    Bufs<RecKP> b;
    b.k[0] = keys;
    b.p = payload;
    SortKP::run(ctx, b, ilog2(n_pad), n_pad, 0);
}

void sort_kv(vo_ctx *ctx, uint64_t *keys, uint64_t *vals, uint64_t n_pad)
{
    Bufs<RecKV> b;
    b.k[0] = keys;
    b.p = vals;
    SortKV::run(ctx, b, ilog2(n_pad), n_pad, 0);
}

int sort_plan_passes(vo_ctx *ctx, uint64_t n_pad) { return SortKP::passes(ctx, ilog2(n_pad)); }

void op_sort_multi(vo_ctx *ctx, const vo_sort_args &a)
{
    VO_CHECK(a.n_sort_columns >= 1 && a.n_sort_columns <= VO_MAX_SORT_COLUMNS, VO_ERR_INVALID,
             "n_sort_columns must be 1..8");
    VO_CHECK(is_pow2(a.n_padded_rows) && a.n_rows <= a.n_padded_rows, VO_ERR_INVALID,
             "n_padded_rows must be a power of two >= n_rows");
    VO_CHECK(a.indices != nullptr, VO_ERR_INVALID, "null indices");
    PackSpec ps{};
    ps.ncols = (uint32_t)a.n_sort_columns;
    uint32_t word = 0, used = 0;
    for (uint32_t c = 0; c < ps.ncols; ++c) {
        const uint32_t bits = type_bits(a.col_types[c]);
        VO_CHECK(bits != 0, VO_ERR_UNSUPPORTED, "unsupported sort column type");
        VO_CHECK(a.sort_cols[c] != nullptr, VO_ERR_INVALID, "null sort column");
        if (used + bits > 64) {
            ++word;
            used = 0;
        }
        ps.cols[c] = a.sort_cols[c];
        ps.types[c] = a.col_types[c];
        ps.bits[c] = bits;
        ps.word[c] = word;
        ps.desc[c] = a.ascending_flags[c] ? 0 : 1;
        ps.fill[c] = pad_raw(a.col_types[c], a.ascending_flags[c] != 0);
        used += bits;
    }
    ps.nwords = word + 1;
    // shifts: MSB-first inside each word
    for (uint32_t w = 0; w < ps.nwords; ++w) {
        uint32_t left = 64;
        for (uint32_t c = 0; c < ps.ncols; ++c)
            if (ps.word[c] == w) {
                left -= ps.bits[c];
                ps.shift[c] = left;
            }
    }
    const uint32_t W = ps.nwords <= 1 ? 1 : ps.nwords <= 2 ? 2 : ps.nwords <= 4 ? 4 : 8;
    ps.nwrite = W;
    const uint64_t n_pad = a.n_padded_rows;
    uint64_t *words = (uint64_t *)scratch(ctx, sizeof(uint64_t) * W * n_pad);
    uint64_t *wp[8];
    for (int w = 0; w < 8; ++w) wp[w] = words + (uint64_t)std::min<uint32_t>(w, W - 1) * n_pad;
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n_pad + 255) / 256, 8192));
    LaunchInfo lp{"sort_pack", dim3(blocks), dim3(256), 0, hash_mix(ps.ncols, ps.nwords)};
    launch(ctx, lp, k_pack_keys, ps, wp[0], wp[1], wp[2], wp[3], wp[4], wp[5], wp[6], wp[7], a.indices, a.n_rows,
           n_pad);
    const int n_log = ilog2(n_pad);
    switch (W) {
    case 1: {
        Bufs<RecM1> b;
        b.k[0] = wp[0];
        b.p = a.indices;
        SortM1::run(ctx, b, n_log, n_pad, 0);
        break;
    }
    case 2: {
        Bufs<RecM2> b;
        for (int w = 0; w < 2; ++w) b.k[w] = wp[w];
        b.p = a.indices;
        SortM2::run(ctx, b, n_log, n_pad, 0);
        break;
    }
    case 4: {
        Bufs<RecM4> b;
        for (int w = 0; w < 4; ++w) b.k[w] = wp[w];
        b.p = a.indices;
        SortM4::run(ctx, b, n_log, n_pad, 0);
        break;
    }
    default: {
        Bufs<RecM8> b;
        for (int w = 0; w < 8; ++w) b.k[w] = wp[w];
        b.p = a.indices;
        SortM8::run(ctx, b, n_log, n_pad, 0);
        break;
    }
    }
    LaunchInfo lu{"sort_unpack", dim3(blocks), dim3(256), 0, hash_mix(ps.ncols, ps.nwords)};
    launch(ctx, lu, k_unpack_keys, ps, wp[0], wp[1], wp[2], wp[3], wp[4], wp[5], wp[6], wp[7], n_pad);
}

} // namespace vo
