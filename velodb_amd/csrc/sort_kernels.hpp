// sort_kernels.hpp -- oblivious bitonic sort kernels and their pass schedule.
//
// Reference: executeBitonicSort (include/cuda/sort.cuh:70-118) runs one
// grid-wide pass per (k, j) stage, S = log2N(log2N+1)/2 passes.  Here the
// same compare-exchange network is grouped into far fewer HBM passes:
//   pass 0       k_sort_local_ct : every tile of 2^T_LOG records fully
//                                  sorted (levels p <= T_LOG), registers+LDS;
//   per level p > T_LOG:
//                k_sort_global   : stages q >= T_LOG, up to S_MAX per pass;
//                k_sort_merge    : stages q < T_LOG in one tile pass.
// The schedule is a function of n_pad only.
//
// Addressing: every block's records sit at one uniform (scalar) base plus a
// per-thread offset below 2^13, so loads and stores use the scalar-base +
// 32-bit vector-offset form and the address costs one VGPR.  Batched sorts
// map a block's uniform virtual base through a per-instance offset table
// (instances are never smaller than a tile on these kernels).
#pragma once

#include "sort_core.hpp"

#include <algorithm>

namespace vo {
namespace dev {

// Uniform virtual index -> physical index (identity without a table).  The
// table entry is made wave-uniform explicitly: where the compiler cannot
// prove the table unwritten (the resident queue kernel's bodies) it would
// load it into vector registers, and every buffer access based on it would
// become a loop over lanes.
__device__ __forceinline__ uint64_t map_uniform(const IMap &m, uint64_t v)
{
    if (m.offsets == nullptr) return v;
    const uint64_t o = m.offsets[v >> m.c];
    const uint64_t ou = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(o >> 32)) << 32) |
                        __builtin_amdgcn_readfirstlane((uint32_t)o);
    return ou + (v & ((1ull << m.c) - 1));
}

// Cache policy of the passes' buffer loads and stores (the instructions'
// aux field): 0 for arrays that fit the 256 MB MALL, where the next pass
// re-reads them from it; kCpStream (non-temporal) for larger ones, which
// every pass streams through once.
constexpr int kCpStream = 2;

// Raw buffer access: the uniform base goes into a buffer resource (scalar
// registers), the per-thread part is a 32-bit byte offset.  Dword 3 of the
// resource selects the 32-bit data format (0x00020000) used for untyped
// loads on gfx9-family parts.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *ubase)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(ubase), (short)0, (int)0xFFFFFFFFu, 0x00020000);
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <int CP = 0, typename T>
__device__ __forceinline__ T ld_u(const T *ubase, uint32_t vi)
{
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(ubase);
    if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, vi * 8u, 0, CP));
    } else {
        static_assert(sizeof(T) == 4, "4- or 8-byte columns");
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vi * 4u, 0, CP));
    }
}

template <int CP = 0, typename T>
__device__ __forceinline__ void st_u(T *ubase, uint32_t vi, T v)
{
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(ubase);
    if constexpr (sizeof(T) == 8) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, vi * 8u, 0, CP);
    } else {
        static_assert(sizeof(T) == 4, "4- or 8-byte columns");
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vi * 4u, 0, CP);
    }
}

// Record at uniform index u plus per-thread offset vi.
template <class R, bool TRACE, int CP = 0>
__device__ __forceinline__ void gload_u(R &r, const Bufs<R> &b, uint64_t u, uint32_t vi, uint64_t &h, uint32_t &n)
{
#pragma unroll
    for (int w = 0; w < R::KW; ++w) r.k[w] = ld_u<CP>(b.k[w] + u, vi);
    r.p = ld_u<CP>(b.p + u, vi);
    if constexpr (TRACE) {
        h = tfold(h, 1, u + vi);
        ++n;
    }
}

template <class R, bool TRACE, int CP = 0>
__device__ __forceinline__ void gstore_u(const R &r, const Bufs<R> &b, uint64_t u, uint32_t vi, uint64_t &h,
                                         uint32_t &n)
{
#pragma unroll
    for (int w = 0; w < R::KW; ++w) st_u<CP>(b.k[w] + u, vi, r.k[w]);
    st_u<CP>(b.p + u, vi, r.p);
    if constexpr (TRACE) {
        h = tfold(h, 2, u + vi);
        ++n;
    }
}


// Tile form: one buffer resource per column for the tile's uniform base u,
// the per-register offset so (elements, uniform, < 2^16) in the scalar
// offset field and the per-thread offset vi in the vector offset.
template <typename T, int CP = 0>
__device__ __forceinline__ T ld_t(__amdgpu_buffer_rsrc_t r, uint32_t vi, uint32_t so)
{
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, vi * 8u, so * 8u, CP));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vi * 4u, so * 4u, CP));
}

template <typename T, int CP = 0>
__device__ __forceinline__ void st_t(__amdgpu_buffer_rsrc_t r, uint32_t vi, uint32_t so, T v)
{
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, vi * 8u, so * 8u, CP);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vi * 4u, so * 4u, CP);
}

template <class R, bool TRACE, int CP = 0>
__device__ __forceinline__ void gload_t(R &r, const Bufs<R> &b, uint64_t u, uint32_t so, uint32_t vi, uint64_t &h,
                                        uint32_t &n)
{
#pragma unroll
    for (int w = 0; w < R::KW; ++w) r.k[w] = ld_t<uint64_t, CP>(buf_rsrc(b.k[w] + u), vi, so);
    r.p = ld_t<typename R::P, CP>(buf_rsrc(b.p + u), vi, so);
    if constexpr (TRACE) {
        h = tfold(h, 1, u + so + vi);
        ++n;
    }
}

template <class R, bool TRACE, int CP = 0>
__device__ __forceinline__ void gstore_t(const R &r, const Bufs<R> &b, uint64_t u, uint32_t so, uint32_t vi,
                                         uint64_t &h, uint32_t &n)
{
#pragma unroll
    for (int w = 0; w < R::KW; ++w) st_t<uint64_t, CP>(buf_rsrc(b.k[w] + u), vi, so, r.k[w]);
    st_t<typename R::P, CP>(buf_rsrc(b.p + u), vi, so, r.p);
    if constexpr (TRACE) {
        h = tfold(h, 2, u + so + vi);
        ++n;
    }
}

// ----------------------------------------------------------- tiny arrays

// n_pad <= 8: one thread runs the network directly.
// Each kernel is a device body taking its block index explicitly plus a thin
// __global__ wrapper; the resident queue kernel (queue.hip) runs the same
// bodies over its own block loop.

template <class R, bool TRACE>
__device__ __forceinline__ void body_tiny(Bufs<R> b, int n_log, uint64_t n_valid, int fill, TraceSlot *slot,
                                          uint32_t blk)
{
    if (threadIdx.x != 0 || blk != 0) return;
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
                if (i < ixj) cmpx_scalar(r[i], r[ixj], dir_asc(i, p, n_log));
            }
    for (uint32_t i = 0; i < 8; ++i)
        if (i < N) gstore<R, TRACE>(r[i], b, i, h, nops);
    if constexpr (TRACE) tcommit(slot, 0, h, nops);
}

template <class R, bool TRACE>
__global__ void k_sort_tiny(Bufs<R> b, int n_log, uint64_t n_valid, int fill, TraceSlot *slot)
{
    body_tiny<R, TRACE>(b, n_log, n_valid, fill, slot, blockIdx.x);
}

// ---------------------------------------------------------- tile kernels

// Tiles of 2^t records for t < T_LOG (small arrays): levels 1 .. p_max with
// runtime layouts.  Loads and stores use the coalesced layout a = t - R_LOG.
template <class R, int T_LOG, int R_LOG, int BLOCK, bool TRACE>
__device__ __forceinline__ void body_local(Bufs<R> b, IMap map, int t, int p_max, int p_top, uint64_t n_valid,
                                           int fill, TraceSlot *slot, char *lds_raw, uint32_t blk)
{
    const uint32_t T = 1u << t;
    const Lds<R> s = lds_view<R>(lds_raw, T + (T >> 4));
    const uint32_t tau = threadIdx.x;
    const bool act = tau < (T >> R_LOG);
    const uint64_t vbase = (uint64_t)blk << t;
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
            if (fill) pad_select(r[m], local_row(map, v), live_rows(map, v, n_valid));
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
        for (int m = 0; m < (1 << R_LOG); ++m)
            gstore<R, TRACE>(r[m], b, phys(map, vbase + (base | ((uint32_t)m << a))), h, nops);
    }
    if constexpr (TRACE) tcommit(slot, (uint64_t)blk * BLOCK + tau, h, nops);
}

template <class R, int T_LOG, int R_LOG, int BLOCK, bool TRACE>
__global__ __launch_bounds__(BLOCK) void k_sort_local(Bufs<R> b, IMap map, int t, int p_max, int p_top, uint64_t n_valid,
                                                      int fill, TraceSlot *slot)
{
    extern __shared__ __align__(16) char lds_raw[];
    body_local<R, T_LOG, R_LOG, BLOCK, TRACE>(b, map, t, p_max, p_top, n_valid, fill, slot, lds_raw, blockIdx.x);
}

// Full 2^T_LOG tiles: all levels p = 1 .. T_LOG with a compile-time schedule.
template <class R, int T_LOG, int R_LOG, int BLOCK, bool TRACE, int CP = 0>
__device__ __forceinline__ void body_local_ct(Bufs<R> b, IMap map, int p_top, uint64_t n_valid, int fill,
                                              TraceSlot *slot, char *lds, uint32_t blk)
{
    using TL = TileCT<R, T_LOG, R_LOG, TRACE>;
    constexpr int A0 = T_LOG - R_LOG;
    const uint32_t tau = threadIdx.x;
    const uint64_t vbase = (uint64_t)blk << T_LOG;
    const uint64_t pbase = map_uniform(map, vbase);
    uint64_t h = 0;
    uint32_t nops = 0;
    R r[1 << R_LOG];
    const uint32_t base = lay_base<R_LOG>(tau, A0);
#pragma unroll
    for (int m = 0; m < (1 << R_LOG); ++m) {
        gload_t<R, TRACE, CP>(r[m], b, pbase, (uint32_t)m << A0, base, h, nops);
        if (fill) {
            const uint64_t v = vbase + (base | ((uint32_t)m << A0));
            pad_select(r[m], local_row(map, v), live_rows(map, vbase, n_valid));
        }
    }
    const bool asc_hi = dir_asc(vbase, T_LOG, p_top);
    TL::template levels<1, A0>(r, lds, tau, asc_hi, p_top, h, nops);
    TL::template relayout<TL::end_all(1, A0), A0>(r, lds, tau, h, nops);
#pragma unroll
    for (int m = 0; m < (1 << R_LOG); ++m) gstore_t<R, TRACE, CP>(r[m], b, pbase, (uint32_t)m << A0, base, h, nops);
    if constexpr (TRACE) tcommit(slot, (uint64_t)blk * BLOCK + tau, h, nops);
}

// 1024-thread blocks: two per CU, 8 waves per SIMD (at most 64 VGPRs; the
// second bound is the minimum waves per execution unit)
template <class R, int T_LOG, int R_LOG, int BLOCK, bool TRACE, int CP = 0>
__global__ __launch_bounds__(BLOCK, BLOCK >= 1024 ? 8 : 1) void k_sort_local_ct(Bufs<R> b, IMap map, int p_top,
                                                                                uint64_t n_valid, int fill,
                                                                                TraceSlot *slot)
{
    extern __shared__ __align__(16) char lds[];
    body_local_ct<R, T_LOG, R_LOG, BLOCK, TRACE, CP>(b, map, p_top, n_valid, fill, slot, lds, blockIdx.x);
}

// Stages (p, T_LOG-1 .. 0) of one level p > T_LOG on tiles of 2^T_LOG.
template <class R, int T_LOG, int R_LOG, int BLOCK, bool TRACE, int CP = 0>
__device__ __forceinline__ void body_merge(Bufs<R> b, IMap map, int p, int p_top, TraceSlot *slot, char *lds,
                                           uint32_t blk)
{
    using TL = TileCT<R, T_LOG, R_LOG, TRACE>;
    constexpr int A0 = T_LOG - R_LOG;
    const uint32_t tau = threadIdx.x;
    const uint64_t vbase = (uint64_t)blk << T_LOG;
    const uint64_t pbase = map_uniform(map, vbase);
    uint64_t h = 0;
    uint32_t nops = 0;
    R r[1 << R_LOG];
    const uint32_t base = lay_base<R_LOG>(tau, A0);
#pragma unroll
    for (int m = 0; m < (1 << R_LOG); ++m) gload_t<R, TRACE, CP>(r[m], b, pbase, (uint32_t)m << A0, base, h, nops);
    const bool asc = dir_asc(vbase, p, p_top);
    TL::template level<-1, T_LOG - 1, A0>(r, lds, tau, asc, p_top, h, nops);
    TL::template relayout<TL::end_layout(T_LOG - 1, A0), A0>(r, lds, tau, h, nops);
#pragma unroll
    for (int m = 0; m < (1 << R_LOG); ++m) gstore_t<R, TRACE, CP>(r[m], b, pbase, (uint32_t)m << A0, base, h, nops);
    if constexpr (TRACE) tcommit(slot, (uint64_t)blk * BLOCK + tau, h, nops);
}

template <class R, int T_LOG, int R_LOG, int BLOCK, bool TRACE, int CP = 0>
__global__ __launch_bounds__(BLOCK) void k_sort_merge(Bufs<R> b, IMap map, int p, int p_top, TraceSlot *slot)
{
    extern __shared__ __align__(16) char lds[];
    body_merge<R, T_LOG, R_LOG, BLOCK, TRACE, CP>(b, map, p, p_top, slot, lds, blockIdx.x);
}

// The same merge with resident blocks: each block walks tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... and issues the HBM loads of its
// next tile before working on the current one, so the loads overlap the LDS
// and compare-exchange phase of the tile in hand.
template <class R, int T_LOG, int R_LOG, int BLOCK, bool TRACE>
__global__ __launch_bounds__(BLOCK) void k_sort_merge_pipe(Bufs<R> b, IMap map, int p, int p_top, uint32_t n_tiles,
                                                           TraceSlot *slot)
{
    using TL = TileCT<R, T_LOG, R_LOG, TRACE>;
    constexpr int A0 = T_LOG - R_LOG;
    constexpr int RR = 1 << R_LOG;
    extern __shared__ __align__(16) char lds[];
    const uint32_t tau = threadIdx.x;
    const uint32_t base = lay_base<R_LOG>(tau, A0);
    uint64_t h = 0;
    uint32_t nops = 0;
    R cur[RR], nxt[RR];
    uint32_t t = blockIdx.x;
    if (t < n_tiles) {
        const uint64_t pb = map_uniform(map, (uint64_t)t << T_LOG);
#pragma unroll
        for (int m = 0; m < RR; ++m) gload_t<R, TRACE>(cur[m], b, pb, (uint32_t)m << A0, base, h, nops);
    }
    for (; t < n_tiles; t += gridDim.x) {
        const uint32_t tn = t + gridDim.x;
        if (tn < n_tiles) {
            const uint64_t pn = map_uniform(map, (uint64_t)tn << T_LOG);
#pragma unroll
            for (int m = 0; m < RR; ++m) gload_t<R, TRACE>(nxt[m], b, pn, (uint32_t)m << A0, base, h, nops);
        }
        const uint64_t vbase = (uint64_t)t << T_LOG;
        const uint64_t pbase = map_uniform(map, vbase);
        const bool asc = dir_asc(vbase, p, p_top);
        TL::template level<-1, T_LOG - 1, A0>(cur, lds, tau, asc, p_top, h, nops);
        TL::template relayout<TL::end_layout(T_LOG - 1, A0), A0>(cur, lds, tau, h, nops);
#pragma unroll
        for (int m = 0; m < RR; ++m) gstore_t<R, TRACE>(cur[m], b, pbase, (uint32_t)m << A0, base, h, nops);
#pragma unroll
        for (int m = 0; m < RR; ++m) cur[m] = nxt[m];
    }
    if constexpr (TRACE) tcommit(slot, (uint64_t)blockIdx.x * BLOCK + tau, h, nops);
}

// ------------------------------------------------------------ hybrid pass

// S fused stages (p, q_lo+S-1 .. q_lo) with S up to 7, staged through LDS.
// A block holds 2^T_LOG records: 2^S runs of C = T_LOG - S contiguous
// records, run r starting at element ubase + r * 2^q_lo.  Tile-local index l
// maps to element (l mod 2^C) + (l >> C) * 2^q_lo (plus the block's uniform
// bits), so tile bits [C, T_LOG) are the stage bits and the stages run as a
// tile schedule on them.  With C >= 6 a wavefront's 64 lanes always read
// and write 64 consecutive records.
template <class R, int T_LOG, int R_LOG, int BLOCK, int S, bool TRACE, int CP = 0>
__device__ __forceinline__ void body_hybrid(Bufs<R> b, IMap map, int p, int q_lo, int p_top, TraceSlot *slot,
                                            char *lds, uint32_t blk32)
{
    using TL = TileCT<R, T_LOG, R_LOG, TRACE>;
    constexpr int C = T_LOG - S;
    constexpr int A0 = T_LOG - R_LOG;
    static_assert(C >= 5 && A0 >= 6, "runs of at least 32 records; lanes cover 64 tile-local records");
    const uint32_t tau = threadIdx.x;
    const uint32_t lane = tau & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tau >> 6);
    const uint64_t blk = blk32;
    const uint64_t b_lo = blk & ((1ull << (q_lo - C)) - 1);
    const uint64_t b_hi = blk >> (q_lo - C);
    const uint64_t ublk = (b_lo << C) | (b_hi << (q_lo + S));
    // element offset of tile-local index l (a bit permutation, additive over disjoint bit fields)
    auto eoff = [&](uint64_t l) { return (l & ((1ull << C) - 1)) | ((l >> C) << q_lo); };
    const uint64_t vw = ublk + eoff((uint64_t)wave << 6);
    const uint64_t pw = map_uniform(map, vw);
    // per-lane offset: with runs of 64 or more the wavefront is one run; with
    // runs of 32 its upper half is the next run, 2^q_lo records further
    // (below 2^29 elements, so the byte offset fits the 32-bit field)
    const uint32_t vl = C >= 6 ? lane : ((lane & 31u) | ((lane >> 5) << q_lo));
    uint64_t h = 0;
    uint32_t nops = 0;
    R r[1 << R_LOG];
#pragma unroll
    for (int m = 0; m < (1 << R_LOG); ++m) gload_u<R, TRACE, CP>(r[m], b, pw + eoff((uint64_t)m << A0), vl, h, nops);
    const bool asc = dir_asc(ublk, p, p_top);
    TL::template level<-1, T_LOG - 1, A0, C>(r, lds, tau, asc, p_top, h, nops);
    TL::template relayout<TL::end_layout(T_LOG - 1, A0, C), A0>(r, lds, tau, h, nops);
#pragma unroll
    for (int m = 0; m < (1 << R_LOG); ++m) gstore_u<R, TRACE, CP>(r[m], b, pw + eoff((uint64_t)m << A0), vl, h, nops);
    if constexpr (TRACE) tcommit(slot, blk * BLOCK + tau, h, nops);
}

template <class R, int T_LOG, int R_LOG, int BLOCK, int S, bool TRACE, int CP = 0>
__global__ __launch_bounds__(BLOCK) void k_sort_hybrid(Bufs<R> b, IMap map, int p, int q_lo, int p_top,
                                                       TraceSlot *slot)
{
    extern __shared__ __align__(16) char lds[];
    body_hybrid<R, T_LOG, R_LOG, BLOCK, S, TRACE, CP>(b, map, p, q_lo, p_top, slot, lds, blockIdx.x);
}

// ------------------------------------------------------------ global pass

// S fused stages (p, q_lo+S-1 .. q_lo): thread g owns the 2^S records
// base(g) + m * 2^q_lo, base(g) = (g mod 2^q_lo) + (g / 2^q_lo) * 2^(q_lo+S).
// With q_lo >= 8 a block's 256 threads share every bit of base(g) above bit 7,
// so base(g) = ubase + tid and consecutive lanes touch consecutive records:
// each wavefront load or store is one contiguous 512-byte (keys) or 256-byte
// (payload) run.
template <class R, int S, bool TRACE, int CP = 0>
__device__ __forceinline__ void body_global(Bufs<R> b, IMap map, int p, int q_lo, int p_top, TraceSlot *slot,
                                            uint32_t blk, uint32_t tid)
{
    const uint64_t g0 = (uint64_t)blk * 256;
    const uint64_t ubase = (g0 & ((1ull << q_lo) - 1)) | ((g0 >> q_lo) << (q_lo + S));
    const uint64_t pbase = map_uniform(map, ubase);
    uint64_t h = 0;
    uint32_t nops = 0;
    R r[1 << S];
#pragma unroll
    for (int m = 0; m < (1 << S); ++m) gload_t<R, TRACE, CP>(r[m], b, pbase, (uint32_t)m << q_lo, tid, h, nops);
    const bool asc = dir_asc(ubase, p, p_top);
#pragma unroll
    for (int qq = S - 1; qq >= 0; --qq) {
#pragma unroll
        for (int m = 0; m < (1 << S); ++m) {
            if (m & (1 << qq)) continue;
            cmpx(r[m], r[m | (1 << qq)], asc);
        }
    }
#pragma unroll
    for (int m = 0; m < (1 << S); ++m) gstore_t<R, TRACE, CP>(r[m], b, pbase, (uint32_t)m << q_lo, tid, h, nops);
    if constexpr (TRACE) tcommit(slot, g0 + tid, h, nops);
}

template <class R, int S, bool TRACE, int CP = 0>
__global__ __launch_bounds__(256) void k_sort_global(Bufs<R> b, IMap map, int p, int q_lo, int p_top,
                                                     TraceSlot *slot)
{
    body_global<R, S, TRACE, CP>(b, map, p, q_lo, p_top, slot, blockIdx.x, threadIdx.x);
}

} // namespace dev

// ================================================================== host

namespace sortimpl {

using dev::Bufs;
using dev::IMap;

// Largest padded row count: 8-byte columns of 2^29 rows keep every byte
// offset of the buffer-resource addressing below 2^32.
constexpr uint64_t kMaxSortRows = 1ull << 29;

struct Pass {
    int kind; // 0 tiny, 1 local, 2 global, 3 merge
    int p, q_lo, s;
};

// Pass schedule for n_log = log2(n_pad) with tiles of 2^t_log and at most
// s_max fused global stages.  Depends on public sizes only.
// With h_max > s_max, groups larger than s_max run as LDS-staged hybrid
// passes (kind 4) of up to h_max stages; a level's stages above the tile are
// split into the fewest passes, as evenly as possible.
// Width of group gi when `count` stages are split into `groups` passes.
__host__ __device__ constexpr int plan_width(int count, int groups, int gi)
{
    return count / groups + (gi < count % groups ? 1 : 0);
}

// Fewest groups of one level's `count` stages such that every group is a
// register pass (width <= s_max) or a hybrid pass (h_min <= width <= h_max):
// a caller passes the widths it has built.
__host__ __device__ inline int plan_groups(int count, int s_max, int h_max, int h_min)
{
    const int g_max = s_max > h_max ? s_max : h_max;
    for (int groups = (count + g_max - 1) / g_max;; ++groups) {
        bool ok = true;
        for (int gi = 0; gi < groups; ++gi) {
            const int s = plan_width(count, groups, gi);
            if (!(s <= s_max || (s >= h_min && s <= h_max))) ok = false;
        }
        if (ok) return groups;
    }
}

inline std::vector<Pass> plan_passes(int n_log, int t_log, int s_max, int r_log, int h_max = 0)
{
    std::vector<Pass> v;
    if (n_log < r_log) {
        v.push_back({0, n_log, 0, 0});
        return v;
    }
    const int t = std::min(n_log, t_log);
    v.push_back({1, t, 0, t});
    for (int p = t_log + 1; p <= n_log; ++p) {
        const int count = p - t_log;
        const int groups = plan_groups(count, s_max, h_max, 4); // launch path: hybrid widths 4..8
        int q = p - 1;
        for (int gi = 0; gi < groups; ++gi) {
            const int s = plan_width(count, groups, gi);
            v.push_back({s <= s_max ? 2 : 4, p, q - s + 1, s});
            q -= s;
        }
        v.push_back({3, p, 0, 0});
    }
    return v;
}

// S_LARGE / LARGE_LOG: from n_pad = 2^LARGE_LOG up, register passes take at
// most S_LARGE stages (wider groups go to the hybrid pass); below, S_MAX.
// Hybrid passes of at most HW_MAX stages run with twice the threads per block
// and half the records per thread (the same tiles, so the same records per
// block): more waves per SIMD where that pass is faster.
template <class R, int T_LOG, int R_LOG, int BLOCK, int S_MAX, int H_MAX = 0, int S_LARGE = S_MAX, int LARGE_LOG = 64,
          int HW_MAX = 0>
struct SortImpl {
    static_assert(T_LOG >= 8, "global passes need q_lo >= 8 (256-thread blocks)");
    static_assert(H_MAX == 0 || (T_LOG - H_MAX >= 5 && BLOCK << R_LOG == 1 << T_LOG), "hybrid tile shape");
    static_assert(S_LARGE >= 1 && S_LARGE <= S_MAX, "register cap for large sorts");

    static int smax(vo_ctx *ctx, int n_log)
    {
        const int cap = n_log >= LARGE_LOG ? S_LARGE : S_MAX;
        return std::max(1, std::min(cap, ctx->tuning.max_fused_stages));
    }
    static int hmax(vo_ctx *ctx) { return std::min(H_MAX, ctx->tuning.max_hybrid_stages); }

    template <bool TRACE, int S, int RL, int BL, int CP>
    static void hybrid_shape(vo_ctx *ctx, Bufs<R> b, IMap map, int p, int q_lo, int p_top, uint64_t n_virtual)
    {
        using TL = dev::TileCT<R, T_LOG, RL, TRACE>;
        const size_t lds = TL::lds_bytes();
        static const char *const names[] = {"sort_hybrid_s4", "sort_hybrid_s5", "sort_hybrid_s6", "sort_hybrid_s7",
                                            "sort_hybrid_s8"};
        LaunchInfo li{names[S - 4], dim3((uint32_t)(n_virtual >> T_LOG)), dim3(BL), (uint32_t)lds,
                      hash_mix(hash_mix(p, q_lo), hash_mix(S + 100, p_top))};
        auto k = dev::k_sort_hybrid<R, T_LOG, RL, BL, S, TRACE, CP>;
        VO_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        launch(ctx, li, k, b, map, p, q_lo, p_top, trace_slot(ctx));
    }

    template <bool TRACE, int S, int CP>
    static void launch_hybrid(vo_ctx *ctx, Bufs<R> b, IMap map, int p, int q_lo, int p_top, uint64_t n_virtual)
    {
        if constexpr (S <= H_MAX && T_LOG - S >= 5) {
            static_assert(S >= 4 && S <= 8, "hybrid pass widths 4..8");
            // tuning may flip the shape (both are built when the wide one is)
            const bool wide = (S <= HW_MAX) != (HW_MAX > 0 && ctx->tuning.flip_hybrid_block);
            if (wide) {
                // the 6-stage pass on 1024-thread blocks measured faster with
                // the default policy even at 2^28 (5.53 vs 5.13 TB/s)
                constexpr int CPS = S == 6 ? 0 : CP;
                if constexpr (HW_MAX > 0)
                    hybrid_shape<TRACE, S, R_LOG - 1, 2 * BLOCK, CPS>(ctx, b, map, p, q_lo, p_top, n_virtual);
            } else {
                hybrid_shape<TRACE, S, R_LOG, BLOCK, CP>(ctx, b, map, p, q_lo, p_top, n_virtual);
            }
        } else {
            throw Error{VO_ERR_INVALID, "hybrid pass width not built"};
        }
    }

    template <bool TRACE, int S, int CP>
    static void launch_global(vo_ctx *ctx, const LaunchInfo &li, Bufs<R> b, IMap map, int p, int q_lo, int p_top)
    {
        if constexpr (S <= S_MAX)
            launch(ctx, li, dev::k_sort_global<R, S, TRACE, CP>, b, map, p, q_lo, p_top, trace_slot(ctx));
    }

    template <bool TRACE, int CP>
    static void global_pass(vo_ctx *ctx, Bufs<R> b, IMap map, int p, int q_lo, int s, int p_top, uint64_t n_virtual)
    {
        const uint64_t threads = n_virtual >> s;
        VO_CHECK(threads >= 256 && q_lo >= 8, VO_ERR_INVALID, "global pass below one block");
        static const char *const names[] = {"sort_global_s0", "sort_global_s1", "sort_global_s2", "sort_global_s3",
                                            "sort_global_s4", "sort_global_s5"};
        VO_CHECK(s >= 1 && s <= 5, VO_ERR_INVALID, "bad fused stage count");
        LaunchInfo li{names[s], dim3((uint32_t)(threads / 256)), dim3(256), 0,
                      hash_mix(hash_mix(p, q_lo), hash_mix(s, p_top))};
        switch (s) {
        case 1: launch_global<TRACE, 1, CP>(ctx, li, b, map, p, q_lo, p_top); break;
        case 2: launch_global<TRACE, 2, CP>(ctx, li, b, map, p, q_lo, p_top); break;
        case 3: launch_global<TRACE, 3, CP>(ctx, li, b, map, p, q_lo, p_top); break;
        case 4: launch_global<TRACE, 4, CP>(ctx, li, b, map, p, q_lo, p_top); break;
        case 5: launch_global<TRACE, 5, CP>(ctx, li, b, map, p, q_lo, p_top); break;
        default: throw Error{VO_ERR_INVALID, "bad fused stage count"};
        }
    }

    template <bool TRACE, int CP>
    static void run_t(vo_ctx *ctx, Bufs<R> b, int n_log, uint64_t n_valid, int fill, IMap map, int p_top,
                      uint64_t n_virtual)
    {
        for (const Pass &ps : plan_passes(n_log, T_LOG, smax(ctx, n_log), R_LOG, hmax(ctx))) {
            switch (ps.kind) {
            case 0: {
                LaunchInfo li{"sort_tiny", dim3(1), dim3(64), 0, hash_mix(n_log, fill)};
                launch(ctx, li, dev::k_sort_tiny<R, TRACE>, b, n_log, n_valid, fill, trace_slot(ctx));
                break;
            }
            case 1: {
                const int t = ps.p;
                const bool ct = (t == T_LOG && p_top >= T_LOG);
                const size_t lds = ct ? dev::TileCT<R, T_LOG, R_LOG, TRACE>::lds_bytes() : dev::lds_bytes<R>(t);
                LaunchInfo li{"sort_local", dim3((uint32_t)(n_virtual >> t)), dim3(BLOCK), (uint32_t)lds,
                              hash_mix(hash_mix(t, p_top), fill)};
                if (ct) {
                    auto k = dev::k_sort_local_ct<R, T_LOG, R_LOG, BLOCK, TRACE, CP>;
                    VO_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    launch(ctx, li, k, b, map, p_top, n_valid, fill, trace_slot(ctx));
                } else {
                    auto k = dev::k_sort_local<R, T_LOG, R_LOG, BLOCK, TRACE>;
                    VO_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    launch(ctx, li, k, b, map, t, std::min(t, p_top), p_top, n_valid, fill, trace_slot(ctx));
                }
                break;
            }
            case 2: global_pass<TRACE, CP>(ctx, b, map, ps.p, ps.q_lo, ps.s, p_top, n_virtual); break;
            case 4:
                switch (ps.s) {
                case 4: launch_hybrid<TRACE, 4, CP>(ctx, b, map, ps.p, ps.q_lo, p_top, n_virtual); break;
                case 5: launch_hybrid<TRACE, 5, CP>(ctx, b, map, ps.p, ps.q_lo, p_top, n_virtual); break;
                case 6: launch_hybrid<TRACE, 6, CP>(ctx, b, map, ps.p, ps.q_lo, p_top, n_virtual); break;
                case 7: launch_hybrid<TRACE, 7, CP>(ctx, b, map, ps.p, ps.q_lo, p_top, n_virtual); break;
                case 8: launch_hybrid<TRACE, 8, CP>(ctx, b, map, ps.p, ps.q_lo, p_top, n_virtual); break;
                default: throw Error{VO_ERR_INVALID, "bad hybrid stage count"};
                }
                break;
            case 3: {
                const size_t lds = dev::TileCT<R, T_LOG, R_LOG, TRACE>::lds_bytes();
                const uint32_t n_tiles = (uint32_t)(n_virtual >> T_LOG);
                if (ctx->tuning.pipe_merge && !ctx->rec) { // resident: plain merge blocks
                    auto k = dev::k_sort_merge_pipe<R, T_LOG, R_LOG, BLOCK, TRACE>;
                    VO_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    int per_cu = 0;
                    VO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k, BLOCK, lds));
                    const uint32_t grid = std::min<uint32_t>(n_tiles, (uint32_t)std::max(1, per_cu) * ctx->num_cus);
                    LaunchInfo li{"sort_merge", dim3(grid), dim3(BLOCK), (uint32_t)lds, hash_mix(ps.p, p_top)};
                    launch(ctx, li, k, b, map, ps.p, p_top, n_tiles, trace_slot(ctx));
                } else {
                    LaunchInfo li{"sort_merge", dim3(n_tiles), dim3(BLOCK), (uint32_t)lds, hash_mix(ps.p, p_top)};
                    auto k = dev::k_sort_merge<R, T_LOG, R_LOG, BLOCK, TRACE, CP>;
                    VO_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    launch(ctx, li, k, b, map, ps.p, p_top, trace_slot(ctx));
                }
                break;
            }
            }
        }
    }

    // Streaming loads and stores (dev::kCpStream) when the arrays a launch
    // covers exceed the 256 MB MALL, so that no pass re-reads them from it.
    // Measured on the 12-byte sort: +4% at 2^26 and +2.7% at 2^28, but -8% at
    // 2^24, which the MALL holds.  Resident programs keep the default (their
    // bodies are built with it); tuning bit 0x800 flips the choice.
    static bool streaming(vo_ctx *ctx, uint64_t n_virtual)
    {
        constexpr uint64_t kMallBytes = 256ull << 20;
        const bool big = n_virtual * (8 * R::KW + sizeof(typename R::P)) > kMallBytes;
        return !ctx->rec && (big != (ctx->tuning.flip_streaming != 0));
    }

    static void dispatch(vo_ctx *ctx, Bufs<R> b, int n_log, uint64_t n_valid, int fill, IMap map, int p_top,
                         uint64_t n_virtual)
    {
        if (ctx->tracing)
            run_t<true, 0>(ctx, b, n_log, n_valid, fill, map, p_top, n_virtual);
        else if (streaming(ctx, n_virtual))
            run_t<false, dev::kCpStream>(ctx, b, n_log, n_valid, fill, map, p_top, n_virtual);
        else
            run_t<false, 0>(ctx, b, n_log, n_valid, fill, map, p_top, n_virtual);
    }

    static void run(vo_ctx *ctx, Bufs<R> b, int n_log, uint64_t n_valid, int fill)
    {
        VO_CHECK((1ull << n_log) <= kMaxSortRows, VO_ERR_INVALID,
                 "n_pad above 2^29 (32-bit buffer offsets of 8-byte columns)");
        dispatch(ctx, b, n_log, n_valid, fill, IMap{nullptr, 0, nullptr}, n_log, 1ull << n_log);
    }

    // `count` independent instances of 2^c rows, instance i at element
    // offsets[i] with ns[i] live rows (both device arrays): one schedule for
    // all of them, every launch covering every instance.
    static void run_mapped(vo_ctx *ctx, Bufs<R> b, int c, uint64_t count, const uint64_t *offsets, const uint64_t *ns)
    {
        VO_CHECK(c >= R_LOG && (1ull << c) <= kMaxSortRows, VO_ERR_INVALID, "batched instance size out of range");
        dispatch(ctx, b, c, 0, 1, IMap{offsets, c, ns}, c, count << c);
    }

    static int passes(vo_ctx *ctx, int n_log) { return (int)plan_passes(n_log, T_LOG, smax(ctx, n_log), R_LOG, hmax(ctx)).size(); }
};

// Record kinds and their tile shapes (LDS = 2^T_LOG * 17/16 * record bytes).
using RecKP = dev::Rec<1, uint32_t, true>; // 12 B: u64 key + u32 payload
using RecKV = dev::Rec<1, uint64_t, true>; // 16 B: u64 key + u64 value
// multi-column: packed key words + the u32 row index (the unpack pass widens
// it into the caller's int64 index column)
using RecM1 = dev::Rec<1, uint32_t, false>;
using RecM2 = dev::Rec<2, uint32_t, false>;
using RecM4 = dev::Rec<4, uint32_t, false>;
using RecM8 = dev::Rec<8, uint32_t, false>;

} // namespace sortimpl
} // namespace vo
