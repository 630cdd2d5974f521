// join_kernels.hpp -- device bodies and __global__ wrappers of the PK/FK
// oblivious join + GROUP BY (see join.hip for the method and the
// reference), for launch mode and the resident queue kernel.  The
// single-block scans and the dense output take their thread count as a
// template parameter and their LDS from the caller.
#pragma once

#include "vo_internal.hpp"

namespace vo {
namespace join_k {
namespace {

constexpr int JT = 4096; // scan tile
constexpr int JB = 256;
constexpr int JPER = JT / JB;
constexpr uint64_t PAD_KEY = ~0ull;

template <bool TRACE>
__device__ __forceinline__ void b_join_build(const int32_t *__restrict__ dk, const int32_t *__restrict__ dg, uint64_t n_dim,
                             const int32_t *__restrict__ fk, const int64_t *__restrict__ fv, uint64_t n_fact,
                             uint64_t n_pad, uint64_t *__restrict__ keys, uint64_t *__restrict__ vals,
                             TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        uint64_t k = PAD_KEY, v = 0;
        if (i < n_dim) { // public sizes only
            k = ((uint64_t)(uint32_t)dk[i] << 32) | (i & 0x7FFFFFFFu);
            v = (uint64_t)(uint32_t)dg[i];
            tr.at(1, i);
        } else if (i < n_dim + n_fact) {
            const uint64_t f = i - n_dim;
            k = ((uint64_t)(uint32_t)fk[f] << 32) | (1ull << 31) | (f & 0x7FFFFFFFu);
            v = (uint64_t)fv[f];
            tr.at(2, f);
        }
        keys[i] = k;
        vals[i] = v;
        tr.at(3, i);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ void k_join_build(const int32_t *__restrict__ dk, const int32_t *__restrict__ dg, uint64_t n_dim,
                             const int32_t *__restrict__ fk, const int64_t *__restrict__ fv, uint64_t n_fact,
                             uint64_t n_pad, uint64_t *__restrict__ keys, uint64_t *__restrict__ vals,
                             TraceSlot *slot)
{
    b_join_build<TRACE>(dk, dg, n_dim, fk, fv, n_fact, n_pad, keys, vals, slot, blockIdx.x, gridDim.x);
}

// The (u64 key, u32 payload) union of the dense GROUP BY, chosen on public
// arguments only (dense_groups, max_groups <= 63, no sorted-union outputs).
// There the row index is not needed (the sums do not depend on the order of
// equal keys), so the 96 bits hold key32 and a 64-bit field L, the key's low
// word then the payload:
//   dimension row  L = group, or 63 for a group outside [0, 63) (not counted,
//                  as max_groups <= 63)
//   fact row       L = 64 + (value + 2^63) mod 2^64
// so each key's dimension row sorts before its facts, in 12 bytes instead of
// 16.  A value above INT64_MAX - 65 does not fit (L would wrap into the
// dimension codes or reach the pad record): the block that reads one sets
// its flag and the output reports UINT64_MAX groups.
constexpr int64_t kKpMaxValue = INT64_MAX - 65;

template <bool TRACE>
__device__ __forceinline__ void b_join_build_kp(const int32_t *__restrict__ dk, const int32_t *__restrict__ dg,
                                                       uint64_t n_dim, const int32_t *__restrict__ fk,
                                                       const int64_t *__restrict__ fv, uint64_t n_fact, uint64_t n_pad,
                                                       uint64_t *__restrict__ keys, uint32_t *__restrict__ pay,
                                                       uint32_t *__restrict__ bad, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    int over = 0;
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        uint64_t k = PAD_KEY;
        uint32_t p = 0xFFFFFFFFu;
        if (i < n_dim) { // public sizes only
            const uint32_t g = (uint32_t)dg[i];
            const uint64_t L = g < 63u ? g : 63u;
            k = ((uint64_t)(uint32_t)dk[i] << 32) | (L >> 32);
            p = (uint32_t)L;
            tr.at(1, i);
        } else if (i < n_dim + n_fact) {
            const uint64_t f = i - n_dim;
            const int64_t v = fv[f];
            over |= (int)(v > kKpMaxValue);
            const uint64_t L = 64u + ((uint64_t)v ^ (1ull << 63));
            k = ((uint64_t)(uint32_t)fk[f] << 32) | (L >> 32);
            p = (uint32_t)L;
            tr.at(2, f);
        }
        keys[i] = k;
        pay[i] = p;
        tr.at(3, i);
    }
    over = __syncthreads_or(over);
    if (threadIdx.x == 0) bad[vb] = (uint32_t)over;
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ __launch_bounds__(256) void k_join_build_kp(const int32_t *__restrict__ dk, const int32_t *__restrict__ dg,
                                                       uint64_t n_dim, const int32_t *__restrict__ fk,
                                                       const int64_t *__restrict__ fv, uint64_t n_fact, uint64_t n_pad,
                                                       uint64_t *__restrict__ keys, uint32_t *__restrict__ pay,
                                                       uint32_t *__restrict__ bad, TraceSlot *slot)
{
    b_join_build_kp<TRACE>(dk, dg, n_dim, fk, fv, n_fact, n_pad, keys, pay, bad, slot, blockIdx.x, gridDim.x);
}

// ------------------------------------------------- forward fill (step 3)

struct Fill {
    uint64_t key;   // key32 of the last dimension row, or ~0 for none
    uint64_t group;
};

__device__ __forceinline__ bool is_dim(uint64_t k) { return k != PAD_KEY && ((k >> 31) & 1) == 0; }

// Union row i in either form: key32, the dimension / fact flags and the
// 64-bit field (the group of a dimension row, the value of a fact row).
template <bool KP>
__device__ __forceinline__ void union_row(const uint64_t *__restrict__ keys, const void *__restrict__ vals, uint64_t i,
                                          uint64_t &key32, bool &dim, bool &fact, uint64_t &v)
{
    const uint64_t k = keys[i];
    if constexpr (KP) {
        const uint32_t p = static_cast<const uint32_t *>(vals)[i];
        const uint64_t L = (k << 32) | p;
        const bool pad = (k == PAD_KEY) & (p == 0xFFFFFFFFu);
        dim = !pad & (L < 64u);
        fact = !pad & (L >= 64u);
        v = dim ? L : ((L - 64u) ^ (1ull << 63));
    } else {
        v = static_cast<const uint64_t *>(vals)[i];
        dim = is_dim(k);
        fact = (k != PAD_KEY) & (((k >> 31) & 1) != 0);
    }
    key32 = k >> 32;
}

// Per tile: the last dimension row (key, group) or none.
template <bool TRACE, bool KP>
__device__ __forceinline__ void b_fill_tiles(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                   Fill *__restrict__ tile_last, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    __shared__ uint64_t wk[JB / 64], wg[JB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * JT + (uint64_t)threadIdx.x * JPER;
    uint64_t lk = PAD_KEY, lg = 0;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        uint64_t k32, v;
        bool d, f;
        union_row<KP>(keys, vals, t0 + k, k32, d, f, v);
        tr.at(1, t0 + k);
        lk = d ? k32 : lk;
        lg = d ? v : lg;
    }
    // each wavefront's last lane that holds a dimension row (a register
    // read, no address depends on it), then the block's last wavefront
    const uint64_t has = __builtin_amdgcn_ballot_w64(lk != PAD_KEY);
    const int src = has ? 63 - __builtin_clzll(has) : 0;
    const uint64_t rk = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(lk >> 32), src) << 32) |
                        (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)lk, src);
    const uint64_t rg = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(lg >> 32), src) << 32) |
                        (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)lg, src);
    if ((threadIdx.x & 63) == 0) {
        wk[threadIdx.x >> 6] = has ? rk : PAD_KEY;
        wg[threadIdx.x >> 6] = has ? rg : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t k = PAD_KEY, g = 0;
        for (int w = 0; w < JB / 64; ++w) {
            const bool h = wk[w] != PAD_KEY;
            k = h ? wk[w] : k;
            g = h ? wg[w] : g;
        }
        tile_last[vb] = Fill{k, g};
        tr.at(2, vb);
    }
    tr.done(dev::gtid());
}

template <bool TRACE, bool KP>
__global__ __launch_bounds__(JB) void k_fill_tiles(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                   Fill *__restrict__ tile_last, TraceSlot *slot)
{
    b_fill_tiles<TRACE, KP>(keys, vals, tile_last, slot, blockIdx.x, gridDim.x);
}

// Exclusive scan over the nt tile summaries with one block of NT threads
// (1024 in launch mode, 256 in the resident kernel): each thread folds a
// contiguous chunk, a Hillis-Steele scan combines the chunk summaries in LDS
// (sh: NT values), then each thread writes its chunk's exclusive prefixes.
// op(a, b) is associative with identity `id`, so NT changes no result.
constexpr uint32_t kScanThreads = 1024;

template <uint32_t NT, class T, class Op, class Tr>
__device__ void block_exclusive_scan(const T *__restrict__ in, uint32_t nt, T *__restrict__ out, T id, Op op, T *sh,
                                     Tr &tr)
{
    const uint32_t per = (nt + NT - 1) / NT;
    const uint32_t lo = threadIdx.x * per;
    // a thread's chunk in groups of 8, loaded before any is combined
    constexpr int G = 8;
    T s = id;
    for (uint32_t k0 = 0; k0 < per; k0 += G) {
        T v[G];
#pragma unroll
        for (int j = 0; j < G; ++j) v[j] = (k0 + j < per && lo + k0 + j < nt) ? in[lo + k0 + j] : id;
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (k0 + j < per && lo + k0 + j < nt) { // public size
                s = op(s, v[j]);
                tr.at(1, lo + k0 + j);
            }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < NT; off <<= 1) {
        const T v = threadIdx.x >= off ? sh[threadIdx.x - off] : id;
        __syncthreads();
        if (threadIdx.x >= off) sh[threadIdx.x] = op(v, sh[threadIdx.x]);
        __syncthreads();
    }
    T run = threadIdx.x ? sh[threadIdx.x - 1] : id; // exclusive prefix of this chunk
    for (uint32_t k0 = 0; k0 < per; k0 += G) {
        T v[G];
#pragma unroll
        for (int j = 0; j < G; ++j) v[j] = (k0 + j < per && lo + k0 + j < nt) ? in[lo + k0 + j] : id;
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (k0 + j < per && lo + k0 + j < nt) {
                out[lo + k0 + j] = run;
                run = op(run, v[j]);
                tr.at(2, lo + k0 + j);
            }
    }
}

struct FillOp {
    __device__ Fill operator()(const Fill &a, const Fill &b) const
    {
        const bool has = b.key != PAD_KEY;
        return Fill{has ? b.key : a.key, has ? b.group : a.group};
    }
};

// Exclusive "last dimension row" scan over tiles (sh: NT values of LDS).
template <uint32_t NT, bool TRACE>
__device__ __forceinline__ void b_fill_carry(const Fill *__restrict__ tile_last, uint32_t nt, Fill *__restrict__ carry,
                                             TraceSlot *slot, Fill *sh)
{
    dev::AtrT<TRACE> tr(slot);
    block_exclusive_scan<NT>(tile_last, nt, carry, Fill{PAD_KEY, 0}, FillOp{}, sh, tr);
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ __launch_bounds__(kScanThreads) void k_fill_carry(const Fill *__restrict__ tile_last, uint32_t nt,
                                                             Fill *__restrict__ carry, TraceSlot *slot)
{
    __shared__ Fill sh[kScanThreads];
    b_fill_carry<kScanThreads, TRACE>(tile_last, nt, carry, slot, sh);
}

// Per row: (group, value) columns and the matched mask.
template <bool TRACE>
__device__ __forceinline__ void b_fill_rows(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                  const Fill *__restrict__ carry, uint64_t *__restrict__ col_g,
                                                  uint64_t *__restrict__ col_v, uint8_t *__restrict__ mask,
                                                  TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    __shared__ uint64_t wtk[JB / 64], wtg[JB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * JT + (uint64_t)threadIdx.x * JPER;
    uint64_t lk = PAD_KEY, lg = 0;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t key = keys[t0 + k], v = vals[t0 + k];
        const bool d = is_dim(key);
        lk = d ? (key >> 32) : lk;
        lg = d ? v : lg;
    }
    // carry into this thread (the last dimension row before its 16 rows):
    // wavefront inclusive scan of "the later dimension row wins", then the
    // wavefront totals through LDS
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t ik = lk, ig = lg;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t pk = __shfl_up(ik, off, 64), pg = __shfl_up(ig, off, 64);
        const bool take = (lane >= (uint32_t)off) & (ik == PAD_KEY);
        ik = take ? pk : ik;
        ig = take ? pg : ig;
    }
    const uint64_t ek = __shfl_up(ik, 1, 64), eg = __shfl_up(ig, 1, 64);
    if (lane == 63) {
        wtk[wv] = ik;
        wtg[wv] = ig;
    }
    __syncthreads();
    Fill c = carry[vb];
    tr.at(4, vb);
    for (uint32_t w = 0; w < wv; ++w) {
        const bool has = wtk[w] != PAD_KEY;
        c.key = has ? wtk[w] : c.key;
        c.group = has ? wtg[w] : c.group;
    }
    {
        const bool has = (lane > 0) & (ek != PAD_KEY);
        c.key = has ? ek : c.key;
        c.group = has ? eg : c.group;
    }
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t i = t0 + k;
        const uint64_t key = keys[i], v = vals[i];
        const bool d = is_dim(key);
        c.key = d ? (key >> 32) : c.key;
        c.group = d ? v : c.group;
        const bool fact = key != PAD_KEY && ((key >> 31) & 1);
        const bool matched = fact & (c.key == (key >> 32));
        col_g[i] = c.group;
        col_v[i] = v;
        mask[i] = matched ? 1 : 0;
        tr.at(5, i); // keys, vals read and the three outputs written at row i
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ __launch_bounds__(JB) void k_fill_rows(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                  const Fill *__restrict__ carry, uint64_t *__restrict__ col_g,
                                                  uint64_t *__restrict__ col_v, uint8_t *__restrict__ mask,
                                                  TraceSlot *slot)
{
    b_fill_rows<TRACE>(keys, vals, carry, col_g, col_v, mask, slot, blockIdx.x, gridDim.x);
}

// ------------------------------------------- group-by keys (step 5)

template <bool TRACE>
__device__ __forceinline__ void b_group_keys(const uint64_t *__restrict__ col_g, const uint64_t *__restrict__ col_v,
                             const uint64_t *__restrict__ count, uint64_t n_pad, uint64_t *__restrict__ keys,
                             uint64_t *__restrict__ vals, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    const uint64_t m = *count;
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        const uint64_t gi = col_g[i]; // read for every row: the select below is on the value only
        const uint64_t g = i < m ? (gi & 0xFFFFFFFFull) : 0xFFFFFFFFull;
        keys[i] = (g << 32) | (i & 0xFFFFFFFFull);
        vals[i] = col_v[i];
        tr.at(1, i);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ void k_group_keys(const uint64_t *__restrict__ col_g, const uint64_t *__restrict__ col_v,
                             const uint64_t *__restrict__ count, uint64_t n_pad, uint64_t *__restrict__ keys,
                             uint64_t *__restrict__ vals, TraceSlot *slot)
{
    b_group_keys<TRACE>(col_g, col_v, count, n_pad, keys, vals, slot, blockIdx.x, gridDim.x);
}

// ------------------------------------- segmented sum by group (step 6)

// Per tile: group of the last row, sum of the tile's trailing run of that
// group, whether the whole tile is that one group.
struct Seg {
    uint64_t last_g;
    int64_t tail_sum;
    uint32_t all_one;
};

__device__ __forceinline__ uint64_t grp(uint64_t k) { return k >> 32; }

template <bool TRACE>
__device__ __forceinline__ void b_seg_tiles(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                  Seg *__restrict__ tile_seg, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    __shared__ uint64_t fg[JB], lg[JB];
    __shared__ int64_t ts[JB];
    __shared__ uint32_t one[JB];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * JT + (uint64_t)threadIdx.x * JPER;
    const uint64_t g0 = grp(keys[t0]);
    uint64_t gl = g0;
    int64_t run = 0;
    uint32_t uni = 1;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t g = grp(keys[t0 + k]);
        const int64_t v = (int64_t)vals[t0 + k];
        tr.at(1, t0 + k);
        run = (g == gl ? run : 0) + v;
        uni &= (g == g0);
        gl = g;
    }
    fg[threadIdx.x] = g0;
    lg[threadIdx.x] = gl;
    ts[threadIdx.x] = run;
    one[threadIdx.x] = uni;
    __syncthreads();
    if (threadIdx.x == 0) {
        // combine thread summaries left to right
        uint64_t last = lg[0];
        int64_t tail = ts[0];
        uint32_t all = one[0];
        for (int t = 1; t < JB; ++t) {
            const bool cont = one[t] & (fg[t] == last);
            // a thread that does not continue the run as a whole ends in a
            // different group, whose run starts inside it: its own tail
            tail = cont ? tail + ts[t] : ts[t];
            all = all & one[t] & (fg[t] == last);
            last = lg[t];
        }
        tile_seg[vb] = Seg{last, tail, all};
        tr.at(2, vb);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ __launch_bounds__(JB) void k_seg_tiles(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                  Seg *__restrict__ tile_seg, TraceSlot *slot)
{
    b_seg_tiles<TRACE>(keys, vals, tile_seg, slot, blockIdx.x, gridDim.x);
}

// Summary of a range of tiles for the segmented carry: group of its first
// and last row, sum of its trailing run of last_g, whether it is one group.
struct SegSum {
    uint64_t first_g, last_g;
    int64_t tail;
    uint32_t all_one, empty;
};

struct SegOp {
    __device__ SegSum operator()(const SegSum &a, const SegSum &b) const
    {
        if (a.empty) return b;
        if (b.empty) return a;
        const bool joins = b.all_one & (a.last_g == b.first_g);
        return SegSum{a.first_g, b.last_g, joins ? a.tail + b.tail : b.tail,
                      (uint32_t)(a.all_one & joins), 0u};
    }
};

template <bool TRACE>
__device__ __forceinline__ void b_seg_summ(const Seg *__restrict__ seg, const uint64_t *__restrict__ keys, uint32_t nt,
                           SegSum *__restrict__ out, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint32_t t = vb * 256 + threadIdx.x; t < nt; t += nvb * 256) {
        const Seg s = seg[t];
        out[t] = SegSum{grp(keys[(uint64_t)t * JT]), s.last_g, s.tail_sum, s.all_one, 0u};
        tr.at(1, t);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ void k_seg_summ(const Seg *__restrict__ seg, const uint64_t *__restrict__ keys, uint32_t nt,
                           SegSum *__restrict__ out, TraceSlot *slot)
{
    b_seg_summ<TRACE>(seg, keys, nt, out, slot, blockIdx.x, gridDim.x);
}

// Carry into each tile: the sum of the run of the tile's first group that
// ends just before the tile (0 when the tiles before end in another group).
// (sh: NT values of LDS)
template <uint32_t NT, bool TRACE>
__device__ __forceinline__ void b_seg_carry(const SegSum *__restrict__ summ, uint32_t nt, SegSum *__restrict__ pre,
                                            int64_t *__restrict__ carry, TraceSlot *slot, SegSum *sh)
{
    dev::AtrT<TRACE> tr(slot);
    block_exclusive_scan<NT>(summ, nt, pre, SegSum{0, 0, 0, 0, 1}, SegOp{}, sh, tr);
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < nt; t += NT) {
        const SegSum p = pre[t];
        carry[t] = (!p.empty && p.last_g == summ[t].first_g) ? p.tail : 0;
        tr.at(3, t);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ __launch_bounds__(kScanThreads) void k_seg_carry(const SegSum *__restrict__ summ, uint32_t nt,
                                                            SegSum *__restrict__ pre, int64_t *__restrict__ carry,
                                                            TraceSlot *slot)
{
    __shared__ SegSum sh[kScanThreads];
    b_seg_carry<kScanThreads, TRACE>(summ, nt, pre, carry, slot, sh);
}

template <bool TRACE>
__device__ __forceinline__ void b_seg_rows(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                 const int64_t *__restrict__ carry, uint64_t n_pad,
                                                 uint64_t *__restrict__ col_g, uint64_t *__restrict__ col_s,
                                                 uint8_t *__restrict__ mask, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    __shared__ SegSum wtot[JB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * JT + (uint64_t)threadIdx.x * JPER;
    // this thread's rows in registers (contiguous: vector loads), and the
    // group of the row after them (a public index; none past n_pad)
    uint64_t kg[JPER];
    int64_t kv[JPER];
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        kg[k] = grp(keys[t0 + k]);
        kv[k] = (int64_t)vals[t0 + k];
    }
    const uint64_t nx = t0 + JPER;
    const uint64_t g_next = nx < n_pad ? grp(keys[nx < n_pad ? nx : t0]) : ~0ull;
    const uint64_t g0 = kg[0];
    uint64_t gl = g0;
    int64_t run = 0;
    uint32_t uni = 1;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t g = kg[k];
        run = (g == gl ? run : 0) + kv[k];
        uni &= (g == g0);
        gl = g;
    }
    // running sum of the group that is open when this thread's rows start:
    // the tile carry combined with the summaries of the threads before this
    // one (wavefront inclusive scan, then the wavefront totals through LDS)
    const SegOp op;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    SegSum x{g0, gl, run, uni, 0u};
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        SegSum y;
        y.first_g = __shfl_up(x.first_g, off, 64);
        y.last_g = __shfl_up(x.last_g, off, 64);
        y.tail = __shfl_up(x.tail, off, 64);
        y.all_one = __shfl_up(x.all_one, off, 64);
        y.empty = 0u;
        const SegSum z = op(y, x);
        if (lane >= (uint32_t)off) x = z;
    }
    SegSum e;
    e.first_g = __shfl_up(x.first_g, 1, 64);
    e.last_g = __shfl_up(x.last_g, 1, 64);
    e.tail = __shfl_up(x.tail, 1, 64);
    e.all_one = __shfl_up(x.all_one, 1, 64);
    e.empty = lane == 0 ? 1u : 0u;
    if (lane == 63) wtot[wv] = x;
    __syncthreads();
    const uint64_t tile_first = grp(keys[(uint64_t)vb * JT]);
    SegSum pre{tile_first, tile_first, carry[vb], 1u, 0u};
    tr.at(4, vb);
    for (uint32_t w = 0; w < wv; ++w) pre = op(pre, wtot[w]);
    pre = op(pre, e);
    int64_t c = pre.tail;
    uint64_t cg = pre.last_g;
    // rows
    uint64_t prev_g = cg;
    int64_t acc = c;
    uint8_t mk[JPER];
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t i = t0 + k;
        const uint64_t g = kg[k];
        acc = (g == prev_g ? acc : 0) + kv[k];
        prev_g = g;
        const uint64_t gn = k + 1 < JPER ? kg[k + 1] : g_next;
        const bool last = gn != g;
        col_g[i] = g;
        col_s[i] = (uint64_t)acc;
        mk[k] = (last & (g != 0xFFFFFFFFull)) ? 1 : 0;
        tr.at(5, i);
    }
    static_assert(JPER == 16, "one 16-byte mask store per thread");
    typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
    u32x4_t mw;
#pragma unroll
    for (int w = 0; w < 4; ++w)
        mw[w] = (uint32_t)mk[4 * w] | ((uint32_t)mk[4 * w + 1] << 8) | ((uint32_t)mk[4 * w + 2] << 16) |
                ((uint32_t)mk[4 * w + 3] << 24);
    *reinterpret_cast<u32x4_t *>(mask + t0) = mw; // t0 is a multiple of 16
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ __launch_bounds__(JB) void k_seg_rows(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                 const int64_t *__restrict__ carry, uint64_t n_pad,
                                                 uint64_t *__restrict__ col_g, uint64_t *__restrict__ col_s,
                                                 uint8_t *__restrict__ mask, TraceSlot *slot)
{
    b_seg_rows<TRACE>(keys, vals, carry, n_pad, col_g, col_s, mask, slot, blockIdx.x, gridDim.x);
}

template <bool TRACE>
__device__ __forceinline__ void b_group_out(const uint64_t *__restrict__ col_g, const uint64_t *__restrict__ col_s,
                            const uint64_t *__restrict__ count, uint64_t n_pad, uint64_t max_groups,
                            int32_t *__restrict__ out_g, int64_t *__restrict__ out_s, uint64_t *__restrict__ out_n,
                            TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    const uint64_t G = *count;
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < max_groups; i += (uint64_t)nvb * 256) {
        const bool live = i < G;
        // read for every output slot (row i, or row 0 past the padded size:
        // both public), so the select below is on values only
        const uint64_t ri = i < n_pad ? i : 0;
        const uint64_t g = col_g[ri], s = col_s[ri];
        out_g[i] = live ? (int32_t)g : -1;
        out_s[i] = live ? (int64_t)s : 0;
        tr.at(1, ri);
    }
    if (vb == 0 && threadIdx.x == 0 && out_n) *out_n = G;
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ void k_group_out(const uint64_t *__restrict__ col_g, const uint64_t *__restrict__ col_s,
                            const uint64_t *__restrict__ count, uint64_t n_pad, uint64_t max_groups,
                            int32_t *__restrict__ out_g, int64_t *__restrict__ out_s, uint64_t *__restrict__ out_n,
                            TraceSlot *slot)
{
    b_group_out<TRACE>(col_g, col_s, count, n_pad, max_groups, out_g, out_s, out_n, slot, blockIdx.x, gridDim.x);
}

// ------------------------------- dense GROUP BY (group ids in [0, G))

__device__ __forceinline__ int64_t wave_sum64(int64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The dense GROUP BY fused with the forward fill: per tile, the fill of
// k_fill_rows, and every row adds its value to all G per-thread sums through
// selects (matched fact rows whose group is g); a presence bitmask stands in
// for the per-group counts.  One partial row per tile; no compaction is
// needed because every union row is visited anyway.
template <int G, bool TRACE, bool KP>
__device__ __forceinline__ void b_fill_agg(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                 const Fill *__restrict__ carry, int64_t *__restrict__ part_sum,
                                                 int64_t *__restrict__ part_cnt, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    static_assert(G <= 64, "presence bits in one word");
    __shared__ uint64_t wtk[JB / 64], wtg[JB / 64];
    __shared__ int64_t rs[JB / 64][G];
    __shared__ uint64_t rb[JB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * JT + (uint64_t)threadIdx.x * JPER;
    uint64_t lk = PAD_KEY, lg = 0;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        uint64_t k32, v;
        bool d, f;
        union_row<KP>(keys, vals, t0 + k, k32, d, f, v);
        lk = d ? k32 : lk;
        lg = d ? v : lg;
    }
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t ik = lk, ig = lg;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t pk = __shfl_up(ik, off, 64), pg = __shfl_up(ig, off, 64);
        const bool take = (lane >= (uint32_t)off) & (ik == PAD_KEY);
        ik = take ? pk : ik;
        ig = take ? pg : ig;
    }
    const uint64_t ek = __shfl_up(ik, 1, 64), eg = __shfl_up(ig, 1, 64);
    if (lane == 63) {
        wtk[wv] = ik;
        wtg[wv] = ig;
    }
    __syncthreads();
    Fill c = carry[vb];
    tr.at(4, vb);
    for (uint32_t w = 0; w < wv; ++w) {
        const bool has = wtk[w] != PAD_KEY;
        c.key = has ? wtk[w] : c.key;
        c.group = has ? wtg[w] : c.group;
    }
    {
        const bool has = (lane > 0) & (ek != PAD_KEY);
        c.key = has ? ek : c.key;
        c.group = has ? eg : c.group;
    }
    int64_t acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = 0;
    uint64_t bits = 0;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t i = t0 + k;
        uint64_t k32, v;
        bool d, fact;
        union_row<KP>(keys, vals, i, k32, d, fact, v);
        c.key = d ? k32 : c.key;
        c.group = d ? v : c.group;
        const bool matched = fact & (c.key == k32);
        const uint32_t grp = (uint32_t)c.group;
        // the row's value where it counts, then one 32-bit compare and a
        // select per group; the presence bit by one shift (no address
        // depends on the group)
        const int64_t vm = matched ? (int64_t)v : 0;
#pragma unroll
        for (int g = 0; g < G; ++g) acc[g] += grp == (uint32_t)g ? vm : 0;
        bits |= (matched & (grp < (uint32_t)G)) ? (1ull << (grp & 63u)) : 0ull;
        tr.at(5, i);
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int64_t sg = wave_sum64(acc[g]);
        if (lane == 0) rs[wv][g] = sg;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) bits |= (uint64_t)__shfl_xor((unsigned long long)bits, off, 64);
    if (lane == 0) rb[wv] = bits;
    __syncthreads();
    if (threadIdx.x < G) {
        const uint32_t g = threadIdx.x;
        int64_t sg = 0;
        uint64_t b = 0;
#pragma unroll
        for (int w = 0; w < JB / 64; ++w) {
            sg += rs[w][g];
            b |= rb[w];
        }
        part_sum[(uint64_t)vb * G + g] = sg;
        part_cnt[(uint64_t)vb * G + g] = (int64_t)((b >> g) & 1);
        tr.at(6, (uint64_t)vb * G + g);
    }
    tr.done(dev::gtid());
}

template <int G, bool TRACE, bool KP>
__global__ __launch_bounds__(JB) void k_fill_agg(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                 const Fill *__restrict__ carry, int64_t *__restrict__ part_sum,
                                                 int64_t *__restrict__ part_cnt, TraceSlot *slot)
{
    b_fill_agg<G, TRACE, KP>(keys, vals, carry, part_sum, part_cnt, slot, blockIdx.x, gridDim.x);
}

// One block of NT threads (1024 in launch mode, 256 in the resident kernel):
// totals per group, then output slot t takes the t-th present group, chosen
// by selects over all groups.  The nparts x G partials are read
// contiguously; thread t only ever meets group t mod G (G divides NT), and
// the sums are exact integer sums, so NT changes no result.  sh: 2 NT + 2 G
// words of LDS.
constexpr uint32_t kOutThreads = 1024;

template <int G, uint32_t NT>
__device__ __forceinline__ void b_dense_out(const int64_t *__restrict__ part_sum, const int64_t *__restrict__ part_cnt,
                                            uint32_t nparts, uint64_t max_groups, int32_t *__restrict__ out_g,
                                            int64_t *__restrict__ out_s, uint64_t *__restrict__ out_n,
                                            const uint32_t *__restrict__ bad, uint32_t n_bad, TraceSlot *slot,
                                            int64_t *sh)
{
    static_assert(NT % G == 0, "G divides the block");
    int64_t *ss = sh, *sc = sh + NT, *ts = sh + 2 * NT, *tc = sh + 2 * NT + G;
    dev::Atr tr(slot);
    const uint32_t t = threadIdx.x;
    int64_t s = 0, c = 0;
    const uint64_t total = (uint64_t)nparts * G;
    for (uint64_t i = t; i < total; i += NT) {
        s += part_sum[i];
        c += part_cnt[i];
        tr.at(1, i);
    }
    ss[t] = s;
    sc[t] = c;
    __syncthreads();
    if (t < G) {
        int64_t a = 0, b = 0;
        for (uint32_t u = t; u < NT; u += G) {
            a += ss[u];
            b += sc[u];
        }
        ts[t] = a;
        tc[t] = (uint64_t)t < max_groups ? b : 0;
    }
    __syncthreads();
    uint64_t present = 0;
#pragma unroll 8
    for (int g = 0; g < G; ++g) present += tc[g] > 0;
    for (uint64_t j = t; j < max_groups; j += NT) {
        int32_t og = -1;
        int64_t os = 0;
        uint64_t rank = 0;
#pragma unroll 8
        for (int g = 0; g < G; ++g) {
            const bool here = (tc[g] > 0) & (rank == j);
            og = here ? g : og;
            os = here ? ts[g] : os;
            rank += tc[g] > 0;
        }
        out_g[j] = og;
        out_s[j] = os;
        tr.at(2, j);
    }
    // a fact value that the 12-byte union cannot hold: report UINT64_MAX
    int over = 0;
    for (uint32_t b = t; b < n_bad; b += NT) over |= (int)(bad[b] != 0);
    over = __syncthreads_or(over);
    if (t == 0 && out_n) *out_n = over ? ~0ull : present;
    tr.done(dev::gtid());
}

template <int G>
__global__ __launch_bounds__(kOutThreads) void k_dense_out(const int64_t *__restrict__ part_sum,
                                                           const int64_t *__restrict__ part_cnt, uint32_t nparts,
                                                           uint64_t max_groups, int32_t *__restrict__ out_g,
                                                           int64_t *__restrict__ out_s, uint64_t *__restrict__ out_n,
                                                           const uint32_t *__restrict__ bad, uint32_t n_bad,
                                                           TraceSlot *slot)
{
    __shared__ int64_t sh[2 * kOutThreads + 2 * G];
    b_dense_out<G, kOutThreads>(part_sum, part_cnt, nparts, max_groups, out_g, out_s, out_n, bad, n_bad, slot, sh);
}

} // namespace
} // namespace join_k
} // namespace vo
