// compact_kernels.hpp -- device bodies and __global__ wrappers of the
// oblivious order-preserving compaction (ORCompact routing network), the
// Q6-style filter + exact int128 SUM and OP_SCATTER, used by compact.hip in
// launch mode and by the resident queue kernel.
//
// Reference method: oblivious_compact_elem / _inner_elem / _2power_inner_elem
// (obliviator/join/enclave/oblivious_compact.c:59-153, 327-338).  For a
// power-of-two n the recursion unrolls into log2(n) levels: at level q
// (block size s = 2^(q+1), half = 2^q) every aligned block [b, b+s) with
// offset o = C[b] mod s (C = exclusive prefix count of the selected rows)
// conditionally swaps rows b+i and b+half+i for all i < half, with
//   s == 2 : swap = (!sel[b] & sel[b+1]) ^ o
//   s  > 2 : m1 = C[b+half] - C[b], om = o mod half, om1 = (o+m1) mod half,
//            swap = ((om + m1 >= half) ^ (o >= half)) ^ (i >= om1)
// (the offset form of :88-105; the chain of :121-153 reduces to o = C[b]
// mod s for the blocks it creates).  Every row is read and written at every
// level; the swap decision is a select, so addresses depend on n_pad only.
//
// Passes: count (selected rows per tile) -> tile scan -> local levels
// q < T_LOG inside each tile (LDS) -> fused global levels, S per HBM pass,
// rows at stride 2^q_lo as in the sort's global passes.
//
// Each body takes its block index explicitly (vb) and its larger LDS arrays
// from the caller, so the resident kernel can run it over its own blocks.
#pragma once

#include "vo_internal.hpp"

namespace vo {
namespace compact_k {
namespace {

constexpr int CT_LOG = 12; // tile: 4096 rows
constexpr int CT = 1 << CT_LOG;
constexpr int CBLOCK = 256;

// A column of the compaction: pointer and element width (1, 2, 4, 8 bytes).
struct Col {
    void *p;
    uint32_t w;
};

template <int NW>
struct Cols {
    Col c[NW];
};

// Row loads and stores of the network passes.  NT: non-temporal, for rows
// beyond the 256 MB MALL, which every pass streams through once (below it
// the next pass re-reads them from the MALL).
template <bool NT, typename T>
__device__ __forceinline__ T ld_row(const T *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <bool NT, typename T>
__device__ __forceinline__ void st_row(T *p, T v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <bool NT = false>
__device__ __forceinline__ uint64_t ld_col(const Col &c, uint64_t i)
{
    switch (c.w) {
    case 8: return ld_row<NT>((const uint64_t *)c.p + i);
    case 4: return ld_row<NT>((const uint32_t *)c.p + i);
    case 2: return ld_row<NT>((const uint16_t *)c.p + i);
    default: return ld_row<NT>((const uint8_t *)c.p + i);
    }
}

template <bool NT = false>
__device__ __forceinline__ void st_col(const Col &c, uint64_t i, uint64_t v)
{
    switch (c.w) {
    case 8: st_row<NT>((uint64_t *)c.p + i, v); break;
    case 4: st_row<NT>((uint32_t *)c.p + i, (uint32_t)v); break;
    case 2: st_row<NT>((uint16_t *)c.p + i, (uint16_t)v); break;
    default: st_row<NT>((uint8_t *)c.p + i, (uint8_t)v); break;
    }
}

// swap flag of the pair at offset i in the block starting at prefix count
// cb, with cbh = prefix count at the block's middle; s2 = half size
__device__ __forceinline__ bool level_flag(uint64_t cb, uint64_t cbh, uint64_t half, uint64_t i)
{
    const uint64_t s = half * 2;
    const uint64_t o = cb & (s - 1);
    const uint64_t m1 = cbh - cb;
    const uint64_t om = o & (half - 1);
    const uint64_t om1 = (o + m1) & (half - 1);
    const bool base = ((om + m1) >= half) ^ (o >= half);
    return base ^ (i >= om1);
}

// The same flag in 32-bit arithmetic, for n_pad <= 2^32 (q <= 31): only the
// low 32 bits of the prefix counts matter (o is their low q+1 bits, and the
// count difference m1 <= half is exact modulo 2^32).
__device__ __forceinline__ bool level_flag32(uint32_t cb, uint32_t cbh, int q, uint32_t i)
{
    const uint32_t half = 1u << q;
    const uint32_t smask = half | (half - 1);
    const uint32_t o = cb & smask;
    const uint32_t m1 = cbh - cb;
    const uint32_t om = o & (half - 1);
    const uint32_t om1 = (o + m1) & (half - 1);
    const bool base = ((om + m1) >= half) ^ (o >= half);
    return base ^ (i >= om1);
}

// Instances of one size class 2^c of a batch: virtual row v of the class is
// row v mod 2^c of instance v >> c, stored at physical row
// off[v >> c] + (v mod 2^c), whose live rows are n[v >> c].  off == nullptr:
// a single compaction, virtual rows are physical rows and the caller's n
// applies.  Every caller passes a v whose instance is the same for the whole
// wavefront, so the table reads are scalar.
struct Inst {
    const uint64_t *off;
    const uint64_t *n;
    int c;
};

__device__ __forceinline__ uint64_t uni64(uint64_t v)
{
    return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
           __builtin_amdgcn_readfirstlane((uint32_t)v);
}

__device__ __forceinline__ uint32_t inst_of(const Inst &in, uint64_t v)
{
    return __builtin_amdgcn_readfirstlane((uint32_t)(v >> in.c));
}

// physical row of virtual row v
__device__ __forceinline__ uint64_t inst_row(const Inst &in, uint64_t v)
{
    if (in.off == nullptr) return v;
    return uni64(in.off[inst_of(in, v)]) + (v & ((1ull << in.c) - 1));
}

// row of v inside its instance, and that instance's live rows
__device__ __forceinline__ uint64_t inst_local(const Inst &in, uint64_t v)
{
    return in.off == nullptr ? v : (v & ((1ull << in.c) - 1));
}

__device__ __forceinline__ uint64_t inst_live(const Inst &in, uint64_t v, uint64_t n)
{
    return in.off == nullptr ? n : uni64(in.n[inst_of(in, v)]);
}

// ---------------------------------------------------------------- counting

// Selected rows per tile: sel[i] = mask[i] != 0 && i < n.
__device__ __forceinline__ void b_count_mask(const uint8_t *__restrict__ mask, uint64_t n,
                                             uint32_t *__restrict__ tile_cnt, Inst in, TraceSlot *slot, uint32_t vb)
{
    __shared__ uint32_t red[CBLOCK / 64];
    dev::Atr tr(slot);
    const uint64_t tv = (uint64_t)vb * CT;
    const uint64_t t0 = inst_row(in, tv), l0 = inst_local(in, tv), nn = inst_live(in, tv, n);
    uint32_t c = 0;
#pragma unroll 4
    for (int k = 0; k < CT / CBLOCK; ++k) {
        const uint32_t i = k * CBLOCK + threadIdx.x;
        c += (mask[t0 + i] != 0) & (l0 + i < nn);
        tr.at(1, t0 + i);
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < CBLOCK / 64; ++w) s += red[w];
        tile_cnt[vb] = s;
        tr.at(2, vb);
    }
    tr.done(dev::gtid());
}

__global__ __launch_bounds__(CBLOCK) void k_count_mask(const uint8_t *__restrict__ mask, uint64_t n,
                                                       uint32_t *__restrict__ tile_cnt, Inst in, TraceSlot *slot)
{
    b_count_mask(mask, n, tile_cnt, in, slot, blockIdx.x);
}

// Q6-style predicate: writes mask[i] and per-tile counts in one pass.
struct Pred {
    int32_t day_lo, day_hi;
    int64_t disc_lo, disc_hi, qty_lt;
};

template <bool NT = false>
__device__ __forceinline__ void b_filter_count(const int32_t *__restrict__ day, const int64_t *__restrict__ qty,
                                               const int64_t *__restrict__ disc, uint64_t n, Pred pr,
                                               uint8_t *__restrict__ mask, uint32_t *__restrict__ tile_cnt,
                                               TraceSlot *slot, uint32_t vb)
{
    __shared__ uint32_t red[CBLOCK / 64];
    dev::Atr tr(slot);
    const uint64_t t0 = (uint64_t)vb * CT;
    uint32_t c = 0;
#pragma unroll 4
    for (int k = 0; k < CT / CBLOCK; ++k) {
        const uint64_t i = t0 + k * CBLOCK + threadIdx.x;
        const int32_t d = ld_row<NT>(day + i);
        const int64_t q = ld_row<NT>(qty + i), ds = ld_row<NT>(disc + i);
        const uint32_t m = (d >= pr.day_lo) & (d < pr.day_hi) & (ds >= pr.disc_lo) & (ds <= pr.disc_hi) &
                           (q < pr.qty_lt) & (i < n);
        st_row<NT>(mask + i, (uint8_t)m);
        c += m;
        tr.at(3, i); // day, qty, disc read and mask written at row i
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < CBLOCK / 64; ++w) s += red[w];
        tile_cnt[vb] = s;
        tr.at(2, vb);
    }
    tr.done(dev::gtid());
}

template <bool NT = false>
__global__ __launch_bounds__(CBLOCK) void k_filter_count(const int32_t *__restrict__ day,
                                                         const int64_t *__restrict__ qty,
                                                         const int64_t *__restrict__ disc, uint64_t n, Pred pr,
                                                         uint8_t *__restrict__ mask, uint32_t *__restrict__ tile_cnt,
                                                         TraceSlot *slot)
{
    b_filter_count<NT>(day, qty, disc, n, pr, mask, tile_cnt, slot, blockIdx.x);
}

// Exclusive scan of the tile counts by one block of NT threads:
// pre[0..nt], pre[nt] = total; part holds NT words of LDS.
template <uint32_t NT, bool TOTAL = true>
__device__ __forceinline__ void b_scan_tiles(const uint32_t *__restrict__ cnt, uint32_t nt, uint64_t *__restrict__ pre,
                                             uint64_t *__restrict__ out_count, TraceSlot *slot, uint64_t *part)
{
    dev::Atr tr(slot);
    const uint32_t per = (nt + NT - 1) / NT;
    const uint32_t lo = threadIdx.x * per;
    // a thread's chunk in groups of 16 counts, loaded before any is used: one
    // memory round trip per group rather than per count
    constexpr int G = 16;
    uint64_t s = 0;
    for (uint32_t k0 = 0; k0 < per; k0 += G) {
        uint32_t v[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const uint32_t k = k0 + j, i = lo + k;
            v[j] = (k < per && i < nt) ? cnt[i] : 0u;
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const uint32_t k = k0 + j, i = lo + k;
            if (k < per && i < nt) tr.at(4, i);
            s += v[j];
        }
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < NT; off <<= 1) {
        const uint64_t v = threadIdx.x >= (uint32_t)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = part[threadIdx.x] - s; // exclusive prefix of this thread's chunk
    for (uint32_t k0 = 0; k0 < per; k0 += G) {
        uint32_t v[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const uint32_t k = k0 + j, i = lo + k;
            v[j] = (k < per && i < nt) ? cnt[i] : 0u;
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const uint32_t k = k0 + j, i = lo + k;
            if (k < per && i < nt) {
                pre[i] = run;
                run += v[j];
                tr.at(5, i);
            }
        }
    }
    if (threadIdx.x == NT - 1) {
        if (TOTAL) pre[nt] = part[NT - 1];
        if (out_count) *out_count = part[NT - 1];
        tr.at(6, nt);
    }
    tr.done(dev::gtid());
}

__global__ __launch_bounds__(1024) void k_scan_tiles(const uint32_t *__restrict__ cnt, uint32_t nt,
                                                     uint64_t *__restrict__ pre, uint64_t *__restrict__ out_count,
                                                     TraceSlot *slot)
{
    __shared__ uint64_t part[1024];
    b_scan_tiles<1024>(cnt, nt, pre, out_count, slot, part);
}

// Tile scans of many tiles (launch mode): segments of kScanSeg counts, one
// 1024-thread block each, kScanPer contiguous counts per thread.
// k_scan_part writes each segment's sum (<= 2^25, so 32 bits), k_scan_tiles
// scans those sums into segment bases, and k_scan_fin writes every tile's
// exclusive prefix from its segment's base.  The sums and the prefixes are
// those of the one-block scan: integer addition in any grouping.
constexpr uint32_t kScanPer = 8;
constexpr uint32_t kScanSeg = 1024 * kScanPer;

// Exclusive prefix of v over a 1024-thread block and the block's total;
// sh: 16 words of LDS.
__device__ __forceinline__ uint64_t block_scan1024(uint64_t v, uint64_t *sh, uint64_t &total)
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t u = __shfl_up(inc, off, 64);
        inc += lane >= (uint32_t)off ? u : 0;
    }
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    uint64_t before = 0, tot = 0;
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        const uint64_t x = sh[k];
        before += k < w ? x : 0;
        tot += x;
    }
    total = tot;
    return before + inc - v;
}

__global__ __launch_bounds__(1024) void k_scan_part(const uint32_t *__restrict__ cnt, uint32_t nt,
                                                    uint32_t *__restrict__ seg_sum, TraceSlot *slot)
{
    __shared__ uint64_t sh[16];
    dev::Atr tr(slot);
    const uint64_t lo = (uint64_t)blockIdx.x * kScanSeg + (uint64_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) v[j] = lo + j < nt ? cnt[lo + j] : 0u;
    uint64_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) {
        s += v[j];
        if (lo + j < nt) tr.at(4, lo + j);
    }
    uint64_t total;
    (void)block_scan1024(s, sh, total);
    if (threadIdx.x == 0) {
        seg_sum[blockIdx.x] = (uint32_t)total;
        tr.at(7, blockIdx.x);
    }
    tr.done(dev::gtid());
}

__global__ __launch_bounds__(1024) void k_scan_fin(const uint32_t *__restrict__ cnt, uint32_t nt,
                                                   const uint64_t *__restrict__ seg_base, uint64_t *__restrict__ pre,
                                                   TraceSlot *slot)
{
    __shared__ uint64_t sh[16];
    dev::Atr tr(slot);
    const uint64_t lo = (uint64_t)blockIdx.x * kScanSeg + (uint64_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) v[j] = lo + j < nt ? cnt[lo + j] : 0u;
    uint64_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) s += v[j];
    uint64_t total;
    uint64_t run = seg_base[blockIdx.x] + block_scan1024(s, sh, total);
    tr.at(7, blockIdx.x);
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) {
        if (lo + j < nt) {
            pre[lo + j] = run;
            tr.at(4, lo + j);
            tr.at(5, lo + j);
        }
        run += v[j];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pre[nt] = seg_base[gridDim.x]; // the total
        tr.at(6, nt);
    }
    tr.done(dev::gtid());
}

// The tile scans of every instance of a batch class, instance vb by block vb:
// its nts tiles' exclusive prefix counts go to pre[vb * nts ..] (each
// instance restarts at zero, so each compacts to its own front) and its
// selected count to out[idx[vb]] when out is given.
__device__ __forceinline__ void b_scan_seg(const uint32_t *__restrict__ cnt, uint32_t nts, uint64_t *__restrict__ pre,
                                           uint64_t *__restrict__ out, const uint64_t *__restrict__ idx,
                                           TraceSlot *slot, uint64_t *part, uint32_t vb)
{
    const uint64_t o = (uint64_t)vb * nts;
    uint64_t *oc = out ? out + uni64(idx[vb]) : nullptr;
    b_scan_tiles<CBLOCK, false>(cnt + o, nts, pre + o, oc, slot, part);
}

__global__ __launch_bounds__(CBLOCK) void k_scan_seg(const uint32_t *__restrict__ cnt, uint32_t nts,
                                                     uint64_t *__restrict__ pre, uint64_t *__restrict__ out,
                                                     const uint64_t *__restrict__ idx, TraceSlot *slot)
{
    __shared__ uint64_t part[CBLOCK];
    b_scan_seg(cnt, nts, pre, out, idx, slot, part, blockIdx.x);
}

// ---------------------------------------------------------- local levels

// Levels q = 0 .. CT_LOG-1 inside each 4096-row tile.  Thread tau holds 16
// rows of each column in registers under layout a: row base(tau, a) | m << a
// for m < 16, base(tau, a) = the 8 bits of tau around the 4-bit hole at a.
// Levels a .. a+3 pair rows inside a thread; layouts a = 0, 4, 8 cover the 12
// levels, with two LDS transposes between them (and one after the coalesced
// load, layout 8).  Rows are read from `src` and written to `dst` (which may
// be the same columns).  cpre (LDS) holds the tile's exclusive prefix counts.
constexpr uint32_t CTPAD = CT + CT / 16; // transpose buffer: row l at l + l/16

__device__ __forceinline__ uint32_t cpos(uint32_t l) { return l + (l >> 4); }

__device__ __forceinline__ uint32_t clay(uint32_t tau, int a) { return (tau & ((1u << a) - 1)) | ((tau >> a) << (a + 4)); }

template <int NW>
__device__ __forceinline__ void ctranspose(uint64_t (&r)[16][NW], uint64_t *buf, uint32_t tau, int a_from, int a_to)
{
    const uint32_t bf = clay(tau, a_from), bt = clay(tau, a_to);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
#pragma unroll
        for (int m = 0; m < 16; ++m) buf[cpos(bf | ((uint32_t)m << a_from))] = r[m][w];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 16; ++m) r[m][w] = buf[cpos(bt | ((uint32_t)m << a_to))];
        __syncthreads();
    }
}

template <int NW, int A>
__device__ __forceinline__ void clevels(uint64_t (&r)[16][NW], const uint16_t *cpre, uint32_t tbase, uint32_t tau)
{
    const uint32_t base = clay(tau, A);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int q = A + k;
        const uint32_t half = 1u << q;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            if (m & (1 << k)) continue;
            const int m2 = m | (1 << k);
            const uint32_t e0 = base | ((uint32_t)m << A);
            const uint32_t b = e0 & ~(2 * half - 1);
            const uint32_t cb = tbase + cpre[b];
            bool sw;
            if (q == 0) {
                const uint32_t e1 = e0 | 1;
                const bool s0 = cpre[e0 + 1] != cpre[e0];
                const bool s1 = cpre[e1 + 1] != cpre[e1];
                sw = ((!s0) & s1) ^ (bool)(cb & 1);
            } else {
                sw = level_flag32(cb, tbase + cpre[b + half], q, e0 & (half - 1));
            }
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint64_t x = r[m][w], y = r[m2][w];
                r[m][w] = sw ? y : x;
                r[m2][w] = sw ? x : y;
            }
        }
    }
}

// W8: every column is 8 bytes wide (no per-load width dispatch)
template <int NW, bool W8, bool TRACE, bool NT = false>
__device__ __forceinline__ void b_compact_local(Cols<NW> src, Cols<NW> dst, const uint8_t *__restrict__ mask,
                                                uint64_t n, const uint64_t *__restrict__ tile_pre, Inst in,
                                                TraceSlot *slot, char *lds_raw, uint32_t vb)
{
    static_assert(CT == 16 * CBLOCK, "16 rows per thread");
    uint64_t *buf = reinterpret_cast<uint64_t *>(lds_raw);          // one column, CTPAD words
    uint16_t *cpre = reinterpret_cast<uint16_t *>(buf + CTPAD);      // CT + 1 counts <= 4096
    __shared__ uint32_t wsum[CBLOCK / 64];
    const uint32_t tid = threadIdx.x;
    const uint64_t tv = (uint64_t)vb * CT;
    const uint64_t t0 = inst_row(in, tv), l0 = inst_local(in, tv), nn = inst_live(in, tv, n);
    const uint64_t tbase = tile_pre[vb];
    constexpr int PER = CT / CBLOCK; // 16 consecutive rows per thread for the scan
    dev::AtrT<TRACE> tr(slot);
    tr.at(7, vb);

    // rows, coalesced (layout 8: row tid + 256 m); issued before the scan
    uint64_t r[16][NW];
#pragma unroll
    for (int m = 0; m < 16; ++m)
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            r[m][w] = W8 ? ld_row<NT>((const uint64_t *)src.c[w].p + t0 + tid + 256u * m)
                         : ld_col<NT>(src.c[w], t0 + tid + 256u * m);
            tr.at(8 + w, t0 + tid + 256u * m);
        }

    // exclusive prefix of sel inside the tile: thread tid owns rows [16 tid, 16 tid + 16)
    uint32_t sel[PER];
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t i = tid * PER + k;
        sel[k] = (mask[t0 + i] != 0) & (l0 + i < nn);
        run += sel[k];
        tr.at(1, t0 + i);
    }
    uint32_t incl = run;
    const uint32_t lane = tid & 63;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off);
        incl += lane >= (uint32_t)off ? v : 0;
    }
    if (lane == 63) wsum[tid >> 6] = incl;
    __syncthreads();
    uint32_t wpre = 0;
    for (uint32_t w = 0; w < (tid >> 6); ++w) wpre += wsum[w];
    uint32_t c = wpre + incl - run;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        cpre[tid * PER + k] = (uint16_t)c;
        c += sel[k];
    }
    if (tid == CBLOCK - 1) cpre[CT] = (uint16_t)c;
    // (the first transpose's barriers also publish cpre)
    ctranspose<NW>(r, buf, tid, 8, 0);
    clevels<NW, 0>(r, cpre, (uint32_t)tbase, tid);
    ctranspose<NW>(r, buf, tid, 0, 4);
    clevels<NW, 4>(r, cpre, (uint32_t)tbase, tid);
    ctranspose<NW>(r, buf, tid, 4, 8);
    clevels<NW, 8>(r, cpre, (uint32_t)tbase, tid);
    uint32_t tid_s = tid; // store addresses recomputed here (register pressure)
    asm volatile("" : "+v"(tid_s));
#pragma unroll
    for (int m = 0; m < 16; ++m)
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            if (W8)
                st_row<NT>((uint64_t *)dst.c[w].p + t0 + tid_s + 256u * m, r[m][w]);
            else
                st_col<NT>(dst.c[w], t0 + tid_s + 256u * m, r[m][w]);
            tr.at(16 + w, t0 + tid_s + 256u * m);
        }
    tr.done(dev::gtid());
}

template <int NW, bool W8, bool TRACE, bool NT = false>
__global__ __launch_bounds__(CBLOCK) void k_compact_local(Cols<NW> src, Cols<NW> dst, const uint8_t *__restrict__ mask,
                                                          uint64_t n, const uint64_t *__restrict__ tile_pre, Inst in,
                                                          TraceSlot *slot)
{
    extern __shared__ __align__(16) char lds_raw[];
    b_compact_local<NW, W8, TRACE, NT>(src, dst, mask, n, tile_pre, in, slot, lds_raw, blockIdx.x);
}

// ---------------------------------------------------------- global levels

// Levels q_lo .. q_lo+S-1 (ascending, q_lo >= CT_LOG): thread g owns the 2^S
// rows base + m * 2^q_lo.  The block start of any pair has every bit below
// q_lo + 1 clear, so its prefix counts are the same for all lanes of a
// wavefront and come from the tile prefix array.
template <int NW, int S, bool TRACE, bool NT = false>
__device__ __forceinline__ void b_compact_global(Cols<NW> cols, const uint64_t *__restrict__ tile_pre, int q_lo,
                                                 Inst in, TraceSlot *slot, uint32_t vb)
{
    dev::AtrT<TRACE> tr(slot);
    const uint64_t g = (uint64_t)vb * 256 + threadIdx.x;
    const uint64_t base = (g & ((1ull << q_lo) - 1)) | ((g >> q_lo) << (q_lo + S));
    // the 2^S rows lie in one instance (q_lo + S <= its size class)
    const uint64_t pb = inst_row(in, base);
    uint64_t r[1 << S][NW];
#pragma unroll
    for (int m = 0; m < (1 << S); ++m)
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            r[m][w] = ld_col<NT>(cols.c[w], pb + ((uint64_t)m << q_lo));
            tr.at(8 + w, pb + ((uint64_t)m << q_lo));
        }
    // uniform part of base (bits >= 8 are shared by the wavefront)
    const uint64_t ubase = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32)) << 32) |
                           __builtin_amdgcn_readfirstlane((uint32_t)base & ~0xFFu);
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int q = q_lo + k;
        const uint64_t half = 1ull << q;
#pragma unroll
        for (int m = 0; m < (1 << S); ++m) {
            if (m & (1 << k)) continue;
            const uint64_t e = base + ((uint64_t)m << q_lo);
            const uint64_t ue = ubase + ((uint64_t)m << q_lo);
            const uint64_t b = ue & ~(2 * half - 1);
            const uint64_t cb = tile_pre[b >> CT_LOG];
            const uint64_t cbh = tile_pre[(b + half) >> CT_LOG];
            tr.at(7, b >> CT_LOG);
            tr.at(7, (b + half) >> CT_LOG);
            const bool sw = level_flag(cb, cbh, half, e & (half - 1));
            const int m2 = m | (1 << k);
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint64_t x = r[m][w], y = r[m2][w];
                r[m][w] = sw ? y : x;
                r[m2][w] = sw ? x : y;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < (1 << S); ++m)
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            st_col<NT>(cols.c[w], pb + ((uint64_t)m << q_lo), r[m][w]);
            tr.at(16 + w, pb + ((uint64_t)m << q_lo));
        }
    tr.done(dev::gtid());
}

template <int NW, int S, bool TRACE, bool NT = false>
__global__ __launch_bounds__(256) void k_compact_global(Cols<NW> cols, const uint64_t *__restrict__ tile_pre, int q_lo,
                                                        Inst in, TraceSlot *slot)
{
    b_compact_global<NW, S, TRACE, NT>(cols, tile_pre, q_lo, in, slot, blockIdx.x);
}

// ------------------------------------------------- LDS-staged global levels

// S = 5..8 levels q_lo .. q_lo+S-1 (q_lo >= CT_LOG) in one pass.  A block
// takes the 4096 rows hibase + m * 2^q_lo + j0 + c (m < 2^S, c < 2^(12-S)):
// local row l = m << (12-S) | c, so level q_lo+k pairs local bit 12-S+k.
// Registers + LDS transposes as in the local pass (layout 4 for local bits
// 12-S .. 7, layout 8 for bits 8 .. 11); the prefix counts at every block
// boundary the levels need, tile_pre at hibase + m * 2^q_lo, go to LDS.
template <int NW, int A, int S>
__device__ __forceinline__ void hlevels(uint64_t (&r)[16][NW], const uint32_t *P, uint32_t tau, int q_lo, uint32_t j0)
{
    constexpr int CB = CT_LOG - S;
    const uint32_t base = clay(tau, A);
#pragma unroll
    for (int bit = A; bit < A + 4; ++bit) {
        if (bit < CB) continue;
        const int k = bit - CB; // level q_lo + k
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            if (m & (1 << (bit - A))) continue;
            const int m2 = m | (1 << (bit - A));
            const uint32_t l0 = base | ((uint32_t)m << A);
            const uint32_t ml = l0 >> CB;
            const uint32_t mb = ml & ~((2u << k) - 1);
            const uint32_t i = ((ml & ((1u << k) - 1)) << q_lo) + j0 + (l0 & ((1u << CB) - 1));
            const bool sw = level_flag32(P[mb], P[mb + (1u << k)], q_lo + k, i);
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint64_t x = r[m][w], y = r[m2][w];
                r[m][w] = sw ? y : x;
                r[m2][w] = sw ? x : y;
            }
        }
    }
}

template <int NW, int S, bool W8, bool TRACE, bool NT = false>
__device__ __forceinline__ void b_compact_hybrid(Cols<NW> cols, const uint64_t *__restrict__ tile_pre, int q_lo,
                                                 Inst in, TraceSlot *slot, char *lds_raw, uint32_t vb)
{
    dev::AtrT<TRACE> tr(slot);
    static_assert(S >= 5 && S <= 8, "hybrid widths 5..8");
    constexpr int CB = CT_LOG - S;
    uint64_t *buf = reinterpret_cast<uint64_t *>(lds_raw);     // CTPAD words
    uint32_t *P = reinterpret_cast<uint32_t *>(buf + CTPAD); // low words of 2^S + 1 prefix counts
    const uint32_t tid = threadIdx.x;
    // row indices fit 32 bits (n_pad <= 2^32)
    const uint32_t runs = (1u << q_lo) >> CB; // blocks per group of 2^(q_lo+S) rows
    const uint32_t hibase = (vb / runs) << (q_lo + S);
    const uint32_t j0 = (vb % runs) << CB;
    // the group's 2^(q_lo+S) rows lie in one instance: physical row of its first
    const uint64_t ph = inst_row(in, hibase);
    for (uint32_t k = tid; k <= (1u << S); k += CBLOCK) {
        P[k] = (uint32_t)tile_pre[((uint64_t)hibase + ((uint64_t)k << q_lo)) >> CT_LOG];
        tr.at(7, ((uint64_t)hibase + ((uint64_t)k << q_lo)) >> CT_LOG);
    }
    uint64_t r[16][NW];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const uint32_t l = tid | ((uint32_t)m << 8);
        const uint32_t g = ((l >> CB) << q_lo) + j0 + (l & ((1u << CB) - 1)); // row in the group
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            r[m][w] = W8 ? ld_row<NT>((const uint64_t *)cols.c[w].p + ph + g) : ld_col<NT>(cols.c[w], ph + g);
            tr.at(8 + w, ph + g);
        }
    }
    ctranspose<NW>(r, buf, tid, 8, 4); // its barriers also publish P
    hlevels<NW, 4, S>(r, P, tid, q_lo, j0);
    ctranspose<NW>(r, buf, tid, 4, 8);
    hlevels<NW, 8, S>(r, P, tid, q_lo, j0);
    // recompute the store addresses here rather than keep the load addresses
    // live across the levels (register pressure)
    uint32_t tid_s = tid;
    asm volatile("" : "+v"(tid_s));
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const uint32_t l = tid_s | ((uint32_t)m << 8);
        const uint32_t g = ((l >> CB) << q_lo) + j0 + (l & ((1u << CB) - 1));
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            if (W8)
                st_row<NT>((uint64_t *)cols.c[w].p + ph + g, r[m][w]);
            else
                st_col<NT>(cols.c[w], ph + g, r[m][w]);
            tr.at(16 + w, ph + g);
        }
    }
    tr.done(dev::gtid());
}

template <int NW, int S, bool W8, bool TRACE, bool NT = false>
__global__ __launch_bounds__(CBLOCK) void k_compact_hybrid(Cols<NW> cols, const uint64_t *__restrict__ tile_pre,
                                                           int q_lo, Inst in, TraceSlot *slot)
{
    extern __shared__ __align__(16) char lds_raw[];
    b_compact_hybrid<NW, S, W8, TRACE, NT>(cols, tile_pre, q_lo, in, slot, lds_raw, blockIdx.x);
}

// ------------------------------------------------------- int128 SUM pass

// Exact SUM(a[i] * b[i]) over rows i < count (count read on the device);
// every row is read, rows >= count contribute a multiplied-out zero.
// red: 256 int128 of LDS.
template <bool NT = false>
__device__ __forceinline__ void b_sum_prod_i128(const int64_t *__restrict__ a, const int64_t *__restrict__ b,
                                                uint64_t n_pad, const uint64_t *__restrict__ count,
                                                __int128 *__restrict__ partial, TraceSlot *slot, __int128 *red,
                                                uint32_t vb, uint32_t nvb)
{
    dev::Atr tr(slot);
    const uint64_t m = *count;
    __int128 acc = 0;
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        const __int128 prod = (__int128)ld_row<NT>(a + i) * (__int128)ld_row<NT>(b + i);
        acc += prod * (__int128)(i < m);
        tr.at(8, i);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[vb] = red[0];
        tr.at(16, vb);
    }
    tr.done(dev::gtid());
}

template <bool NT = false>
__global__ __launch_bounds__(256) void k_sum_prod_i128(const int64_t *__restrict__ a, const int64_t *__restrict__ b,
                                                       uint64_t n_pad, const uint64_t *__restrict__ count,
                                                       __int128 *__restrict__ partial, TraceSlot *slot)
{
    __shared__ __int128 red[256];
    b_sum_prod_i128<NT>(a, b, n_pad, count, partial, slot, red, blockIdx.x, gridDim.x);
}

// One 256-thread block: strided partial sums, then a tree in LDS (exact:
// int128 addition is associative).
__device__ __forceinline__ void b_sum_final(const __int128 *__restrict__ partial, uint32_t np,
                                            const uint64_t *__restrict__ count, uint64_t *__restrict__ out,
                                            TraceSlot *slot, __int128 *red)
{
    dev::Atr tr(slot);
    __int128 s = 0;
    for (uint32_t i = threadIdx.x; i < np; i += 256) {
        s += partial[i];
        tr.at(8, i);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const __int128 t = red[0];
        out[0] = *count;
        out[1] = 0;
        out[2] = (uint64_t)t;
        out[3] = (uint64_t)(t >> 64);
        tr.at(16, 0);
    }
    tr.done(dev::gtid());
}

__global__ __launch_bounds__(256) void k_sum_final(const __int128 *__restrict__ partial, uint32_t np,
                                                   const uint64_t *__restrict__ count, uint64_t *__restrict__ out,
                                                   TraceSlot *slot)
{
    __shared__ __int128 red[256];
    b_sum_final(partial, np, count, out, slot, red);
}

// ------------------------------------------------------------ OP_SCATTER

__device__ __forceinline__ void b_scatter_positions(const uint8_t *__restrict__ mask, uint64_t n,
                                                    const uint64_t *__restrict__ tile_pre, int32_t *__restrict__ out,
                                                    TraceSlot *slot, uint32_t vb)
{
    __shared__ uint32_t wsum[CBLOCK / 64];
    dev::Atr tr(slot);
    const uint32_t tid = threadIdx.x;
    const uint64_t t0 = (uint64_t)vb * CT;
    constexpr int PER = CT / CBLOCK;
    uint32_t sel[PER], run = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint64_t i = t0 + tid * PER + k;
        if (i < n) {
            sel[k] = mask[i] != 0;
            tr.at(1, i);
        } else {
            sel[k] = 0;
        }
        run += sel[k];
    }
    uint32_t incl = run;
    const uint32_t lane = tid & 63;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off);
        incl += lane >= (uint32_t)off ? v : 0;
    }
    if (lane == 63) wsum[tid >> 6] = incl;
    __syncthreads();
    uint32_t wpre = 0;
    for (uint32_t w = 0; w < (tid >> 6); ++w) wpre += wsum[w];
    uint64_t c = tile_pre[vb] + wpre + incl - run;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint64_t i = t0 + tid * PER + k;
        if (sel[k]) { // reference semantics: only selected rows are written (not oblivious)
            out[i] = (int32_t)c;
            tr.at(16, i);
        }
        c += sel[k];
    }
    tr.done(dev::gtid());
}

__global__ __launch_bounds__(CBLOCK) void k_scatter_positions(const uint8_t *__restrict__ mask, uint64_t n,
                                                              const uint64_t *__restrict__ tile_pre,
                                                              int32_t *__restrict__ out, TraceSlot *slot)
{
    b_scatter_positions(mask, n, tile_pre, out, slot, blockIdx.x);
}

} // namespace
} // namespace compact_k
} // namespace vo
