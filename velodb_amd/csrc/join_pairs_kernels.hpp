// join_pairs_kernels.hpp -- device bodies and __global__ wrappers of the
// oblivious equi-join into row pairs (see join_pairs.hip for the method and
// the reference), for launch mode and the resident queue kernel.
#pragma once

#include "vo_internal.hpp"

#include <algorithm>


namespace vo {
namespace jp_k {
namespace {

constexpr int PT = 4096; // scan tile (rows)
constexpr int PB = 256;
constexpr int PPER = PT / PB;
constexpr uint32_t INF32 = 0xFFFFFFFFu;
constexpr uint64_t EMPTY = ~0ull; // element word 0 of an empty slot (target field 0xFFFFFFFF)

constexpr int DT = 2048;         // distribution / fill tile (slots); the resident kernel uses 1024
constexpr int DT_RES = 1024;
constexpr int DH = 256;          // halo: levels 128 .. 1 run in the tile kernel


// ------------------------------------------------------------ union build

__device__ __forceinline__ uint64_t key_code(const void *p, uint32_t type, uint64_t i)
{
    switch (type) {
    case VO_T_BOOLEAN: return ((const uint8_t *)p)[i];
    case VO_T_TINYINT: return (uint64_t)(int64_t)((const int8_t *)p)[i] ^ (1ull << 63);
    case VO_T_SMALLINT: return (uint64_t)(int64_t)((const int16_t *)p)[i] ^ (1ull << 63);
    case VO_T_INTEGER: return (uint64_t)(int64_t)((const int32_t *)p)[i] ^ (1ull << 63);
    case VO_T_BIGINT: return (uint64_t)((const int64_t *)p)[i] ^ (1ull << 63);
    case VO_T_DATE: return ((const uint32_t *)p)[i];
    default: return ((const uint64_t *)p)[i]; // VARCHAR ordinal, UBIGINT
    }
}

template <bool TRACE>
__device__ __forceinline__ void b_jp_build(const void *__restrict__ lk, uint64_t nl, const void *__restrict__ rk, uint64_t nr,
                           uint32_t type, uint64_t n_pad, uint64_t *__restrict__ keys, uint64_t *__restrict__ vals,
                           TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        uint64_t k = ~0ull, v = (3ull << 62) | i;
        if (i < nl) { // public sizes only
            k = key_code(lk, type, i);
            v = i;
            tr.at(1, i);
        } else if (i < nl + nr) {
            k = key_code(rk, type, i - nl);
            v = (2ull << 62) | (i - nl);
            tr.at(2, i - nl);
        }
        keys[i] = k;
        vals[i] = v;
        tr.at(3, i);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ void k_jp_build(const void *__restrict__ lk, uint64_t nl, const void *__restrict__ rk, uint64_t nr,
                           uint32_t type, uint64_t n_pad, uint64_t *__restrict__ keys, uint64_t *__restrict__ vals,
                           TraceSlot *slot)
{
    b_jp_build<TRACE>(lk, nl, rk, nr, type, n_pad, keys, vals, slot, blockIdx.x, gridDim.x);
}

// Key types up to 32 bits: the union as 12-byte (u64 key, u32 payload)
// records, key = code32 << 32 | class << 30 (class 0 left, 2 right, 3 pad),
// payload = row.  code32 orders like key_code, so the sort order (key,
// class, row) is that of the 16-byte form, and ld_union below decodes a row
// to the same (code, class << 62 | row) pair.
__device__ __forceinline__ uint32_t key_code32(const void *p, uint32_t type, uint64_t i)
{
    switch (type) {
    case VO_T_BOOLEAN: return ((const uint8_t *)p)[i];
    case VO_T_TINYINT: return (uint32_t)(int32_t)((const int8_t *)p)[i] ^ 0x80000000u;
    case VO_T_SMALLINT: return (uint32_t)(int32_t)((const int16_t *)p)[i] ^ 0x80000000u;
    case VO_T_INTEGER: return (uint32_t)((const int32_t *)p)[i] ^ 0x80000000u;
    default: return ((const uint32_t *)p)[i]; // DATE
    }
}

bool key_fits_32(uint32_t type)
{
    return type == VO_T_BOOLEAN || type == VO_T_TINYINT || type == VO_T_SMALLINT || type == VO_T_INTEGER ||
           type == VO_T_DATE;
}

template <bool TRACE>
__device__ __forceinline__ void b_jp_build_kp(const void *__restrict__ lk, uint64_t nl, const void *__restrict__ rk, uint64_t nr,
                              uint32_t type, uint64_t n_pad, uint64_t *__restrict__ keys, uint32_t *__restrict__ pay,
                              TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        uint64_t k = ~0ull;
        uint32_t r = (uint32_t)i;
        if (i < nl) { // public sizes only
            k = (uint64_t)key_code32(lk, type, i) << 32;
            tr.at(1, i);
        } else if (i < nl + nr) {
            r = (uint32_t)(i - nl);
            k = ((uint64_t)key_code32(rk, type, i - nl) << 32) | (2ull << 30);
            tr.at(2, i - nl);
        }
        keys[i] = k;
        pay[i] = r;
        tr.at(3, i);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ void k_jp_build_kp(const void *__restrict__ lk, uint64_t nl, const void *__restrict__ rk, uint64_t nr,
                              uint32_t type, uint64_t n_pad, uint64_t *__restrict__ keys, uint32_t *__restrict__ pay,
                              TraceSlot *slot)
{
    b_jp_build_kp<TRACE>(lk, nl, rk, nr, type, n_pad, keys, pay, slot, blockIdx.x, gridDim.x);
}

// Union row i as (key code, class << 62 | row) in either form.
template <bool KP>
__device__ __forceinline__ void ld_union(const uint64_t *__restrict__ keys, const void *__restrict__ vals, uint64_t i,
                                         uint64_t &k, uint64_t &v)
{
    if constexpr (KP) {
        const uint64_t kk = keys[i];
        const uint64_t c = (kk >> 30) & 3;
        k = c == 3 ? ~0ull : (kk >> 32);
        v = (c << 62) | static_cast<const uint32_t *>(vals)[i];
    } else {
        k = keys[i];
        v = static_cast<const uint64_t *>(vals)[i];
    }
}

__device__ __forceinline__ uint32_t cls(uint64_t v) { return (uint32_t)(v >> 62); } // 0 left, 2 right, 3 pad

__device__ __forceinline__ bool same_seg(uint64_t ka, uint64_t va, uint64_t kb, uint64_t vb)
{
    return (int)(ka == kb) & (int)(cls(va) != 3) & (int)(cls(vb) != 3);
}

// ------------------------------------------------------ tile-level scans

// Exclusive scan of nt tile summaries with one block of NT threads (1024 in
// launch mode, 256 in the resident kernel; forward, or backward when REV):
// contiguous chunks per thread, a Hillis-Steele scan of the chunk totals in
// LDS (sh: NT values).  The operators are associative, so NT changes no
// result.
constexpr uint32_t kScanThreads = 1024;

template <uint32_t NT, bool REV, class T, class Op, class Tr>
__device__ void tile_scan(const T *__restrict__ in, uint32_t nt, T *__restrict__ out, T id, Op op, T *sh, Tr &tr)
{
    const uint32_t per = (nt + NT - 1) / NT;
    const uint32_t lo = threadIdx.x * per;
    auto at = [&](uint32_t x) { return REV ? nt - 1 - x : x; };
    // a thread's chunk in groups of 8, loaded before any is combined
    constexpr int G = 8;
    T s = id;
    for (uint32_t k0 = 0; k0 < per; k0 += G) {
        T v[G];
#pragma unroll
        for (int j = 0; j < G; ++j) v[j] = (k0 + j < per && lo + k0 + j < nt) ? in[at(lo + k0 + j)] : id;
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (k0 + j < per && lo + k0 + j < nt) { // public size
                s = op(s, v[j]);
                tr.at(1, at(lo + k0 + j));
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
    T run = threadIdx.x ? sh[threadIdx.x - 1] : id;
    for (uint32_t k0 = 0; k0 < per; k0 += G) {
        T v[G];
#pragma unroll
        for (int j = 0; j < G; ++j) v[j] = (k0 + j < per && lo + k0 + j < nt) ? in[at(lo + k0 + j)] : id;
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (k0 + j < per && lo + k0 + j < nt) {
                out[at(lo + k0 + j)] = run;
                run = op(run, v[j]);
                tr.at(2, at(lo + k0 + j));
            }
    }
}

struct Min2 {
    uint32_t e, q;
};
struct Min2Op {
    __device__ Min2 operator()(const Min2 &a, const Min2 &b) const { return Min2{min(a.e, b.e), min(a.q, b.q)}; }
};

struct Fwd {
    uint32_t s, r; // latest segment start / first right row (maxima)
    uint64_t x;    // sum of m1 over left rows
};
struct FwdOp {
    __device__ Fwd operator()(const Fwd &a, const Fwd &b) const { return Fwd{max(a.s, b.s), max(a.r, b.r), a.x + b.x}; }
};

template <int NW>
struct Elem {
    uint64_t w[NW];
};
template <int NW>
struct LastOp { // the later non-empty element wins
    __device__ Elem<NW> operator()(const Elem<NW> &a, const Elem<NW> &b) const { return b.w[0] != EMPTY ? b : a; }
};

// The three tile carries (sh: NT values of LDS).
template <uint32_t NT, bool TRACE>
__device__ __forceinline__ void b_jp_bwd_carry(const Min2 *__restrict__ tiles, uint32_t nt, Min2 *__restrict__ carry,
                                               TraceSlot *slot, Min2 *sh)
{
    dev::AtrT<TRACE> tr(slot);
    tile_scan<NT, true>(tiles, nt, carry, Min2{INF32, INF32}, Min2Op{}, sh, tr);
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ __launch_bounds__(kScanThreads) void k_jp_bwd_carry(const Min2 *__restrict__ tiles, uint32_t nt,
                                                               Min2 *__restrict__ carry, TraceSlot *slot)
{
    __shared__ Min2 sh[kScanThreads];
    b_jp_bwd_carry<kScanThreads, TRACE>(tiles, nt, carry, slot, sh);
}

template <uint32_t NT, bool TRACE>
__device__ __forceinline__ void b_jp_fwd_carry(const Fwd *__restrict__ tiles, uint32_t nt, Fwd *__restrict__ carry,
                                               TraceSlot *slot, Fwd *sh)
{
    dev::AtrT<TRACE> tr(slot);
    tile_scan<NT, false>(tiles, nt, carry, Fwd{0, 0, 0}, FwdOp{}, sh, tr);
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ __launch_bounds__(kScanThreads) void k_jp_fwd_carry(const Fwd *__restrict__ tiles, uint32_t nt,
                                                               Fwd *__restrict__ carry, TraceSlot *slot)
{
    __shared__ Fwd sh[kScanThreads];
    b_jp_fwd_carry<kScanThreads, TRACE>(tiles, nt, carry, slot, sh);
}

template <uint32_t NT, int NW, bool TRACE>
__device__ __forceinline__ void b_jp_fill_carry(const Elem<NW> *__restrict__ tiles, uint32_t nt,
                                                Elem<NW> *__restrict__ carry, TraceSlot *slot, Elem<NW> *sh)
{
    dev::AtrT<TRACE> tr(slot);
    Elem<NW> id;
#pragma unroll
    for (int w = 0; w < NW; ++w) id.w[w] = EMPTY;
    tile_scan<NT, false>(tiles, nt, carry, id, LastOp<NW>{}, sh, tr);
    tr.done(dev::gtid());
}

template <int NW, bool TRACE>
__global__ __launch_bounds__(kScanThreads) void k_jp_fill_carry(const Elem<NW> *__restrict__ tiles, uint32_t nt,
                                                                Elem<NW> *__restrict__ carry, TraceSlot *slot)
{
    __shared__ Elem<NW> sh[kScanThreads];
    b_jp_fill_carry<kScanThreads, NW, TRACE>(tiles, nt, carry, slot, sh);
}

// --------------------------------------------- backward scan (step 2)

// Row events of the backward scan: e = end of the key segment (i+1 at the
// segment's last row), q = i at a right row, else e.
struct RowsBuf {
    uint64_t k[PPER + 1], v[PPER + 1]; // the thread's rows and the next one
};

template <bool KP>
__device__ __forceinline__ void load_rows(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                          uint64_t t0, uint64_t n_pad, uint64_t *k, uint64_t *v)
{
#pragma unroll
    for (int r = 0; r < PPER; ++r) ld_union<KP>(keys, vals, t0 + r, k[r], v[r]);
    const bool has_next = t0 + PPER < n_pad; // public
    uint64_t nk, nv;
    ld_union<KP>(keys, vals, has_next ? t0 + PPER : t0, nk, nv);
    k[PPER] = has_next ? nk : ~0ull;
    v[PPER] = has_next ? nv : (3ull << 62);
}

template <bool TRACE, bool KP>
__device__ __forceinline__ void b_jp_bwd_tiles(const uint64_t *__restrict__ keys,
                                                     const void *__restrict__ vals, uint64_t n_pad,
                                                     Min2 *__restrict__ tiles, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    __shared__ Min2 red[PB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * PT + (uint64_t)threadIdx.x * PPER;
    uint64_t k[PPER + 1], v[PPER + 1];
    load_rows<KP>(keys, vals, t0, n_pad, k, v);
    tr.at(1, t0);
    uint32_t me = INF32, mq = INF32;
#pragma unroll
    for (int r = 0; r < PPER; ++r) {
        const uint32_t i = (uint32_t)(t0 + r);
        const bool end = !same_seg(k[r], v[r], k[r + 1], v[r + 1]);
        const uint32_t e = end ? i + 1 : INF32;
        const uint32_t q = cls(v[r]) == 2 ? i : e;
        me = min(me, e);
        mq = min(mq, q);
    }
    for (int off = 32; off > 0; off >>= 1) {
        me = min(me, (uint32_t)__shfl_xor(me, off));
        mq = min(mq, (uint32_t)__shfl_xor(mq, off));
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = Min2{me, mq};
    __syncthreads();
    if (threadIdx.x == 0) {
        Min2 m = red[0];
        for (int w = 1; w < PB / 64; ++w) m = Min2Op{}(m, red[w]);
        tiles[vb] = m;
        tr.at(2, vb);
    }
    tr.done(dev::gtid());
}

template <bool TRACE, bool KP>
__global__ __launch_bounds__(PB) void k_jp_bwd_tiles(const uint64_t *__restrict__ keys,
                                                     const void *__restrict__ vals, uint64_t n_pad,
                                                     Min2 *__restrict__ tiles, TraceSlot *slot)
{
    b_jp_bwd_tiles<TRACE, KP>(keys, vals, n_pad, tiles, slot, blockIdx.x, gridDim.x);
}

// Per row: E and Q (suffix minima), aux = m1 for left rows, E for right
// rows; per tile: the forward-scan summary (latest start, latest first right
// row, sum of m1 over left rows).
template <bool TRACE, bool KP>
__device__ __forceinline__ void b_jp_bwd_rows(const uint64_t *__restrict__ keys,
                                                    const void *__restrict__ vals, uint64_t n_pad,
                                                    const Min2 *__restrict__ carry, uint32_t *__restrict__ aux,
                                                    Fwd *__restrict__ ftiles, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    __shared__ Min2 wt[PB / 64];
    __shared__ Fwd fr[PB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * PT + (uint64_t)threadIdx.x * PPER;
    uint64_t k[PPER + 1], v[PPER + 1];
    load_rows<KP>(keys, vals, t0, n_pad, k, v);
    const bool has_prev = t0 > 0; // public
    uint64_t pk, pv;
    ld_union<KP>(keys, vals, has_prev ? t0 - 1 : 0, pk, pv);
    pk = has_prev ? pk : ~0ull;
    pv = has_prev ? pv : (3ull << 62);
    tr.at(1, t0);
    uint32_t me = INF32, mq = INF32;
#pragma unroll
    for (int r = 0; r < PPER; ++r) {
        const uint32_t i = (uint32_t)(t0 + r);
        const bool end = !same_seg(k[r], v[r], k[r + 1], v[r + 1]);
        const uint32_t e = end ? i + 1 : INF32;
        me = min(me, e);
        mq = min(mq, cls(v[r]) == 2 ? i : e);
    }
    // minima over the later lanes, waves and tiles
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t se = me, sq = mq;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t oe = __shfl_down(se, off, 64), oq = __shfl_down(sq, off, 64);
        const bool take = lane + off < 64;
        se = take ? min(se, oe) : se;
        sq = take ? min(sq, oq) : sq;
    }
    uint32_t ce = __shfl_down(se, 1, 64), cq = __shfl_down(sq, 1, 64);
    ce = lane == 63 ? INF32 : ce;
    cq = lane == 63 ? INF32 : cq;
    if (lane == 0) wt[wv] = Min2{se, sq};
    __syncthreads();
    const Min2 c = carry[vb];
    tr.at(2, vb);
    ce = min(ce, c.e);
    cq = min(cq, c.q);
    for (uint32_t w = wv + 1; w < PB / 64; ++w) {
        ce = min(ce, wt[w].e);
        cq = min(cq, wt[w].q);
    }
    uint32_t E[PPER], Q[PPER];
#pragma unroll
    for (int r = PPER - 1; r >= 0; --r) {
        const uint32_t i = (uint32_t)(t0 + r);
        const bool end = !same_seg(k[r], v[r], k[r + 1], v[r + 1]);
        const uint32_t e = end ? i + 1 : INF32;
        ce = min(ce, e);
        cq = min(cq, cls(v[r]) == 2 ? i : e);
        E[r] = ce;
        Q[r] = cq;
    }
    // forward summary of this thread's rows
    Fwd f{0, 0, 0};
#pragma unroll
    for (int r = 0; r < PPER; ++r) {
        const uint32_t i = (uint32_t)(t0 + r);
        const uint64_t kp = r ? k[r - 1] : pk, vp = r ? v[r - 1] : pv;
        const bool start = !same_seg(kp, vp, k[r], v[r]);
        const uint32_t c0 = cls(v[r]);
        const bool rstart = (c0 == 2) & (start | (cls(vp) != 2));
        const uint32_t m1 = E[r] - Q[r];
        f.s = start ? i : f.s;
        f.r = rstart ? i : f.r;
        f.x += c0 == 0 ? m1 : 0;
        aux[t0 + r] = c0 == 0 ? m1 : (c0 == 2 ? E[r] : 0);
    }
    tr.at(3, t0);
    for (int off = 32; off > 0; off >>= 1) {
        const Fwd o{(uint32_t)__shfl_xor(f.s, off), (uint32_t)__shfl_xor(f.r, off),
                    (uint64_t)__shfl_xor((unsigned long long)f.x, off)};
        f = FwdOp{}(f, o);
    }
    if (lane == 0) fr[wv] = f;
    __syncthreads();
    if (threadIdx.x == 0) {
        Fwd t = fr[0];
        for (int w = 1; w < PB / 64; ++w) t = FwdOp{}(t, fr[w]);
        ftiles[vb] = t;
        tr.at(4, vb);
    }
    tr.done(dev::gtid());
}

template <bool TRACE, bool KP>
__global__ __launch_bounds__(PB) void k_jp_bwd_rows(const uint64_t *__restrict__ keys,
                                                    const void *__restrict__ vals, uint64_t n_pad,
                                                    const Min2 *__restrict__ carry, uint32_t *__restrict__ aux,
                                                    Fwd *__restrict__ ftiles, TraceSlot *slot)
{
    b_jp_bwd_rows<TRACE, KP>(keys, vals, n_pad, carry, aux, ftiles, slot, blockIdx.x, gridDim.x);
}

// --------------------------------------------- forward scan (step 3)

// Per row: targets and the compaction columns / masks; the last thread of
// the last tile writes n_out.
template <bool TRACE, bool KP>
__device__ __forceinline__ void b_jp_fwd_rows(const uint64_t *__restrict__ keys,
                                                    const void *__restrict__ vals,
                                                    const uint32_t *__restrict__ aux, uint64_t n_pad,
                                                    const Fwd *__restrict__ carry, uint64_t *__restrict__ cl,
                                                    uint8_t *__restrict__ ml, uint64_t *__restrict__ cr0,
                                                    uint64_t *__restrict__ cr1, uint8_t *__restrict__ mr,
                                                    uint64_t *__restrict__ n_out, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    __shared__ Fwd wt[PB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * PT + (uint64_t)threadIdx.x * PPER;
    uint64_t k[PPER], v[PPER];
    uint32_t ax[PPER];
#pragma unroll
    for (int r = 0; r < PPER; ++r) {
        ld_union<KP>(keys, vals, t0 + r, k[r], v[r]);
        ax[r] = aux[t0 + r];
    }
    const bool has_prev = t0 > 0;
    uint64_t pk, pv;
    ld_union<KP>(keys, vals, has_prev ? t0 - 1 : 0, pk, pv);
    pk = has_prev ? pk : ~0ull;
    pv = has_prev ? pv : (3ull << 62);
    tr.at(1, t0);
    bool st[PPER], rs[PPER];
    Fwd f{0, 0, 0};
#pragma unroll
    for (int r = 0; r < PPER; ++r) {
        const uint32_t i = (uint32_t)(t0 + r);
        const uint64_t kp = r ? k[r - 1] : pk, vp = r ? v[r - 1] : pv;
        st[r] = !same_seg(kp, vp, k[r], v[r]);
        rs[r] = (int)(cls(v[r]) == 2) & ((int)st[r] | (int)(cls(vp) != 2));
        f.s = st[r] ? i : f.s;
        f.r = rs[r] ? i : f.r;
        f.x += cls(v[r]) == 0 ? ax[r] : 0;
    }
    // exclusive prefix over earlier lanes, waves and tiles
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    Fwd in = f;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const Fwd o{(uint32_t)__shfl_up(in.s, off, 64), (uint32_t)__shfl_up(in.r, off, 64),
                    (uint64_t)__shfl_up((unsigned long long)in.x, off, 64)};
        in = lane >= (uint32_t)off ? FwdOp{}(o, in) : in;
    }
    Fwd ex{(uint32_t)__shfl_up(in.s, 1, 64), (uint32_t)__shfl_up(in.r, 1, 64),
           (uint64_t)__shfl_up((unsigned long long)in.x, 1, 64)};
    ex = lane == 0 ? Fwd{0, 0, 0} : ex;
    if (lane == 63) wt[wv] = in;
    __syncthreads();
    Fwd c = carry[vb];
    tr.at(2, vb);
    for (uint32_t w = 0; w < wv; ++w) c = FwdOp{}(c, wt[w]);
    c = FwdOp{}(c, ex);
    uint32_t S = c.s, R = c.r;
    uint64_t X = c.x;
#pragma unroll
    for (int r = 0; r < PPER; ++r) {
        const uint64_t i = t0 + r;
        S = st[r] ? (uint32_t)i : S;
        R = rs[r] ? (uint32_t)i : R;
        const uint32_t c0 = cls(v[r]);
        const uint64_t row = v[r] & 0xFFFFFFFFull;
        // left row: ax = m1, target X
        const uint64_t tl = X;
        // right row: ax = E
        const uint64_t m0 = (uint64_t)(R - S), m1 = (uint64_t)(ax[r] - R), b = i - R;
        const uint64_t B = X - m0 * m1;
        const uint64_t tr_ = B + b * m0;
        cl[i] = (tl << 32) | row;
        ml[i] = (c0 == 0) & (ax[r] != 0);
        cr0[i] = (tr_ << 32) | row;
        cr1[i] = ((B + b) & 0xFFFFFFFFull) | (m1 << 32);
        mr[i] = (c0 == 2) & (m0 != 0);
        X += c0 == 0 ? ax[r] : 0;
    }
    tr.at(3, t0);
    if (vb == nvb - 1 && threadIdx.x == PB - 1) {
        n_out[0] = X;
        tr.at(4, 0);
    }
    tr.done(dev::gtid());
}

template <bool TRACE, bool KP>
__global__ __launch_bounds__(PB) void k_jp_fwd_rows(const uint64_t *__restrict__ keys,
                                                    const void *__restrict__ vals,
                                                    const uint32_t *__restrict__ aux, uint64_t n_pad,
                                                    const Fwd *__restrict__ carry, uint64_t *__restrict__ cl,
                                                    uint8_t *__restrict__ ml, uint64_t *__restrict__ cr0,
                                                    uint64_t *__restrict__ cr1, uint8_t *__restrict__ mr,
                                                    uint64_t *__restrict__ n_out, TraceSlot *slot)
{
    b_jp_fwd_rows<TRACE, KP>(keys, vals, aux, n_pad, carry, cl, ml, cr0, cr1, mr, n_out, slot, blockIdx.x, gridDim.x);
}

// ------------------------------------------------ distribution (step 5)

__device__ __forceinline__ uint32_t tgt(uint64_t w0) { return (uint32_t)(w0 >> 32); }

// Copy of the compacted rows into the first slots; slots >= k are empty.
template <int NW, bool TRACE>
__device__ __forceinline__ void b_jp_dist_init(const uint64_t *__restrict__ s0, const uint64_t *__restrict__ s1, uint64_t src_rows,
                               const uint64_t *__restrict__ kdev, uint64_t m_pad, uint64_t *__restrict__ d0,
                               uint64_t *__restrict__ d1, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    const uint64_t kk = kdev[0];
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < m_pad; i += (uint64_t)nvb * 256) {
        uint64_t a = EMPTY, b = EMPTY;
        if (i < src_rows) { // public
            a = s0[i];
            if constexpr (NW == 2) b = s1[i];
            tr.at(1, i);
        }
        const bool valid = i < kk;
        d0[i] = valid ? a : EMPTY;
        if constexpr (NW == 2) d1[i] = valid ? b : EMPTY;
        tr.at(2, i);
    }
    tr.done(dev::gtid());
}

template <int NW, bool TRACE>
__global__ void k_jp_dist_init(const uint64_t *__restrict__ s0, const uint64_t *__restrict__ s1, uint64_t src_rows,
                               const uint64_t *__restrict__ kdev, uint64_t m_pad, uint64_t *__restrict__ d0,
                               uint64_t *__restrict__ d1, TraceSlot *slot)
{
    b_jp_dist_init<NW, TRACE>(s0, s1, src_rows, kdev, m_pad, d0, d1, slot, blockIdx.x, gridDim.x);
}

// S levels j_top, j_top/2, .., j_top/2^(S-1) in one pass: the value of slot
// p after them depends only on the 2^S slots p - m * (j_top >> (S-1)),
// m < 2^S, so each output reads those (the neighbours' reads are the same
// rows and come from the caches) and applies the levels in order.
#ifndef VO_DIST_FUSE
#define VO_DIST_FUSE 3
#endif
constexpr int kDistFuse = VO_DIST_FUSE;
static_assert(kDistFuse >= 1 && kDistFuse <= 4, "1 to 4 distribution levels per pass");

template <int NW>
__device__ __forceinline__ Elem<NW> dist_step(const Elem<NW> &cur, const Elem<NW> &prv, int64_t p, uint64_t j)
{
    const bool in = (prv.w[0] != EMPTY) & ((int64_t)tgt(prv.w[0]) >= p);
    const bool out = (cur.w[0] != EMPTY) & ((int64_t)tgt(cur.w[0]) >= p + (int64_t)j);
    Elem<NW> r;
    r.w[0] = in ? prv.w[0] : (out ? EMPTY : cur.w[0]);
    if constexpr (NW == 2) r.w[1] = in ? prv.w[1] : cur.w[1];
    return r;
}

template <int S, int NW, class Tr>
__device__ __forceinline__ Elem<NW> dist_val(const uint64_t *__restrict__ i0, const uint64_t *__restrict__ i1,
                                             int64_t p, uint64_t j_top, Tr &tr)
{
    if constexpr (S == 0) {
        Elem<NW> e;
        e.w[0] = EMPTY;
        if constexpr (NW == 2) e.w[1] = EMPTY;
        if (p >= 0) { // public: the position only
            e.w[0] = i0[p];
            if constexpr (NW == 2) e.w[1] = i1[p];
            tr.at(1, (uint64_t)p);
        }
        return e;
    } else {
        const uint64_t j = j_top >> (S - 1);
        const Elem<NW> a = dist_val<S - 1, NW>(i0, i1, p, j_top, tr);
        const Elem<NW> b = dist_val<S - 1, NW>(i0, i1, p - (int64_t)j, j_top, tr);
        return dist_step<NW>(a, b, p, j);
    }
}

template <int NW, int S, bool TRACE>
__device__ __forceinline__ void b_jp_dist_levels(const uint64_t *__restrict__ i0, const uint64_t *__restrict__ i1, uint64_t m_pad,
                                 uint64_t j_top, uint64_t *__restrict__ o0, uint64_t *__restrict__ o1,
                                 TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t p = (uint64_t)vb * 256 + threadIdx.x; p < m_pad; p += (uint64_t)nvb * 256) {
        const Elem<NW> e = dist_val<S, NW>(i0, i1, (int64_t)p, j_top, tr);
        o0[p] = e.w[0];
        if constexpr (NW == 2) o1[p] = e.w[1];
        tr.at(2, p);
    }
    tr.done(dev::gtid());
}

template <int NW, int S, bool TRACE>
__global__ void k_jp_dist_levels(const uint64_t *__restrict__ i0, const uint64_t *__restrict__ i1, uint64_t m_pad,
                                 uint64_t j_top, uint64_t *__restrict__ o0, uint64_t *__restrict__ o1,
                                 TraceSlot *slot)
{
    b_jp_dist_levels<NW, S, TRACE>(i0, i1, m_pad, j_top, o0, o1, slot, blockIdx.x, gridDim.x);
}

// The same levels on the lattice of the smallest stride u = j_top >> (S-1):
// slot p = q + k u with residue q < u and row k.  Output row k reads rows
// k, k-1, .., k-(2^S-1) of the same residue only, so a block takes 256
// consecutive residues and K rows plus the 2^S - 1 rows before them into
// registers and runs the S levels there, largest stride first (each slot
// read (K + 2^S - 1) / K times instead of 2^S).  A level with lattice
// stride s updates rows from the top down, so row r - s still holds the
// previous level; rows below the running sum of strides hold partial values
// that no output row reads.  Rows k < 0 (p < 0) are empty, as in dist_val.
// Needs u >= 256 and K <= m_pad / u.
template <int NW, int S, int K, bool TRACE>
__device__ __forceinline__ void b_jp_dist_lattice(const uint64_t *__restrict__ i0,
                                                         const uint64_t *__restrict__ i1, uint64_t u,
                                                         uint64_t *__restrict__ o0, uint64_t *__restrict__ o1,
                                                         TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    constexpr int H = (1 << S) - 1;
    constexpr int W = K + H;
    dev::AtrT<TRACE> tr(slot);
    const uint64_t res_blocks = u / 256;
    const uint64_t q = (vb % res_blocks) * 256 + threadIdx.x;
    const int64_t k0 = (int64_t)(vb / res_blocks) * K - H;
    Elem<NW> v[W];
#pragma unroll
    for (int r = 0; r < W; ++r) {
        const int64_t k = k0 + r;
#pragma unroll
        for (int w = 0; w < NW; ++w) v[r].w[w] = EMPTY;
        if (k >= 0) { // public: the row only
            const uint64_t pp = q + (uint64_t)k * u;
            v[r].w[0] = i0[pp];
            if constexpr (NW == 2) v[r].w[1] = i1[pp];
            tr.at(1, pp);
        }
    }
#pragma unroll
    for (int l = 0; l < S; ++l) {
        const int st = 1 << (S - 1 - l); // lattice stride of this level
        const uint64_t j = (uint64_t)st * u;
#pragma unroll
        for (int r = W - 1; r >= 1; --r) {
            if (r < st) continue; // compile-time after unrolling
            const int64_t pp = (int64_t)q + (k0 + r) * (int64_t)u;
            v[r] = dist_step<NW>(v[r], v[r - st], pp, j);
        }
    }
#pragma unroll
    for (int r = H; r < W; ++r) {
        const uint64_t pp = q + (uint64_t)(k0 + r) * u;
        o0[pp] = v[r].w[0];
        if constexpr (NW == 2) o1[pp] = v[r].w[1];
        tr.at(2, pp);
    }
    tr.done(dev::gtid());
}

template <int NW, int S, int K, bool TRACE>
__global__ __launch_bounds__(256) void k_jp_dist_lattice(const uint64_t *__restrict__ i0,
                                                         const uint64_t *__restrict__ i1, uint64_t u,
                                                         uint64_t *__restrict__ o0, uint64_t *__restrict__ o1,
                                                         TraceSlot *slot)
{
    b_jp_dist_lattice<NW, S, K, TRACE>(i0, i1, u, o0, o1, slot, blockIdx.x, gridDim.x);
}

// Levels 128 .. 1 of one TS-slot tile in LDS (window of the 256 slots
// before it), then the tile's last non-empty element for the fill.  Every
// output slot depends on the 256 slots before it only, so the tile size
// changes no result.  lds: [2][NW][TS + 256] words.
template <int NW, int TS, bool TRACE>
__device__ __forceinline__ void b_jp_dist_tile(const uint64_t *__restrict__ i0, const uint64_t *__restrict__ i1,
                                               uint64_t *__restrict__ o0, uint64_t *__restrict__ o1,
                                               Elem<NW> *__restrict__ last, TraceSlot *slot, uint64_t *lds,
                                               uint32_t vb)
{
    constexpr int DW = TS + DH, DPER = DW / PB, FPER = TS / PB;
    dev::AtrT<TRACE> tr(slot);
    const int64_t base = (int64_t)vb * TS - DH;
#pragma unroll
    for (int r = 0; r < DPER; ++r) {
        const int x = threadIdx.x + r * PB;
        const int64_t g = base + x;
        uint64_t a = EMPTY, b = EMPTY;
        if (g >= 0) { // public
            a = i0[g];
            if constexpr (NW == 2) b = i1[g];
            tr.at(1, (uint64_t)g);
        }
        lds[x] = a;
        if constexpr (NW == 2) lds[DW + x] = b;
    }
    __syncthreads();
    int cur = 0;
#pragma unroll
    for (int j = DH / 2; j >= 1; j >>= 1) {
        const uint64_t *s = lds + cur * NW * DW;
        uint64_t *d = lds + (cur ^ 1) * NW * DW;
#pragma unroll
        for (int r = 0; r < DPER; ++r) {
            const int x = threadIdx.x + r * PB;
            const int64_t g = base + x;
            const uint64_t c0 = s[x];
            const int xp = x >= j ? x - j : x; // window-relative; x < j only in the halo
            const uint64_t q0 = x >= j ? s[xp] : EMPTY;
            const bool in = (q0 != EMPTY) & ((int64_t)tgt(q0) >= g);
            const bool out = (c0 != EMPTY) & ((int64_t)tgt(c0) >= g + j);
            d[x] = in ? q0 : (out ? EMPTY : c0);
            if constexpr (NW == 2) d[DW + x] = in ? s[DW + xp] : s[DW + x];
        }
        cur ^= 1;
        __syncthreads();
    }
    const uint64_t *s = lds + cur * NW * DW;
    Elem<NW> lst;
#pragma unroll
    for (int w = 0; w < NW; ++w) lst.w[w] = EMPTY;
#pragma unroll
    for (int r = 0; r < FPER; ++r) {
        const int x = DH + threadIdx.x + r * PB;
        const uint64_t g = (uint64_t)vb * TS + threadIdx.x + r * PB;
        Elem<NW> e;
        e.w[0] = s[x];
        if constexpr (NW == 2) e.w[1] = s[DW + x];
        o0[g] = e.w[0];
        if constexpr (NW == 2) o1[g] = e.w[1];
        tr.at(2, g);
        lst = LastOp<NW>{}(lst, e);
    }
    // the latest non-empty element over threads: slots are interleaved
    // (thread t holds t, t+256, ...), so compare slot positions
    __shared__ int wbest[PB / 64];
    __shared__ Elem<NW> win;
    int pos = -1;
#pragma unroll
    for (int r = 0; r < FPER; ++r) pos = s[DH + threadIdx.x + r * PB] != EMPTY ? (int)(threadIdx.x + r * PB) : pos;
    int wmax = pos;
    for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off));
    if ((threadIdx.x & 63) == 0) wbest[threadIdx.x >> 6] = wmax;
    __syncthreads();
    int best = wbest[0];
    for (int w = 1; w < PB / 64; ++w) best = max(best, wbest[w]);
    if (pos >= 0 && pos == best) win = lst; // one thread: slot positions are distinct
    __syncthreads();
    if (threadIdx.x == 0) {
        Elem<NW> r;
#pragma unroll
        for (int w = 0; w < NW; ++w) r.w[w] = best >= 0 ? win.w[w] : EMPTY;
        last[vb] = r;
        tr.at(3, vb);
    }
    tr.done(dev::gtid());
}

template <int NW, int TS, bool TRACE>
__global__ __launch_bounds__(PB) void k_jp_dist_tile(const uint64_t *__restrict__ i0, const uint64_t *__restrict__ i1,
                                                     uint64_t *__restrict__ o0, uint64_t *__restrict__ o1,
                                                     Elem<NW> *__restrict__ last, TraceSlot *slot)
{
    extern __shared__ uint64_t lds[];
    b_jp_dist_tile<NW, TS, TRACE>(i0, i1, o0, o1, last, slot, lds, blockIdx.x);
}

// ---------------------------------------------------- fill (step 6)

// SIDE 0: out_left[p] = left row of slot p (p < n_out) or -1 (p < out_cap).
// SIDE 1: keys[p] = final row of the right copy in slot p, payload = right
//         row; slots >= n_out get (~0, p).
template <int NW, int SIDE, int TS, bool TRACE>
__device__ __forceinline__ void b_jp_fill_rows(const uint64_t *__restrict__ i0, const uint64_t *__restrict__ i1,
                                               const Elem<NW> *__restrict__ carry, uint64_t n_out, uint64_t out_cap,
                                               int64_t *__restrict__ out_left, uint64_t *__restrict__ skeys,
                                               uint32_t *__restrict__ spay, TraceSlot *slot, uint32_t vb)
{
    constexpr int FPER = TS / PB;
    __shared__ Elem<NW> wt[PB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t g0 = (uint64_t)vb * TS + (uint64_t)threadIdx.x * FPER;
    Elem<NW> e[FPER];
    Elem<NW> lst;
#pragma unroll
    for (int w = 0; w < NW; ++w) lst.w[w] = EMPTY;
#pragma unroll
    for (int r = 0; r < FPER; ++r) {
        e[r].w[0] = i0[g0 + r];
        if constexpr (NW == 2) e[r].w[1] = i1[g0 + r];
        lst = LastOp<NW>{}(lst, e[r]);
    }
    tr.at(1, g0);
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    Elem<NW> in = lst;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        Elem<NW> o;
#pragma unroll
        for (int w = 0; w < NW; ++w) o.w[w] = __shfl_up((unsigned long long)in.w[w], off, 64);
        in = lane >= (uint32_t)off ? LastOp<NW>{}(o, in) : in;
    }
    Elem<NW> ex;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        ex.w[w] = __shfl_up((unsigned long long)in.w[w], 1, 64);
        ex.w[w] = lane == 0 ? EMPTY : ex.w[w];
    }
    if (lane == 63) wt[wv] = in;
    __syncthreads();
    Elem<NW> c = carry[vb];
    tr.at(2, vb);
    for (uint32_t w = 0; w < wv; ++w) c = LastOp<NW>{}(c, wt[w]);
    c = LastOp<NW>{}(c, ex);
#pragma unroll
    for (int r = 0; r < FPER; ++r) {
        const uint64_t g = g0 + r;
        c = LastOp<NW>{}(c, e[r]);
        const bool live = g < n_out; // public
        if constexpr (SIDE == 0) {
            if (g < out_cap) out_left[g] = live ? (int64_t)(c.w[0] & 0xFFFFFFFFull) : -1;
        } else {
            const uint64_t a = g - tgt(c.w[0]);
            const uint64_t fin = (c.w[1] & 0xFFFFFFFFull) + a * (c.w[1] >> 32);
            skeys[g] = live ? fin : ~0ull;
            spay[g] = live ? (uint32_t)c.w[0] : (uint32_t)g;
        }
    }
    tr.at(3, g0);
    tr.done(dev::gtid());
}

template <int NW, int SIDE, int TS, bool TRACE>
__global__ __launch_bounds__(PB) void k_jp_fill_rows(const uint64_t *__restrict__ i0, const uint64_t *__restrict__ i1,
                                                     const Elem<NW> *__restrict__ carry, uint64_t n_out,
                                                     uint64_t out_cap, int64_t *__restrict__ out_left,
                                                     uint64_t *__restrict__ skeys, uint32_t *__restrict__ spay,
                                                     TraceSlot *slot)
{
    b_jp_fill_rows<NW, SIDE, TS, TRACE>(i0, i1, carry, n_out, out_cap, out_left, skeys, spay, slot, blockIdx.x);
}

// out_right[p] = sorted right row (p < n_out) or -1; out_left[p] = -1 for
// p >= m_pad.
template <bool TRACE>
__device__ __forceinline__ void b_jp_out(const uint32_t *__restrict__ spay, uint64_t m_pad, uint64_t n_out, uint64_t out_cap,
                         int64_t *__restrict__ out_left, int64_t *__restrict__ out_right, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t g = (uint64_t)vb * 256 + threadIdx.x; g < out_cap; g += (uint64_t)nvb * 256) {
        int64_t r = -1;
        if (g < m_pad) { // public
            const uint32_t p = spay[g];
            r = g < n_out ? (int64_t)p : -1;
            tr.at(1, g);
        } else {
            out_left[g] = -1;
        }
        out_right[g] = r;
        tr.at(2, g);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ void k_jp_out(const uint32_t *__restrict__ spay, uint64_t m_pad, uint64_t n_out, uint64_t out_cap,
                         int64_t *__restrict__ out_left, int64_t *__restrict__ out_right, TraceSlot *slot)
{
    b_jp_out<TRACE>(spay, m_pad, n_out, out_cap, out_left, out_right, slot, blockIdx.x, gridDim.x);
}

} // namespace
} // namespace jp_k
} // namespace vo
