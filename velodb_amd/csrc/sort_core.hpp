// sort_core.hpp -- device building blocks of the oblivious bitonic network.
//
// The network is the reference's: for k = 2, 4, ..., N and j = k/2, ..., 1,
// element i is compare-exchanged with i ^ j, ascending when (i & k) == 0
// (include/cuda/sort.cuh:104-117).  Stage (k = 2^p, j = 2^q) is written
// (p, q).  The schedule here executes exactly those compare-exchanges in the
// same per-element order, but groups stages so that one HBM pass applies many
// of them:
//   * tile kernels hold a 2^t-element tile per block; each thread keeps
//     2^R_LOG elements in registers whose indices differ in R_LOG adjacent
//     bits [a, a + R_LOG) ("layout a"); stages on those bits run in
//     registers, and a change of layout goes through LDS;
//   * global kernels apply S consecutive stages (p, q_lo + S - 1 .. q_lo) to
//     groups of 2^S elements at stride 2^q_lo, one group per thread.
// Every compare-exchange reads and writes both slots and selects with
// v_cndmask, so the address stream depends only on the public sizes.
#pragma once

#include "vo_internal.hpp"

namespace vo {
namespace dev {

// ---------------------------------------------------------------- records

// A record: KW 64-bit key words compared lexicographically, then (when TIE)
// the payload.  Stored as KW + 1 separate arrays (structure of arrays).
template <int KW_, typename P_, bool TIE_>
struct Rec {
    static constexpr int KW = KW_;
    using P = P_;
    static constexpr bool TIE = TIE_;
    uint64_t k[KW];
    P p;
};

template <class R>
struct Bufs {
    uint64_t *k[R::KW];
    typename R::P *p;
};

template <class R>
__device__ __forceinline__ bool rec_gt(const R &a, const R &b)
{
    bool gt = false, eq = true;
#pragma unroll
    for (int w = 0; w < R::KW; ++w) {
        gt = gt | (eq & (a.k[w] > b.k[w]));
        eq = eq & (a.k[w] == b.k[w]);
    }
    if constexpr (R::TIE) gt = gt | (eq & (a.p > b.p));
    return gt;
}

// Compare-exchange: afterwards a <= b when asc, a >= b otherwise.  Ties
// (equal over every compared field) are left in place, as needSwap does
// (sort.cuh:31-40).
template <class R>
__device__ __forceinline__ void cmpx(R &a, R &b, bool asc)
{
    const bool sw = asc ? rec_gt(a, b) : rec_gt(b, a);
#pragma unroll
    for (int w = 0; w < R::KW; ++w) {
        const uint64_t x = a.k[w], y = b.k[w];
        a.k[w] = sw ? y : x;
        b.k[w] = sw ? x : y;
    }
    const typename R::P x = a.p, y = b.p;
    a.p = sw ? y : x;
    b.p = sw ? x : y;
}

// Direction of stage level p for (virtual) element index g: ascending when
// bit p of g is clear, and always at the top level p_top of an instance.
__device__ __forceinline__ bool dir_asc(uint64_t g, int p, int p_top)
{
    return (p >= p_top) | (((g >> p) & 1) == 0);
}

// Virtual-to-physical index map.  Plain sorts use the identity; batched
// sorts place instance v >> c at offsets[v >> c].
struct IMap {
    const uint64_t *offsets; // nullptr -> identity
    int c;                   // log2 instance size when offsets != nullptr
};

__device__ __forceinline__ uint64_t phys(const IMap &m, uint64_t v)
{
    if (m.offsets == nullptr) return v;
    return m.offsets[v >> m.c] + (v & ((1ull << m.c) - 1));
}

template <class R, bool TRACE>
__device__ __forceinline__ void gload(R &r, const Bufs<R> &b, uint64_t i, uint64_t &h, uint32_t &n)
{
#pragma unroll
    for (int w = 0; w < R::KW; ++w) r.k[w] = b.k[w][i];
    r.p = b.p[i];
    if constexpr (TRACE) {
        h = tfold(h, 1, i);
        ++n;
    }
}

template <class R, bool TRACE>
__device__ __forceinline__ void gstore(const R &r, const Bufs<R> &b, uint64_t i, uint64_t &h, uint32_t &n)
{
#pragma unroll
    for (int w = 0; w < R::KW; ++w) b.k[w][i] = r.k[w];
    b.p[i] = r.p;
    if constexpr (TRACE) {
        h = tfold(h, 2, i);
        ++n;
    }
}

// Pad record for sorts that fill [n, n_pad): all key bits set, payload = index.
template <class R>
__device__ __forceinline__ void pad_select(R &r, uint64_t g, uint64_t n_valid)
{
    const bool pad = g >= n_valid;
#pragma unroll
    for (int w = 0; w < R::KW; ++w) r.k[w] = pad ? ~0ull : r.k[w];
    r.p = pad ? (typename R::P)g : r.p;
}

// ------------------------------------------------------------ tile layout

// LDS position of tile element l: one pad slot per 16 elements keeps the
// strided layouts free of bank conflicts for 4- and 8-byte columns.
__device__ __forceinline__ uint32_t lpos(uint32_t l) { return l + (l >> 4); }

// Element held in register m by thread tau under layout a.
template <int R_LOG>
__device__ __forceinline__ uint32_t lay_base(uint32_t tau, int a)
{
    return (tau & ((1u << a) - 1)) | ((tau >> a) << (a + R_LOG));
}

template <class R>
struct Lds {
    uint64_t *k; // KW arrays of tpad words
    typename R::P *p;
    uint32_t tpad;
};

template <class R>
__device__ __forceinline__ Lds<R> lds_view(char *base, uint32_t tpad)
{
    Lds<R> s;
    s.k = reinterpret_cast<uint64_t *>(base);
    s.p = reinterpret_cast<typename R::P *>(base + sizeof(uint64_t) * R::KW * tpad);
    s.tpad = tpad;
    return s;
}

template <class R>
constexpr size_t lds_bytes(int t)
{
    return (size_t)((1u << t) + ((1u << t) >> 4)) * (sizeof(uint64_t) * R::KW + sizeof(typename R::P));
}

template <class R, int R_LOG, bool TRACE>
__device__ __forceinline__ void lds_put(const R (&r)[1 << R_LOG], const Lds<R> &s, uint32_t tau, int a,
                                        uint64_t &h, uint32_t &n)
{
    const uint32_t base = lay_base<R_LOG>(tau, a);
#pragma unroll
    for (int m = 0; m < (1 << R_LOG); ++m) {
        const uint32_t l = lpos(base | ((uint32_t)m << a));
#pragma unroll
        for (int w = 0; w < R::KW; ++w) s.k[w * s.tpad + l] = r[m].k[w];
        s.p[l] = r[m].p;
        if constexpr (TRACE) {
            h = tfold(h, 3, l);
            ++n;
        }
    }
}

template <class R, int R_LOG, bool TRACE>
__device__ __forceinline__ void lds_get(R (&r)[1 << R_LOG], const Lds<R> &s, uint32_t tau, int a, uint64_t &h,
                                        uint32_t &n)
{
    const uint32_t base = lay_base<R_LOG>(tau, a);
#pragma unroll
    for (int m = 0; m < (1 << R_LOG); ++m) {
        const uint32_t l = lpos(base | ((uint32_t)m << a));
#pragma unroll
        for (int w = 0; w < R::KW; ++w) r[m].k[w] = s.k[w * s.tpad + l];
        r[m].p = s.p[l];
        if constexpr (TRACE) {
            h = tfold(h, 4, l);
            ++n;
        }
    }
}

// Stages (p, q) for q = q_hi down to q_lo, all inside the register bits of
// layout a (a <= q_lo <= q_hi < a + R_LOG).  `vbase` is the virtual index of
// tile element 0.
template <class R, int R_LOG>
__device__ __forceinline__ void reg_stages(R (&r)[1 << R_LOG], uint32_t tau, int a, int q_hi, int q_lo, int p,
                                           int p_top, uint64_t vbase)
{
    const uint32_t base = lay_base<R_LOG>(tau, a);
#pragma unroll
    for (int qq = R_LOG - 1; qq >= 0; --qq) {
        const int q = a + qq;
        if (q > q_hi || q < q_lo) continue;
#pragma unroll
        for (int m = 0; m < (1 << R_LOG); ++m) {
            if (m & (1 << qq)) continue;
            const uint64_t g = vbase + (base | ((uint32_t)m << a));
            cmpx(r[m], r[m | (1 << qq)], dir_asc(g, p, p_top));
        }
    }
}

// Moves the tile from layout a_old to layout a_new through LDS.
template <class R, int R_LOG, bool TRACE>
__device__ __forceinline__ void relayout(R (&r)[1 << R_LOG], const Lds<R> &s, uint32_t tau, bool act, int a_old,
                                         int a_new, uint64_t &h, uint32_t &n)
{
    __syncthreads();
    if (act) lds_put<R, R_LOG, TRACE>(r, s, tau, a_old, h, n);
    __syncthreads();
    if (act) lds_get<R, R_LOG, TRACE>(r, s, tau, a_new, h, n);
}

// Runs stages (p, q_start .. 0) of one level on the tile in registers,
// changing layout whenever the next stage bit is not a register bit.
template <class R, int R_LOG, bool TRACE>
__device__ __forceinline__ void tile_level(R (&r)[1 << R_LOG], const Lds<R> &s, uint32_t tau, bool act, int &a,
                                           int p, int q_start, int p_top, uint64_t vbase, uint64_t &h, uint32_t &n)
{
    int q = q_start;
    while (q >= 0) {
        if (!(a <= q && q < a + R_LOG)) {
            const int a_new = q - R_LOG + 1 > 0 ? q - R_LOG + 1 : 0;
            relayout<R, R_LOG, TRACE>(r, s, tau, act, a, a_new, h, n);
            a = a_new;
        }
        if (act) reg_stages<R, R_LOG>(r, tau, a, q, a, p, p_top, vbase);
        q = a - 1;
    }
}

} // namespace dev
} // namespace vo
