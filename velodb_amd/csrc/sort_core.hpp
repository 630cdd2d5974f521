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
// The comparison runs on wavefront lane masks: each field costs one v_cmp
// for > and one for <, the lexicographic combination and the direction are
// scalar mask operations, and the swap is v_cndmask on the resulting mask
// (2 per 32-bit word).  `ascm` holds the lanes whose pair sorts ascending.
// Lanes where (a2:a1:a0) < (b2:b1:b0) as unsigned 96-bit numbers: the borrow
// out of a subtract-with-borrow chain (three VALU ops, no scalar logic).
__device__ __forceinline__ uint64_t lt_mask96(uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1, uint32_t a2,
                                              uint32_t b2)
{
    uint64_t m;
    uint32_t t0, t1, t2;
    asm("v_sub_co_u32_e64 %0, %3, %4, %5\n\t"
        "v_subb_co_u32_e64 %1, %3, %6, %7, %3\n\t"
        "v_subb_co_u32_e64 %2, %3, %8, %9, %3"
        : "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&s"(m)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2));
    return m;
}

__device__ __forceinline__ uint64_t lt_mask128(uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1, uint32_t a2,
                                               uint32_t b2, uint32_t a3, uint32_t b3)
{
    uint64_t m;
    uint32_t t0, t1, t2, t3;
    asm("v_sub_co_u32_e64 %0, %4, %5, %6\n\t"
        "v_subb_co_u32_e64 %1, %4, %7, %8, %4\n\t"
        "v_subb_co_u32_e64 %2, %4, %9, %10, %4\n\t"
        "v_subb_co_u32_e64 %3, %4, %11, %12, %4"
        : "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&v"(t3), "=&s"(m)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2), "v"(a3), "v"(b3));
    return m;
}

template <class R>
__device__ __forceinline__ void cmpx_m(R &a, R &b, uint64_t ascm)
{
    if constexpr (R::TIE && R::KW == 1) {
        // (key, payload) records compare as one integer, key high.  Records
        // equal in every field are identical, so exchanging them on a tie
        // changes nothing and the swap is simply lt(a, b) xor ascending.
        uint64_t lt;
        if constexpr (sizeof(typename R::P) == 4)
            lt = lt_mask96((uint32_t)a.p, (uint32_t)b.p, (uint32_t)a.k[0], (uint32_t)b.k[0],
                           (uint32_t)(a.k[0] >> 32), (uint32_t)(b.k[0] >> 32));
        else
            lt = lt_mask128((uint32_t)a.p, (uint32_t)b.p, (uint32_t)((uint64_t)a.p >> 32),
                            (uint32_t)((uint64_t)b.p >> 32), (uint32_t)a.k[0], (uint32_t)b.k[0],
                            (uint32_t)(a.k[0] >> 32), (uint32_t)(b.k[0] >> 32));
        const bool sw = __builtin_amdgcn_inverse_ballot_w64(lt ^ ascm);
        const uint64_t x = a.k[0], y = b.k[0];
        a.k[0] = sw ? y : x;
        b.k[0] = sw ? x : y;
        const typename R::P u = a.p, v = b.p;
        a.p = sw ? v : u;
        b.p = sw ? u : v;
        return;
    }
    uint64_t gt = 0, lt = 0, eq = ~0ull;
#pragma unroll
    for (int w = 0; w < R::KW; ++w) {
        const uint64_t g = __builtin_amdgcn_ballot_w64(a.k[w] > b.k[w]);
        const uint64_t l = __builtin_amdgcn_ballot_w64(a.k[w] < b.k[w]);
        gt |= eq & g;
        lt |= eq & l;
        eq &= ~(g | l);
    }
    if constexpr (R::TIE) {
        const uint64_t g = __builtin_amdgcn_ballot_w64(a.p > b.p);
        const uint64_t l = __builtin_amdgcn_ballot_w64(a.p < b.p);
        gt |= eq & g;
        lt |= eq & l;
    }
    const bool sw = __builtin_amdgcn_inverse_ballot_w64((ascm & gt) | (~ascm & lt));
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

template <class R>
__device__ __forceinline__ void cmpx(R &a, R &b, bool asc)
{
    cmpx_m(a, b, __builtin_amdgcn_ballot_w64(asc));
}

// Scalar reference form of the same exchange (used by the one-thread kernel,
// where lane masks would not describe the per-iteration pairs).
template <class R>
__device__ __forceinline__ void cmpx_scalar(R &a, R &b, bool asc)
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
    const uint64_t *ns;      // live rows per instance (pad fill), or nullptr
};

__device__ __forceinline__ uint64_t phys(const IMap &m, uint64_t v)
{
    if (m.offsets == nullptr) return v;
    return m.offsets[v >> m.c] + (v & ((1ull << m.c) - 1));
}

// Pad fill for virtual row v: its row index inside its instance and the
// instance's live row count.
__device__ __forceinline__ uint64_t local_row(const IMap &m, uint64_t v)
{
    return m.offsets == nullptr ? v : (v & ((1ull << m.c) - 1));
}

__device__ __forceinline__ uint64_t live_rows(const IMap &m, uint64_t v, uint64_t n_valid)
{
    return m.ns == nullptr ? n_valid : m.ns[v >> m.c];
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

// ------------------------------------------------- compile-time tile schedule

// Same schedule as tile_level, with the tile size, every layout and every
// level fixed at compile time, so that all LDS offsets are immediates and the
// stage loops fully unroll.  Used for full 2^T_LOG tiles.
template <class R, int T_LOG, int R_LOG, bool TRACE>
struct TileCT {
    static constexpr int RR = 1 << R_LOG;
    static constexpr uint32_t T = 1u << T_LOG;
    static constexpr uint32_t TPAD = T + (T >> 4);
    using P = typename R::P;

    // layout reached after stages (q_start .. 0) starting from layout a;
    // stage 0 in layout 1 runs across lane pairs and keeps the layout
    static constexpr int end_layout(int q, int a, int qlo = 0)
    {
        while (q >= qlo) {
            if (q == 0 && a == 1) break;
            if (!(a <= q && q < a + R_LOG)) a = q - R_LOG + 1 > 0 ? q - R_LOG + 1 : 0;
            q = (a > qlo ? a : qlo) - 1;
        }
        return a;
    }

    // The record in register m of the lane whose index differs in bit 0: a
    // DPP quad permutation (lanes 1 0 3 2) per 32-bit word, no LDS.
    static __device__ __forceinline__ uint32_t xor1(uint32_t v)
    {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    }
    static __device__ __forceinline__ R lane_partner(const R &a)
    {
        R o;
#pragma unroll
        for (int w = 0; w < R::KW; ++w)
            o.k[w] = (uint64_t)xor1((uint32_t)a.k[w]) | ((uint64_t)xor1((uint32_t)(a.k[w] >> 32)) << 32);
        if constexpr (sizeof(P) == 8)
            o.p = (P)((uint64_t)xor1((uint32_t)a.p) | ((uint64_t)xor1((uint32_t)((uint64_t)a.p >> 32)) << 32));
        else
            o.p = (P)xor1((uint32_t)a.p);
        return o;
    }

    // Stage 0 of level P in layout 1: record index bit 0 is thread index
    // bit 0, so the pair of register m is register m of lanes (2t, 2t+1).
    // Each lane keeps the smaller or the larger record of its pair, as the
    // compare-exchange would leave it (ties stay in place).
    template <int P>
    static __device__ __forceinline__ void lane_stage0(R (&r)[RR], uint32_t tau, bool asc_hi, int p_top)
    {
        const uint32_t base = lay_base<R_LOG>(tau, 1);
        const bool lower = (tau & 1u) == 0;
#pragma unroll
        for (int m = 0; m < RR; ++m) {
            bool asc;
            if constexpr (P < 0) {
                asc = asc_hi;
            } else {
                const uint32_t l = base | ((uint32_t)m << 1);
                asc = (P >= p_top) | (P >= T_LOG ? asc_hi : (((l >> P) & 1u) == 0));
            }
            const R o = lane_partner(r[m]);
            // swap when this lane should keep the smaller and holds the
            // larger, or the other way round
            const bool keep_small = asc == lower;
            const bool sw = keep_small ? rec_gt(r[m], o) : rec_gt(o, r[m]);
#pragma unroll
            for (int w = 0; w < R::KW; ++w) r[m].k[w] = sw ? o.k[w] : r[m].k[w];
            r[m].p = sw ? o.p : r[m].p;
        }
    }

    template <int A>
    static __device__ __forceinline__ void put(const R (&r)[RR], char *lds, uint32_t tau, uint64_t &h, uint32_t &n)
    {
        uint64_t *sk = reinterpret_cast<uint64_t *>(lds);
        P *sp = reinterpret_cast<P *>(lds + sizeof(uint64_t) * R::KW * TPAD);
        const uint32_t lb = lpos(lay_base<R_LOG>(tau, A));
#pragma unroll
        for (int m = 0; m < RR; ++m) {
            const uint32_t l = lb + lpos((uint32_t)m << A);
#pragma unroll
            for (int w = 0; w < R::KW; ++w) sk[w * TPAD + l] = r[m].k[w];
            sp[l] = r[m].p;
            if constexpr (TRACE) {
                h = tfold(h, 3, l);
                ++n;
            }
        }
    }

    template <int A>
    static __device__ __forceinline__ void get(R (&r)[RR], char *lds, uint32_t tau, uint64_t &h, uint32_t &n)
    {
        const uint64_t *sk = reinterpret_cast<const uint64_t *>(lds);
        const P *sp = reinterpret_cast<const P *>(lds + sizeof(uint64_t) * R::KW * TPAD);
        const uint32_t lb = lpos(lay_base<R_LOG>(tau, A));
#pragma unroll
        for (int m = 0; m < RR; ++m) {
            const uint32_t l = lb + lpos((uint32_t)m << A);
#pragma unroll
            for (int w = 0; w < R::KW; ++w) r[m].k[w] = sk[w * TPAD + l];
            r[m].p = sp[l];
            if constexpr (TRACE) {
                h = tfold(h, 4, l);
                ++n;
            }
        }
    }

    // LDS bytes of one block: a single column buffer (the widest column)
    // that every column passes through in turn.  Keeping it to 8 B per slot
    // lets two 2^13-record tiles share a CU's 160 KB.
    static constexpr size_t lds_bytes() { return (size_t)TPAD * sizeof(uint64_t); }

    template <int A, typename C>
    static __device__ __forceinline__ void put_col(const C (&v)[RR], C *buf, uint32_t tau, uint64_t &h, uint32_t &n)
    {
        const uint32_t lb = lpos(lay_base<R_LOG>(tau, A));
#pragma unroll
        for (int m = 0; m < RR; ++m) {
            const uint32_t l = lb + lpos((uint32_t)m << A);
            buf[l] = v[m];
            if constexpr (TRACE) {
                h = tfold(h, 3, l);
                ++n;
            }
        }
    }

    template <int A, typename C>
    static __device__ __forceinline__ void get_col(C (&v)[RR], const C *buf, uint32_t tau, uint64_t &h, uint32_t &n)
    {
        const uint32_t lb = lpos(lay_base<R_LOG>(tau, A));
#pragma unroll
        for (int m = 0; m < RR; ++m) {
            const uint32_t l = lb + lpos((uint32_t)m << A);
            v[m] = buf[l];
            if constexpr (TRACE) {
                h = tfold(h, 4, l);
                ++n;
            }
        }
    }

    template <int A_OLD, int A_NEW>
    static __device__ __forceinline__ void relayout(R (&r)[RR], char *lds, uint32_t tau, uint64_t &h, uint32_t &n)
    {
        if constexpr (A_OLD != A_NEW) {
#pragma unroll
            for (int w = 0; w < R::KW; ++w) {
                uint64_t v[RR];
#pragma unroll
                for (int m = 0; m < RR; ++m) v[m] = r[m].k[w];
                __syncthreads();
                put_col<A_OLD>(v, reinterpret_cast<uint64_t *>(lds), tau, h, n);
                __syncthreads();
                get_col<A_NEW>(v, reinterpret_cast<const uint64_t *>(lds), tau, h, n);
#pragma unroll
                for (int m = 0; m < RR; ++m) r[m].k[w] = v[m];
            }
            P v[RR];
#pragma unroll
            for (int m = 0; m < RR; ++m) v[m] = r[m].p;
            __syncthreads();
            put_col<A_OLD>(v, reinterpret_cast<P *>(lds), tau, h, n);
            __syncthreads();
            get_col<A_NEW>(v, reinterpret_cast<const P *>(lds), tau, h, n);
#pragma unroll
            for (int m = 0; m < RR; ++m) r[m].p = v[m];
        }
    }

    // Stages q = QHI .. QLO of level P (P >= 0 compile-time; P < 0 means a
    // level above the tile whose direction `asc_hi` is uniform per block).
    template <int P, int A, int QHI, int QLO>
    static __device__ __forceinline__ void stages(R (&r)[RR], uint32_t tau, bool asc_hi, int p_top)
    {
        const uint32_t base = lay_base<R_LOG>(tau, A);
#pragma unroll
        for (int q = QHI; q >= QLO; --q) {
            const int qq = q - A;
#pragma unroll
            for (int m = 0; m < RR; ++m) {
                if (m & (1 << qq)) continue;
                bool asc;
                if constexpr (P < 0) {
                    asc = asc_hi;
                } else {
                    const uint32_t l = base | ((uint32_t)m << A);
                    asc = (P >= p_top) | (P >= T_LOG ? asc_hi : (((l >> P) & 1u) == 0));
                }
                cmpx(r[m], r[m | (1 << qq)], asc);
            }
        }
    }

    // Stages (P, Q .. QLO) from layout A, changing layout as needed.
    template <int P, int Q, int A, int QLO = 0>
    static __device__ __forceinline__ void level(R (&r)[RR], char *lds, uint32_t tau, bool asc_hi, int p_top,
                                                 uint64_t &h, uint32_t &n)
    {
        if constexpr (Q == 0 && A == 1 && QLO == 0) {
            lane_stage0<P>(r, tau, asc_hi, p_top);
        } else if constexpr (Q >= QLO) {
            constexpr bool covered = (A <= Q && Q < A + R_LOG);
            constexpr int AN = covered ? A : (Q - R_LOG + 1 > 0 ? Q - R_LOG + 1 : 0);
            constexpr int LO = AN > QLO ? AN : QLO;
            relayout<A, AN>(r, lds, tau, h, n);
            stages<P, AN, Q, LO>(r, tau, asc_hi, p_top);
            level<P, LO - 1, AN, QLO>(r, lds, tau, asc_hi, p_top, h, n);
        }
    }

    // Levels P .. T_LOG of a full local sort from layout A; returns through
    // the template chain and leaves the tile in layout end_all(P, A).
    static constexpr int end_all(int p, int a)
    {
        for (; p <= T_LOG; ++p) a = end_layout(p - 1, a);
        return a;
    }

    template <int P, int A>
    static __device__ __forceinline__ void levels(R (&r)[RR], char *lds, uint32_t tau, bool asc_hi, int p_top,
                                                  uint64_t &h, uint32_t &n)
    {
        if constexpr (P <= T_LOG) {
            level<P, P - 1, A>(r, lds, tau, asc_hi, p_top, h, n);
            levels<P + 1, end_layout(P - 1, A)>(r, lds, tau, asc_hi, p_top, h, n);
        }
    }
};

} // namespace dev
} // namespace vo
