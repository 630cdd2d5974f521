// gather_obl_kernels.hpp -- device bodies and __global__ wrappers of the
// oblivious gather (the materialization offload; see gather_obl.hip for the
// method and the reference), for launch mode and the resident queue kernel.
#pragma once

#include "vo_internal.hpp"

namespace vo {
namespace og_k {
namespace {

constexpr int GT = 4096; // scan tile
constexpr int GB = 256;
constexpr int GPER = GT / GB;
constexpr uint64_t NONE = ~0ull;


__device__ __forceinline__ uint64_t load_elem(const void *p, uint32_t eb, uint64_t i)
{
    switch (eb) {
    case 1: return ((const uint8_t *)p)[i];
    case 2: return ((const uint16_t *)p)[i];
    case 4: return ((const uint32_t *)p)[i];
    default: return ((const uint64_t *)p)[i];
    }
}

__device__ __forceinline__ void store_elem(void *p, uint32_t eb, uint64_t i, uint64_t v)
{
    switch (eb) {
    case 1: ((uint8_t *)p)[i] = (uint8_t)v; break;
    case 2: ((uint16_t *)p)[i] = (uint16_t)v; break;
    case 4: ((uint32_t *)p)[i] = (uint32_t)v; break;
    default: ((uint64_t *)p)[i] = v; break;
    }
}

template <bool TRACE, bool KP>
__device__ __forceinline__ void b_og_build(const void *__restrict__ table, uint64_t n_table, uint32_t eb,
                           const int64_t *__restrict__ ids, uint64_t n, uint64_t n_pad, uint64_t *__restrict__ keys,
                           void *__restrict__ vals, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        uint64_t k = NONE, v = i;
        if (i < n_table) { // public sizes only
            k = i << 33;
            v = load_elem(table, eb, i);
            tr.at(1, i);
        } else if (i < n_table + n) {
            const uint64_t r = i - n_table;
            const int64_t id = ids[r];
            const bool ok = (id >= 0) & ((uint64_t)id < n_table);
            // an id out of range sorts after every table row and finds nothing
            k = ((ok ? (uint64_t)id : 0x7FFFFFFFull) << 33) | (1ull << 32) | r;
            v = 0;
            tr.at(2, r);
        }
        keys[i] = k;
        if constexpr (KP) static_cast<uint32_t *>(vals)[i] = (uint32_t)v; // values of at most 4 bytes
        else static_cast<uint64_t *>(vals)[i] = v;
        tr.at(3, i);
    }
    tr.done(dev::gtid());
}

template <bool TRACE, bool KP>
__global__ void k_og_build(const void *__restrict__ table, uint64_t n_table, uint32_t eb,
                           const int64_t *__restrict__ ids, uint64_t n, uint64_t n_pad, uint64_t *__restrict__ keys,
                           void *__restrict__ vals, TraceSlot *slot)
{
    b_og_build<TRACE, KP>(table, n_table, eb, ids, n, n_pad, keys, vals, slot, blockIdx.x, gridDim.x);
}

// Union row value: u64 column, or the u32 payload column of the 12-byte
// records used for values of at most 4 bytes (the keys are unique, pads
// included through their row, so the order is that of the 16-byte form).
template <bool KP>
__device__ __forceinline__ uint64_t union_val(const void *__restrict__ vals, uint64_t i)
{
    if constexpr (KP) return static_cast<const uint32_t *>(vals)[i];
    else return static_cast<const uint64_t *>(vals)[i];
}

// last table row of each tile: (row, value) or NONE
struct Last {
    uint64_t row, val;
};
struct LastOp {
    __device__ Last operator()(const Last &a, const Last &b) const { return b.row != NONE ? b : a; }
};

__device__ __forceinline__ bool is_table(uint64_t k) { return k != NONE && ((k >> 32) & 1) == 0; }

template <bool TRACE, bool KP>
__device__ __forceinline__ void b_og_tiles(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                 Last *__restrict__ tiles, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    __shared__ Last red[GB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * GT + (uint64_t)threadIdx.x * GPER;
    Last l{NONE, 0};
#pragma unroll
    for (int r = 0; r < GPER; ++r) {
        const uint64_t k = keys[t0 + r], v = union_val<KP>(vals, t0 + r);
        const bool t = is_table(k);
        l.row = t ? (k >> 33) : l.row;
        l.val = t ? v : l.val;
    }
    tr.at(1, t0);
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) { // inclusive scan: lane 63 holds the wave's last
        const Last o{(uint64_t)__shfl_up((unsigned long long)l.row, off, 64),
                     (uint64_t)__shfl_up((unsigned long long)l.val, off, 64)};
        l = lane >= (uint32_t)off ? LastOp{}(o, l) : l;
    }
    if (lane == 63) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        Last t = red[0];
        for (int w = 1; w < GB / 64; ++w) t = LastOp{}(t, red[w]);
        tiles[vb] = t;
        tr.at(2, vb);
    }
    tr.done(dev::gtid());
}

template <bool TRACE, bool KP>
__global__ __launch_bounds__(GB) void k_og_tiles(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                 Last *__restrict__ tiles, TraceSlot *slot)
{
    b_og_tiles<TRACE, KP>(keys, vals, tiles, slot, blockIdx.x, gridDim.x);
}

constexpr uint32_t kScanThreads = 1024;

// exclusive "last table row" scan over the tiles, one block of NT threads
// (1024 in launch mode, 256 in the resident kernel; the operator is
// associative, so NT changes no result); sh: NT values of LDS
template <uint32_t NT, bool TRACE>
__device__ __forceinline__ void b_og_carry(const Last *__restrict__ in, uint32_t nt, Last *__restrict__ out,
                                           TraceSlot *slot, Last *sh)
{
    dev::AtrT<TRACE> tr(slot);
    const uint32_t per = (nt + NT - 1) / NT;
    const uint32_t lo = threadIdx.x * per;
    // a thread's chunk in groups of 8, loaded before any is combined
    constexpr int G = 8;
    Last s{NONE, 0};
    for (uint32_t k0 = 0; k0 < per; k0 += G) {
        Last v[G];
#pragma unroll
        for (int j = 0; j < G; ++j) v[j] = (k0 + j < per && lo + k0 + j < nt) ? in[lo + k0 + j] : Last{NONE, 0};
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (k0 + j < per && lo + k0 + j < nt) { // public size
                s = LastOp{}(s, v[j]);
                tr.at(1, lo + k0 + j);
            }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < NT; off <<= 1) {
        const Last v = threadIdx.x >= off ? sh[threadIdx.x - off] : Last{NONE, 0};
        __syncthreads();
        if (threadIdx.x >= off) sh[threadIdx.x] = LastOp{}(v, sh[threadIdx.x]);
        __syncthreads();
    }
    Last run = threadIdx.x ? sh[threadIdx.x - 1] : Last{NONE, 0};
    for (uint32_t k0 = 0; k0 < per; k0 += G) {
        Last v[G];
#pragma unroll
        for (int j = 0; j < G; ++j) v[j] = (k0 + j < per && lo + k0 + j < nt) ? in[lo + k0 + j] : Last{NONE, 0};
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (k0 + j < per && lo + k0 + j < nt) {
                out[lo + k0 + j] = run;
                run = LastOp{}(run, v[j]);
                tr.at(2, lo + k0 + j);
            }
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ __launch_bounds__(kScanThreads) void k_og_carry(const Last *__restrict__ in, uint32_t nt,
                                                           Last *__restrict__ out, TraceSlot *slot)
{
    __shared__ Last sh[kScanThreads];
    b_og_carry<kScanThreads, TRACE>(in, nt, out, slot, sh);
}

// per row: request rows get (index << 1 | found, value) and mask 1
template <bool TRACE, bool KP>
__device__ __forceinline__ void b_og_rows(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                const Last *__restrict__ carry, uint64_t *__restrict__ c0,
                                                uint64_t *__restrict__ c1, uint8_t *__restrict__ mask,
                                                TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    __shared__ Last wt[GB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)vb * GT + (uint64_t)threadIdx.x * GPER;
    uint64_t k[GPER], v[GPER];
    Last l{NONE, 0};
#pragma unroll
    for (int r = 0; r < GPER; ++r) {
        k[r] = keys[t0 + r];
        v[r] = union_val<KP>(vals, t0 + r);
        const bool t = is_table(k[r]);
        l.row = t ? (k[r] >> 33) : l.row;
        l.val = t ? v[r] : l.val;
    }
    tr.at(1, t0);
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    Last in = l;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const Last o{(uint64_t)__shfl_up((unsigned long long)in.row, off, 64),
                     (uint64_t)__shfl_up((unsigned long long)in.val, off, 64)};
        in = lane >= (uint32_t)off ? LastOp{}(o, in) : in;
    }
    Last ex{(uint64_t)__shfl_up((unsigned long long)in.row, 1, 64),
            (uint64_t)__shfl_up((unsigned long long)in.val, 1, 64)};
    ex = lane == 0 ? Last{NONE, 0} : ex;
    if (lane == 63) wt[wv] = in;
    __syncthreads();
    Last c = carry[vb];
    tr.at(2, vb);
    for (uint32_t w = 0; w < wv; ++w) c = LastOp{}(c, wt[w]);
    c = LastOp{}(c, ex);
#pragma unroll
    for (int r = 0; r < GPER; ++r) {
        const uint64_t i = t0 + r;
        const bool t = is_table(k[r]);
        c.row = t ? (k[r] >> 33) : c.row;
        c.val = t ? v[r] : c.val;
        const bool req = k[r] != NONE && ((k[r] >> 32) & 1);
        const bool found = req & (c.row == (k[r] >> 33));
        c0[i] = ((k[r] & 0xFFFFFFFFull) << 1) | (found ? 1 : 0);
        c1[i] = found ? c.val : 0;
        mask[i] = req ? 1 : 0;
    }
    tr.at(3, t0);
    tr.done(dev::gtid());
}

template <bool TRACE, bool KP>
__global__ __launch_bounds__(GB) void k_og_rows(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                const Last *__restrict__ carry, uint64_t *__restrict__ c0,
                                                uint64_t *__restrict__ c1, uint8_t *__restrict__ mask,
                                                TraceSlot *slot)
{
    b_og_rows<TRACE, KP>(keys, vals, carry, c0, c1, mask, slot, blockIdx.x, gridDim.x);
}

// keys of the final sort: request rows (the first n) by index, pads last
// Requests back in request order by the (u64 key, u32 payload) sort: c0 =
// index << 1 | found is unique and below 2^30 (table + requests <= 2^29), so
// key = c0 << 32 | value low word orders by index alone, and the value's high
// word rides as the payload: 12 bytes per record instead of the 16 of a
// (u64, u64) sort.  Pad rows get distinct keys above every request key.
template <bool TRACE>
__device__ __forceinline__ void b_og_order(const uint64_t *__restrict__ c0, const uint64_t *__restrict__ c1, uint64_t n,
                           uint64_t m_pad, uint64_t *__restrict__ keys, uint32_t *__restrict__ hi,
                           TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < m_pad; i += (uint64_t)nvb * 256) {
        uint64_t k = NONE - m_pad + i;
        uint32_t h = 0;
        if (i < n) { // public
            const uint64_t v = c1[i];
            k = (c0[i] << 32) | (v & 0xFFFFFFFFull);
            h = (uint32_t)(v >> 32);
            tr.at(1, i);
        }
        keys[i] = k;
        hi[i] = h;
        tr.at(2, i);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ void k_og_order(const uint64_t *__restrict__ c0, const uint64_t *__restrict__ c1, uint64_t n,
                           uint64_t m_pad, uint64_t *__restrict__ keys, uint32_t *__restrict__ hi,
                           TraceSlot *slot)
{
    b_og_order<TRACE>(c0, c1, n, m_pad, keys, hi, slot, blockIdx.x, gridDim.x);
}

template <bool TRACE>
__device__ __forceinline__ void b_og_out(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ hi, uint64_t n,
                         uint32_t eb, void *__restrict__ out, uint8_t *__restrict__ found, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n; i += (uint64_t)nvb * 256) {
        const uint64_t k = keys[i];
        store_elem(out, eb, i, ((uint64_t)hi[i] << 32) | (k & 0xFFFFFFFFull));
        if (found) found[i] = (uint8_t)((k >> 32) & 1);
        tr.at(1, i);
    }
    tr.done(dev::gtid());
}

template <bool TRACE>
__global__ void k_og_out(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ hi, uint64_t n,
                         uint32_t eb, void *__restrict__ out, uint8_t *__restrict__ found, TraceSlot *slot)
{
    b_og_out<TRACE>(keys, hi, n, eb, out, found, slot, blockIdx.x, gridDim.x);
}

} // namespace
} // namespace og_k
} // namespace vo
