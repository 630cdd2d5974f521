// gather_obl.hip -- oblivious gather of table values by record id, the
// device form of the host oblivious materialization of
// src/operator/materialization_operator.cpp:362-517 (obliviousRecoverSingle:
// bitonic sort of (record id, position, pairing index) tuples, reads in
// sorted order with a dummy read per duplicate, restore by the pairing
// index).
//
// Here no address depends on a record id:
//   1. union of the table rows (key = row << 33) and the requests
//      (key = id << 33 | 1 << 32 | request index), padded to N = 2^k;
//   2. sort by key: each request follows the table row it asks for (values
//      of at most 4 bytes ride as the u32 payload of 12-byte records, wider
//      ones as a u64 column);
//   3. forward fill of the last table row (row, value) to the requests;
//   4. ORCompact of the request rows: the first n rows hold
//      (request index << 1 | found, value) in record-id order;
//   5. sort of those n rows (padded to 2^k) by request index, as
//      (u64 key, u32 payload) records: the value's low word rides below the
//      index in the key and its high word is the payload.
// Every pass reads and writes every row of its padded size; the record ids
// and their repetitions never reach an address.
#include "vo_internal.hpp"

#include <algorithm>

namespace vo {
namespace {

constexpr int GT = 4096; // scan tile
constexpr int GB = 256;
constexpr int GPER = GT / GB;
constexpr uint64_t NONE = ~0ull;

dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65536))); }

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
__global__ void k_og_build(const void *__restrict__ table, uint64_t n_table, uint32_t eb,
                           const int64_t *__restrict__ ids, uint64_t n, uint64_t n_pad, uint64_t *__restrict__ keys,
                           void *__restrict__ vals, TraceSlot *slot)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
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
__global__ __launch_bounds__(GB) void k_og_tiles(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                 Last *__restrict__ tiles, TraceSlot *slot)
{
    __shared__ Last red[GB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)blockIdx.x * GT + (uint64_t)threadIdx.x * GPER;
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
        tiles[blockIdx.x] = t;
        tr.at(2, blockIdx.x);
    }
    tr.done(dev::gtid());
}

constexpr uint32_t kScanThreads = 1024;

// exclusive "last table row" scan over the tiles, one 1024-thread block
template <bool TRACE>
__global__ __launch_bounds__(kScanThreads) void k_og_carry(const Last *__restrict__ in, uint32_t nt,
                                                           Last *__restrict__ out, TraceSlot *slot)
{
    __shared__ Last sh[kScanThreads];
    dev::AtrT<TRACE> tr(slot);
    const uint32_t per = (nt + kScanThreads - 1) / kScanThreads;
    const uint32_t lo = threadIdx.x * per;
    Last s{NONE, 0};
    for (uint32_t k = 0; k < per; ++k)
        if (lo + k < nt) { // public size
            s = LastOp{}(s, in[lo + k]);
            tr.at(1, lo + k);
        }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        const Last v = threadIdx.x >= off ? sh[threadIdx.x - off] : Last{NONE, 0};
        __syncthreads();
        if (threadIdx.x >= off) sh[threadIdx.x] = LastOp{}(v, sh[threadIdx.x]);
        __syncthreads();
    }
    Last run = threadIdx.x ? sh[threadIdx.x - 1] : Last{NONE, 0};
    for (uint32_t k = 0; k < per; ++k)
        if (lo + k < nt) {
            out[lo + k] = run;
            run = LastOp{}(run, in[lo + k]);
            tr.at(2, lo + k);
        }
    tr.done(dev::gtid());
}

// per row: request rows get (index << 1 | found, value) and mask 1
template <bool TRACE, bool KP>
__global__ __launch_bounds__(GB) void k_og_rows(const uint64_t *__restrict__ keys, const void *__restrict__ vals,
                                                const Last *__restrict__ carry, uint64_t *__restrict__ c0,
                                                uint64_t *__restrict__ c1, uint8_t *__restrict__ mask,
                                                TraceSlot *slot)
{
    __shared__ Last wt[GB / 64];
    dev::AtrT<TRACE> tr(slot);
    const uint64_t t0 = (uint64_t)blockIdx.x * GT + (uint64_t)threadIdx.x * GPER;
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
    Last c = carry[blockIdx.x];
    tr.at(2, blockIdx.x);
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

// keys of the final sort: request rows (the first n) by index, pads last
// Requests back in request order by the (u64 key, u32 payload) sort: c0 =
// index << 1 | found is unique and below 2^30 (table + requests <= 2^29), so
// key = c0 << 32 | value low word orders by index alone, and the value's high
// word rides as the payload: 12 bytes per record instead of the 16 of a
// (u64, u64) sort.  Pad rows get distinct keys above every request key.
template <bool TRACE>
__global__ void k_og_order(const uint64_t *__restrict__ c0, const uint64_t *__restrict__ c1, uint64_t n,
                           uint64_t m_pad, uint64_t *__restrict__ keys, uint32_t *__restrict__ hi,
                           TraceSlot *slot)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m_pad; i += (uint64_t)gridDim.x * blockDim.x) {
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
__global__ void k_og_out(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ hi, uint64_t n,
                         uint32_t eb, void *__restrict__ out, uint8_t *__restrict__ found, TraceSlot *slot)
{
    dev::AtrT<TRACE> tr(slot);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        store_elem(out, eb, i, ((uint64_t)hi[i] << 32) | (k & 0xFFFFFFFFull));
        if (found) found[i] = (uint8_t)((k >> 32) & 1);
        tr.at(1, i);
    }
    tr.done(dev::gtid());
}

} // namespace

void op_gather_oblivious(vo_ctx *ctx, const vo_gather_oblivious_args &a)
{
    const uint32_t eb = a.elem_bytes;
    VO_CHECK(eb == 1 || eb == 2 || eb == 4 || eb == 8, VO_ERR_INVALID, "elem_bytes must be 1, 2, 4 or 8");
    VO_CHECK(a.n_table < (1ull << 31) && a.n < (1ull << 32), VO_ERR_INVALID, "table below 2^31 rows, requests below 2^32");
    VO_CHECK((a.table || a.n_table == 0) && (a.record_ids || a.n == 0) && (a.out || a.n == 0), VO_ERR_INVALID,
             "null buffer");
    if (a.n == 0) return;
    uint64_t n_pad = GT;
    while (n_pad < a.n_table + a.n) n_pad <<= 1;
    uint64_t m_pad = 1;
    while (m_pad < a.n) m_pad <<= 1;
    VO_CHECK(n_pad <= (1ull << 29), VO_ERR_INVALID, "table + requests above 2^29 rows");
    const uint32_t nt = (uint32_t)(n_pad / GT);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_k = take(n_pad * 8), o_v = take(n_pad * 8), o_c0 = take(n_pad * 8), o_c1 = take(n_pad * 8),
                 o_m = take(n_pad), o_t = take(nt * sizeof(Last)), o_c = take(nt * sizeof(Last)),
                 o_cnt = take(compact_rows2_scratch(n_pad));
    char *base = (char *)scratch(ctx, off);
    uint64_t *keys = (uint64_t *)(base + o_k), *vals = (uint64_t *)(base + o_v);
    uint64_t *c0 = (uint64_t *)(base + o_c0), *c1 = (uint64_t *)(base + o_c1);
    uint8_t *mask = (uint8_t *)(base + o_m);
    Last *tiles = (Last *)(base + o_t), *carry = (Last *)(base + o_c);
    uint64_t *cnt = (uint64_t *)(base + o_cnt);
    const bool tr = ctx->tracing;
    // values of at most 4 bytes (a public width) ride as the u32 payload of
    // 12-byte union records
    const bool kp = eb <= 4;
    auto pick = [&](auto f0, auto f1, auto k0, auto k1) { return kp ? (tr ? k1 : k0) : (tr ? f1 : f0); };
    launch(ctx, LaunchInfo{"gather_obl_build", grid_for(n_pad), dim3(256), 0, hash_mix(hash_mix(eb, a.n_table), a.n)},
           pick(k_og_build<false, false>, k_og_build<true, false>, k_og_build<false, true>, k_og_build<true, true>),
           a.table, a.n_table, eb, a.record_ids, a.n, n_pad, keys, (void *)vals, trace_slot(ctx));
    if (kp) sort_kp_full(ctx, keys, reinterpret_cast<uint32_t *>(vals), n_pad);
    else sort_kv(ctx, keys, vals, n_pad);
    launch(ctx, LaunchInfo{"gather_obl_tiles", dim3(nt), dim3(GB), 0, n_pad},
           pick(k_og_tiles<false, false>, k_og_tiles<true, false>, k_og_tiles<false, true>, k_og_tiles<true, true>),
           (const uint64_t *)keys, (const void *)vals, tiles, trace_slot(ctx));
    launch(ctx, LaunchInfo{"gather_obl_carry", dim3(1), dim3(kScanThreads), 0, nt},
           tr ? k_og_carry<true> : k_og_carry<false>, (const Last *)tiles, nt, carry, trace_slot(ctx));
    launch(ctx, LaunchInfo{"gather_obl_rows", dim3(nt), dim3(GB), 0, n_pad},
           pick(k_og_rows<false, false>, k_og_rows<true, false>, k_og_rows<false, true>, k_og_rows<true, true>),
           (const uint64_t *)keys, (const void *)vals, (const Last *)carry, c0, c1, mask, trace_slot(ctx));
    compact_rows2(ctx, c0, c1, mask, n_pad, cnt);
    // the first n rows are the requests; reuse keys (and vals as the u32
    // payload column) for the final sort
    uint32_t *hi = reinterpret_cast<uint32_t *>(vals);
    launch(ctx, LaunchInfo{"gather_obl_order", grid_for(m_pad), dim3(256), 0, hash_mix(a.n, m_pad)},
           tr ? k_og_order<true> : k_og_order<false>, (const uint64_t *)c0, (const uint64_t *)c1, a.n, m_pad, keys,
           hi, trace_slot(ctx));
    sort_kp_full(ctx, keys, hi, m_pad);
    launch(ctx, LaunchInfo{"gather_obl_out", grid_for(a.n), dim3(256), 0, hash_mix(a.n, eb)},
           tr ? k_og_out<true> : k_og_out<false>, (const uint64_t *)keys, (const uint32_t *)hi, a.n, eb, a.out,
           a.out_found, trace_slot(ctx));
}

} // namespace vo
