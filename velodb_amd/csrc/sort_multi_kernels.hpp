// sort_multi_kernels.hpp -- pack / unpack passes of the multi-column OP_SORT
// (reference semantics, include/cuda/sort.cuh:11-118) as device bodies plus
// __global__ wrappers, for launch mode (sort_multi.hip) and the resident
// queue kernel.
#pragma once

#include "vo_internal.hpp"

namespace vo {
namespace smulti_k {
namespace {

// Order-preserving unsigned image of a column value.  DESC columns are
// bit-inverted, so lexicographic ascending order of the packed words is the
// reference's column-by-column order with per-column direction
// (sort.cuh:31-40).  FLOAT/DOUBLE map -0.0 to the image of +0.0, since the
// reference compares them equal (compareSingle, sort.cuh:11-29); the sign of
// a zero travels in a payload bit the comparison ignores, so the unpack pass
// writes back the row's own bits.  NaN has no place in the reference's order
// (it compares unequal and not less than everything); here it takes the IEEE
// total-order image (above +inf, or below -inf for a negative NaN).
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
    uint32_t negz[VO_MAX_SORT_COLUMNS];  // FLOAT/DOUBLE: payload bit that marks a -0.0 (0 for other types)
    uint32_t idx_mask;                   // payload bits holding the row index
    uint32_t ncols;
    uint32_t nwords;  // words holding columns
    uint32_t nwrite;  // words written (>= nwords; the rest are zero)
};

// Packs the normalised columns into key words (MSB-first per word) and
// writes index = row.  Pad rows take the image of max for ASC columns and
// of lowest for DESC columns, the values sort.cuh:79-100 writes.
__device__ __forceinline__ void b_pack_keys(const PackSpec &ps, uint64_t *w0, uint64_t *w1, uint64_t *w2,
                                            uint64_t *w3, uint64_t *w4, uint64_t *w5, uint64_t *w6, uint64_t *w7,
                                            uint32_t *idx, uint64_t n, uint64_t n_pad, TraceSlot *slot, uint32_t vb,
                                            uint32_t nvb)
{
    uint64_t *ws[8] = {w0, w1, w2, w3, w4, w5, w6, w7};
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        uint64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const bool pad = i >= n;
        const uint64_t src = pad ? 0 : i;
        uint32_t flags = 0;
        for (uint32_t c = 0; c < ps.ncols; ++c) {
            const uint64_t mask = ps.bits[c] == 64 ? ~0ull : ((1ull << ps.bits[c]) - 1);
            const uint64_t raw = pad ? ps.fill[c] : load_raw(ps.cols[c], ps.types[c], src);
            // the sign of a zero: -0.0 is the sign bit alone
            const uint64_t sign = ps.bits[c] == 64 ? 0x8000000000000000ull : 0x80000000ull;
            flags |= (raw == sign) ? ps.negz[c] : 0u;
            uint64_t v = norm_bits(ps.types[c], raw);
            v = ps.desc[c] ? (~v & mask) : v;
            acc[ps.word[c]] |= v << ps.shift[c];
        }
        for (uint32_t w = 0; w < ps.nwrite; ++w) ws[w][i] = acc[w];
        idx[i] = (uint32_t)i | flags;
        tr.at(1, src); // columns read at row src (row 0 for pad rows: n is public)
        tr.at(2, i);   // key words and index written at row i
    }
    tr.done((uint64_t)vb * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_pack_keys(PackSpec ps, uint64_t *w0, uint64_t *w1, uint64_t *w2,
                                                   uint64_t *w3, uint64_t *w4, uint64_t *w5, uint64_t *w6,
                                                   uint64_t *w7, uint32_t *idx, uint64_t n, uint64_t n_pad,
                                                   TraceSlot *slot)
{
    b_pack_keys(ps, w0, w1, w2, w3, w4, w5, w6, w7, idx, n, n_pad, slot, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void b_unpack_keys(const PackSpec &ps, const uint64_t *w0, const uint64_t *w1,
                                              const uint64_t *w2, const uint64_t *w3, const uint64_t *w4,
                                              const uint64_t *w5, const uint64_t *w6, const uint64_t *w7,
                                              const uint32_t *idx, int64_t *out_idx, uint64_t n_pad, TraceSlot *slot,
                                              uint32_t vb, uint32_t nvb)
{
    const uint64_t *ws[8] = {w0, w1, w2, w3, w4, w5, w6, w7};
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        const uint32_t pay = idx[i];
        for (uint32_t c = 0; c < ps.ncols; ++c) {
            const uint64_t mask = ps.bits[c] == 64 ? ~0ull : ((1ull << ps.bits[c]) - 1);
            uint64_t v = (ws[ps.word[c]][i] >> ps.shift[c]) & mask;
            v = ps.desc[c] ? (~v & mask) : v;
            // a -0.0 row: the image of +0.0 with the sign bit put back (the
            // image of -0.0 is that of +0.0 with the top bit flipped: 0x0..0
            // -> 0x80..0 denormalises to +0.0, 0x7F..F to -0.0)
            if (pay & ps.negz[c]) v = ps.bits[c] == 64 ? 0x7FFFFFFFFFFFFFFFull : 0x7FFFFFFFull;
            denorm_store(const_cast<void *>(ps.cols[c]), ps.types[c], i, v);
        }
        out_idx[i] = (int64_t)(pay & ps.idx_mask);
        tr.at(1, i);
    }
    tr.done((uint64_t)vb * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_unpack_keys(PackSpec ps, const uint64_t *w0, const uint64_t *w1,
                                                     const uint64_t *w2, const uint64_t *w3, const uint64_t *w4,
                                                     const uint64_t *w5, const uint64_t *w6, const uint64_t *w7,
                                                     const uint32_t *idx, int64_t *out_idx, uint64_t n_pad,
                                                     TraceSlot *slot)
{
    b_unpack_keys(ps, w0, w1, w2, w3, w4, w5, w6, w7, idx, out_idx, n_pad, slot, blockIdx.x, gridDim.x);
}


} // namespace
} // namespace smulti_k
} // namespace vo
