// sort_multi.hip -- multi-column OP_SORT with the reference's semantics
// (include/cuda/sort.cuh:11-118): columns are packed into order-preserving
// key words, sorted with the row index as payload (ties not broken, so the
// network decides them exactly as the reference does), then unpacked.
#include "sort_kernels.hpp"

namespace vo {

using namespace sortimpl;

namespace {

// 1- and 2-word keys: register passes of at most 3 stages, wider groups as
// LDS-staged hybrid passes (ORDER BY of two keys at 2^24: 5.66 -> 5.39 ms)
using SortM1 = SortImpl<RecM1, 12, 4, 256, 3, 7>;
using SortM2 = SortImpl<RecM2, 12, 4, 256, 3, 7>;
using SortM4 = SortImpl<RecM4, 11, 3, 256, 3, 6>;
using SortM8 = SortImpl<RecM8, 10, 2, 256, 2>;

// ------------------------------------------------- multi-column (OP_SORT)

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
__global__ void k_pack_keys(PackSpec ps, uint64_t *w0, uint64_t *w1, uint64_t *w2, uint64_t *w3, uint64_t *w4,
                            uint64_t *w5, uint64_t *w6, uint64_t *w7, uint32_t *idx, uint64_t n, uint64_t n_pad,
                            TraceSlot *slot)
{
    uint64_t *ws[8] = {w0, w1, w2, w3, w4, w5, w6, w7};
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
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
    tr.done(dev::gtid());
}

__global__ void k_unpack_keys(PackSpec ps, const uint64_t *w0, const uint64_t *w1, const uint64_t *w2,
                              const uint64_t *w3, const uint64_t *w4, const uint64_t *w5, const uint64_t *w6,
                              const uint64_t *w7, const uint32_t *idx, int64_t *out_idx, uint64_t n_pad,
                              TraceSlot *slot)
{
    const uint64_t *ws[8] = {w0, w1, w2, w3, w4, w5, w6, w7};
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
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
    tr.done(dev::gtid());
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

void op_sort_multi(vo_ctx *ctx, const vo_sort_args &a)
{
    VO_CHECK(a.n_sort_columns >= 1 && a.n_sort_columns <= VO_MAX_SORT_COLUMNS, VO_ERR_INVALID,
             "n_sort_columns must be 1..8");
    VO_CHECK(is_pow2(a.n_padded_rows) && a.n_rows <= a.n_padded_rows, VO_ERR_INVALID,
             "n_padded_rows must be a power of two >= n_rows");
    VO_CHECK(a.n_padded_rows <= kMaxSortRows, VO_ERR_INVALID, "n_padded_rows above 2^29");
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
    // The u32 payload holds the row index in its low log2(n_pad) bits; the
    // bits above carry, per FLOAT/DOUBLE key column, whether the row's value
    // is -0.0, so the unpack pass restores the sign the comparison ignored.
    const int idx_bits = ilog2(n_pad);
    ps.idx_mask = idx_bits >= 32 ? ~0u : (uint32_t)((1ull << idx_bits) - 1);
    uint32_t n_fp = 0;
    for (uint32_t c = 0; c < ps.ncols; ++c) {
        if (a.col_types[c] != VO_T_FLOAT && a.col_types[c] != VO_T_DOUBLE) continue;
        VO_CHECK((int)n_fp < 32 - idx_bits, VO_ERR_UNSUPPORTED,
                 "more FLOAT/DOUBLE key columns than 32 - log2(n_padded_rows) payload bits can carry the sign "
                 "of a zero for");
        ps.negz[c] = 1u << (31 - n_fp);
        ++n_fp;
    }
    // key words, then the u32 row indices the sort carries
    uint64_t *words = (uint64_t *)scratch(ctx, sizeof(uint64_t) * W * n_pad + sizeof(uint32_t) * n_pad);
    uint32_t *idx = reinterpret_cast<uint32_t *>(words + (uint64_t)W * n_pad);
    uint64_t *wp[8];
    for (int w = 0; w < 8; ++w) wp[w] = words + (uint64_t)std::min<uint32_t>(w, W - 1) * n_pad;
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n_pad + 255) / 256, 8192));
    LaunchInfo lp{"sort_pack", dim3(blocks), dim3(256), 0, hash_mix(ps.ncols, ps.nwords)};
    launch(ctx, lp, k_pack_keys, ps, wp[0], wp[1], wp[2], wp[3], wp[4], wp[5], wp[6], wp[7], idx, a.n_rows,
           n_pad, trace_slot(ctx));
    const int n_log = ilog2(n_pad);
    switch (W) {
    case 1: {
        Bufs<RecM1> b;
        b.k[0] = wp[0];
        b.p = idx;
        SortM1::run(ctx, b, n_log, n_pad, 0);
        break;
    }
    case 2: {
        Bufs<RecM2> b;
        for (int w = 0; w < 2; ++w) b.k[w] = wp[w];
        b.p = idx;
        SortM2::run(ctx, b, n_log, n_pad, 0);
        break;
    }
    case 4: {
        Bufs<RecM4> b;
        for (int w = 0; w < 4; ++w) b.k[w] = wp[w];
        b.p = idx;
        SortM4::run(ctx, b, n_log, n_pad, 0);
        break;
    }
    default: {
        Bufs<RecM8> b;
        for (int w = 0; w < 8; ++w) b.k[w] = wp[w];
        b.p = idx;
        SortM8::run(ctx, b, n_log, n_pad, 0);
        break;
    }
    }
    LaunchInfo lu{"sort_unpack", dim3(blocks), dim3(256), 0, hash_mix(ps.ncols, ps.nwords)};
    launch(ctx, lu, k_unpack_keys, ps, wp[0], wp[1], wp[2], wp[3], wp[4], wp[5], wp[6], wp[7],
           (const uint32_t *)idx, a.indices, n_pad, trace_slot(ctx));
}

} // namespace vo
