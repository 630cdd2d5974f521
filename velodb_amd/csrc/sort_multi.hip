// sort_multi.hip -- multi-column OP_SORT with the reference's semantics
// (include/cuda/sort.cuh:11-118): columns are packed into order-preserving
// key words, sorted with the row index as payload (ties not broken, so the
// network decides them exactly as the reference does), then unpacked.
#include "resident_tables.hpp"
#include "sort_kernels.hpp"
#include "sort_multi_kernels.hpp"

namespace vo {

using namespace sortimpl;

namespace {

using namespace smulti_k;

const std::pair<const void *, BodyInfo> kBodies[] = {VO_SORT_M_BODIES(VO_BODY_ENTRY) VO_SORT_PACK_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

// 1- and 2-word keys: register passes of at most 3 stages, wider groups as
// LDS-staged hybrid passes (ORDER BY of two keys at 2^24: 5.66 -> 5.39 ms)
using SortM1 = SortImpl<RecM1, 12, 4, 256, 3, 7>;
using SortM2 = SortImpl<RecM2, 12, 4, 256, 3, 7>;
using SortM4 = SortImpl<RecM4, 11, 3, 256, 3, 6>;
using SortM8 = SortImpl<RecM8, 10, 2, 256, 2>;

// ------------------------------------------------- multi-column (OP_SORT)

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
