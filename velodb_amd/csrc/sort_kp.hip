// sort_kp.hip -- the (u64 key, u32 payload) oblivious sort (the headline
// path) and the (u64 key, u64 value) sort used by the join.
#include "sort_kernels.hpp"

namespace vo {

using namespace sortimpl;

namespace {
using SortKP = SortImpl<RecKP, 13, 4, 512, 5, 7>;
using SortKV = SortImpl<RecKV, 12, 4, 256, 4, 6>;
} // namespace

void op_sort_kp(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n, uint64_t n_pad)
{
    VO_CHECK(is_pow2(n_pad) && n <= n_pad, VO_ERR_INVALID, "n_pad must be a power of two >= n");
    VO_CHECK(n_pad <= kMaxSortRows, VO_ERR_INVALID, "n_pad above 2^29 (32-bit buffer offsets of 8-byte columns)");
    VO_CHECK(keys && payload, VO_ERR_INVALID, "null buffer");
    Bufs<RecKP> b;
    b.k[0] = keys;
    b.p = payload;
    SortKP::run(ctx, b, ilog2(n_pad), n, 1);
}

void sort_kp_full(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n_pad)
{
    Bufs<RecKP> b;
    b.k[0] = keys;
    b.p = payload;
    SortKP::run(ctx, b, ilog2(n_pad), n_pad, 0);
}

void sort_kv(vo_ctx *ctx, uint64_t *keys, uint64_t *vals, uint64_t n_pad)
{
    Bufs<RecKV> b;
    b.k[0] = keys;
    b.p = vals;
    SortKV::run(ctx, b, ilog2(n_pad), n_pad, 0);
}

int sort_plan_passes(vo_ctx *ctx, uint64_t n_pad) { return SortKP::passes(ctx, ilog2(n_pad)); }

} // namespace vo
