// misc.hip -- the reference's simple element-wise device ops and the
// synthetic workload generators.
//   OP_PERMUTE             include/cuda/permute.cuh:10-42
//   OP_GATHER              include/cuda/gather.cuh:11-45
//   OP_COMPARE_EQ          include/cuda/compare.cuh:25-91
//   OP_COMPOSE_POSITION_MAP / OP_ENRICH_POSITIONS
//                          include/cuda/oblivious_shuffle.cuh:10-34
// PERMUTE, GATHER, COMPOSE and ENRICH read or write at data-dependent
// addresses exactly as the reference does; the oblivious paths (sort,
// compaction, shuffle) never call them.
#include "misc_kernels.hpp"
#include "resident_tables.hpp"

#include <algorithm>

namespace vo {
namespace {

using namespace misc_k;

const std::pair<const void *, BodyInfo> kBodies[] = {VO_MISC_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

size_t type_bytes(uint32_t t)
{
    switch (t) {
    case VO_T_BOOLEAN: case VO_T_TINYINT: return 1;
    case VO_T_SMALLINT: return 2;
    case VO_T_INTEGER: case VO_T_FLOAT: case VO_T_DATE: return 4;
    case VO_T_BIGINT: case VO_T_DOUBLE: case VO_T_VARCHAR: case VO_T_UBIGINT: return 8;
    default: return 0;
    }
}

dim3 grid_for(uint64_t n, uint32_t block = 256, uint64_t cap = 16384)
{
    uint64_t b = (n + block - 1) / block;
    return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap)));
}

__global__ void k_gen_sort(uint64_t seed, uint64_t *__restrict__ keys, uint32_t *__restrict__ payload, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        keys[i] = dev::rand64(seed, 0, i);
        payload[i] = (uint32_t)i;
    }
}

__global__ void k_gen_fact(uint64_t seed, int32_t *__restrict__ day, int64_t *__restrict__ qty,
                           int64_t *__restrict__ price, int64_t *__restrict__ disc, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        day[i] = (int32_t)dev::rand_range(seed, 1, i, 8036, 10561 - 8036 + 1);
        qty[i] = dev::rand_range(seed, 2, i, 1, 50);
        price[i] = dev::rand_range(seed, 3, i, 90000, 10495000 - 90000 + 1);
        disc[i] = dev::rand_range(seed, 4, i, 0, 11);
    }
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// 16 bytes per lane, 4 per thread, non-temporal (streaming) loads and stores.
__global__ __launch_bounds__(256) void k_hbm_copy(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, uint64_t n16)
{
    const uint64_t base = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
    u32x4 v[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const uint64_t i = base + m * 256;
        v[m] = i < n16 ? __builtin_nontemporal_load(src + i) : u32x4{0, 0, 0, 0};
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const uint64_t i = base + m * 256;
        if (i < n16) __builtin_nontemporal_store(v[m], dst + i);
    }
}

#define VO_TYPE_SWITCH(type_id, BODY)                                                                \
    switch (type_bytes(type_id)) {                                                                     \
    case 1: { using T = uint8_t; BODY; break; }                                                        \
    case 2: { using T = uint16_t; BODY; break; }                                                       \
    case 4: { using T = uint32_t; BODY; break; }                                                       \
    case 8: { using T = uint64_t; BODY; break; }                                                       \
    default: throw Error{VO_ERR_UNSUPPORTED, "unsupported column type"};                              \
    }

} // namespace

void op_permute(vo_ctx *ctx, const vo_permute_args &a)
{
    if (a.n == 0) return;
    VO_CHECK(a.out_data && a.in_data && a.in_indices, VO_ERR_INVALID, "null buffer");
    LaunchInfo li{"permute", grid_for(a.n), dim3(256), 0, hash_mix(a.n, a.type_id)};
    VO_TYPE_SWITCH(a.type_id, launch(ctx, li, k_permute<T>, (T *)a.out_data, (const T *)a.in_data, a.in_indices, a.n));
}

void op_gather(vo_ctx *ctx, const vo_gather_args &a)
{
    if (a.n == 0) return;
    VO_CHECK(a.out_data && a.in_data && a.in_indices, VO_ERR_INVALID, "null buffer");
    LaunchInfo li{"gather", grid_for(a.n), dim3(256), 0, hash_mix(a.n, a.type_id)};
    VO_TYPE_SWITCH(a.type_id,
                   launch(ctx, li, k_gather<T>, (T *)a.out_data, (const T *)a.in_data, a.in_indices, a.in_mask, a.n));
}

void op_compare_eq(vo_ctx *ctx, const vo_compare_eq_args &a)
{
    if (a.n == 0) return;
    VO_CHECK(a.left_data && a.right_data && a.out_mask, VO_ERR_INVALID, "null buffer");
    LaunchInfo li{"compare_eq", grid_for(a.n), dim3(256), 0, hash_mix(a.n, a.type_id)};
    if (a.type_id == VO_T_FLOAT) {
        launch(ctx, li, k_compare_eq<float>, (const float *)a.left_data, (const float *)a.right_data, a.out_mask, a.n);
        return;
    }
    if (a.type_id == VO_T_DOUBLE) {
        launch(ctx, li, k_compare_eq<double>, (const double *)a.left_data, (const double *)a.right_data, a.out_mask,
               a.n);
        return;
    }
    VO_TYPE_SWITCH(a.type_id,
                   launch(ctx, li, k_compare_eq<T>, (const T *)a.left_data, (const T *)a.right_data, a.out_mask, a.n));
}

void op_compose(vo_ctx *ctx, const vo_compose_args &a)
{
    if (a.n == 0) return;
    VO_CHECK(a.old_map && a.sigma && a.new_map, VO_ERR_INVALID, "null buffer");
    LaunchInfo li{"compose_position_map", grid_for(a.n), dim3(256), 0, a.n};
    launch(ctx, li, k_compose, a.old_map, a.sigma, a.new_map, a.n);
}

void op_enrich(vo_ctx *ctx, const vo_enrich_args &a)
{
    if (a.n == 0) return;
    VO_CHECK(a.record_ids && a.position_map && a.positions, VO_ERR_INVALID, "null buffer");
    LaunchInfo li{"enrich_positions", grid_for(a.n), dim3(256), 0, a.n};
    launch(ctx, li, k_enrich, a.record_ids, a.position_map, a.positions, a.n);
}

void op_hbm_copy(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    VO_CHECK(dst && src && bytes % 16 == 0, VO_ERR_INVALID, "copy needs non-null buffers and a multiple of 16 bytes");
    const uint64_t n16 = bytes / 16;
    if (n16 == 0) return;
    LaunchInfo li{"hbm_copy", dim3((uint32_t)((n16 + 1023) / 1024)), dim3(256), 0, bytes};
    launch(ctx, li, k_hbm_copy, (const u32x4 *)src, (u32x4 *)dst, n16);
}

void op_gen_sort_input(vo_ctx *ctx, uint64_t seed, uint64_t *keys, uint32_t *payload, uint64_t n)
{
    if (n == 0) return;
    VO_CHECK(keys && payload, VO_ERR_INVALID, "null buffer");
    LaunchInfo li{"gen_sort", grid_for(n, 256, 65536), dim3(256), 0, n};
    launch(ctx, li, k_gen_sort, seed, keys, payload, n);
}

void op_gen_fact(vo_ctx *ctx, uint64_t seed, int32_t *day, int64_t *qty, int64_t *price, int64_t *disc, uint64_t n)
{
    if (n == 0) return;
    VO_CHECK(day && qty && price && disc, VO_ERR_INVALID, "null buffer");
    LaunchInfo li{"gen_fact", grid_for(n, 256, 65536), dim3(256), 0, n};
    launch(ctx, li, k_gen_fact, seed, day, qty, price, disc, n);
}

} // namespace vo
