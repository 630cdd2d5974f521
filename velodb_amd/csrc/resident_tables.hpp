// resident_tables.hpp -- the kernels the resident queue kernel can run.
//
// One X-macro table per source file.  An entry is
//   X(ID, KERNEL, CALL, LDS_EXTRA)
// KERNEL names the __global__ kernel launch mode uses (the host side maps its
// address to ID); CALL is the device body's call with the kernel's own
// parameters `a...` followed by the body's extra arguments: `vb` (virtual
// block), `nvb` (virtual grid), `lds` (the resident block's dynamic LDS);
// LDS_EXTRA is the LDS the body carves from `lds` beyond the launch's dynamic
// LDS.  Every body runs with 256 threads per block.
//
// Only ids are needed where a table is expanded for the id enum; the host
// table of a source file names its own kernels; queue.hip includes every
// kernel header and expands every table into its dispatch.
#pragma once

#include "vo_internal.hpp"

// ------------------------------------------------------------ element-wise
#define VO_MISC_TYPED(X, T, S)                                                                                     \
    X(kB_permute_##S, (misc_k::k_permute<T>), (misc_k::b_permute<T>(a..., vb, nvb)), 0)                          \
    X(kB_gather_##S, (misc_k::k_gather<T>), (misc_k::b_gather<T>(a..., vb, nvb)), 0)                             \
    X(kB_compare_eq_##S, (misc_k::k_compare_eq<T>), (misc_k::b_compare_eq<T>(a..., vb, nvb)), 0)

#define VO_MISC_BODIES(X)                                                                                          \
    VO_MISC_TYPED(X, uint8_t, u8)                                                                                  \
    VO_MISC_TYPED(X, uint16_t, u16)                                                                                \
    VO_MISC_TYPED(X, uint32_t, u32)                                                                                \
    VO_MISC_TYPED(X, uint64_t, u64)                                                                                \
    X(kB_compare_eq_f32, (misc_k::k_compare_eq<float>), (misc_k::b_compare_eq<float>(a..., vb, nvb)), 0)          \
    X(kB_compare_eq_f64, (misc_k::k_compare_eq<double>), (misc_k::b_compare_eq<double>(a..., vb, nvb)), 0)        \
    X(kB_compose, (misc_k::k_compose), (misc_k::b_compose(a..., vb, nvb)), 0)                                     \
    X(kB_enrich, (misc_k::k_enrich), (misc_k::b_enrich(a..., vb, nvb)), 0)

// ------------------------------------------------------------------ sorts
// A record kind's kernels for tiles of 2^T records, 2^RL per thread, register
// passes of 1..SM stages and LDS-staged passes of 4..HM stages.
#define VO_SORT_TILE(X, TAG, R, T, RL)                                                                             \
    X(kB_##TAG##_tiny, (dev::k_sort_tiny<R, false>), (dev::body_tiny<R, false>(a..., vb)), 0)                      \
    X(kB_##TAG##_local, (dev::k_sort_local<R, T, RL, 256, false>),                                                 \
      (dev::body_local<R, T, RL, 256, false>(a..., lds, vb)), 0)                                                   \
    X(kB_##TAG##_local_ct, (dev::k_sort_local_ct<R, T, RL, 256, false>),                                           \
      (dev::body_local_ct<R, T, RL, 256, false>(a..., lds, vb)), 0)                                                \
    X(kB_##TAG##_merge, (dev::k_sort_merge<R, T, RL, 256, false>),                                                 \
      (dev::body_merge<R, T, RL, 256, false>(a..., lds, vb)), 0)
#define VO_SORT_G(X, TAG, R, S)                                                                                    \
    X(kB_##TAG##_g##S, (dev::k_sort_global<R, S, false>), (dev::body_global<R, S, false>(a..., vb, threadIdx.x)), 0)
#define VO_SORT_H(X, TAG, R, T, RL, S)                                                                             \
    X(kB_##TAG##_h##S, (dev::k_sort_hybrid<R, T, RL, 256, S, false>),                                              \
      (dev::body_hybrid<R, T, RL, 256, S, false>(a..., lds, vb)), 0)

// (u64 key, u32 payload) at the resident tile 2^12 (launch mode: 2^13 with
// 512 threads; the network is the same); register passes 1..5, LDS 4..7.
#define VO_SORT_KP_BODIES(X)                                                                                       \
    VO_SORT_TILE(X, kp, sortimpl::RecKP, 12, 4)                                                                    \
    VO_SORT_G(X, kp, sortimpl::RecKP, 1) VO_SORT_G(X, kp, sortimpl::RecKP, 2) VO_SORT_G(X, kp, sortimpl::RecKP, 3) \
    VO_SORT_G(X, kp, sortimpl::RecKP, 4) VO_SORT_G(X, kp, sortimpl::RecKP, 5)                                      \
    VO_SORT_H(X, kp, sortimpl::RecKP, 12, 4, 4) VO_SORT_H(X, kp, sortimpl::RecKP, 12, 4, 5)                        \
    VO_SORT_H(X, kp, sortimpl::RecKP, 12, 4, 6) VO_SORT_H(X, kp, sortimpl::RecKP, 12, 4, 7)

// (u64 key, u64 value): as launch mode (tile 2^12, passes 1..5 and 4..7)
#define VO_SORT_KV_BODIES(X)                                                                                       \
    VO_SORT_TILE(X, kv, sortimpl::RecKV, 12, 4)                                                                    \
    VO_SORT_G(X, kv, sortimpl::RecKV, 1) VO_SORT_G(X, kv, sortimpl::RecKV, 2) VO_SORT_G(X, kv, sortimpl::RecKV, 3) \
    VO_SORT_G(X, kv, sortimpl::RecKV, 4) VO_SORT_G(X, kv, sortimpl::RecKV, 5)                                      \
    VO_SORT_H(X, kv, sortimpl::RecKV, 12, 4, 4) VO_SORT_H(X, kv, sortimpl::RecKV, 12, 4, 5)                        \
    VO_SORT_H(X, kv, sortimpl::RecKV, 12, 4, 6) VO_SORT_H(X, kv, sortimpl::RecKV, 12, 4, 7)

// multi-column records: as launch mode
#define VO_SORT_M_BODIES(X)                                                                                        \
    VO_SORT_TILE(X, m1, sortimpl::RecM1, 12, 4)                                                                    \
    VO_SORT_G(X, m1, sortimpl::RecM1, 1) VO_SORT_G(X, m1, sortimpl::RecM1, 2) VO_SORT_G(X, m1, sortimpl::RecM1, 3) \
    VO_SORT_H(X, m1, sortimpl::RecM1, 12, 4, 4) VO_SORT_H(X, m1, sortimpl::RecM1, 12, 4, 5)                        \
    VO_SORT_H(X, m1, sortimpl::RecM1, 12, 4, 6) VO_SORT_H(X, m1, sortimpl::RecM1, 12, 4, 7)                        \
    VO_SORT_TILE(X, m2, sortimpl::RecM2, 12, 4)                                                                    \
    VO_SORT_G(X, m2, sortimpl::RecM2, 1) VO_SORT_G(X, m2, sortimpl::RecM2, 2) VO_SORT_G(X, m2, sortimpl::RecM2, 3) \
    VO_SORT_H(X, m2, sortimpl::RecM2, 12, 4, 4) VO_SORT_H(X, m2, sortimpl::RecM2, 12, 4, 5)                        \
    VO_SORT_H(X, m2, sortimpl::RecM2, 12, 4, 6) VO_SORT_H(X, m2, sortimpl::RecM2, 12, 4, 7)                        \
    VO_SORT_TILE(X, m4, sortimpl::RecM4, 11, 3)                                                                    \
    VO_SORT_G(X, m4, sortimpl::RecM4, 1) VO_SORT_G(X, m4, sortimpl::RecM4, 2) VO_SORT_G(X, m4, sortimpl::RecM4, 3) \
    VO_SORT_H(X, m4, sortimpl::RecM4, 11, 3, 4) VO_SORT_H(X, m4, sortimpl::RecM4, 11, 3, 5)                        \
    VO_SORT_H(X, m4, sortimpl::RecM4, 11, 3, 6)                                                                    \
    VO_SORT_TILE(X, m8, sortimpl::RecM8, 10, 2)                                                                    \
    VO_SORT_G(X, m8, sortimpl::RecM8, 1) VO_SORT_G(X, m8, sortimpl::RecM8, 2)

#define VO_SORT_PACK_BODIES(X)                                                                                     \
    X(kB_sort_pack, (smulti_k::k_pack_keys), (smulti_k::b_pack_keys(a..., vb, nvb)), 0)                           \
    X(kB_sort_unpack, (smulti_k::k_unpack_keys), (smulti_k::b_unpack_keys(a..., vb, nvb)), 0)

#define VO_SORT_BODIES(X) VO_SORT_KP_BODIES(X) VO_SORT_KV_BODIES(X) VO_SORT_M_BODIES(X) VO_SORT_PACK_BODIES(X)

// ------------------------------------------------------------- compaction
#define VO_COMPACT_G(X, NW, S)                                                                                     \
    X(kB_compact_g##NW##_##S, (compact_k::k_compact_global<NW, S, false>),                                         \
      (compact_k::b_compact_global<NW, S, false>(a..., vb)), 0)
#define VO_COMPACT_H(X, NW, S, W8, TAG)                                                                            \
    X(kB_compact_h##NW##_##S##_##TAG, (compact_k::k_compact_hybrid<NW, S, W8, false>),                             \
      (compact_k::b_compact_hybrid<NW, S, W8, false>(a..., lds, vb)), 0)
#define VO_COMPACT_NW(X, NW)                                                                                       \
    X(kB_compact_local##NW##_w8, (compact_k::k_compact_local<NW, true, false>),                                    \
      (compact_k::b_compact_local<NW, true, false>(a..., lds, vb)), 0)                                             \
    X(kB_compact_local##NW##_wx, (compact_k::k_compact_local<NW, false, false>),                                   \
      (compact_k::b_compact_local<NW, false, false>(a..., lds, vb)), 0)                                            \
    VO_COMPACT_G(X, NW, 1) VO_COMPACT_G(X, NW, 2) VO_COMPACT_G(X, NW, 3) VO_COMPACT_G(X, NW, 4)                    \
    VO_COMPACT_H(X, NW, 5, true, w8) VO_COMPACT_H(X, NW, 6, true, w8) VO_COMPACT_H(X, NW, 7, true, w8)             \
    VO_COMPACT_H(X, NW, 8, true, w8) VO_COMPACT_H(X, NW, 5, false, wx) VO_COMPACT_H(X, NW, 6, false, wx)           \
    VO_COMPACT_H(X, NW, 7, false, wx) VO_COMPACT_H(X, NW, 8, false, wx)

// the 1024-thread tile scan runs as a 256-thread body on the resident grid;
// LDS_EXTRA: the scan's 256 words, the int128 reductions' 256 values
#define VO_COMPACT_BODIES(X)                                                                                       \
    VO_COMPACT_NW(X, 1) VO_COMPACT_NW(X, 2) VO_COMPACT_G(X, 1, 5)                                                  \
    X(kB_count_mask, (compact_k::k_count_mask), (compact_k::b_count_mask(a..., vb)), 0)                           \
    X(kB_filter_count, (compact_k::k_filter_count<false>), (compact_k::b_filter_count<false>(a..., vb)), 0)       \
    X(kB_scan_tiles, (compact_k::k_scan_tiles), (compact_k::b_scan_tiles<256>(a..., (uint64_t *)lds)), 2048)      \
    X(kB_scan_seg, (compact_k::k_scan_seg), (compact_k::b_scan_seg(a..., (uint64_t *)lds, vb)), 2048)             \
    X(kB_sum_prod, (compact_k::k_sum_prod_i128<false>),                                                            \
      (compact_k::b_sum_prod_i128<false>(a..., (__int128 *)lds, vb, nvb)),                                         \
      4096)                                                                                                        \
    X(kB_sum_final, (compact_k::k_sum_final), (compact_k::b_sum_final(a..., (__int128 *)lds)), 4096)              \
    X(kB_scatter_positions, (compact_k::k_scatter_positions), (compact_k::b_scatter_positions(a..., vb)), 0)

// -------------------------------------------------------- predicate, shuffle
#define VO_PRED_BODIES(X)                                                                                          \
    X(kB_predicate4, (pred_k::k_predicate_mask<4>), (pred_k::b_predicate_mask<4>(a..., vb, nvb)), 0)               \
    X(kB_predicate1, (pred_k::k_predicate_mask<1>), (pred_k::b_predicate_mask<1>(a..., vb, nvb)), 0)

#define VO_SHUFFLE_BODIES(X)                                                                                       \
    X(kB_shuffle_tags, (shuffle_k::k_shuffle_tags), (shuffle_k::b_shuffle_tags(a..., vb, nvb)), 0)                 \
    X(kB_shuffle_unpack, (shuffle_k::k_shuffle_unpack), (shuffle_k::b_shuffle_unpack(a..., vb, nvb)), 0)           \
    X(kB_shuffle_keys, (shuffle_k::k_shuffle_keys), (shuffle_k::b_shuffle_keys(a..., vb, nvb)), 0)

// ------------------------------------------------------ join + GROUP BY
// the single-block 1024-thread scans and output run as 256-thread bodies
#define VO_JOIN_AGG(X, G)                                                                                          \
    X(kB_join_agg##G##_kp, (join_k::k_fill_agg<G, false, true>), (join_k::b_fill_agg<G, false, true>(a..., vb, nvb)), \
      0)                                                                                                           \
    X(kB_join_agg##G##_kv, (join_k::k_fill_agg<G, false, false>),                                                  \
      (join_k::b_fill_agg<G, false, false>(a..., vb, nvb)), 0)                                                     \
    X(kB_join_dense_out##G, (join_k::k_dense_out<G>), (join_k::b_dense_out<G, 256>(a..., (int64_t *)lds)),         \
      (2 * 256 + 2 * G) * 8)

#define VO_JOIN_BODIES(X)                                                                                          \
    X(kB_join_build, (join_k::k_join_build<false>), (join_k::b_join_build<false>(a..., vb, nvb)), 0)               \
    X(kB_join_build_kp, (join_k::k_join_build_kp<false>), (join_k::b_join_build_kp<false>(a..., vb, nvb)), 0)      \
    X(kB_join_fill_tiles_kp, (join_k::k_fill_tiles<false, true>), (join_k::b_fill_tiles<false, true>(a..., vb, nvb)), \
      0)                                                                                                           \
    X(kB_join_fill_tiles_kv, (join_k::k_fill_tiles<false, false>),                                                 \
      (join_k::b_fill_tiles<false, false>(a..., vb, nvb)), 0)                                                      \
    X(kB_join_fill_carry, (join_k::k_fill_carry<false>),                                                           \
      (join_k::b_fill_carry<256, false>(a..., (join_k::Fill *)lds)), 256 * sizeof(join_k::Fill))                   \
    X(kB_join_fill_rows, (join_k::k_fill_rows<false>), (join_k::b_fill_rows<false>(a..., vb, nvb)), 0)             \
    X(kB_join_group_keys, (join_k::k_group_keys<false>), (join_k::b_group_keys<false>(a..., vb, nvb)), 0)          \
    X(kB_join_seg_tiles, (join_k::k_seg_tiles<false>), (join_k::b_seg_tiles<false>(a..., vb, nvb)), 0)             \
    X(kB_join_seg_summ, (join_k::k_seg_summ<false>), (join_k::b_seg_summ<false>(a..., vb, nvb)), 0)                \
    X(kB_join_seg_carry, (join_k::k_seg_carry<false>),                                                             \
      (join_k::b_seg_carry<256, false>(a..., (join_k::SegSum *)lds)), 256 * sizeof(join_k::SegSum))                \
    X(kB_join_seg_rows, (join_k::k_seg_rows<false>), (join_k::b_seg_rows<false>(a..., vb, nvb)), 0)                \
    X(kB_join_group_out, (join_k::k_group_out<false>), (join_k::b_group_out<false>(a..., vb, nvb)), 0)             \
    VO_JOIN_AGG(X, 32) VO_JOIN_AGG(X, 64)

// ---------------------------------------------------- equi-join into pairs
// resident distribution / fill tiles of 1024 slots (launch mode 2048; every
// slot depends on the 256 before it only)
#define VO_JP_LAT(X, NW, S, K)                                                                                     \
    X(kB_jp_lat##NW##_##S##_##K, (jp_k::k_jp_dist_lattice<NW, S, K, false>),                                       \
      (jp_k::b_jp_dist_lattice<NW, S, K, false>(a..., vb, nvb)), 0)
#define VO_JP_NW(X, NW)                                                                                            \
    X(kB_jp_dist_init##NW, (jp_k::k_jp_dist_init<NW, false>), (jp_k::b_jp_dist_init<NW, false>(a..., vb, nvb)), 0) \
    X(kB_jp_dist_levels##NW, (jp_k::k_jp_dist_levels<NW, 4, false>),                                              \
      (jp_k::b_jp_dist_levels<NW, 4, false>(a..., vb, nvb)), 0)                                                    \
    X(kB_jp_dist_tile##NW, (jp_k::k_jp_dist_tile<NW, 1024, false>),                                                \
      (jp_k::b_jp_dist_tile<NW, 1024, false>(a..., (uint64_t *)lds, vb)), 0)                                       \
    X(kB_jp_fill_carry##NW, (jp_k::k_jp_fill_carry<NW, false>),                                                    \
      (jp_k::b_jp_fill_carry<256, NW, false>(a..., (jp_k::Elem<NW> *)lds)), 256 * 8 * NW)                          \
    VO_JP_LAT(X, NW, 1, 2) VO_JP_LAT(X, NW, 1, 4) VO_JP_LAT(X, NW, 2, 4) VO_JP_LAT(X, NW, 1, 8)                    \
    VO_JP_LAT(X, NW, 2, 8) VO_JP_LAT(X, NW, 3, 8) VO_JP_LAT(X, NW, 1, 16) VO_JP_LAT(X, NW, 2, 16)                  \
    VO_JP_LAT(X, NW, 3, 16)
#define VO_JP_KP(X, KP, TAG)                                                                                       \
    X(kB_jp_bwd_tiles_##TAG, (jp_k::k_jp_bwd_tiles<false, KP>), (jp_k::b_jp_bwd_tiles<false, KP>(a..., vb, nvb)),  \
      0)                                                                                                           \
    X(kB_jp_bwd_rows_##TAG, (jp_k::k_jp_bwd_rows<false, KP>), (jp_k::b_jp_bwd_rows<false, KP>(a..., vb, nvb)), 0)  \
    X(kB_jp_fwd_rows_##TAG, (jp_k::k_jp_fwd_rows<false, KP>), (jp_k::b_jp_fwd_rows<false, KP>(a..., vb, nvb)), 0)

#define VO_JP_BODIES(X)                                                                                            \
    X(kB_jp_build, (jp_k::k_jp_build<false>), (jp_k::b_jp_build<false>(a..., vb, nvb)), 0)                         \
    X(kB_jp_build_kp, (jp_k::k_jp_build_kp<false>), (jp_k::b_jp_build_kp<false>(a..., vb, nvb)), 0)               \
    VO_JP_KP(X, true, kp) VO_JP_KP(X, false, kv)                                                                   \
    X(kB_jp_bwd_carry, (jp_k::k_jp_bwd_carry<false>), (jp_k::b_jp_bwd_carry<256, false>(a..., (jp_k::Min2 *)lds)), \
      256 * sizeof(jp_k::Min2))                                                                                    \
    X(kB_jp_fwd_carry, (jp_k::k_jp_fwd_carry<false>), (jp_k::b_jp_fwd_carry<256, false>(a..., (jp_k::Fwd *)lds)),  \
      256 * sizeof(jp_k::Fwd))                                                                                     \
    VO_JP_NW(X, 1) VO_JP_NW(X, 2) VO_JP_LAT(X, 1, 1, 32) VO_JP_LAT(X, 1, 2, 32) VO_JP_LAT(X, 1, 3, 32)             \
    X(kB_jp_fill_left, (jp_k::k_jp_fill_rows<1, 0, 1024, false>),                                                  \
      (jp_k::b_jp_fill_rows<1, 0, 1024, false>(a..., vb)), 0)                                                      \
    X(kB_jp_fill_right, (jp_k::k_jp_fill_rows<2, 1, 1024, false>),                                                 \
      (jp_k::b_jp_fill_rows<2, 1, 1024, false>(a..., vb)), 0)                                                      \
    X(kB_jp_out, (jp_k::k_jp_out<false>), (jp_k::b_jp_out<false>(a..., vb, nvb)), 0)

// ------------------------------------------------------- oblivious gather
#define VO_OG_KP(X, KP, TAG)                                                                                       \
    X(kB_og_build_##TAG, (og_k::k_og_build<false, KP>), (og_k::b_og_build<false, KP>(a..., vb, nvb)), 0)           \
    X(kB_og_tiles_##TAG, (og_k::k_og_tiles<false, KP>), (og_k::b_og_tiles<false, KP>(a..., vb, nvb)), 0)           \
    X(kB_og_rows_##TAG, (og_k::k_og_rows<false, KP>), (og_k::b_og_rows<false, KP>(a..., vb, nvb)), 0)

#define VO_OG_BODIES(X)                                                                                            \
    VO_OG_KP(X, true, kp) VO_OG_KP(X, false, kv)                                                                   \
    X(kB_og_carry, (og_k::k_og_carry<false>), (og_k::b_og_carry<256, false>(a..., (og_k::Last *)lds)),             \
      256 * sizeof(og_k::Last))                                                                                    \
    X(kB_og_order, (og_k::k_og_order<false>), (og_k::b_og_order<false>(a..., vb, nvb)), 0)                         \
    X(kB_og_out, (og_k::k_og_out<false>), (og_k::b_og_out<false>(a..., vb, nvb)), 0)

// ------------------------------------------------------------ serving loop
#define VO_SERVE_BODIES(X)                                                                                         \
    X(kB_sort_checksum, (serve_k::k_sort_checksum), (serve_k::b_sort_checksum(a..., (uint64_t *)lds, vb, nvb)),   \
      512 * 8)

// --------------------------------------------------------------- all tables
#define VO_ALL_BODIES(X)                                                                                           \
    VO_MISC_BODIES(X) VO_SORT_BODIES(X) VO_COMPACT_BODIES(X) VO_PRED_BODIES(X) VO_SHUFFLE_BODIES(X)                \
    VO_JOIN_BODIES(X) VO_JP_BODIES(X) VO_OG_BODIES(X) VO_SERVE_BODIES(X)

namespace vo {

#define VO_BODY_ENUM(ID, K, CALL, L) ID,
enum BodyId : uint32_t {
    kBodyBase = kFirstBody - 1,
    VO_ALL_BODIES(VO_BODY_ENUM) kBodyEnd
};
#undef VO_BODY_ENUM

} // namespace vo

// Host-side registry entry of one table row (expanded in each source file).
#define VO_BODY_ENTRY(ID, K, CALL, L) {(const void *)(K), ::vo::BodyInfo{::vo::ID, (uint32_t)(L)}},
