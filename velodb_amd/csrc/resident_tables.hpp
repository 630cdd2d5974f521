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

// --------------------------------------------------------------- all tables
#define VO_ALL_BODIES(X) VO_MISC_BODIES(X) VO_SORT_BODIES(X)

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
