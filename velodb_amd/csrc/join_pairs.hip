// join_pairs.hip -- oblivious equi-join of two key columns into row-index
// pairs: OP_SORT_MERGE_JOIN_COUNT / OP_SORT_MERGE_JOIN_WRITE, the device
// work of SortMergeJoinOperator (src/operator/sort_merge_join_operator.cpp:27-257).
//
// Reference method: the general m0 x m1 join of
// obliviator/join/enclave/scalable_oblivious_join.c:159-317 with the
// oblivious distribution of obliviator/join_kks_INT_INT/Include/sort.h:87-126.
//
// COUNT (over the union, n_pad = 2^k rows):
//   1. union records key = order-preserving code of the key,
//      val = (class << 62) | row, class 0 left, 2 right, 3 pad, and a sort by
//      (key, val): a key's left rows precede its right rows;
//   2. a backward scan (suffix minima) gives every row the end E of its key
//      segment and, for left rows, the first right row Q of the segment:
//      m1 = E - Q;
//   3. a forward scan gives the segment start S, the first right row R and
//      the exclusive sum X of m1 over the left rows before the row:
//        left row : target T = X                    (= B + a*m1)
//        right row: m0 = R - S, m1 = E - R, b = i - R,
//                   B = X - m0*m1, target T = B + b*m0,
//      B being the segment's first output row and n_out the total of m1;
//   4. ORCompact of the left rows with m1 > 0 and of the right rows with
//      m0 > 0 (the control bits of :104-122).
// WRITE (over M_pad = 2^k >= n_out output rows):
//   5. distribution of each compacted copy to its targets: levels
//      j = M_pad/2 .. 1, an element moves from p to p + j when its target is
//      >= p + j, and the slot it enters is always empty (the order of the
//      targets guarantees it), so every level is one branch-free pass
//      new[p] = moves(old[p-j]) ? old[p-j] : moves(old[p]) ? empty : old[p];
//      levels j >= 256 run one per HBM pass, the last 8 in LDS per 2048-slot
//      tile with a 256-slot halo (an output slot depends only on the 255
//      slots before it in those levels);
//   6. forward fill of the empty slots (:305-314): slot p holds the left row
//      of its pair; for the right copy, copy a = p - T goes to final row
//      B + a*m1 + b (the j_order of aggregation_tree.c:200-221) by a sort of
//      (final row, right row).
// Every launch reads and writes every row of its padded size; only n_out
// leaves the device.
#include "join_pairs_kernels.hpp"
#include "resident_tables.hpp"

#include <algorithm>

namespace vo {
namespace {

using namespace jp_k;

const std::pair<const void *, BodyInfo> kBodies[] = {VO_JP_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65536))); }

// --------------------------------------------------------- host driver

bool key_type_ok(uint32_t t)
{
    switch (t) {
    case VO_T_BOOLEAN: case VO_T_TINYINT: case VO_T_SMALLINT: case VO_T_INTEGER: case VO_T_BIGINT:
    case VO_T_DATE: case VO_T_VARCHAR: case VO_T_UBIGINT: return true;
    default: return false;
    }
}

struct Region {
    size_t off = 0;
    size_t take(size_t bytes)
    {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    }
};

// The routing state a COUNT leaves for its WRITE: the compacted left and
// right copies and their counts (24 B per union row and the tile arrays).
struct StateLayout {
    size_t o_cl, o_cr0, o_cr1, o_nout, o_cntl, o_cntr, bytes;
};

StateLayout state_layout(uint64_t n_pad)
{
    Region sr;
    StateLayout l;
    l.o_cl = sr.take(n_pad * 8);
    l.o_cr0 = sr.take(n_pad * 8);
    l.o_cr1 = sr.take(n_pad * 8);
    l.o_nout = sr.take(8);
    l.o_cntl = sr.take(compact_rows2_scratch(n_pad));
    l.o_cntr = sr.take(compact_rows2_scratch(n_pad));
    l.bytes = sr.off;
    return l;
}

uint64_t union_pad(uint64_t n_left, uint64_t n_right)
{
    uint64_t n_pad = PT;
    while (n_pad < n_left + n_right) n_pad <<= 1;
    return n_pad;
}

// The context-owned state (when the caller passes no state buffer).
struct JoinState {
    char *base = nullptr;
    size_t bytes = 0;
    uint64_t *n_out_host = nullptr; // pinned: the COUNT's pair count, written by the device
    const void *lk = nullptr, *rk = nullptr;
    uint64_t nl = 0, nr = 0, n_pad = 0, n_out = 0;
    uint64_t count_seq = 0; // resident queue: the COUNT command whose count n_out_host receives
    uint32_t type = 0;
    bool valid = false;
};

struct StatePtrs {
    uint64_t *cl, *cr0, *cr1, *cnt_l, *cnt_r, *n_out_dev;
};

StatePtrs state_ptrs(char *base, const StateLayout &l)
{
    return StatePtrs{(uint64_t *)(base + l.o_cl),   (uint64_t *)(base + l.o_cr0),  (uint64_t *)(base + l.o_cr1),
                     (uint64_t *)(base + l.o_cntl), (uint64_t *)(base + l.o_cntr), (uint64_t *)(base + l.o_nout)};
}

// k_jp_dist_lattice<NW, S, K> for runtime S in 1..3 and K in {2, .., 32}
// (K >= 2^S always holds: rows = m_pad / u >= 2^S).  Two-word elements stop
// at K = 16: at 32 the rows take every VGPR and one wave per SIMD remains.
template <int NW>
constexpr uint64_t kLatticeRows = NW == 1 ? 32 : 16;
template <int NW, int S, int K, class F>
void lattice_k(bool tr, F &&f)
{
    if (tr) f(k_jp_dist_lattice<NW, S, K, true>);
    else f(k_jp_dist_lattice<NW, S, K, false>);
}

template <int NW, int S, class F>
void lattice_s(int K, bool tr, F &&f)
{
    switch (K) {
    case 2: if constexpr (S <= 1) { lattice_k<NW, S, 2>(tr, f); return; } break;
    case 4: if constexpr (S <= 2) { lattice_k<NW, S, 4>(tr, f); return; } break;
    case 8: lattice_k<NW, S, 8>(tr, f); return;
    case 16: lattice_k<NW, S, 16>(tr, f); return;
    case 32: if constexpr (NW == 1) { lattice_k<NW, S, 32>(tr, f); return; } break;
    default: break;
    }
    throw Error{VO_ERR_INVALID, "distribution lattice: bad row count"};
}

template <int NW, class F>
void dist_lattice_launch(int S, int K, bool tr, F &&f)
{
    switch (S) {
    case 1: lattice_s<NW, 1>(K, tr, f); break;
    case 2: lattice_s<NW, 2>(K, tr, f); break;
    default: lattice_s<NW, 3>(K, tr, f); break;
    }
}

// Distribution tile (slots per tile): launch mode 2048, the resident kernel
// 1024 (its LDS holds 48 KiB per block).
int dist_tile(vo_ctx *ctx) { return ctx->rec ? DT_RES : DT; }

template <int NW, int TS>
void distribute(vo_ctx *ctx, const uint64_t *s0, const uint64_t *s1, uint64_t src_rows, const uint64_t *kdev,
                uint64_t m_pad, uint64_t *a0, uint64_t *a1, uint64_t *b0, uint64_t *b1, Elem<NW> *tl,
                Elem<NW> *tc, uint64_t **res0, uint64_t **res1)
{
    const bool tr = ctx->tracing;
    launch(ctx, LaunchInfo{NW == 1 ? "join_pairs_dist_init_w1" : "join_pairs_dist_init_w2", grid_for(m_pad), dim3(256),
                           0, hash_mix(NW, m_pad)},
           tr ? k_jp_dist_init<NW, true> : k_jp_dist_init<NW, false>, s0, s1, src_rows, kdev, m_pad, a0, a1,
           trace_slot(ctx));
    uint64_t *c0 = a0, *c1 = a1, *n0 = b0, *n1 = b1;
    // levels m_pad/2 .. 256, up to kDistFuse per pass, as even as possible
    const int count = ilog2(m_pad / 2) - ilog2(DH) + 1;
    const int groups = (count + kDistFuse - 1) / kDistFuse;
    uint64_t j = m_pad / 2;
    for (int gi = 0; gi < groups; ++gi) {
        const int sl = count / groups + (gi < count % groups ? 1 : 0);
        LaunchInfo li{NW == 1 ? "join_pairs_dist_level_w1" : "join_pairs_dist_level_w2", grid_for(m_pad), dim3(256), 0,
                      hash_mix(hash_mix(NW, m_pad), hash_mix(j, sl))};
        auto go = [&](auto k) {
            launch(ctx, li, k, (const uint64_t *)c0, (const uint64_t *)c1, m_pad, j, n0, n1, trace_slot(ctx));
        };
        if (sl <= 3) {
            // lattice form: u = smallest stride of the group (>= DH = 256),
            // rows = m_pad / u >= 2^sl, K = min(kLatticeRows, rows)
            const uint64_t u = j >> (sl - 1);
            const uint64_t rows = m_pad / u;
            const int K = (int)std::min<uint64_t>(kLatticeRows<NW>, rows);
            LaunchInfo ll = li;
            ll.grid = dim3((uint32_t)(m_pad / (256ull * (uint64_t)K)));
            ll.block = dim3(256);
            auto lat = [&](auto k) {
                launch(ctx, ll, k, (const uint64_t *)c0, (const uint64_t *)c1, u, n0, n1, trace_slot(ctx));
            };
            dist_lattice_launch<NW>(sl, K, tr, lat);
        } else {
            go(tr ? k_jp_dist_levels<NW, 4, true> : k_jp_dist_levels<NW, 4, false>);
        }
        j >>= sl;
        std::swap(c0, n0);
        std::swap(c1, n1);
    }
    const uint32_t nt = (uint32_t)(m_pad / TS);
    const size_t lds = (size_t)2 * NW * (TS + DH) * 8;
    auto kt = tr ? k_jp_dist_tile<NW, TS, true> : k_jp_dist_tile<NW, TS, false>;
    VO_HIP(hipFuncSetAttribute((const void *)kt, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    launch(ctx, LaunchInfo{NW == 1 ? "join_pairs_dist_tile_w1" : "join_pairs_dist_tile_w2", dim3(nt), dim3(PB),
                           (uint32_t)lds, hash_mix(NW, m_pad)},
           kt,
           (const uint64_t *)c0, (const uint64_t *)c1, n0, n1, tl, trace_slot(ctx));
    launch(ctx, LaunchInfo{"join_pairs_fill_carry", dim3(1), dim3(kScanThreads), 0, hash_mix(NW, nt)},
           tr ? k_jp_fill_carry<NW, true> : k_jp_fill_carry<NW, false>, (const Elem<NW> *)tl, nt, tc,
           trace_slot(ctx));
    *res0 = n0;
    *res1 = n1;
}

JoinState &state_of(vo_ctx *ctx)
{
    if (!ctx->join_pairs_state) ctx->join_pairs_state = new JoinState();
    return *(JoinState *)ctx->join_pairs_state;
}

// Waits until no submitted work may still use a buffer about to be replaced.
void quiesce(vo_ctx *ctx)
{
    if (ctx->rec)
        queue_drain_locked(ctx);
    else
        VO_HIP(hipStreamSynchronize(ctx->stream));
}

void check_keys(const vo_join_pairs_args &a)
{
    VO_CHECK(key_type_ok(a.type_id), a.type_id == VO_T_FLOAT || a.type_id == VO_T_DOUBLE ? VO_ERR_UNSUPPORTED
                                                                                          : VO_ERR_INVALID,
             "join keys must be an integer, date or ordinal type");
    VO_CHECK((a.left_keys || a.n_left == 0) && (a.right_keys || a.n_right == 0), VO_ERR_INVALID, "null key column");
    VO_CHECK(a.n_left < (1ull << 31) && a.n_right < (1ull << 31), VO_ERR_INVALID, "row counts must be below 2^31");
    VO_CHECK(union_pad(a.n_left, a.n_right) <= (1ull << 29), VO_ERR_INVALID, "join union above 2^29 rows");
}

// The WRITE's expansion of a COUNT state into the output pairs.
template <int TS>
void write_pairs(vo_ctx *ctx, const vo_join_pairs_args &a, const StatePtrs &sp, uint64_t n_pad, uint64_t n_out,
                 uint64_t out_cap)
{
    const bool tr = ctx->tracing;
    if (n_out == 0) {
        if (out_cap)
            launch(ctx, LaunchInfo{"join_pairs_out", grid_for(out_cap), dim3(256), 0, hash_mix(0, out_cap)},
                   tr ? k_jp_out<true> : k_jp_out<false>, (const uint32_t *)nullptr, (uint64_t)0, (uint64_t)0,
                   out_cap, a.out_left, a.out_right, trace_slot(ctx));
        return;
    }
    uint64_t m_pad = 8192;
    while (m_pad < n_out) m_pad <<= 1;
    const uint32_t nt = (uint32_t)(m_pad / TS);
    Region r;
    const size_t o_a0 = r.take(m_pad * 8), o_a1 = r.take(m_pad * 8), o_b0 = r.take(m_pad * 8),
                 o_b1 = r.take(m_pad * 8), o_sk = r.take(m_pad * 8), o_sp = r.take(m_pad * 4),
                 o_tl = r.take(nt * sizeof(Elem<2>)), o_tc = r.take(nt * sizeof(Elem<2>));
    char *base = (char *)scratch(ctx, r.off);
    uint64_t *a0 = (uint64_t *)(base + o_a0), *a1 = (uint64_t *)(base + o_a1);
    uint64_t *b0 = (uint64_t *)(base + o_b0), *b1 = (uint64_t *)(base + o_b1);
    uint64_t *skeys = (uint64_t *)(base + o_sk);
    uint32_t *spay = (uint32_t *)(base + o_sp);
    const uint64_t src_rows = std::min(n_pad, m_pad);
    uint64_t *d0 = nullptr, *d1 = nullptr;
    // left copy: one word (target << 32 | row)
    {
        Elem<1> *tl = (Elem<1> *)(base + o_tl), *tc = (Elem<1> *)(base + o_tc);
        distribute<1, TS>(ctx, sp.cl, nullptr, src_rows, sp.cnt_l, m_pad, a0, nullptr, b0, nullptr, tl, tc, &d0,
                          &d1);
        launch(ctx, LaunchInfo{"join_pairs_fill_left", dim3(nt), dim3(PB), 0, hash_mix(hash_mix(1, m_pad), out_cap)},
               tr ? k_jp_fill_rows<1, 0, TS, true> : k_jp_fill_rows<1, 0, TS, false>, (const uint64_t *)d0,
               (const uint64_t *)nullptr, (const Elem<1> *)tc, n_out, out_cap, a.out_left, (uint64_t *)nullptr,
               (uint32_t *)nullptr, trace_slot(ctx));
    }
    // right copy: (target << 32 | row, (B + b) | m1 << 32)
    {
        Elem<2> *tl = (Elem<2> *)(base + o_tl), *tc = (Elem<2> *)(base + o_tc);
        distribute<2, TS>(ctx, sp.cr0, sp.cr1, src_rows, sp.cnt_r, m_pad, a0, a1, b0, b1, tl, tc, &d0, &d1);
        launch(ctx, LaunchInfo{"join_pairs_fill_right", dim3(nt), dim3(PB), 0, hash_mix(2, m_pad)},
               tr ? k_jp_fill_rows<2, 1, TS, true> : k_jp_fill_rows<2, 1, TS, false>, (const uint64_t *)d0,
               (const uint64_t *)d1, (const Elem<2> *)tc, n_out, out_cap, (int64_t *)nullptr, skeys, spay,
               trace_slot(ctx));
    }
    sort_kp_full(ctx, skeys, spay, m_pad);
    launch(ctx, LaunchInfo{"join_pairs_out", grid_for(out_cap), dim3(256), 0, hash_mix(m_pad, out_cap)},
           tr ? k_jp_out<true> : k_jp_out<false>, (const uint32_t *)spay, m_pad, n_out, out_cap, a.out_left,
           a.out_right, trace_slot(ctx));
}

} // namespace

void join_pairs_state_free(vo_ctx *ctx)
{
    if (!ctx->join_pairs_state) return;
    JoinState *s = (JoinState *)ctx->join_pairs_state;
    if (s->base) (void)hipFree(s->base);
    if (s->n_out_host) (void)hipHostFree(s->n_out_host);
    delete s;
    ctx->join_pairs_state = nullptr;
}

uint64_t join_pairs_last_count(vo_ctx *ctx) { return ctx->join_pairs_state ? *state_of(ctx).n_out_host : 0; }

uint64_t join_pairs_state_bytes(uint64_t n_left, uint64_t n_right)
{
    return state_layout(union_pad(n_left, n_right)).bytes;
}

uint64_t join_pairs_count_offset(uint64_t n_left, uint64_t n_right)
{
    return state_layout(union_pad(n_left, n_right)).o_nout;
}

uint64_t op_join_pairs_count(vo_ctx *ctx, const vo_join_pairs_args &a)
{
    check_keys(a);
    const uint64_t n_pad = union_pad(a.n_left, a.n_right);
    const uint32_t nt = (uint32_t)(n_pad / PT);
    const StateLayout sl = state_layout(n_pad);
    JoinState *st = nullptr;
    char *sbase;
    if (a.state) {
        // caller-owned state (MergeJoinBlock buffers, commands.hpp:82-98)
        sbase = (char *)a.state;
    } else {
        st = &state_of(ctx);
        st->valid = false;
        if (!st->n_out_host) VO_HIP(hipHostMalloc((void **)&st->n_out_host, 8, hipHostMallocCoherent | hipHostMallocMapped));
        if (st->bytes < sl.bytes) {
            quiesce(ctx);
            if (st->base) VO_HIP(hipFree(st->base));
            st->base = nullptr;
            st->bytes = 0;
            VO_HIP(hipMalloc((void **)&st->base, sl.bytes));
            st->bytes = sl.bytes;
        }
        sbase = st->base;
    }
    const StatePtrs sp = state_ptrs(sbase, sl);
    // per-call scratch: sorted union, aux, masks, tile summaries
    Region r;
    const size_t o_k = r.take(n_pad * 8), o_v = r.take(n_pad * 8), o_aux = r.take(n_pad * 4), o_ml = r.take(n_pad),
                 o_mr = r.take(n_pad), o_bt = r.take(nt * sizeof(Min2)), o_bc = r.take(nt * sizeof(Min2)),
                 o_ft = r.take(nt * sizeof(Fwd)), o_fc = r.take(nt * sizeof(Fwd));
    char *base = (char *)scratch(ctx, r.off);
    uint64_t *keys = (uint64_t *)(base + o_k), *vals = (uint64_t *)(base + o_v);
    uint32_t *aux = (uint32_t *)(base + o_aux);
    uint8_t *ml = (uint8_t *)(base + o_ml), *mr = (uint8_t *)(base + o_mr);
    Min2 *bt = (Min2 *)(base + o_bt), *bc = (Min2 *)(base + o_bc);
    Fwd *ft = (Fwd *)(base + o_ft), *fc = (Fwd *)(base + o_fc);
    const bool tr = ctx->tracing;

    // key types up to 32 bits sort the union as 12-byte records (public: the type)
    const bool kp = key_fits_32(a.type_id);
    if (kp) {
        launch(ctx, LaunchInfo{"join_pairs_build", grid_for(n_pad), dim3(256), 0, hash_mix(a.type_id, n_pad)},
               tr ? k_jp_build_kp<true> : k_jp_build_kp<false>, a.left_keys, a.n_left, a.right_keys, a.n_right,
               a.type_id, n_pad, keys, (uint32_t *)vals, trace_slot(ctx));
        sort_kp_full(ctx, keys, (uint32_t *)vals, n_pad);
    } else {
        launch(ctx, LaunchInfo{"join_pairs_build", grid_for(n_pad), dim3(256), 0, hash_mix(a.type_id, n_pad)},
               tr ? k_jp_build<true> : k_jp_build<false>, a.left_keys, a.n_left, a.right_keys, a.n_right, a.type_id,
               n_pad, keys, vals, trace_slot(ctx));
        sort_kv(ctx, keys, vals, n_pad);
    }
    auto pick = [&](auto kt_f, auto kt_t, auto kp_f, auto kp_t) { return kp ? (tr ? kp_t : kp_f) : (tr ? kt_t : kt_f); };
    launch(ctx, LaunchInfo{"join_pairs_bwd_tiles", dim3(nt), dim3(PB), 0, n_pad},
           pick(k_jp_bwd_tiles<false, false>, k_jp_bwd_tiles<true, false>, k_jp_bwd_tiles<false, true>,
                k_jp_bwd_tiles<true, true>),
           (const uint64_t *)keys, (const void *)vals, n_pad, bt, trace_slot(ctx));
    launch(ctx, LaunchInfo{"join_pairs_bwd_carry", dim3(1), dim3(kScanThreads), 0, nt},
           tr ? k_jp_bwd_carry<true> : k_jp_bwd_carry<false>, (const Min2 *)bt, nt, bc, trace_slot(ctx));
    launch(ctx, LaunchInfo{"join_pairs_bwd_rows", dim3(nt), dim3(PB), 0, n_pad},
           pick(k_jp_bwd_rows<false, false>, k_jp_bwd_rows<true, false>, k_jp_bwd_rows<false, true>,
                k_jp_bwd_rows<true, true>),
           (const uint64_t *)keys, (const void *)vals, n_pad, (const Min2 *)bc, aux, ft, trace_slot(ctx));
    launch(ctx, LaunchInfo{"join_pairs_fwd_carry", dim3(1), dim3(kScanThreads), 0, nt},
           tr ? k_jp_fwd_carry<true> : k_jp_fwd_carry<false>, (const Fwd *)ft, nt, fc, trace_slot(ctx));
    launch(ctx, LaunchInfo{"join_pairs_fwd_rows", dim3(nt), dim3(PB), 0, n_pad},
           pick(k_jp_fwd_rows<false, false>, k_jp_fwd_rows<true, false>, k_jp_fwd_rows<false, true>,
                k_jp_fwd_rows<true, true>),
           (const uint64_t *)keys, (const void *)vals, (const uint32_t *)aux, n_pad, (const Fwd *)fc, sp.cl, ml,
           sp.cr0, sp.cr1, mr, sp.n_out_dev, trace_slot(ctx));
    compact_rows1(ctx, sp.cl, ml, n_pad, sp.cnt_l);
    compact_rows2(ctx, sp.cr0, sp.cr1, mr, n_pad, sp.cnt_r);
    if (a.out_count) copy_async(ctx, a.out_count, sp.n_out_dev, 8);
    if (!st) return 0;
    // the context state: the count goes to pinned host memory for the WRITE
    if (ctx->rec) {
        Step s{};
        s.id = kStepCopy;
        s.grid = 1;
        s.slots = 1;
        const CopyArgs ca{st->n_out_host, sp.n_out_dev, 8};
        std::memcpy(s.blob, &ca, sizeof ca);
        ctx->rec->steps.push_back(s);
        st->count_seq = queue_next_seq(ctx);
    } else {
        VO_HIP(hipMemcpyAsync(st->n_out_host, sp.n_out_dev, 8, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipStreamSynchronize(ctx->stream));
        st->count_seq = 0;
    }
    st->lk = a.left_keys;
    st->rk = a.right_keys;
    st->nl = a.n_left;
    st->nr = a.n_right;
    st->type = a.type_id;
    st->n_pad = n_pad;
    st->valid = true;
    return ctx->rec ? 0 : *st->n_out_host;
}

void op_join_pairs_write(vo_ctx *ctx, const vo_join_pairs_args &a)
{
    uint64_t n_out, n_pad;
    StatePtrs sp;
    if (a.state) {
        check_keys(a);
        n_pad = union_pad(a.n_left, a.n_right);
        sp = state_ptrs((char *)a.state, state_layout(n_pad));
        n_out = a.n_out;
        VO_CHECK(n_out <= (uint64_t)a.n_left * a.n_right, VO_ERR_INVALID, "n_out above n_left * n_right");
    } else {
        JoinState &st = state_of(ctx);
        VO_CHECK(st.valid && st.lk == a.left_keys && st.rk == a.right_keys && st.nl == a.n_left &&
                     st.nr == a.n_right && (a.type_id == 0 || st.type == a.type_id),
                 VO_ERR_INVALID, "join WRITE needs the COUNT of the same key columns first");
        // resident queue: the count is known once the COUNT has finished
        if (st.count_seq) queue_wait_seq_locked(ctx, st.count_seq);
        n_out = *st.n_out_host;
        n_pad = st.n_pad;
        sp = state_ptrs(st.base, state_layout(n_pad));
    }
    VO_CHECK(n_out <= (1ull << 29), VO_ERR_UNSUPPORTED, "join output above 2^29 pairs");
    // out_cap 0: the reference's allocation, the power of two >= n_out
    uint64_t out_cap = a.out_cap;
    if (out_cap == 0 && n_out > 0)
        for (out_cap = 1; out_cap < n_out; out_cap <<= 1) {
        }
    VO_CHECK(out_cap >= n_out, VO_ERR_INVALID, "out_cap below the pair count");
    VO_CHECK(out_cap == 0 || (a.out_left && a.out_right), VO_ERR_INVALID, "null output buffer");
    if (dist_tile(ctx) == DT_RES)
        write_pairs<DT_RES>(ctx, a, sp, n_pad, n_out, out_cap);
    else
        write_pairs<DT>(ctx, a, sp, n_pad, n_out, out_cap);
}

} // namespace vo
