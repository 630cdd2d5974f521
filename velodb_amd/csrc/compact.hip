// compact.hip -- host drivers of the oblivious order-preserving compaction
// (ORCompact), the Q6-style filter + exact int128 SUM and OP_SCATTER; the
// device bodies and the method are in compact_kernels.hpp.
#include "compact_kernels.hpp"
#include "resident_tables.hpp"

#include <algorithm>
#include <atomic>
#include <mutex>

// Fixed device scratch of vo_compact_noalloc: the tile counts, the tile
// prefix array and the count of one compaction of at most 2^29 rows
// (1.53 MiB at 2^29; the array keeps 2 MiB).  Each device holds its own copy
// for the life of the process.  Host code never passes the array's name to a
// kernel (in host code it is a stand-in, not a device address): it uses the
// address hipGetSymbolAddress returned on that device (noalloc_scratch_init).
static __device__ __align__(256) unsigned char scratch[2u << 20];

namespace vo {
namespace {

using namespace compact_k;

const std::pair<const void *, BodyInfo> kBodies[] = {VO_COMPACT_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

// --------------------------------------------------------------- drivers

struct Workspace {
    uint32_t *tile_cnt;
    uint64_t *tile_pre;
    uint64_t *count;
    uint8_t *mask;
    __int128 *partial;
};

constexpr uint32_t kSumBlocks = 2048;

// The workspace laid out from `base` (nullptr: only the size); returns its
// bytes.
size_t workspace_at(char *base, uint64_t n_pad, bool want_mask, Workspace *ws)
{
    const uint64_t nt = n_pad / CT;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_cnt = take(nt * 4), o_pre = take((nt + 1) * 8), o_count = take(8),
                 o_part = take(kSumBlocks * sizeof(__int128)), o_mask = take(want_mask ? n_pad : 0);
    if (ws)
        *ws = Workspace{(uint32_t *)(base + o_cnt), (uint64_t *)(base + o_pre), (uint64_t *)(base + o_count),
                        want_mask ? (uint8_t *)(base + o_mask) : nullptr, (__int128 *)(base + o_part)};
    return off;
}

Workspace workspace(vo_ctx *ctx, uint64_t n_pad, bool want_mask)
{
    Workspace ws;
    char *base = (char *)scratch(ctx, workspace_at(nullptr, n_pad, want_mask, nullptr));
    workspace_at(base, n_pad, want_mask, &ws);
    return ws;
}

// Exclusive scan of nt tile counts into pre[0..nt] (pre[nt] and *out_count:
// the total).  One 1024-thread block up to kScanSeg tiles and in resident
// programs; above, in launch mode, the three-pass scan (one block's
// latency-bound walk took 201 us over the 131072 tiles of 2^29 rows).  tmp:
// 2 KB of device memory for the segment sums and bases.
void scan_tiles(vo_ctx *ctx, const uint32_t *cnt, uint32_t nt, uint64_t *pre, uint64_t *out_count, void *tmp)
{
    if (nt <= kScanSeg || ctx->rec) {
        launch(ctx, LaunchInfo{"compact_scan", dim3(1), dim3(1024), 0, nt}, k_scan_tiles, cnt, nt, pre, out_count,
               trace_slot(ctx));
        return;
    }
    const uint32_t nb = (nt + kScanSeg - 1) / kScanSeg;
    VO_CHECK(nb <= 128, VO_ERR_INVALID, "tile scan above 2^32 rows");
    uint32_t *seg_sum = (uint32_t *)tmp;
    uint64_t *seg_base = (uint64_t *)((char *)tmp + 1024);
    launch(ctx, LaunchInfo{"compact_scan_part", dim3(nb), dim3(1024), 0, nt}, k_scan_part, cnt, nt, seg_sum,
           trace_slot(ctx));
    launch(ctx, LaunchInfo{"compact_scan", dim3(1), dim3(1024), 0, nb}, k_scan_tiles, (const uint32_t *)seg_sum, nb,
           seg_base, out_count, trace_slot(ctx));
    launch(ctx, LaunchInfo{"compact_scan_fin", dim3(nb), dim3(1024), 0, nt}, k_scan_fin, cnt, nt,
           (const uint64_t *)seg_base, pre, trace_slot(ctx));
}

// The network's passes with one tracing mode (TR) and cache policy (NT).
template <int NW, bool TR, bool NT>
void run_network_t(vo_ctx *ctx, Cols<NW> src, Cols<NW> cols, const uint8_t *mask, uint64_t n,
                   const uint64_t *tile_pre, Inst in, uint64_t rows, int n_log, bool w8)
{
    const uint32_t nt = (uint32_t)(rows / CT);
    const size_t lds = (size_t)CTPAD * 8 + (CT + 1) * 2;
    auto kl = w8 ? k_compact_local<NW, true, TR, NT> : k_compact_local<NW, false, TR, NT>;
    VO_HIP(hipFuncSetAttribute((const void *)kl, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    launch(ctx, LaunchInfo{"compact_local", dim3(nt), dim3(CBLOCK), (uint32_t)lds, hash_mix(NW, n_log)}, kl, src,
           cols, mask, n, tile_pre, in, trace_slot(ctx));
    // levels CT_LOG .. n_log-1: passes of at most 8 levels, as even as
    // possible; 5..8 levels run LDS-staged, fewer in registers
    const int count = n_log - CT_LOG;
    const int groups = (count + 7) / 8;
    int q = CT_LOG;
    for (int gi = 0; gi < groups; ++gi) {
        const int s = count / groups + (gi < count % groups ? 1 : 0);
        if (s >= 5) {
            const size_t hl = (size_t)CTPAD * 8 + 257 * 8;
            LaunchInfo li{"compact_hybrid", dim3(nt), dim3(CBLOCK), (uint32_t)hl, hash_mix(hash_mix(NW, q), s + 100)};
            auto pick = [&](auto k5, auto k6, auto k7, auto k8) {
                auto k = s == 5 ? k5 : s == 6 ? k6 : s == 7 ? k7 : k8;
                VO_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)hl));
                launch(ctx, li, k, cols, tile_pre, q, in, trace_slot(ctx));
            };
            if (w8)
                pick(k_compact_hybrid<NW, 5, true, TR, NT>, k_compact_hybrid<NW, 6, true, TR, NT>,
                     k_compact_hybrid<NW, 7, true, TR, NT>, k_compact_hybrid<NW, 8, true, TR, NT>);
            else
                pick(k_compact_hybrid<NW, 5, false, TR, NT>, k_compact_hybrid<NW, 6, false, TR, NT>,
                     k_compact_hybrid<NW, 7, false, TR, NT>, k_compact_hybrid<NW, 8, false, TR, NT>);
            q += s;
            continue;
        }
        const uint32_t blocks = (uint32_t)((rows >> s) / 256);
        LaunchInfo li{"compact_global", dim3(blocks), dim3(256), 0, hash_mix(hash_mix(NW, q), s)};
        switch (s) {
        case 1: launch(ctx, li, k_compact_global<NW, 1, TR, NT>, cols, tile_pre, q, in, trace_slot(ctx)); break;
        case 2: launch(ctx, li, k_compact_global<NW, 2, TR, NT>, cols, tile_pre, q, in, trace_slot(ctx)); break;
        case 3: launch(ctx, li, k_compact_global<NW, 3, TR, NT>, cols, tile_pre, q, in, trace_slot(ctx)); break;
        case 4: launch(ctx, li, k_compact_global<NW, 4, TR, NT>, cols, tile_pre, q, in, trace_slot(ctx)); break;
        default:
            if constexpr (NW == 1) {
                launch(ctx, li, k_compact_global<NW, 5, TR, NT>, cols, tile_pre, q, in, trace_slot(ctx));
                break;
            }
            throw Error{VO_ERR_INVALID, "compaction pass width"};
        }
        q += s;
    }
}

// The network's passes.  One compaction: n_pad rows, levels below
// log2(n_pad).  A batch class (in.off set): `rows` virtual rows of instances
// of 2^in.c rows each, levels below in.c, every pass over all of them.
// Streaming (non-temporal) row accesses when the rows a pass moves exceed
// the 256 MB MALL: filter + SUM at 3e8 rows / 2^29 took 16.2 ms against
// 17.0; at 2^23 (134 MB, held by the MALL) they cost 5%.  Resident programs
// keep the default (their bodies are built with it); tuning bit 0x800 flips
// the choice.
template <int NW>
void run_network(vo_ctx *ctx, Cols<NW> src, Cols<NW> cols, const uint8_t *mask, uint64_t n, uint64_t n_pad,
                 const uint64_t *tile_pre, Inst in = Inst{nullptr, nullptr, 0}, uint64_t rows = 0)
{
    const int n_log = in.off ? in.c : ilog2(n_pad);
    if (!in.off) rows = n_pad;
    bool w8 = true;
    uint64_t row_bytes = 0;
    for (int w = 0; w < NW; ++w) {
        w8 = w8 && src.c[w].w == 8 && cols.c[w].w == 8;
        row_bytes += cols.c[w].w;
    }
    constexpr uint64_t kMallBytes = 256ull << 20;
    if (ctx->tracing)
        run_network_t<NW, true, false>(ctx, src, cols, mask, n, tile_pre, in, rows, n_log, w8);
    else if (!ctx->rec && ((rows * row_bytes > kMallBytes) != (ctx->tuning.flip_streaming != 0)))
        run_network_t<NW, false, true>(ctx, src, cols, mask, n, tile_pre, in, rows, n_log, w8);
    else
        run_network_t<NW, false, false>(ctx, src, cols, mask, n, tile_pre, in, rows, n_log, w8);
}

// Count + scan for a mask already in device memory.
void count_and_scan(vo_ctx *ctx, const uint8_t *mask, uint64_t n, uint64_t n_pad, const Workspace &ws,
                    uint64_t *out_count)
{
    const uint32_t nt = (uint32_t)(n_pad / CT);
    launch(ctx, LaunchInfo{"compact_count", dim3(nt), dim3(CBLOCK), 0, n_pad}, k_count_mask, mask, n, ws.tile_cnt,
           Inst{nullptr, nullptr, 0}, trace_slot(ctx));
    // the sum partials' region is free until a sum pass: the scan's temporaries
    scan_tiles(ctx, ws.tile_cnt, nt, ws.tile_pre, out_count ? out_count : ws.count, ws.partial);
}

void check_pad(uint64_t n, uint64_t n_pad)
{
    VO_CHECK(is_pow2(n_pad) && n <= n_pad, VO_ERR_INVALID, "n_pad must be a power of two >= n");
    VO_CHECK(n_pad >= (uint64_t)CT, VO_ERR_INVALID, "compaction needs n_pad >= 4096 (pad further)");
    VO_CHECK(n_pad <= (1ull << 32), VO_ERR_INVALID, "compaction n_pad above 2^32");
}

} // namespace

// Compaction of two 8-byte columns by a device mask of n_pad rows (all rows
// live); used by the join and group-by.  Leaves the count in *count_dev.
void compact_rows2(vo_ctx *ctx, uint64_t *c0, uint64_t *c1, const uint8_t *mask, uint64_t n_pad, uint64_t *count_dev)
{
    check_pad(n_pad, n_pad);
    // the workspace shares the context arena with the caller's buffers, so
    // the caller passes its own regions; here only tile arrays are needed
    const uint32_t nt = (uint32_t)(n_pad / CT);
    uint32_t *tile_cnt = (uint32_t *)(count_dev + 1);
    uint64_t *tile_pre = (uint64_t *)(((uintptr_t)(tile_cnt + nt) + 255) & ~(uintptr_t)255);
    launch(ctx, LaunchInfo{"compact_count", dim3(nt), dim3(CBLOCK), 0, n_pad}, k_count_mask, mask, n_pad, tile_cnt,
           Inst{nullptr, nullptr, 0}, trace_slot(ctx));
    launch(ctx, LaunchInfo{"compact_scan", dim3(1), dim3(1024), 0, nt}, k_scan_tiles, (const uint32_t *)tile_cnt, nt,
           tile_pre, count_dev, trace_slot(ctx));
    Cols<2> cs{{{c0, 8}, {c1, 8}}};
    run_network<2>(ctx, cs, cs, mask, n_pad, n_pad, tile_pre);
}

// The same for one 8-byte column.
void compact_rows1(vo_ctx *ctx, uint64_t *c0, const uint8_t *mask, uint64_t n_pad, uint64_t *count_dev)
{
    check_pad(n_pad, n_pad);
    const uint32_t nt = (uint32_t)(n_pad / CT);
    uint32_t *tile_cnt = (uint32_t *)(count_dev + 1);
    uint64_t *tile_pre = (uint64_t *)(((uintptr_t)(tile_cnt + nt) + 255) & ~(uintptr_t)255);
    launch(ctx, LaunchInfo{"compact_count", dim3(nt), dim3(CBLOCK), 0, n_pad}, k_count_mask, mask, n_pad, tile_cnt,
           Inst{nullptr, nullptr, 0}, trace_slot(ctx));
    launch(ctx, LaunchInfo{"compact_scan", dim3(1), dim3(1024), 0, nt}, k_scan_tiles, (const uint32_t *)tile_cnt, nt,
           tile_pre, count_dev, trace_slot(ctx));
    Cols<1> cs{{{c0, 8}}};
    run_network<1>(ctx, cs, cs, mask, n_pad, n_pad, tile_pre);
}

// Bytes of device memory compact_rows1/2 need after count_dev.
size_t compact_rows2_scratch(uint64_t n_pad)
{
    const uint64_t nt = n_pad / CT;
    return 8 + nt * 4 + 256 + (nt + 1) * 8 + 256;
}

namespace {

// The compaction's passes over a workspace the caller provides.
void compact_with(vo_ctx *ctx, const vo_compact_args &a, const Workspace &ws)
{
    VO_CHECK(a.mask && a.n_cols >= 1 && a.n_cols <= VO_MAX_SORT_COLUMNS, VO_ERR_INVALID, "bad compaction columns");
    for (uint32_t c = 0; c < a.n_cols; ++c) {
        const uint32_t w = a.col_bytes[c];
        VO_CHECK(a.cols[c] && (w == 1 || w == 2 || w == 4 || w == 8), VO_ERR_INVALID, "bad column");
    }
    count_and_scan(ctx, a.mask, a.n, a.n_pad, ws, a.out_count);
    // columns in pairs of two words per network run; the network depends on
    // the mask only, so each run applies the same routing
    uint32_t c = 0;
    while (c < a.n_cols) {
        const uint32_t take = std::min<uint32_t>(2, a.n_cols - c);
        if (take == 2) {
            Cols<2> cs{{{a.cols[c], a.col_bytes[c]}, {a.cols[c + 1], a.col_bytes[c + 1]}}};
            run_network<2>(ctx, cs, cs, a.mask, a.n, a.n_pad, ws.tile_pre);
        } else {
            Cols<1> cs{{{a.cols[c], a.col_bytes[c]}}};
            run_network<1>(ctx, cs, cs, a.mask, a.n, a.n_pad, ws.tile_pre);
        }
        c += take;
    }
}

} // namespace

void op_compact(vo_ctx *ctx, const vo_compact_args &a)
{
    check_pad(a.n, a.n_pad);
    compact_with(ctx, a, workspace(ctx, a.n_pad, false));
}

// Independent compactions bucketed by power-of-two size class.  A class's
// items run as one compaction over its virtual rows (item k of the class at
// virtual rows [k 2^c, (k+1) 2^c), physical rows from its offset): one count
// pass, one scan block per item (each item's prefix counts restart at zero),
// then the network's passes for levels below c.  A level below c never pairs
// rows of two items, and each item sees the prefix counts it would see
// alone, so its rows end where vo_compact on it alone leaves them.  Classes
// are independent pieces of the command (in the resident queue they share
// the lane between barriers).
void op_compact_batch(vo_ctx *ctx, const vo_compact_batch_args &a)
{
    VO_CHECK(a.mask && (a.items || a.n_items == 0), VO_ERR_INVALID, "null buffer");
    VO_CHECK(a.n_cols >= 1 && a.n_cols <= VO_MAX_SORT_COLUMNS, VO_ERR_INVALID, "bad compaction columns");
    for (uint32_t c = 0; c < a.n_cols; ++c) {
        const uint32_t w = a.col_bytes[c];
        VO_CHECK(a.cols[c] && (w == 1 || w == 2 || w == 4 || w == 8), VO_ERR_INVALID, "bad column");
    }
    std::vector<std::vector<uint64_t>> off(33), ns(33), idx(33);
    for (uint64_t i = 0; i < a.n_items; ++i) {
        const vo_batch_item &it = a.items[i];
        VO_CHECK(is_pow2(it.n_pad) && it.n_pad >= (uint64_t)CT && it.n_pad <= (1ull << 32) && it.n <= it.n_pad &&
                     it.offset % it.n_pad == 0,
                 VO_ERR_INVALID,
                 "compaction batch item: n_pad must be a power of two in [4096, 2^32], >= n, dividing offset");
        const int c = ilog2(it.n_pad);
        off[c].push_back(it.offset);
        ns[c].push_back(it.n);
        idx[c].push_back(i);
    }
    std::vector<int> classes;
    std::vector<uint64_t> host;
    std::vector<size_t> at(33, 0), cnt_at(33, 0), pre_at(33, 0);
    for (int c = 32; c >= CT_LOG; --c) {
        if (off[c].empty()) continue;
        VO_CHECK(off[c].size() <= ((1ull << 32) >> c), VO_ERR_INVALID,
                 "compaction batch: the items of one size class exceed 2^32 rows");
        classes.push_back(c);
        at[c] = host.size();
        host.insert(host.end(), off[c].begin(), off[c].end());
        host.insert(host.end(), ns[c].begin(), ns[c].end());
        host.insert(host.end(), idx[c].begin(), idx[c].end());
    }
    if (classes.empty()) return;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t bytes = al(host.size() * 8);
    for (int c : classes) {
        const uint64_t tiles = (uint64_t)off[c].size() << (c - CT_LOG);
        cnt_at[c] = bytes;
        bytes += al(tiles * 4);
        pre_at[c] = bytes;
        bytes += al((tiles + 1) * 8); // the hybrid pass reads one prefix past its group
    }
    char *base = (char *)scratch(ctx, bytes);
    uint64_t *dtab = (uint64_t *)base;
    upload_async(ctx, dtab, host.data(), host.size() * 8);
    run_tracks(ctx, classes.size(), [&](size_t t) {
        const int c = classes[t];
        const uint64_t k = off[c].size();
        const Inst in{dtab + at[c], dtab + at[c] + k, c};
        const uint64_t *ix = dtab + at[c] + 2 * k;
        const uint32_t nts = 1u << (c - CT_LOG);
        const uint64_t tiles = k * nts;
        uint32_t *tile_cnt = (uint32_t *)(base + cnt_at[c]);
        uint64_t *tile_pre = (uint64_t *)(base + pre_at[c]);
        launch(ctx, LaunchInfo{"compact_count", dim3((uint32_t)tiles), dim3(CBLOCK), 0, hash_mix(k, c)}, k_count_mask,
               a.mask, (uint64_t)0, tile_cnt, in, trace_slot(ctx));
        launch(ctx, LaunchInfo{"compact_scan_seg", dim3((uint32_t)k), dim3(CBLOCK), 0, hash_mix(k, c)}, k_scan_seg,
               (const uint32_t *)tile_cnt, nts, tile_pre, a.out_counts, ix, trace_slot(ctx));
        const uint64_t rows = k << c;
        uint32_t j = 0;
        while (j < a.n_cols) {
            const uint32_t take = std::min<uint32_t>(2, a.n_cols - j);
            if (take == 2) {
                Cols<2> cs{{{a.cols[j], a.col_bytes[j]}, {a.cols[j + 1], a.col_bytes[j + 1]}}};
                run_network<2>(ctx, cs, cs, a.mask, 0, 0, tile_pre, in, rows);
            } else {
                Cols<1> cs{{{a.cols[j], a.col_bytes[j]}}};
                run_network<1>(ctx, cs, cs, a.mask, 0, 0, tile_pre, in, rows);
            }
            j += take;
        }
    });
    // the host table must outlive its upload (recorded: it travels with the program)
    if (!ctx->rec) VO_HIP(hipStreamSynchronize(ctx->stream));
}

namespace {

constexpr uint64_t kNoAllocMaxRows = 1ull << 29;
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
std::atomic<char *> g_noalloc_ptr[kMaxDevices];

} // namespace

// Called by vo_ctx_create with `device` current: looks up this device's copy
// of the fixed scratch once and keeps its address.
void noalloc_scratch_init(int device)
{
    VO_CHECK(device >= 0 && device < kMaxDevices, VO_ERR_INVALID, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device].load(std::memory_order_acquire)) return;
    void *p = nullptr;
    VO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    VO_CHECK(p != nullptr, VO_ERR_HIP, "hipGetSymbolAddress returned no address for the fixed scratch");
    g_noalloc_ptr[device].store((char *)p, std::memory_order_release);
}

// op_compact over the fixed scratch instead of the context arena: nothing is
// allocated.  Every array in it is written by the count and scan passes
// before it is read, so nothing needs clearing between calls.
void op_compact_noalloc(vo_ctx *ctx, const vo_compact_args &a)
{
    check_pad(a.n, a.n_pad);
    VO_CHECK(a.n_pad <= kNoAllocMaxRows, VO_ERR_INVALID,
             "vo_compact_noalloc: n_pad above 2^29 (its fixed scratch holds the tile arrays of at most 2^29 rows; "
             "use vo_compact)");
    VO_CHECK(!ctx->tracing && !(ctx->flags & VO_CFG_KERNEL_TIMING), VO_ERR_INVALID,
             "vo_compact_noalloc: tracing and kernel timing allocate; turn them off for this entry point");
    VO_CHECK(ctx->device >= 0 && ctx->device < kMaxDevices, VO_ERR_INVALID, "device ordinal above 63");
    char *base = g_noalloc_ptr[ctx->device].load(std::memory_order_acquire);
    VO_CHECK(base != nullptr, VO_ERR_INVALID, "vo_compact_noalloc: fixed scratch not set up on this device");
    VO_CHECK(workspace_at(nullptr, a.n_pad, false, nullptr) <= sizeof(::scratch), VO_ERR_INVALID,
             "vo_compact_noalloc: workspace exceeds the fixed scratch");
    Workspace ws;
    workspace_at(base, a.n_pad, false, &ws);
    compact_with(ctx, a, ws);
}

namespace {

void check_filter(const vo_filter_args &a)
{
    check_pad(a.n, a.n_pad);
    VO_CHECK(a.day && a.qty && a.price && a.disc && a.out, VO_ERR_INVALID, "null buffer");
    VO_CHECK((a.out_price == nullptr) == (a.out_disc == nullptr), VO_ERR_INVALID,
             "out_price and out_disc are given together");
}

// The filter's passes over a workspace the caller provides.
void filter_sum_with(vo_ctx *ctx, const vo_filter_args &a, const Workspace &ws);

} // namespace

void op_filter_sum(vo_ctx *ctx, const vo_filter_args &a)
{
    check_filter(a);
    filter_sum_with(ctx, a, workspace(ctx, a.n_pad, true));
}

namespace {

void filter_sum_with(vo_ctx *ctx, const vo_filter_args &a, const Workspace &ws)
{
    const bool in_place_copy = a.out_price != nullptr;
    const uint32_t nt = (uint32_t)(a.n_pad / CT);
    Pred pr{a.day_lo, a.day_hi, a.disc_lo, a.disc_hi, a.qty_lt};
    // the count pass reads 20 B per row, the sum pass 16: streaming loads
    // beyond the MALL, as in the network passes
    const bool stream = !ctx->rec && !ctx->tracing &&
                        ((a.n_pad * 16 > (256ull << 20)) != (ctx->tuning.flip_streaming != 0));
    launch(ctx, LaunchInfo{"filter_count", dim3(nt), dim3(CBLOCK), 0, a.n_pad},
           stream ? k_filter_count<true> : k_filter_count<false>, a.day, a.qty, a.disc, a.n, pr, ws.mask,
           ws.tile_cnt, trace_slot(ctx));
    scan_tiles(ctx, ws.tile_cnt, nt, ws.tile_pre, ws.count, ws.partial); // partials are written after the scan
    // the network's first pass reads the caller's price/disc and writes the
    // output columns (or the inputs in place when no outputs are given)
    int64_t *op = in_place_copy ? a.out_price : const_cast<int64_t *>(a.price);
    int64_t *od = in_place_copy ? a.out_disc : const_cast<int64_t *>(a.disc);
    Cols<2> in{{{const_cast<int64_t *>(a.price), 8}, {const_cast<int64_t *>(a.disc), 8}}};
    Cols<2> cs{{{op, 8}, {od, 8}}};
    run_network<2>(ctx, in, cs, ws.mask, a.n, a.n_pad, ws.tile_pre);
    // at least 8 rows per thread, at most kSumBlocks partial sums
    const uint32_t sb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(kSumBlocks, a.n_pad / 2048));
    launch(ctx, LaunchInfo{"filter_sum", dim3(sb), dim3(256), 0, a.n_pad},
           stream ? k_sum_prod_i128<true> : k_sum_prod_i128<false>, (const int64_t *)op,
           (const int64_t *)od, a.n_pad, (const uint64_t *)ws.count, ws.partial, trace_slot(ctx));
    launch(ctx, LaunchInfo{"filter_sum_final", dim3(1), dim3(256), 0, 0}, k_sum_final, (const __int128 *)ws.partial,
           sb, (const uint64_t *)ws.count, a.out, trace_slot(ctx));
}

} // namespace

void op_scatter(vo_ctx *ctx, const vo_scatter_args &a)
{
    VO_CHECK(a.in_mask && a.out_indices, VO_ERR_INVALID, "null buffer");
    if (a.n == 0) {
        if (a.out_count) fill_async(ctx, a.out_count, 0, 8);
        return;
    }
    // pad the count/scan grid to whole tiles; rows >= n are never read
    const uint64_t n_pad = ((a.n + CT - 1) / CT) * CT;
    Workspace ws = workspace(ctx, n_pad, true);
    const uint32_t nt = (uint32_t)(n_pad / CT);
    fill_async(ctx, ws.mask, 0, n_pad);
    copy_async(ctx, ws.mask, a.in_mask, a.n);
    count_and_scan(ctx, ws.mask, a.n, n_pad, ws, a.out_count);
    launch(ctx, LaunchInfo{"scatter_positions", dim3(nt), dim3(CBLOCK), 0, n_pad}, k_scatter_positions,
           (const uint8_t *)ws.mask, a.n, (const uint64_t *)ws.tile_pre, a.out_indices, trace_slot(ctx));
}

} // namespace vo
