// join.hip -- oblivious primary-key / foreign-key join + GROUP BY SUM.
//
// Reference method: the union -> sort -> expand -> compact join of
// obliviator/join/enclave/scalable_oblivious_join.c:159-317, specialised to
// a PK/FK join (every fact row matches at most one dimension row, so the
// m0 x m1 expansion of :211-224 / oblivious_distribute.c is a forward fill),
// followed by the sort + segmented-scan GROUP BY of
// obliviator/tpch_q3/our_3/enclave/scalable_oblivious_join.c:257-321
// (one row per group is kept by the compaction network instead of a second
// sort).  Sums are exact int64.
//
//   1. union of n_dim + n_fact rows padded to n_pad:
//        key = (key32 << 32) | (is_fact << 31) | row,  val = group | value
//   2. oblivious sort of (key, val)
//   3. forward fill of the last dimension row's (key, group); a fact row is
//      matched when its key equals it
//   4. ORCompact of the matched rows' (group, value)
//   5. sort by (group << 32 | position), rows past the match count get
//      group 0xFFFFFFFF
//   6. segmented inclusive sum by group; the last row of each group selected
//   7. ORCompact -> the first G rows are (group, sum), ascending by group.
// Every pass reads and writes every row; only the group count G and the
// match count leave the device.
#include "vo_internal.hpp"

#include <algorithm>

namespace vo {
namespace {

constexpr int JT = 4096; // scan tile
constexpr int JB = 256;
constexpr int JPER = JT / JB;
constexpr uint64_t PAD_KEY = ~0ull;

__global__ void k_join_build(const int32_t *__restrict__ dk, const int32_t *__restrict__ dg, uint64_t n_dim,
                             const int32_t *__restrict__ fk, const int64_t *__restrict__ fv, uint64_t n_fact,
                             uint64_t n_pad, uint64_t *__restrict__ keys, uint64_t *__restrict__ vals)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k = PAD_KEY, v = 0;
        if (i < n_dim) {
            k = ((uint64_t)(uint32_t)dk[i] << 32) | (i & 0x7FFFFFFFu);
            v = (uint64_t)(uint32_t)dg[i];
        } else if (i < n_dim + n_fact) {
            const uint64_t f = i - n_dim;
            k = ((uint64_t)(uint32_t)fk[f] << 32) | (1ull << 31) | (f & 0x7FFFFFFFu);
            v = (uint64_t)fv[f];
        }
        keys[i] = k;
        vals[i] = v;
    }
}

// ------------------------------------------------- forward fill (step 3)

struct Fill {
    uint64_t key;   // key32 of the last dimension row, or ~0 for none
    uint64_t group;
};

__device__ __forceinline__ bool is_dim(uint64_t k) { return k != PAD_KEY && ((k >> 31) & 1) == 0; }

// Per tile: the last dimension row (key, group) or none.
__global__ __launch_bounds__(JB) void k_fill_tiles(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                   Fill *__restrict__ tile_last)
{
    __shared__ uint64_t wk[JB], wg[JB];
    const uint64_t t0 = (uint64_t)blockIdx.x * JT + (uint64_t)threadIdx.x * JPER;
    uint64_t lk = PAD_KEY, lg = 0;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t key = keys[t0 + k], v = vals[t0 + k];
        const bool d = is_dim(key);
        lk = d ? (key >> 32) : lk;
        lg = d ? v : lg;
    }
    wk[threadIdx.x] = lk;
    wg[threadIdx.x] = lg;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t k = PAD_KEY, g = 0;
        for (int t = 0; t < JB; ++t) {
            const bool has = wk[t] != PAD_KEY;
            k = has ? wk[t] : k;
            g = has ? wg[t] : g;
        }
        tile_last[blockIdx.x] = Fill{k, g};
    }
}

// Exclusive "last dimension row" scan over tiles (one thread, nt entries).
__global__ void k_fill_carry(const Fill *__restrict__ tile_last, uint32_t nt, Fill *__restrict__ carry)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Fill c{PAD_KEY, 0};
    for (uint32_t t = 0; t < nt; ++t) {
        carry[t] = c;
        const Fill f = tile_last[t];
        const bool has = f.key != PAD_KEY;
        c.key = has ? f.key : c.key;
        c.group = has ? f.group : c.group;
    }
}

// Per row: (group, value) columns and the matched mask.
__global__ __launch_bounds__(JB) void k_fill_rows(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                  const Fill *__restrict__ carry, uint64_t *__restrict__ col_g,
                                                  uint64_t *__restrict__ col_v, uint8_t *__restrict__ mask)
{
    __shared__ uint64_t wk[JB], wg[JB];
    const uint64_t t0 = (uint64_t)blockIdx.x * JT + (uint64_t)threadIdx.x * JPER;
    uint64_t lk = PAD_KEY, lg = 0;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t key = keys[t0 + k], v = vals[t0 + k];
        const bool d = is_dim(key);
        lk = d ? (key >> 32) : lk;
        lg = d ? v : lg;
    }
    wk[threadIdx.x] = lk;
    wg[threadIdx.x] = lg;
    __syncthreads();
    // carry into this thread: the last dimension row before its 16 rows
    Fill c = carry[blockIdx.x];
    for (uint32_t t = 0; t < threadIdx.x; ++t) {
        const bool has = wk[t] != PAD_KEY;
        c.key = has ? wk[t] : c.key;
        c.group = has ? wg[t] : c.group;
    }
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t i = t0 + k;
        const uint64_t key = keys[i], v = vals[i];
        const bool d = is_dim(key);
        c.key = d ? (key >> 32) : c.key;
        c.group = d ? v : c.group;
        const bool fact = key != PAD_KEY && ((key >> 31) & 1);
        const bool matched = fact & (c.key == (key >> 32));
        col_g[i] = c.group;
        col_v[i] = v;
        mask[i] = matched ? 1 : 0;
    }
}

// ------------------------------------------- group-by keys (step 5)

__global__ void k_group_keys(const uint64_t *__restrict__ col_g, const uint64_t *__restrict__ col_v,
                             const uint64_t *__restrict__ count, uint64_t n_pad, uint64_t *__restrict__ keys,
                             uint64_t *__restrict__ vals)
{
    const uint64_t m = *count;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t g = i < m ? (col_g[i] & 0xFFFFFFFFull) : 0xFFFFFFFFull;
        keys[i] = (g << 32) | (i & 0xFFFFFFFFull);
        vals[i] = col_v[i];
    }
}

// ------------------------------------- segmented sum by group (step 6)

// Per tile: group of the last row, sum of the tile's trailing run of that
// group, whether the whole tile is that one group.
struct Seg {
    uint64_t last_g;
    int64_t tail_sum;
    uint32_t all_one;
};

__device__ __forceinline__ uint64_t grp(uint64_t k) { return k >> 32; }

__global__ __launch_bounds__(JB) void k_seg_tiles(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                  Seg *__restrict__ tile_seg)
{
    __shared__ uint64_t fg[JB], lg[JB];
    __shared__ int64_t ts[JB];
    __shared__ uint32_t one[JB];
    const uint64_t t0 = (uint64_t)blockIdx.x * JT + (uint64_t)threadIdx.x * JPER;
    const uint64_t g0 = grp(keys[t0]);
    uint64_t gl = g0;
    int64_t run = 0;
    uint32_t uni = 1;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t g = grp(keys[t0 + k]);
        const int64_t v = (int64_t)vals[t0 + k];
        run = (g == gl ? run : 0) + v;
        uni &= (g == g0);
        gl = g;
    }
    fg[threadIdx.x] = g0;
    lg[threadIdx.x] = gl;
    ts[threadIdx.x] = run;
    one[threadIdx.x] = uni;
    __syncthreads();
    if (threadIdx.x == 0) {
        // combine thread summaries left to right
        uint64_t last = lg[0];
        int64_t tail = ts[0];
        uint32_t all = one[0];
        for (int t = 1; t < JB; ++t) {
            const bool cont = one[t] & (fg[t] == last);
            // a thread that does not continue the run as a whole ends in a
            // different group, whose run starts inside it: its own tail
            tail = cont ? tail + ts[t] : ts[t];
            all = all & one[t] & (fg[t] == last);
            last = lg[t];
        }
        tile_seg[blockIdx.x] = Seg{last, tail, all};
    }
}

// Carry into each tile: the sum of the run of the tile's first group that
// ends just before the tile (0 when the previous tile ends another group).
__global__ void k_seg_carry(const Seg *__restrict__ seg, const uint64_t *__restrict__ keys, uint32_t nt,
                            int64_t *__restrict__ carry)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t run = 0;  // running sum of the trailing group up to the end of tile t-1
    uint64_t rg = ~0ull;
    for (uint32_t t = 0; t < nt; ++t) {
        const uint64_t first = grp(keys[(uint64_t)t * JT]);
        carry[t] = (first == rg) ? run : 0;
        const Seg s = seg[t];
        run = (s.all_one & (s.last_g == rg)) ? run + s.tail_sum : s.tail_sum;
        rg = s.last_g;
    }
}

__global__ __launch_bounds__(JB) void k_seg_rows(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                 const int64_t *__restrict__ carry, uint64_t n_pad,
                                                 uint64_t *__restrict__ col_g, uint64_t *__restrict__ col_s,
                                                 uint8_t *__restrict__ mask)
{
    __shared__ uint64_t fg[JB], lg[JB];
    __shared__ int64_t ts[JB];
    __shared__ uint32_t one[JB];
    const uint64_t t0 = (uint64_t)blockIdx.x * JT + (uint64_t)threadIdx.x * JPER;
    const uint64_t g0 = grp(keys[t0]);
    uint64_t gl = g0;
    int64_t run = 0;
    uint32_t uni = 1;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t g = grp(keys[t0 + k]);
        run = (g == gl ? run : 0) + (int64_t)vals[t0 + k];
        uni &= (g == g0);
        gl = g;
    }
    fg[threadIdx.x] = g0;
    lg[threadIdx.x] = gl;
    ts[threadIdx.x] = run;
    one[threadIdx.x] = uni;
    __syncthreads();
    // running sum of the group that is open when this thread's rows start
    const uint64_t tile_first = grp(keys[(uint64_t)blockIdx.x * JT]);
    int64_t c = carry[blockIdx.x];
    uint64_t cg = tile_first;
    for (uint32_t t = 0; t < threadIdx.x; ++t) {
        const bool cont = one[t] & (fg[t] == cg);
        c = cont ? c + ts[t] : ts[t];
        cg = lg[t];
    }
    // rows
    uint64_t prev_g = cg;
    int64_t acc = c;
#pragma unroll
    for (int k = 0; k < JPER; ++k) {
        const uint64_t i = t0 + k;
        const uint64_t g = grp(keys[i]);
        acc = (g == prev_g ? acc : 0) + (int64_t)vals[i];
        prev_g = g;
        const uint64_t gn = (i + 1 < n_pad) ? grp(keys[i + 1]) : ~0ull;
        const bool last = gn != g;
        col_g[i] = g;
        col_s[i] = (uint64_t)acc;
        mask[i] = (last & (g != 0xFFFFFFFFull)) ? 1 : 0;
    }
}

__global__ void k_group_out(const uint64_t *__restrict__ col_g, const uint64_t *__restrict__ col_s,
                            const uint64_t *__restrict__ count, uint64_t max_groups, int32_t *__restrict__ out_g,
                            int64_t *__restrict__ out_s, uint64_t *__restrict__ out_n)
{
    const uint64_t G = *count;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < max_groups; i += (uint64_t)gridDim.x * blockDim.x) {
        const bool live = i < G;
        out_g[i] = live ? (int32_t)col_g[i] : -1;
        out_s[i] = live ? (int64_t)col_s[i] : 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && out_n) *out_n = G;
}

dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65536))); }

} // namespace

void op_join_groupsum(vo_ctx *ctx, const vo_join_args &a)
{
    const uint64_t n_pad = a.n_pad;
    VO_CHECK(is_pow2(n_pad) && a.n_dim + a.n_fact <= n_pad, VO_ERR_INVALID, "n_pad must be a power of two >= n_dim+n_fact");
    VO_CHECK(n_pad >= (uint64_t)JT, VO_ERR_INVALID, "join needs n_pad >= 4096");
    VO_CHECK(a.n_dim < (1ull << 31) && a.n_fact < (1ull << 31), VO_ERR_INVALID, "row counts must be below 2^31");
    VO_CHECK(a.dim_key && a.dim_group && a.fact_fk && a.fact_val && a.out_group && a.out_sum, VO_ERR_INVALID,
             "null buffer");
    const uint32_t nt = (uint32_t)(n_pad / JT);
    // arena: keys, vals, col_a, col_b (8 B each), mask (1 B), tile arrays, compaction scratch
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_k = take(n_pad * 8), o_v = take(n_pad * 8), o_a = take(n_pad * 8), o_b = take(n_pad * 8),
                 o_m = take(n_pad), o_fl = take(nt * sizeof(Fill)), o_fc = take(nt * sizeof(Fill)),
                 o_sg = take(nt * sizeof(Seg)), o_sc = take(nt * 8), o_c1 = take(compact_rows2_scratch(n_pad)),
                 o_c2 = take(compact_rows2_scratch(n_pad));
    char *base = (char *)scratch(ctx, off);
    uint64_t *keys = (uint64_t *)(base + o_k), *vals = (uint64_t *)(base + o_v);
    uint64_t *col_a = (uint64_t *)(base + o_a), *col_b = (uint64_t *)(base + o_b);
    uint8_t *mask = (uint8_t *)(base + o_m);
    Fill *tile_last = (Fill *)(base + o_fl), *fcarry = (Fill *)(base + o_fc);
    Seg *seg = (Seg *)(base + o_sg);
    int64_t *scarry = (int64_t *)(base + o_sc);
    uint64_t *cnt1 = (uint64_t *)(base + o_c1), *cnt2 = (uint64_t *)(base + o_c2);

    // 1-2. union and sort
    launch(ctx, LaunchInfo{"join_build", grid_for(n_pad), dim3(256), 0, n_pad}, k_join_build, a.dim_key, a.dim_group,
           a.n_dim, a.fact_fk, a.fact_val, a.n_fact, n_pad, keys, vals);
    sort_kv(ctx, keys, vals, n_pad);
    if (a.sorted_keys) VO_HIP(hipMemcpyAsync(a.sorted_keys, keys, n_pad * 8, hipMemcpyDeviceToDevice, ctx->stream));
    if (a.sorted_vals) VO_HIP(hipMemcpyAsync(a.sorted_vals, vals, n_pad * 8, hipMemcpyDeviceToDevice, ctx->stream));
    // 3. forward fill
    launch(ctx, LaunchInfo{"join_fill_tiles", dim3(nt), dim3(JB), 0, n_pad}, k_fill_tiles, (const uint64_t *)keys,
           (const uint64_t *)vals, tile_last);
    launch(ctx, LaunchInfo{"join_fill_carry", dim3(1), dim3(64), 0, nt}, k_fill_carry, (const Fill *)tile_last, nt,
           fcarry);
    launch(ctx, LaunchInfo{"join_fill_rows", dim3(nt), dim3(JB), 0, n_pad}, k_fill_rows, (const uint64_t *)keys,
           (const uint64_t *)vals, (const Fill *)fcarry, col_a, col_b, mask);
    // 4. compaction of matched rows
    compact_rows2(ctx, col_a, col_b, mask, n_pad, cnt1);
    // 5. group-by sort
    launch(ctx, LaunchInfo{"group_keys", grid_for(n_pad), dim3(256), 0, n_pad}, k_group_keys, (const uint64_t *)col_a,
           (const uint64_t *)col_b, (const uint64_t *)cnt1, n_pad, keys, vals);
    sort_kv(ctx, keys, vals, n_pad);
    // 6. segmented sum
    launch(ctx, LaunchInfo{"group_seg_tiles", dim3(nt), dim3(JB), 0, n_pad}, k_seg_tiles, (const uint64_t *)keys,
           (const uint64_t *)vals, seg);
    launch(ctx, LaunchInfo{"group_seg_carry", dim3(1), dim3(64), 0, nt}, k_seg_carry, (const Seg *)seg,
           (const uint64_t *)keys, nt, scarry);
    launch(ctx, LaunchInfo{"group_seg_rows", dim3(nt), dim3(JB), 0, n_pad}, k_seg_rows, (const uint64_t *)keys,
           (const uint64_t *)vals, (const int64_t *)scarry, n_pad, col_a, col_b, mask);
    // 7. one row per group
    compact_rows2(ctx, col_a, col_b, mask, n_pad, cnt2);
    launch(ctx, LaunchInfo{"group_out", grid_for(std::max<uint64_t>(a.max_groups, 1)), dim3(256), 0, a.max_groups},
           k_group_out, (const uint64_t *)col_a, (const uint64_t *)col_b, (const uint64_t *)cnt2, a.max_groups,
           a.out_group, a.out_sum, a.out_n_groups);
}

} // namespace vo
