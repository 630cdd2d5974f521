/*
 * oracle.c -- CPU restatement of the reference's oblivious-operator
 * algorithms.  TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the
 * HIP path, never linked into or called by the product library.
 */
#include "oracle.h"

#include <limits.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- gen --- */

#define GOLDEN64 0x9E3779B97F4A7C15ull

uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_rand64(uint64_t seed, uint32_t stream, uint64_t i)
{
    uint64_t s0 = orc_mix64(seed ^ (GOLDEN64 * ((uint64_t)stream + 1)));
    return orc_mix64(s0 + (i + 1) * GOLDEN64);
}

int64_t orc_rand_range(uint64_t seed, uint32_t stream, uint64_t i, int64_t lo, uint64_t range)
{
    uint64_t r32 = orc_rand64(seed, stream, i) >> 32;
    return lo + (int64_t)((r32 * range) >> 32);
}

static inline uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t *hi)
{
    uint64_t p = (uint64_t)a * b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    /* Round multipliers and Weyl key increments from the SC'11 paper. */
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, hi1;
        uint32_t lo0 = mulhilo32(M0, c0, &hi0);
        uint32_t lo1 = mulhilo32(M1, c2, &hi1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += W0; k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint64_t orc_philox_tag(uint64_t seed, uint64_t i)
{
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0u, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return ((uint64_t)out[1] << 32) | out[0];
}

void orc_gen_sort_input(uint64_t seed, size_t n, uint64_t *keys, uint32_t *payload)
{
    for (size_t i = 0; i < n; ++i) {
        keys[i] = orc_rand64(seed, 0, i);
        payload[i] = (uint32_t)i;
    }
}

/* ------------------------------------------------------------ bitonic --- */

static inline void trace_fold(orc_trace *t, int64_t v)
{
    /* FNV-1a over the 8 bytes of v */
    uint64_t h = t->hash;
    for (int b = 0; b < 8; ++b) {
        h ^= (uint64_t)((v >> (8 * b)) & 0xff);
        h *= 0x100000001B3ull;
    }
    t->hash = h;
    t->n_ops++;
}

void orc_bitonic_sort_kp(uint64_t *keys, uint32_t *payload, size_t n, size_t npad, orc_trace *trace)
{
    if (trace) {
        trace->hash = 0xCBF29CE484222325ull;
        trace->n_ops = 0;
        trace->n_stages = 0;
    }
    for (size_t i = n; i < npad; ++i) {
        keys[i] = UINT64_MAX;
        payload[i] = (uint32_t)i;
        if (trace) trace_fold(trace, -1 - (int64_t)i);
    }
    for (size_t k = 2; k <= npad; k <<= 1) {
        for (size_t j = k >> 1; j > 0; j >>= 1) {
            if (trace) trace->n_stages++;
            for (size_t i = 0; i < npad; ++i) {
                size_t ixj = i ^ j;
                if (i < ixj) {
                    int ascending = ((i & k) == 0);
                    uint64_t ka = keys[i], kb = keys[ixj];
                    uint32_t pa = payload[i], pb = payload[ixj];
                    int gt = (ka > kb) || (ka == kb && pa > pb);
                    int lt = (ka < kb) || (ka == kb && pa < pb);
                    int sw = ascending ? gt : lt;
                    /* branch-free select, both slots always written */
                    uint64_t m64 = (uint64_t)0 - (uint64_t)sw;
                    uint32_t m32 = (uint32_t)0 - (uint32_t)sw;
                    keys[i] = (ka & ~m64) | (kb & m64);
                    keys[ixj] = (kb & ~m64) | (ka & m64);
                    payload[i] = (pa & ~m32) | (pb & m32);
                    payload[ixj] = (pb & ~m32) | (pa & m32);
                    if (trace) {
                        trace_fold(trace, (int64_t)i);
                        trace_fold(trace, (int64_t)ixj);
                        trace_fold(trace, -1 - (int64_t)i);
                        trace_fold(trace, -1 - (int64_t)ixj);
                    }
                }
            }
        }
    }
}

void orc_bitonic_sort_kv(uint64_t *keys, uint64_t *vals, size_t npad)
{
    for (size_t k = 2; k <= npad; k <<= 1) {
        for (size_t j = k >> 1; j > 0; j >>= 1) {
            for (size_t i = 0; i < npad; ++i) {
                size_t ixj = i ^ j;
                if (i < ixj) {
                    int ascending = ((i & k) == 0);
                    uint64_t ka = keys[i], kb = keys[ixj];
                    uint64_t va = vals[i], vb = vals[ixj];
                    int gt = (ka > kb) || (ka == kb && va > vb);
                    int lt = (ka < kb) || (ka == kb && va < vb);
                    if (ascending ? gt : lt) {
                        keys[i] = kb; keys[ixj] = ka;
                        vals[i] = vb; vals[ixj] = va;
                    }
                }
            }
        }
    }
}

static size_t type_size(uint32_t t)
{
    switch (t) {
    case ORC_T_BOOLEAN: case ORC_T_TINYINT: return 1;
    case ORC_T_SMALLINT: return 2;
    case ORC_T_INTEGER: case ORC_T_FLOAT: case ORC_T_DATE: return 4;
    case ORC_T_BIGINT: case ORC_T_DOUBLE: case ORC_T_VARCHAR: case ORC_T_UBIGINT: return 8;
    default: return 0;
    }
}

/* compareSingle (sort.cuh:11-29): 0 on equality, -1 if a < b, else 1 */
static int compare_single(const void *col, uint32_t t, size_t a, size_t b)
{
#define CMP(T)                                                   \
    do {                                                         \
        T x = ((const T *)col)[a], y = ((const T *)col)[b];      \
        if (x == y) return 0;                                    \
        return (x < y) ? -1 : 1;                                 \
    } while (0)
    switch (t) {
    case ORC_T_BOOLEAN: CMP(uint8_t);
    case ORC_T_TINYINT: CMP(int8_t);
    case ORC_T_SMALLINT: CMP(int16_t);
    case ORC_T_INTEGER: CMP(int32_t);
    case ORC_T_BIGINT: CMP(int64_t);
    case ORC_T_FLOAT: CMP(float);
    case ORC_T_DOUBLE: CMP(double);
    case ORC_T_VARCHAR: CMP(uint64_t);
    case ORC_T_DATE: CMP(uint32_t);
    case ORC_T_UBIGINT: CMP(uint64_t);
    default: return 0;
    }
#undef CMP
}

static void fill_pad(void *col, uint32_t t, int asc, size_t i)
{
    switch (t) {
    case ORC_T_BOOLEAN: ((uint8_t *)col)[i] = asc ? UINT8_MAX : 0; break;
    case ORC_T_TINYINT: ((int8_t *)col)[i] = asc ? INT8_MAX : INT8_MIN; break;
    case ORC_T_SMALLINT: ((int16_t *)col)[i] = asc ? INT16_MAX : INT16_MIN; break;
    case ORC_T_INTEGER: ((int32_t *)col)[i] = asc ? INT32_MAX : INT32_MIN; break;
    case ORC_T_BIGINT: ((int64_t *)col)[i] = asc ? INT64_MAX : INT64_MIN; break;
    case ORC_T_FLOAT: {
        /* numeric_limits<float>::max() / lowest() */
        float v = asc ? 3.40282347e+38f : -3.40282347e+38f;
        ((float *)col)[i] = v;
        break;
    }
    case ORC_T_DOUBLE: ((double *)col)[i] = asc ? 1.7976931348623157e+308 : -1.7976931348623157e+308; break;
    case ORC_T_VARCHAR: ((uint64_t *)col)[i] = asc ? UINT64_MAX : 0; break;
    case ORC_T_DATE: ((uint32_t *)col)[i] = asc ? UINT32_MAX : 0; break;
    case ORC_T_UBIGINT: ((uint64_t *)col)[i] = asc ? UINT64_MAX : 0; break;
    default: break;
    }
}

static void swap_bytes(void *col, size_t sz, size_t a, size_t b)
{
    unsigned char tmp[16];
    unsigned char *p = (unsigned char *)col;
    memcpy(tmp, p + a * sz, sz);
    memcpy(p + a * sz, p + b * sz, sz);
    memcpy(p + b * sz, tmp, sz);
}

int orc_bitonic_sort_multi(void *const *cols, const uint32_t *types, const uint8_t *asc,
                           size_t ncols, int64_t *indices, size_t n, size_t npad)
{
    if (ncols == 0 || ncols > 8 || n > npad || (npad & (npad - 1)) != 0) return -1;
    for (size_t c = 0; c < ncols; ++c)
        if (type_size(types[c]) == 0) return -1;
    for (size_t i = 0; i < npad; ++i) {
        indices[i] = (int64_t)i;
        if (i >= n)
            for (size_t c = 0; c < ncols; ++c) fill_pad(cols[c], types[c], asc[c], i);
    }
    for (size_t k = 2; k <= npad; k <<= 1) {
        for (size_t j = k >> 1; j > 0; j >>= 1) {
            for (size_t i = 0; i < npad; ++i) {
                size_t ixj = i ^ j;
                if (i < ixj) {
                    int ascending = ((i & k) == 0);
                    int need = 0;
                    for (size_t c = 0; c < ncols; ++c) {
                        int cmp = compare_single(cols[c], types[c], i, ixj);
                        if (cmp != 0) {
                            need = (ascending ^ (asc[c] != 0)) ? (cmp < 0) : (cmp > 0);
                            break;
                        }
                    }
                    if (need) {
                        int64_t t = indices[i];
                        indices[i] = indices[ixj];
                        indices[ixj] = t;
                        for (size_t c = 0; c < ncols; ++c) swap_bytes(cols[c], type_size(types[c]), i, ixj);
                    }
                }
            }
        }
    }
    return 0;
}

void orc_permute(void *out, const void *in, const int64_t *idx, size_t n, size_t elem_size)
{
    unsigned char *o = (unsigned char *)out;
    const unsigned char *s = (const unsigned char *)in;
    for (size_t i = 0; i < n; ++i) memcpy(o + i * elem_size, s + (size_t)idx[i] * elem_size, elem_size);
}

/* ---------------------------------------------------------- ORCompact --- */

/* o_memswap(a, b, size, cond): conditional swap of whole elements */
static void cswap(unsigned char *a, unsigned char *b, size_t sz, int cond)
{
    unsigned char m = (unsigned char)(0 - (cond != 0));
    for (size_t i = 0; i < sz; ++i) {
        unsigned char x = (unsigned char)((a[i] ^ b[i]) & m);
        a[i] ^= x;
        b[i] ^= x;
    }
}

/* largest power of two strictly below N for N > 1 (common/util.c:58-65) */
static size_t pow2_lt(size_t N)
{
    size_t N1 = 1;
    while (N1 < N) N1 <<= 1;
    return N1 >> 1;
}

/* oblivious_compact_2power_inner_elem (oblivious_compact.c:59-106) */
static void compact_2power(unsigned char *buf, size_t N, size_t sz, size_t offset,
                           const uint8_t *sel, const size_t *cnt)
{
    if (N == 1) return;
    if (N == 2) {
        int swap = (((!sel[0]) & sel[1]) != 0) ^ (int)(offset & 1);
        cswap(buf, buf + sz, sz, swap);
        return;
    }
    size_t half = N / 2;
    size_t m1 = cnt[half] - cnt[0];
    size_t offset_mod = offset & (half - 1);
    size_t offset_m1_mod = (offset + m1) & (half - 1);
    int offset_right = offset >= half;
    int left_wrapped = (offset_mod + m1) >= half;
    compact_2power(buf, half, sz, offset_mod, sel, cnt);
    compact_2power(buf + half * sz, half, sz, offset_m1_mod, sel + half, cnt + half);
    int swap_flag = left_wrapped ^ offset_right;
    for (size_t i = 0; i < half; ++i) {
        swap_flag = swap_flag ^ (i == offset_m1_mod);
        cswap(buf + i * sz, buf + (half + i) * sz, sz, swap_flag);
    }
}

/* oblivious_compact_inner_elem (oblivious_compact.c:108-153) */
static void compact_inner(unsigned char *buf, size_t N, size_t sz, const uint8_t *sel, const size_t *cnt)
{
    if (N <= 1) return;
    if (N == 2) {
        cswap(buf, buf + sz, sz, ((!sel[0]) & sel[1]) != 0);
        return;
    }
    size_t gt_pow2 = pow2_lt(N);
    size_t split = N - gt_pow2;
    size_t mL = cnt[split] - cnt[0];
    compact_inner(buf, split, sz, sel, cnt);
    compact_2power(buf + split * sz, gt_pow2, sz, (gt_pow2 - split + mL) % gt_pow2, sel + split, cnt + split);
    unsigned char *L = buf, *R = buf + gt_pow2 * sz;
    for (size_t i = 0; i < split; ++i) cswap(L + i * sz, R + i * sz, sz, i >= mL);
}

void orc_compact(void *buf, size_t elem_size, const uint8_t *selected, size_t n)
{
    size_t *cnt = (size_t *)calloc(n + 1, sizeof(size_t));
    for (size_t i = 0; i < n; ++i) cnt[i + 1] = cnt[i] + (selected[i] != 0);
    uint8_t *sel = (uint8_t *)malloc(n ? n : 1);
    for (size_t i = 0; i < n; ++i) sel[i] = selected[i] != 0;
    compact_inner((unsigned char *)buf, n, elem_size, sel, cnt);
    free(sel);
    free(cnt);
}

void orc_compact_levels(void *buf, size_t elem_size, const uint8_t *selected, size_t n)
{
    size_t *C = (size_t *)calloc(n + 1, sizeof(size_t));
    for (size_t i = 0; i < n; ++i) C[i + 1] = C[i] + (selected[i] != 0);
    unsigned char *p = (unsigned char *)buf;
    for (size_t s = 2; s <= n; s <<= 1) {
        size_t half = s / 2;
        for (size_t b = 0; b < n; b += s) {
            size_t o = C[b] & (s - 1);
            if (s == 2) {
                int sw = ((selected[b] == 0) & (selected[b + 1] != 0)) ^ (int)(o & 1);
                cswap(p + b * elem_size, p + (b + 1) * elem_size, elem_size, sw);
                continue;
            }
            size_t m1 = C[b + half] - C[b];
            size_t om = o & (half - 1);
            size_t om1 = (o + m1) & (half - 1);
            int base = ((om + m1) >= half) ^ (o >= half);
            for (size_t i = 0; i < half; ++i)
                cswap(p + (b + i) * elem_size, p + (b + half + i) * elem_size, elem_size, base ^ (i >= om1));
        }
    }
    free(C);
}

/* ------------------------------------------------------ filter + sum --- */

void orc_gen_fact(uint64_t seed, size_t n, int32_t *day, int64_t *qty, int64_t *price, int64_t *disc)
{
    for (size_t i = 0; i < n; ++i) {
        day[i] = (int32_t)orc_rand_range(seed, 1, i, 8036, 10561 - 8036 + 1);
        qty[i] = orc_rand_range(seed, 2, i, 1, 50);
        price[i] = orc_rand_range(seed, 3, i, 90000, 10495000 - 90000 + 1);
        disc[i] = orc_rand_range(seed, 4, i, 0, 11);
    }
}

void orc_filter_mask(const int32_t *day, const int64_t *qty, const int64_t *disc,
                     size_t n, size_t npad, const orc_filter_spec *spec, uint8_t *mask)
{
    for (size_t i = 0; i < npad; ++i) {
        int live = i < n;
        size_t r = live ? i : 0;
        uint8_t m = (uint8_t)((day[r] >= spec->day_lo) & (day[r] < spec->day_hi) &
                              (disc[r] >= spec->disc_lo) & (disc[r] <= spec->disc_hi) &
                              (qty[r] < spec->qty_lt) & live);
        mask[i] = m;
    }
}

size_t orc_filter_sum(const int32_t *day, const int64_t *qty, const int64_t *price, const int64_t *disc,
                      size_t n, size_t npad, const orc_filter_spec *spec,
                      int64_t *out_price, int64_t *out_disc, uint64_t *sum_lo, int64_t *sum_hi)
{
    uint8_t *mask = (uint8_t *)malloc(npad);
    int64_t *pairs = (int64_t *)malloc(npad * 2 * sizeof(int64_t));
    orc_filter_mask(day, qty, disc, n, npad, spec, mask);
    size_t m = 0;
    for (size_t i = 0; i < npad; ++i) {
        m += mask[i];
        pairs[2 * i] = i < n ? price[i] : 0;
        pairs[2 * i + 1] = i < n ? disc[i] : 0;
    }
    orc_compact(pairs, 2 * sizeof(int64_t), mask, npad);
    __int128 acc = 0;
    for (size_t i = 0; i < npad; ++i) {
        __int128 prod = (__int128)pairs[2 * i] * pairs[2 * i + 1];
        __int128 keep = (__int128)(i < m);
        acc += prod * keep;
        if (out_price) out_price[i] = pairs[2 * i];
        if (out_disc) out_disc[i] = pairs[2 * i + 1];
    }
    *sum_lo = (uint64_t)acc;
    *sum_hi = (int64_t)(acc >> 64);
    free(mask);
    free(pairs);
    return m;
}

/* ------------------------------------------------------ join + group --- */

typedef struct { uint64_t tag; uint64_t idx; } tag_pair;

static int cmp_tag_pair(const void *a, const void *b)
{
    const tag_pair *x = (const tag_pair *)a, *y = (const tag_pair *)b;
    if (x->tag != y->tag) return x->tag < y->tag ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

void orc_gen_join(uint64_t seed, size_t n_dim, size_t n_fact, uint32_t n_groups,
                  int64_t vlo, int64_t vhi,
                  int32_t *dim_key, int32_t *dim_group, int32_t *fact_fk, int64_t *fact_val)
{
    tag_pair *tp = (tag_pair *)malloc((n_dim ? n_dim : 1) * sizeof(tag_pair));
    for (size_t i = 0; i < n_dim; ++i) {
        tp[i].tag = orc_rand64(seed, 5, i);
        tp[i].idx = i;
    }
    qsort(tp, n_dim, sizeof(tag_pair), cmp_tag_pair);
    for (size_t i = 0; i < n_dim; ++i) {
        dim_key[i] = (int32_t)(tp[i].idx + 1);
        dim_group[i] = (int32_t)orc_rand_range(seed, 6, i, 0, n_groups);
    }
    for (size_t i = 0; i < n_fact; ++i) {
        fact_fk[i] = (int32_t)orc_rand_range(seed, 7, i, 1, n_dim);
        fact_val[i] = orc_rand_range(seed, 8, i, vlo, (uint64_t)(vhi - vlo + 1));
    }
    free(tp);
}

uint64_t orc_union_key(int32_t key, int is_fact, uint32_t row)
{
    return ((uint64_t)(uint32_t)key << 32) | ((uint64_t)(is_fact != 0) << 31) | (row & 0x7FFFFFFFu);
}

size_t orc_join_groupsum(const int32_t *dim_key, const int32_t *dim_group, size_t n_dim,
                         const int32_t *fact_fk, const int64_t *fact_val, size_t n_fact,
                         size_t npad,
                         uint64_t *sorted_keys, uint64_t *sorted_vals,
                         int32_t *out_group, int64_t *out_sum, size_t max_groups)
{
    uint64_t *k = (uint64_t *)malloc(npad * sizeof(uint64_t));
    uint64_t *v = (uint64_t *)malloc(npad * sizeof(uint64_t));
    /* 1. union: dimension rows then fact rows, pads with key = max */
    for (size_t i = 0; i < npad; ++i) {
        if (i < n_dim) {
            k[i] = orc_union_key(dim_key[i], 0, (uint32_t)i);
            v[i] = (uint64_t)(uint32_t)dim_group[i];
        } else if (i < n_dim + n_fact) {
            size_t f = i - n_dim;
            k[i] = orc_union_key(fact_fk[f], 1, (uint32_t)f);
            v[i] = (uint64_t)fact_val[f];
        } else {
            k[i] = UINT64_MAX;
            v[i] = 0;
        }
    }
    /* 2. oblivious sort by union key */
    orc_bitonic_sort_kv(k, v, npad);
    if (sorted_keys) memcpy(sorted_keys, k, npad * sizeof(uint64_t));
    if (sorted_vals) memcpy(sorted_vals, v, npad * sizeof(uint64_t));
    /* 3. forward fill of the last dimension row (branch-free selects) */
    uint8_t *sel = (uint8_t *)malloc(npad);
    int64_t *rows = (int64_t *)malloc(npad * 2 * sizeof(int64_t));
    uint64_t cur_key = UINT64_MAX, cur_group = 0;
    for (size_t i = 0; i < npad; ++i) {
        uint64_t key32 = k[i] >> 32;
        int is_pad = k[i] == UINT64_MAX;
        int is_fact = (int)((k[i] >> 31) & 1) & !is_pad;
        int is_dim = !is_fact & !is_pad;
        cur_key = is_dim ? key32 : cur_key;
        cur_group = is_dim ? v[i] : cur_group;
        int matched = is_fact & (cur_key == key32);
        sel[i] = (uint8_t)matched;
        rows[2 * i] = (int64_t)cur_group;
        rows[2 * i + 1] = (int64_t)v[i];
    }
    /* 4. compaction of matched fact rows */
    size_t m = 0;
    for (size_t i = 0; i < npad; ++i) m += sel[i];
    orc_compact(rows, 2 * sizeof(int64_t), sel, npad);
    /* 5. sort join rows by (group, position); rows >= m get group 0xFFFFFFFF */
    for (size_t i = 0; i < npad; ++i) {
        uint64_t g = i < m ? (uint64_t)rows[2 * i] : 0xFFFFFFFFull;
        k[i] = (g << 32) | (uint64_t)i;
        v[i] = (uint64_t)rows[2 * i + 1];
    }
    orc_bitonic_sort_kv(k, v, npad);
    /* 6. segmented inclusive sum, last row of each group selected */
    int64_t run = 0;
    for (size_t i = 0; i < npad; ++i) {
        uint64_t g = k[i] >> 32;
        int same = i > 0 && (k[i - 1] >> 32) == g;
        run = (same ? run : 0) + (int64_t)v[i];
        int last = (i + 1 == npad) || ((k[i + 1] >> 32) != g);
        sel[i] = (uint8_t)(last & (g != 0xFFFFFFFFull));
        rows[2 * i] = (int64_t)g;
        rows[2 * i + 1] = run;
    }
    /* 7. compaction -> one row per group */
    size_t G = 0;
    for (size_t i = 0; i < npad; ++i) G += sel[i];
    orc_compact(rows, 2 * sizeof(int64_t), sel, npad);
    for (size_t i = 0; i < G && i < max_groups; ++i) {
        out_group[i] = (int32_t)rows[2 * i];
        out_sum[i] = rows[2 * i + 1];
    }
    free(k);
    free(v);
    free(sel);
    free(rows);
    return G;
}

/* ------------------------------------------------------------------------ */
/* General equi-join into row-index pairs                                    */
/* ------------------------------------------------------------------------ */

/* sort.h:18-22: the largest power of two strictly below x (0 for x <= 1) */
static size_t prev_pow_two(size_t x)
{
    size_t y = 1;
    while (y < x) y <<= 1;
    return y >> 1;
}

/* The shifting loop of obliv_distribute (sort.h:103-121) over elements of sz
 * bytes already ordered by target (valid first, targets ascending), with
 * tgt[i] = -1 for dummies; buf and tgt hold m slots (n_in given, the rest
 * dummies).  Returns -1 if a move would land on a non-dummy. */
static int kks_shift(unsigned char *buf, int64_t *tgt, size_t m, size_t sz)
{
    unsigned char *tmp = (unsigned char *)malloc(sz ? sz : 1);
    int bad = 0;
    for (size_t j = prev_pow_two(m); j >= 1; j >>= 1) {
        for (size_t ii = m - j; ii-- > 0;) {
            int64_t dest = tgt[ii];
            if (dest >= (int64_t)(ii + j)) {
                bad |= tgt[ii + j] != -1;
                memcpy(tmp, buf + ii * sz, sz);
                memcpy(buf + ii * sz, buf + (ii + j) * sz, sz);
                memcpy(buf + (ii + j) * sz, tmp, sz);
                tgt[ii] = tgt[ii + j];
                tgt[ii + j] = dest;
            }
        }
        if (j == 1) break;
    }
    free(tmp);
    return bad ? -1 : 0;
}

typedef struct {
    int64_t t;
    uint64_t p;
} dist_pair;

/* ind_func_comp (sort.h:78-84): dummies last, then by target */
static int cmp_dist(const void *a, const void *b)
{
    const dist_pair *x = (const dist_pair *)a, *y = (const dist_pair *)b;
    int dx = x->t == -1, dy = y->t == -1;
    if (dx != dy) return dx - dy;
    return (x->t > y->t) - (x->t < y->t);
}

int orc_distribute(const int64_t *target, const uint64_t *payload, size_t n, size_t m,
                   int64_t *out_target, uint64_t *out_payload)
{
    size_t cap = n > m ? n : m;
    dist_pair *d = (dist_pair *)calloc(cap ? cap : 1, sizeof(dist_pair));
    int64_t *tg = (int64_t *)malloc((cap ? cap : 1) * sizeof(int64_t));
    for (size_t i = 0; i < n; ++i) {
        d[i].t = target[i] < 0 ? -1 : target[i];
        d[i].p = payload[i];
    }
    qsort(d, n, sizeof(dist_pair), cmp_dist);
    for (size_t i = n; i < cap; ++i) d[i].t = -1;
    int rc = 0;
    for (size_t i = 0; i < m; ++i) {
        rc |= d[i].t >= (int64_t)m ? -1 : 0;
        tg[i] = d[i].t;
    }
    if (rc == 0) rc = kks_shift((unsigned char *)d, tg, m, sizeof(dist_pair));
    for (size_t i = 0; i < m; ++i) {
        out_target[i] = tg[i];
        out_payload[i] = d[i].t == -1 ? 0 : d[i].p;
    }
    free(d);
    free(tg);
    return rc;
}

/* one row of the join union (the reference's elem_t, common/elem_t.h:15-23,
 * with the data bytes replaced by the row index) */
typedef struct {
    uint64_t key;
    uint32_t tab; /* 0 left (table_0), 1 right, 2 pad */
    uint32_t row;
    int64_t m0, m1, j_order;
    int64_t has_value;
} jp_elem;

static int jp_less(const jp_elem *a, const jp_elem *b)
{
    if (a->key != b->key) return a->key < b->key;
    if (a->tab != b->tab) return a->tab < b->tab;
    return a->row < b->row;
}

static void jp_bitonic(jp_elem *a, size_t npad)
{
    for (size_t k = 2; k <= npad; k <<= 1)
        for (size_t j = k >> 1; j > 0; j >>= 1)
            for (size_t i = 0; i < npad; ++i) {
                size_t x = i ^ j;
                if (x > i) {
                    int asc = (i & k) == 0;
                    int sw = asc ? jp_less(&a[x], &a[i]) : jp_less(&a[i], &a[x]);
                    cswap((unsigned char *)&a[i], (unsigned char *)&a[x], sizeof(jp_elem), sw);
                }
            }
}

static int cmp_jorder(const void *a, const void *b)
{
    const jp_elem *x = (const jp_elem *)a, *y = (const jp_elem *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->j_order > y->j_order) - (x->j_order < y->j_order);
}

/* Steps 5 of the join for one copy: the compacted rows [0, k) with targets,
 * distributed over n_out slots, then forward fill (:305-314). */
static int jp_expand(const jp_elem *c, const int64_t *tgt, size_t k, size_t n_out, jp_elem *out)
{
    int64_t *tg = (int64_t *)malloc((n_out ? n_out : 1) * sizeof(int64_t));
    size_t take = k < n_out ? k : n_out;
    memset(out, 0, n_out * sizeof(jp_elem));
    for (size_t i = 0; i < n_out; ++i) {
        tg[i] = -1;
        if (i < take) {
            out[i] = c[i];
            out[i].has_value = 1;
            tg[i] = tgt[i];
        }
    }
    int rc = kks_shift((unsigned char *)out, tg, n_out, sizeof(jp_elem));
    for (size_t i = 0; i < n_out; ++i) out[i].has_value = tg[i] != -1;
    for (size_t i = 1; i < n_out; ++i)
        if (!out[i].has_value) out[i] = out[i - 1];
    free(tg);
    return rc;
}

size_t orc_join_pairs(const uint64_t *lkey, size_t n_left, const uint64_t *rkey, size_t n_right,
                      int64_t *out_left, int64_t *out_right, size_t cap, int *status)
{
    size_t n = n_left + n_right, npad = 1;
    while (npad < n) npad <<= 1;
    jp_elem *arr = (jp_elem *)calloc(npad, sizeof(jp_elem));
    jp_elem *arr_ = (jp_elem *)calloc(npad, sizeof(jp_elem));
    *status = 0;
    /* 1. union, sort */
    for (size_t i = 0; i < npad; ++i) {
        if (i < n_left) {
            arr[i].key = lkey[i];
            arr[i].tab = 0;
            arr[i].row = (uint32_t)i;
        } else if (i < n) {
            arr[i].key = rkey[i - n_left];
            arr[i].tab = 1;
            arr[i].row = (uint32_t)(i - n_left);
        } else {
            arr[i].key = UINT64_MAX;
            arr[i].tab = 2;
            arr[i].row = (uint32_t)i;
        }
    }
    jp_bitonic(arr, npad);
    /* 2. m0 / m1: forward counts, backward propagation (:211-224); pads
     * belong to no key */
    for (size_t i = 0; i < npad; ++i) {
        int t0 = arr[i].tab == 0, t1 = arr[i].tab == 1;
        int c = i > 0 && arr[i].key == arr[i - 1].key && arr[i].tab != 2 && arr[i - 1].tab != 2;
        int64_t pm0 = i > 0 ? arr[i - 1].m0 : 0, pm1 = i > 0 ? arr[i - 1].m1 : 0;
        arr[i].m0 = (t0 && c) * (pm0 + 1) + (t1 && c) * pm0 + (t0 && !c);
        arr[i].m1 = (t0 && c) * pm1 + (t1 && c) * (pm1 + 1) + (t1 && !c);
    }
    for (size_t i = npad; i-- > 1;) {
        int c = arr[i - 1].key == arr[i].key && arr[i].tab != 2 && arr[i - 1].tab != 2;
        arr[i - 1].m0 = c ? arr[i].m0 : arr[i - 1].m0;
        arr[i - 1].m1 = c ? arr[i].m1 : arr[i - 1].m1;
    }
    /* 3. control bits, two copies, compaction */
    uint8_t *cb1 = (uint8_t *)calloc(npad, 1), *cb2 = (uint8_t *)calloc(npad, 1);
    size_t k1 = 0, k2 = 0;
    for (size_t i = 0; i < npad; ++i) {
        cb1[i] = arr[i].tab == 0 && arr[i].m1 > 0;
        cb2[i] = arr[i].tab == 1 && arr[i].m0 > 0;
        k1 += cb1[i];
        k2 += cb2[i];
        arr_[i] = arr[i];
    }
    orc_compact(arr, sizeof(jp_elem), cb1, npad);
    orc_compact(arr_, sizeof(jp_elem), cb2, npad);
    /* 4. index targets */
    int64_t *it1 = (int64_t *)malloc((k1 ? k1 : 1) * sizeof(int64_t));
    int64_t *it2 = (int64_t *)malloc((k2 ? k2 : 1) * sizeof(int64_t));
    size_t n_out = 0, n_out2 = 0;
    for (size_t i = 0; i < k1; ++i) {
        it1[i] = (int64_t)n_out;
        n_out += (size_t)arr[i].m1;
    }
    for (size_t i = 0; i < k2; ++i) {
        it2[i] = (int64_t)n_out2;
        n_out2 += (size_t)arr_[i].m0;
    }
    if (n_out2 != n_out) *status = -1;
    if (*status == 0 && n_out <= cap && n_out > 0) {
        /* 5. distribution + fill of both copies */
        jp_elem *e1 = (jp_elem *)malloc(n_out * sizeof(jp_elem));
        jp_elem *e2 = (jp_elem *)malloc(n_out * sizeof(jp_elem));
        *status |= jp_expand(arr, it1, k1, n_out, e1);
        *status |= jp_expand(arr_, it2, k2, n_out, e2);
        /* 6. j_order of the right copy, sort by (key, j_order) */
        int64_t j = 0;
        for (size_t i = 0; i < n_out; ++i) {
            j = (i > 0 && e2[i].key == e2[i - 1].key) ? j + 1 : 0;
            int64_t m0 = e2[i].m0 > 1 ? e2[i].m0 : 1;
            e2[i].j_order = (j % m0) * e2[i].m1 + j / m0;
        }
        qsort(e2, n_out, sizeof(jp_elem), cmp_jorder);
        for (size_t i = 0; i < n_out; ++i) {
            out_left[i] = e1[i].row;
            out_right[i] = e2[i].row;
            if (e1[i].key != e2[i].key) *status = -1;
        }
        free(e1);
        free(e2);
    }
    free(arr);
    free(arr_);
    free(cb1);
    free(cb2);
    free(it1);
    free(it2);
    return n_out;
}

/* ------------------------------------------------------------------------ */
/* Oblivious materialization                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t record_id;
    uint32_t position;
    uint32_t pairing_idx;
} sortable_tuple; /* SortableTuple of include/operator/materialization_operator.hpp */

void orc_gather_recover(const uint64_t *table, size_t n_table, const int64_t *ids, size_t n,
                        uint64_t *out, uint8_t *found)
{
    size_t npad = 1;
    while (npad < n) npad <<= 1;
    sortable_tuple *t = (sortable_tuple *)malloc((npad ? npad : 1) * sizeof(sortable_tuple));
    for (size_t i = 0; i < npad; ++i) {
        if (i < n) {
            t[i].record_id = ids[i];
            t[i].position = (ids[i] >= 0 && (uint64_t)ids[i] < n_table) ? (uint32_t)ids[i] : 0;
            t[i].pairing_idx = (uint32_t)i;
        } else { /* sentinels sort last (:147-152) */
            t[i].record_id = INT64_MAX;
            t[i].position = 0;
            t[i].pairing_idx = UINT32_MAX;
        }
    }
    /* bitonic sort by record_id (:156-169) */
    for (size_t k = 2; k <= npad; k <<= 1)
        for (size_t j = k >> 1; j > 0; j >>= 1)
            for (size_t i = 0; i < npad; ++i) {
                size_t x = i ^ j;
                if (x > i) {
                    int asc = (i & k) == 0;
                    int sw = asc ? (t[i].record_id > t[x].record_id) : (t[i].record_id < t[x].record_id);
                    cswap((unsigned char *)&t[i], (unsigned char *)&t[x], sizeof(sortable_tuple), sw);
                }
            }
    /* gather in sorted order with a one-entry cache (:476-503) */
    uint64_t *sv = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint8_t *sf = (uint8_t *)malloc(n ? n : 1);
    int64_t cached_id = INT64_MIN;
    uint64_t cached = 0;
    uint8_t cached_found = 0;
    for (size_t i = 0; i < n; ++i) {
        if (t[i].record_id != cached_id) {
            int ok = t[i].record_id >= 0 && (uint64_t)t[i].record_id < n_table;
            cached = ok ? table[t[i].position] : 0;
            cached_found = (uint8_t)ok;
            cached_id = t[i].record_id;
        }
        sv[i] = cached;
        sf[i] = cached_found;
    }
    /* restore the request order (:505-516) */
    for (size_t i = 0; i < n; ++i) {
        out[t[i].pairing_idx] = sv[i];
        if (found) found[t[i].pairing_idx] = sf[i];
    }
    free(t);
    free(sv);
    free(sf);
}

/* ------------------------------------------------------------------------ */
/* Multi-threaded bitonic sort (CPU baseline)                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t *k;
    uint32_t *p;
    size_t lo, hi, c;
    int asc, threads;
} mt_args;

static inline void mt_cmpx(uint64_t *k, uint32_t *p, size_t i, size_t j, int asc)
{
    uint64_t ka = k[i], kb = k[j];
    uint32_t pa = p[i], pb = p[j];
    int gt = (ka > kb) | ((ka == kb) & (pa > pb));
    int sw = asc ? gt : !gt;
    uint64_t mk = (uint64_t)0 - (uint64_t)sw;
    uint32_t mp = (uint32_t)0 - (uint32_t)sw;
    uint64_t xk = (ka ^ kb) & mk;
    uint32_t xp = (pa ^ pb) & mp;
    k[i] = ka ^ xk;
    k[j] = kb ^ xk;
    p[i] = pa ^ xp;
    p[j] = pb ^ xp;
}

/* bitonic_merge_2 (bitonic.c:57-68): compares i, i + c for i in [lo, hi) */
static void *mt_pass(void *v)
{
    mt_args *a = (mt_args *)v;
    for (size_t i = a->lo; i < a->hi; ++i) mt_cmpx(a->k, a->p, i, i + a->c, a->asc);
    return NULL;
}

static void *mt_merge(void *v);

/* bitonic_merge (bitonic.c:70-155) */
static void *mt_merge(void *v)
{
    mt_args *a = (mt_args *)v;
    if (a->hi <= a->lo + 1) return NULL;
    size_t mid = prev_pow_two(a->hi - a->lo);
    size_t len = a->hi - mid - a->lo;
    if (a->threads <= 1) {
        for (size_t i = a->lo; i < a->hi - mid; ++i) mt_cmpx(a->k, a->p, i, i + mid, a->asc);
    } else {
        int t = a->threads;
        pthread_t th[64];
        mt_args sub[64];
        size_t start = a->lo;
        for (int i = 0; i < t; ++i) {
            size_t l = len / t + ((size_t)i < len % t);
            sub[i] = (mt_args){a->k, a->p, start, start + l, mid, a->asc, 1};
            start += l;
            if (i < t - 1) pthread_create(&th[i], NULL, mt_pass, &sub[i]);
        }
        mt_pass(&sub[t - 1]);
        for (int i = 0; i < t - 1; ++i) pthread_join(th[i], NULL);
    }
    int tl = a->threads / 2, tr = a->threads - tl;
    mt_args left = {a->k, a->p, a->lo, a->lo + mid, 0, a->asc, a->threads > 1 ? tl : 1};
    mt_args right = {a->k, a->p, a->lo + mid, a->hi, 0, a->asc, a->threads > 1 ? tr : 1};
    if (a->threads > 1) {
        pthread_t th;
        pthread_create(&th, NULL, mt_merge, &left);
        mt_merge(&right);
        pthread_join(th, NULL);
    } else {
        mt_merge(&left);
        mt_merge(&right);
    }
    return NULL;
}

/* bitonic_sort_new (bitonic.c:157-230) */
static void *mt_sort(void *v)
{
    mt_args *a = (mt_args *)v;
    size_t mid = a->lo + (a->hi - a->lo) / 2;
    if (mid == a->lo) return NULL;
    int tl = a->threads / 2, tr = a->threads - tl;
    mt_args left = {a->k, a->p, a->lo, mid, 0, !a->asc, a->threads > 1 ? tl : 1};
    mt_args right = {a->k, a->p, mid, a->hi, 0, a->asc, a->threads > 1 ? tr : 1};
    if (a->threads > 1) {
        pthread_t th;
        pthread_create(&th, NULL, mt_sort, &left);
        mt_sort(&right);
        pthread_join(th, NULL);
    } else {
        mt_sort(&left);
        mt_sort(&right);
    }
    mt_args m = {a->k, a->p, a->lo, a->hi, 0, a->asc, a->threads};
    mt_merge(&m);
    return NULL;
}

void orc_bitonic_sort_kp_mt(uint64_t *keys, uint32_t *payload, size_t n, size_t npad, int nthreads)
{
    for (size_t i = n; i < npad; ++i) {
        keys[i] = UINT64_MAX;
        payload[i] = (uint32_t)i;
    }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    mt_args a = {keys, payload, 0, npad, 0, 1, nthreads};
    mt_sort(&a);
}

/* ------------------------------------------------------------------------ */
/* Predicate mask (comparison_expression.cpp:13-92, logical_expression.cpp:61-150) */
/* ------------------------------------------------------------------------ */

#define ORC_CMP(v, k, c)                                                                                \
    ((c) == 0 ? (v) == (k) : (c) == 1 ? (v) != (k) : (c) == 2 ? (v) < (k) : (c) == 3 ? (v) <= (k)    \
     : (c) == 4 ? (v) > (k) : (v) >= (k))

/* One comparison of row i; *ok = 0 for an unsupported type. */
static int orc_cmp_row(const void *col, uint32_t type, uint32_t cmp, uint64_t kb, size_t i, int *ok)
{
    switch (type) {
    case 1: { uint8_t v = ((const uint8_t *)col)[i], k = (uint8_t)kb; return ORC_CMP(v, k, cmp); }
    case 2: { int8_t v = ((const int8_t *)col)[i], k = (int8_t)(uint8_t)kb; return ORC_CMP(v, k, cmp); }
    case 3: { int16_t v = ((const int16_t *)col)[i], k = (int16_t)(uint16_t)kb; return ORC_CMP(v, k, cmp); }
    case 4: { int32_t v = ((const int32_t *)col)[i], k = (int32_t)(uint32_t)kb; return ORC_CMP(v, k, cmp); }
    case 10: { uint32_t v = ((const uint32_t *)col)[i], k = (uint32_t)kb; return ORC_CMP(v, k, cmp); }
    case 5: { int64_t v = ((const int64_t *)col)[i], k = (int64_t)kb; return ORC_CMP(v, k, cmp); }
    case 6: {
        float v = ((const float *)col)[i], k;
        uint32_t b = (uint32_t)kb;
        memcpy(&k, &b, 4);
        return ORC_CMP(v, k, cmp);
    }
    case 7: {
        double v = ((const double *)col)[i], k;
        memcpy(&k, &kb, 8);
        return ORC_CMP(v, k, cmp);
    }
    case 8: case 64: { uint64_t v = ((const uint64_t *)col)[i]; return ORC_CMP(v, kb, cmp); }
    default: *ok = 0; return 0;
    }
}

int orc_predicate_mask(const void *const *cols, const uint8_t *const *valid, const uint32_t *types,
                       const uint32_t *cmps, const uint64_t *consts, uint32_t n_terms, uint32_t conn,
                       size_t n, size_t npad, uint8_t *mask)
{
    if (n_terms == 0 || conn > 1) return -1;
    for (uint32_t j = 0; j < n_terms; ++j)
        if (cmps[j] > 5) return -1;
    int ok = 1;
    for (size_t i = 0; i < npad; ++i) {
        if (i >= n) {
            mask[i] = 0;
            continue;
        }
        /* left-deep chain: (data, null) of the running result */
        int data = 0, null = 0;
        for (uint32_t j = 0; j < n_terms; ++j) {
            const int tnull = valid[j] != NULL && valid[j][i] == 0;
            const int tdata = tnull ? 0 : orc_cmp_row(cols[j], types[j], cmps[j], consts[j], i, &ok);
            if (j == 0) {
                data = tdata;
                null = tnull;
            } else if (null || tnull) {
                data = 0;
                null = 1;
            } else {
                data = conn == 0 ? (data & tdata) : (data | tdata);
            }
        }
        mask[i] = (uint8_t)(null ? 0 : data);
    }
    return ok ? 0 : -1;
}
