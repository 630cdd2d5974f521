"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each check is independent of the oracle's own formula: published known-answer
vectors, the reference's own known-answer programs and test expectations,
closed forms (numpy sorts, stable partitions, brute-force sums) and an
independent formulation of the same network.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

M64 = (1 << 64) - 1
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _mix64_py(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


# ------------------------------------------------------------- generators

def test_splitmix64_published_sequence():
    # SplitMix64 from state 0 (Vigna's splitmix64.c): first three outputs.
    lib = oracle.lib()
    g = 0x9E3779B97F4A7C15
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert [lib.orc_mix64((g * (i + 1)) & M64) for i in range(3)] == want


def test_rand64_stream_definition():
    # counter-based stream = SplitMix64 started at mix64(seed ^ g*(stream+1))
    g = 0x9E3779B97F4A7C15
    for seed, stream in [(1, 0), (2, 0), (123456789, 3), (M64, 8)]:
        s0 = _mix64_py(seed ^ ((g * (stream + 1)) & M64))
        for i in [0, 1, 2, 1000, 2**31 + 5]:
            assert oracle.rand64(seed, stream, i) == _mix64_py((s0 + (i + 1) * g) & M64)


def test_philox4x32_10_random123_kat():
    # Random123 kat_vectors, philox4x32 with 10 rounds.
    assert oracle.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                       0x6D5451FD]
    assert oracle.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_fact_generator_ranges():
    day, qty, price, disc = oracle.gen_fact(7, 200000)
    assert day.min() >= 8036 and day.max() <= 10561
    assert qty.min() >= 1 and qty.max() <= 50
    assert price.min() >= 90000 and price.max() <= 10495000
    assert disc.min() >= 0 and disc.max() <= 10
    # every value of the small ranges appears
    assert len(np.unique(disc)) == 11 and len(np.unique(qty)) == 50


# ---------------------------------------------------------------- bitonic

def _lex_sorted(keys, pay):
    order = np.lexsort((pay, keys))
    return keys[order], pay[order]


@pytest.mark.parametrize("n,npad", [(0, 1), (1, 1), (1, 2), (3, 4), (5, 8), (8, 8), (11, 16), (1000, 1024),
                                    (1024, 1024), (3000, 4096), (65536, 65536)])
def test_bitonic_kp_equals_unique_sorted_order(n, npad):
    k, p = oracle.gen_sort_input(11 + n, n)
    sk, sp = oracle.bitonic_sort_kp(k, p, n, npad)
    ek = np.concatenate([k, np.full(npad - n, M64, np.uint64)])
    ep = np.concatenate([p, np.arange(n, npad, dtype=np.uint32)])
    ek, ep = _lex_sorted(ek, ep)
    assert np.array_equal(sk, ek) and np.array_equal(sp, ep)


def test_bitonic_kp_with_duplicate_keys_ties_by_payload():
    rng = np.random.default_rng(5)
    k = rng.integers(0, 7, 4096).astype(np.uint64)
    p = rng.permutation(4096).astype(np.uint32)
    sk, sp = oracle.bitonic_sort_kp(k, p, 4096, 4096)
    ek, ep = _lex_sorted(k, p)
    assert np.array_equal(sk, ek) and np.array_equal(sp, ep)


def test_bitonic_kks_known_answers():
    # obliviator/join_kks_INT_INT/test_sort/test_sorts.cpp:12-44
    a = np.array([3, 7, 4, 6, 0, 2, 1, 5], np.uint64)
    b = np.array([7, 10, 8, 1, 3, 6, 9, 2, 4, 5, 0], np.uint64)
    sa, _ = oracle.bitonic_sort_kp(a, np.arange(8, dtype=np.uint32), 8, 8)
    sb, _ = oracle.bitonic_sort_kp(b, np.arange(11, dtype=np.uint32), 11, 16)
    assert sa.tolist() == list(range(8))
    assert sb[:11].tolist() == list(range(11))


def test_bitonic_matches_reference_built_kat_program():
    """oracle/_ref/kks_test_sorts is compiled from the reference's own sources
    (make -C oracle ref); its printed outputs must equal the oracle's."""
    if not os.path.exists(oracle.REF_KAT):
        if os.path.isdir("/root/reference"):
            subprocess.check_call(["make", "-s", "-C", os.path.dirname(oracle.REF_KAT), "ref"])
        else:
            pytest.skip("reference KAT program not built and /root/reference absent")
    out = subprocess.check_output([oracle.REF_KAT], timeout=30).decode().split("\n")
    rows = [list(map(int, out[i + 1].split())) for i, line in enumerate(out) if line.endswith("sorted:")]
    a = np.array([3, 7, 4, 6, 0, 2, 1, 5], np.uint64)
    b = np.array([7, 10, 8, 1, 3, 6, 9, 2, 4, 5, 0], np.uint64)
    sa, _ = oracle.bitonic_sort_kp(a, np.arange(8, dtype=np.uint32), 8, 8)
    sb, _ = oracle.bitonic_sort_kp(b, np.arange(11, dtype=np.uint32), 11, 16)
    assert rows == [sa.tolist(), sb[:11].tolist()]


def test_bitonic_trace_is_value_independent_and_counts_stages():
    for npad in [16, 1024]:
        t = []
        for seed in (1, 2):
            k, p = oracle.gen_sort_input(seed, npad - 3)
            _, _, tr = oracle.bitonic_sort_kp(k, p, npad - 3, npad, trace=True)
            t.append((tr.hash, tr.n_ops, tr.n_stages))
        lg = int(np.log2(npad))
        assert t[0] == t[1]
        assert t[0][2] == lg * (lg + 1) // 2
        # every stage touches every element: 2 reads + 2 writes per pair
        assert t[0][1] == 3 + t[0][2] * npad * 2


def test_multi_column_sort_reference_order_by_expectations():
    g = json.load(open(os.path.join(GOLDEN, "order_by_employees.json")))
    cols = g["columns"]
    types = {"id": 4, "name": 8, "salary": 7, "department": 8, "age": 4, "active": 1}
    dt = {4: np.int32, 8: np.uint64, 7: np.float64, 1: np.uint8}
    n, npad = 8, 8
    for case in g["cases"]:
        names = [c for c, _ in case["order_by"]]
        arrs = [np.array(cols[c], dt[types[c]]) for c in names]
        out, idx = oracle.bitonic_sort_multi(arrs, [types[c] for c in names], [a for _, a in case["order_by"]],
                                             n, npad)
        for col, want in case["expect_prefix"].items():
            got = out[names.index(col)] if col in names else np.array(cols[col])[idx]
            assert got[:len(want)].tolist() == want


def test_multi_column_sort_unique_keys_equal_lexsort():
    rng = np.random.default_rng(3)
    n, npad = 1000, 1024
    a = rng.integers(-5, 5, npad).astype(np.int32)
    b = rng.permutation(npad).astype(np.int64)  # unique second key
    out, idx = oracle.bitonic_sort_multi([a, b], [4, 5], [True, False], n, npad)
    # expected: (a asc, b desc) over real rows, then pad rows (max, lowest)
    order = np.lexsort((-b[:n], a[:n]))
    assert np.array_equal(idx[:n], order)
    assert np.all(out[0][n:] == np.iinfo(np.int32).max) and np.all(out[1][n:] == np.iinfo(np.int64).min)


def test_multi_column_sort_permute_consistency():
    # sorted columns equal the input columns permuted by the index output (permute.cuh:10-22)
    rng = np.random.default_rng(9)
    npad = 512
    a = rng.integers(0, 4, npad).astype(np.int16)
    f = rng.standard_normal(npad).astype(np.float32)
    out, idx = oracle.bitonic_sort_multi([a, f], [3, 6], [False, True], npad, npad)
    assert np.array_equal(out[0], a[idx]) and np.array_equal(out[1], f[idx])
    assert np.all(np.diff(out[0].astype(np.int32)) <= 0)


# ------------------------------------------------------------- ORCompact

def _stable(buf, sel):
    return np.concatenate([buf[sel != 0], buf[sel == 0]])


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8])
def test_orcompact_exhaustive_small(n):
    buf = np.arange(n, dtype=np.int64) + 100
    for mask in range(1 << n):
        sel = np.array([(mask >> i) & 1 for i in range(n)], np.uint8)
        out = oracle.compact(buf, sel)
        m = int(sel.sum())
        assert np.array_equal(out[:m], buf[sel != 0])
        assert sorted(out[m:].tolist()) == sorted(buf[sel == 0].tolist())
        if n & (n - 1) == 0:
            assert np.array_equal(out, oracle.compact_levels(buf, sel))


@pytest.mark.parametrize("n,p", [(1 << 12, 0.02), (1 << 13, 0.5), (1 << 14, 0.97), (3000, 0.3), (12345, 0.1)])
def test_orcompact_stable_prefix_and_closed_form(n, p):
    rng = np.random.default_rng(n)
    buf = np.stack([np.arange(n, dtype=np.int64), rng.integers(0, 1 << 40, n)], axis=1)
    sel = (rng.random(n) < p).astype(np.uint8)
    out = oracle.compact(buf, sel)
    m = int(sel.sum())
    assert np.array_equal(out[:m], buf[sel != 0])
    assert sorted(out[m:, 0].tolist()) == sorted(buf[sel == 0, 0].tolist())
    if n & (n - 1) == 0:
        assert np.array_equal(out, oracle.compact_levels(buf, sel))


# ------------------------------------------------------- filter + int128

@pytest.mark.parametrize("n,npad", [(0, 1), (1000, 1024), (60000, 65536), (1 << 16, 1 << 16)])
def test_filter_sum_brute_force(n, npad):
    day, qty, price, disc = oracle.gen_fact(3, n)
    m, s, op, od = oracle.filter_sum(day, qty, price, disc, n, npad, want_rows=True)
    sel = (day >= 8766) & (day < 9131) & (disc >= 5) & (disc <= 7) & (qty < 24)
    assert m == int(sel.sum())
    assert s == sum(int(a) * int(b) for a, b in zip(price[sel], disc[sel]))
    assert np.array_equal(op[:m], price[sel]) and np.array_equal(od[:m], disc[sel])


def test_filter_selectivity_near_expected():
    day, qty, price, disc = oracle.gen_fact(1, 400000)
    sel = oracle.filter_mask(day, qty, disc, 400000, 400000)
    frac = sel.mean()
    expect = 365 / 2526 * 3 / 11 * 23 / 50
    assert abs(frac - expect) < 0.002


def test_filter_int128_exceeds_int64():
    # the int128 path must hold sums beyond 2^63: scale a copy of the prices
    n = 1024
    day = np.full(n, 8800, np.int32)
    qty = np.ones(n, np.int64)
    price = np.full(n, (1 << 62) - 1, np.int64)
    disc = np.full(n, 6, np.int64)
    m, s = oracle.filter_sum(day, qty, price, disc, n, n)
    assert m == n and s == n * ((1 << 62) - 1) * 6


# --------------------------------------------------------- join + group

@pytest.mark.parametrize("n_dim,n_fact,npad", [(100, 400, 512), (1500, 6000, 8192), (5000, 3000, 8192)])
def test_join_groupsum_brute_force(n_dim, n_fact, npad):
    dk, dg, fk, fv = oracle.gen_join(4, n_dim, n_fact)
    assert sorted(dk.tolist()) == list(range(1, n_dim + 1))
    groups, sums, sk, sv = oracle.join_groupsum(dk, dg, fk, fv, npad, want_sorted=True)
    g_of_key = np.zeros(n_dim + 1, np.int64)
    g_of_key[dk] = dg
    want = {}
    for k, v in zip(fk.tolist(), fv.tolist()):
        g = int(g_of_key[k])
        want[g] = want.get(g, 0) + v
    assert groups.tolist() == sorted(want)
    assert sums.tolist() == [want[g] for g in sorted(want)]
    # sorted union equals numpy's sort of the union keys
    lib = oracle.lib()
    keys = [lib.orc_union_key(int(k), 0, i) for i, k in enumerate(dk)] + \
           [lib.orc_union_key(int(k), 1, i) for i, k in enumerate(fk)]
    keys = np.array(keys + [M64] * (npad - n_dim - n_fact), np.uint64)
    assert np.array_equal(sk, np.sort(keys))


def test_join_with_unmatched_fact_rows():
    # fact keys outside the dimension never join
    dk = np.array([1, 2, 3], np.int32)
    dg = np.array([0, 1, 0], np.int32)
    fk = np.array([1, 4, 3, 3, 9, 2], np.int32)
    fv = np.array([10, 1000, 5, 7, 1000, 3], np.int64)
    groups, sums = oracle.join_groupsum(dk, dg, fk, fv, 16)
    assert groups.tolist() == [0, 1] and sums.tolist() == [22, 3]


# -------------------------------------------------- equi-join into row pairs

def _brute_pairs(lk, rk):
    """Nested loops: every (l, r) with equal keys, ordered by (key, l, r)."""
    out = sorted((int(lk[i]), i, j) for i in range(len(lk)) for j in range(len(rk)) if lk[i] == rk[j])
    return [o[1] for o in out], [o[2] for o in out]


def _read_join_file(path):
    """obliviator's join input format: "n1 n2", a blank line, then n1 and n2
    rows of "key data"."""
    lines = open(path).read().split("\n")
    n1, n2 = map(int, lines[0].split())
    rows = [ln.split() for ln in lines[1:] if ln.strip()]
    return rows[:n1], rows[n1:n1 + n2]


def test_distribute_matches_reference_built_kat_program():
    """oracle/_ref/kks_test_obliv_distribute is the reference's own
    distribution known-answer program (join_kks_INT_INT/test_sort/
    test_obliv_distribute.cpp) compiled from its sources; the oracle's
    distribution of the same three inputs must print the same slots."""
    if not os.path.exists(oracle.REF_DIST):
        if os.path.isdir("/root/reference"):
            subprocess.check_call(["make", "-s", "-C", os.path.dirname(oracle.REF_DIST), "ref"])
        else:
            pytest.skip("reference KAT program not built and /root/reference absent")
    out = subprocess.check_output([oracle.REF_DIST], timeout=30).decode().split("\n")
    got, cur = {}, None
    for ln in out:
        if ln.startswith("Result for "):
            cur = ln[len("Result for "):-1]
            got[cur] = []
        elif cur and ":" in ln:
            v = ln.split(":", 1)[1].strip()
            got[cur].append(None if v == "D" else tuple(int(x) for x in v.split(" -> ")))
    # (size, m, {slot: (id, target)}) of the program's inputs A, B, C
    cases = {
        "A": (16, 16, {2: (2, 8), 5: (5, 14), 6: (6, 1), 10: (10, 4), 13: (13, 15)}),
        "B": (17, 31, {3: (0, 0), 7: (7, 17), 8: (8, 21), 10: (10, 4), 14: (14, 29), 15: (1, 5)}),
        "C": (14, 100, {0: (0, 8), 1: (1, 10), 2: (2, 4), 3: (3, 3), 4: (4, 9), 5: (5, 92), 6: (6, 44),
                        7: (7, 45), 12: (12, 39)}),
    }
    for name, (size, m, elems) in cases.items():
        tg = np.full(size, -1, np.int64)
        pay = np.zeros(size, np.uint64)
        for slot, (eid, t) in elems.items():
            tg[slot], pay[slot] = t, eid
        ot, op = oracle.distribute(tg, pay, m)
        mine = [None if t == -1 else (int(p), int(t)) for t, p in zip(ot, op)]
        assert mine == got[name], name
        assert all(t == -1 or t == i for i, t in enumerate(ot))  # every element at its target


def test_distribute_random_invariant():
    """Sorted distinct targets from any positions land on their targets; the
    shifting network never overwrites an element (the reference's assert)."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        m = int(rng.integers(1, 300))
        k = int(rng.integers(0, m + 1))
        t = np.sort(rng.choice(m, size=k, replace=False)).astype(np.int64)
        n = int(rng.integers(k, k + 20))
        tg = np.full(n, -1, np.int64)
        pos = np.sort(rng.choice(n, size=k, replace=False))
        tg[pos] = t
        ot, op = oracle.distribute(tg, np.arange(n, dtype=np.uint64), m)
        want = np.full(m, -1, np.int64)
        want[t] = t
        assert np.array_equal(ot, want)
        assert np.array_equal(op[t], pos.astype(np.uint64))


def test_join_pairs_reference_golden_output():
    """obliviator/join/test.txt -> test_output.txt: 512 left rows of one key
    and 512 right rows of distinct keys, one of them equal, give 512 pairs.
    The multiset of (key1, data1, key2, data2) output lines is equal (the
    reference's line order follows its unstable sort)."""
    left, right = _read_join_file(os.path.join(GOLDEN, "obliviator_join_input.txt"))
    expected = sorted(tuple(ln.split()) for ln in open(os.path.join(GOLDEN, "obliviator_join_expected.txt"))
                      if ln.strip())
    lk = oracle.key_codes([int(r[0]) for r in left], 4)
    rk = oracle.key_codes([int(r[0]) for r in right], 4)
    ol, or_ = oracle.join_pairs(lk, rk)
    got = sorted((left[a][0], left[a][1], right[b][0], right[b][1]) for a, b in zip(ol, or_))
    assert len(got) == 512 and got == expected


def test_join_pairs_fk_join_fixture():
    """obliviator/fk_join/test.txt (10 primary keys x 1024 foreign keys):
    equal to nested loops, one pair per fact row."""
    left, right = _read_join_file(os.path.join(GOLDEN, "obliviator_fk_join_input.txt"))
    lk = [int(r[0]) for r in left]
    rk = [int(r[0]) for r in right]
    ol, or_ = oracle.join_pairs(oracle.key_codes(lk, 4), oracle.key_codes(rk, 4))
    bl, br = _brute_pairs(lk, rk)
    assert ol.tolist() == bl and or_.tolist() == br and len(bl) == 1024


@pytest.mark.parametrize("shape", ["uniform", "skew_left", "skew_right", "diagonal", "zipf_blocks", "disjoint"])
def test_join_pairs_brute_force(shape):
    """Random inputs of the shapes of obliviator's example generator
    (join_kks_INT_INT/test_scripts/gen_example.py:33-55) and more."""
    rng = np.random.default_rng(hash(shape) % 1000)
    for trial in range(40):
        n1 = int(rng.integers(1, 60))
        if shape == "uniform":
            lk = rng.integers(-5, 5, n1)
            rk = rng.integers(-5, 5, int(rng.integers(0, 60)))
        elif shape == "skew_left":        # gen_example class 1
            lk = np.zeros(n1, np.int64)
            rk = np.array([0] + [1] * (n1 - 1))
        elif shape == "skew_right":       # class 2
            lk = np.array([0] + [1] * (n1 - 1))
            rk = np.zeros(n1, np.int64)
        elif shape == "diagonal":         # class 3
            lk = rng.permutation(n1)
            rk = rng.permutation(n1)
        elif shape == "zipf_blocks":      # class 4
            lk, rk = [], []
            for key in range(int(rng.integers(1, 8))):
                lk += [key] * int(min(rng.zipf(2), 9))
                rk += [key] * int(min(rng.zipf(2), 9))
            lk, rk = rng.permutation(lk), rng.permutation(rk)
        else:
            lk = rng.integers(0, 100, n1)
            rk = rng.integers(100, 200, int(rng.integers(0, 60)))
        for t in (4, 5):
            ol, or_ = oracle.join_pairs(oracle.key_codes(lk, t), oracle.key_codes(rk, t))
            bl, br = _brute_pairs(list(lk), list(rk))
            assert ol.tolist() == bl and or_.tolist() == br, (shape, trial)


def test_key_codes_preserve_order_and_equality():
    rng = np.random.default_rng(9)
    for t, dt in ((2, np.int8), (3, np.int16), (4, np.int32), (5, np.int64), (1, np.uint8), (10, np.uint32),
                  (64, np.uint64)):
        info = np.iinfo(dt)
        v = rng.integers(info.min, info.max, 500, dtype=dt, endpoint=True)
        v[:4] = [info.min, info.max, 0, info.max]
        c = oracle.key_codes(v, t)
        order = np.argsort(v, kind="stable")
        assert np.array_equal(np.argsort(c, kind="stable"), order)
        assert np.array_equal(c[:, None] == c[None, :], v[:, None] == v[None, :])


# ------------------------------------------------- oblivious materialization

@pytest.mark.parametrize("n_table,n", [(1, 1), (10, 0), (1000, 1), (1000, 5000), (3000, 777), (0, 40)])
def test_gather_recover_equals_take(n_table, n):
    """The restated materialization (tuple bitonic sort, cached reads,
    restore) equals numpy take, with repeated and out-of-range ids."""
    rng = np.random.default_rng(n_table * 7 + n)
    table = rng.integers(0, 1 << 63, n_table, dtype=np.uint64)
    ids = rng.integers(-3, max(n_table, 1) + 3, n).astype(np.int64)
    if n > 10:
        ids[:10] = ids[0]  # a run of one repeated id
    out, found = oracle.gather_recover(table, ids)
    ok = (ids >= 0) & (ids < n_table)
    want = np.zeros(n, np.uint64)
    want[ok] = table[ids[ok]]
    assert np.array_equal(found.astype(bool), ok)
    assert np.array_equal(out, want)


@pytest.mark.parametrize("threads", [1, 2, 3, 8, 16])
def test_parallel_bitonic_equals_single_thread(threads):
    """The multi-threaded recursive bitonic (the multi-core CPU baseline)
    gives the unique sorted order of the single-thread network."""
    for n, npad in [(1000, 1024), (3000, 4096), (1 << 15, 1 << 15)]:
        k, p = oracle.gen_sort_input(5, n)
        ek, ep = oracle.bitonic_sort_kp(k, p, n, npad)
        gk, gp = oracle.bitonic_sort_kp_mt(k, p, n, npad, threads)
        assert np.array_equal(gk, ek) and np.array_equal(gp, ep)
        if n == npad:  # the in-place form the bench times
            wk, wp = k.copy(), p.copy()
            oracle.bitonic_sort_kp_mt_inplace(wk, wp, n, threads)
            assert np.array_equal(wk, ek) and np.array_equal(wp, ep)


# ---------------------------------------------------------------- predicates

def test_predicate_reference_vectors():
    """tests/expression/test_expression_batch.cpp:46-48, 97-118, 145-184 of
    the reference: val > 20 -> F F F T T; (val > 10) AND active -> F F F T F."""
    val = np.array([10, 20, 5, 30, 25], np.int32)
    act = np.array([1, 0, 1, 1, 0], np.uint8)
    assert oracle.predicate_mask([(val, 4, 4, 20)], 5, 5).tolist() == [0, 0, 0, 1, 1]
    assert oracle.predicate_mask([(val, 4, 4, 10), (act, 1, 1, 0)], 5, 5).tolist() == [0, 0, 0, 1, 0]


def test_predicate_q6_equals_filter_mask():
    """The Q6 clause as five generic terms equals the oracle's fused Q6 mask."""
    n, npad = 50_000, 1 << 16
    day, qty, _, disc = oracle.gen_fact(3, n)
    terms = [(day, 4, 5, 8766), (day, 4, 2, 9131), (disc, 5, 5, 5), (disc, 5, 3, 7), (qty, 5, 2, 24)]
    got = oracle.predicate_mask(terms, n, npad)
    assert np.array_equal(got, oracle.filter_mask(day, qty, disc, n, npad))
    assert 0 < got.sum() < n


@pytest.mark.parametrize("type_id,dtype", [(1, np.uint8), (2, np.int8), (3, np.int16), (4, np.int32),
                                           (10, np.uint32), (5, np.int64), (6, np.float32), (7, np.float64),
                                           (8, np.uint64), (64, np.uint64)])
def test_predicate_comparisons_follow_the_type(type_id, dtype):
    """Each comparison in the column's own type (signedness, IEEE order with
    NaN and -0.0) against numpy's typed comparison operators."""
    import operator
    rng = np.random.default_rng(type_id)
    n = 4000
    if np.dtype(dtype).kind == "f":
        col = (rng.integers(-8, 8, n) / 2).astype(dtype)
        col[::37] = np.nan
        col[::41] = -0.0
        k = dtype(0.0)
    else:
        info = np.iinfo(dtype)
        col = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
        col[::5] = info.max // 2
        k = dtype(info.max // 2)
    ops = [operator.eq, operator.ne, operator.lt, operator.le, operator.gt, operator.ge]
    for c, op in enumerate(ops):
        got = oracle.predicate_mask([(col, type_id, c, k.item())], n, n)
        assert np.array_equal(got, op(col, k).astype(np.uint8)), c


def test_predicate_nulls_and_chain_order():
    """A NULL operand makes the row NULL (unselected) under AND and OR
    (logical_expression.cpp:138-143 sets NULL, data 0): TRUE OR NULL and
    FALSE AND NULL both give 0; rows past n are 0."""
    a = np.array([1, 1, 0, 0, 1], np.int32)
    b = np.array([1, 0, 1, 0, 1], np.int32)
    bv = np.array([1, 0, 0, 1, 1], np.uint8)
    t = [(a, 4, 0, 1), (b, 4, 0, 1, bv)]
    assert oracle.predicate_mask(t, 4, 6, conn=1).tolist() == [1, 0, 0, 0, 0, 0]
    assert oracle.predicate_mask(t, 4, 6, conn=0).tolist() == [1, 0, 0, 0, 0, 0]
    assert oracle.predicate_mask([(b, 4, 0, 1, bv)], 5, 5).tolist() == [1, 0, 0, 0, 1]
    with pytest.raises(ValueError):
        oracle.predicate_mask([(a, 99, 0, 1)], 5, 5)


# ---- closed forms that check the configs' full sizes, pinned to the restatements

@pytest.mark.parametrize("n,npad", [(3000, 4096), (20000, 1 << 15)])
def test_filter_closed_form_equals_restatement(n, npad):
    day, qty, price, disc = oracle.gen_fact(21 + n, n)
    m, s, op, od = oracle.filter_sum(day, qty, price, disc, n, npad, want_rows=True)
    cm, cs, cp, cd = oracle.filter_closed_form(day, qty, price, disc)
    assert (cm, cs) == (m, s)
    assert np.array_equal(cp, op[:m]) and np.array_equal(cd, od[:m])
    # a planted mistake (the first selected row dropped) is seen
    first = int(np.flatnonzero(oracle.filter_mask(day, qty, disc, n, npad)[:n])[0])
    keep = np.arange(n) != first
    assert oracle.filter_closed_form(day[keep], qty[keep], price[keep], disc[keep])[:2] != (m, s)


@pytest.mark.parametrize("n_dim,n_fact", [(500, 2000), (3000, 9000)])
def test_join_closed_form_equals_restatement(n_dim, n_fact):
    dk, dg, fk, fv = oracle.gen_join(5 + n_dim, n_dim, n_fact)
    npad = 1 << int(np.ceil(np.log2(n_dim + n_fact)))
    g, s = oracle.join_groupsum(dk, dg, fk, fv, npad)
    cg, cs = oracle.join_groupsum_closed_form(dk, dg, fk, fv)
    assert np.array_equal(cg, g) and np.array_equal(cs, s)
    # unmatched fact rows and groups beyond the bound drop out of both
    fk2 = fk.copy()
    fk2[::7] = n_dim * 10 + 3
    g2, s2 = oracle.join_groupsum(dk, dg, fk2, fv, npad)
    cg2, cs2 = oracle.join_groupsum_closed_form(dk, dg, fk2, fv)
    assert np.array_equal(cg2, g2) and np.array_equal(cs2, s2)
    cg3, _ = oracle.join_groupsum_closed_form(dk, dg, fk, fv, max_groups=10)
    assert cg3.max() < 10


def test_sort_closed_form_equals_restatement():
    n = 5000
    k, p = oracle.gen_sort_input(9, n)
    k[::5] = k[1::5][: len(k[::5])]  # duplicate keys: ties by payload
    ek, ep = oracle.bitonic_sort_kp(k, p, n, 8192)
    ck, cp = oracle.sort_closed_form(k, p)
    assert np.array_equal(ck, ek[:n]) and np.array_equal(cp, ep[:n])
