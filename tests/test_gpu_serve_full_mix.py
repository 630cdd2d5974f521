"""The served workload at its full shape: tables log-uniform in [1e4, 1e7]
rows (one of each kind at 1e7), the 50/30/20 filter+SUM / join+group-by /
sort mix, Poisson arrivals at 800 QPS and a 200 GB admission budget.  Every
query's whole result is checked against the oracle's closed forms of its
table (filter count, exact SUM and compacted rows; the join's groups and
sums; the sorted keys and payloads), for the concurrent mix, for the plain
one-query-at-a-time loop over shipped inputs and for the mix run through the
resident command queue, so each result is the same alone and in the mix."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

RATE_QPS = 800
N_QUERIES = 240


def _tables(srv, S):
    rng = np.random.default_rng(5)
    sizes = [10_000_000] + [int(np.exp(rng.uniform(np.log(1e4), np.log(1e7)))) for _ in range(2)]
    t = {S.FILTER_SUM: [], S.JOIN_GROUPSUM: [], S.SORT: []}
    for i, n in enumerate(sizes):
        t[S.FILTER_SUM].append(srv.add_fact_table(500 + i, n))
        t[S.JOIN_GROUPSUM].append(srv.add_join_tables(600 + i, max(1, n // 5), n - n // 5))
        t[S.SORT].append(srv.add_sort_table(700 + i, n))
    return t, sizes


def _mix(S, tables, seed):
    r = np.random.default_rng(seed)
    t, qs = 0.0, []
    for _ in range(N_QUERIES):
        t += r.exponential(1.0 / RATE_QPS)
        u = r.random()
        kind = S.FILTER_SUM if u < 0.5 else S.JOIN_GROUPSUM if u < 0.8 else S.SORT
        qs.append((kind, int(r.choice(tables[kind])), t))
    return qs


def _expected(srv, S, kind, table):
    """The oracle's result for one table, from the table's own host columns."""
    if kind == S.FILTER_SUM:
        day, qty, price, disc = (srv.table_column(kind, table, c) for c in range(4))
        return oracle.filter_closed_form(day, qty, price, disc)
    if kind == S.JOIN_GROUPSUM:
        dk, dg, fk, fv = (srv.table_column(kind, table, c) for c in range(4))
        return oracle.join_groupsum_closed_form(dk, dg, fk, fv, max_groups=32)
    k, p = srv.table_column(kind, table, 0), srv.table_column(kind, table, 1)
    return oracle.sort_closed_form(k, p)


def _check(S, kind, out, want, n_rows):
    if kind == S.FILTER_SUM:
        m, s, p, d = want
        head = out[:32].view(np.uint64)
        got_s = int(head[2]) | (int(head[3]) << 64)
        assert int(head[0]) == m and got_s == s
        cols = out[32:].view(np.int64)
        assert np.array_equal(cols[:m], p) and np.array_equal(cols[n_rows:n_rows + m], d)
    elif kind == S.JOIN_GROUPSUM:
        g, s = want
        G = int(out[768:776].view(np.uint64)[0])
        assert G == len(g)
        assert np.array_equal(out[:256].view(np.int32)[:G], g)
        assert np.array_equal(out[256:768].view(np.int64)[:G], s)
    else:
        k, p = want
        assert np.array_equal(out[:8 * n_rows].view(np.uint64), k)
        assert np.array_equal(out[8 * n_rows:].view(np.uint32), p)


def _run_checked(srv, S, qs, want, rows, **kw):
    outs = [np.zeros(srv.query_out_bytes(k, t), np.uint8) for k, t, _ in qs]
    res, st = srv.run(qs, outputs=outs, **kw)
    assert all(r.status == 0 for r in res)
    for (k, t, _), o in zip(qs, outs):
        _check(S, k, o, want[(k, t)], rows[(k, t)])
    return res, st


def test_served_mix_tables_to_1e7_at_800qps_matches_the_oracle():
    import velodb_amd
    S = velodb_amd.Server
    srv = S(n_slots=16, hbm_budget_bytes=200 << 30)
    try:
        tables, sizes = _tables(srv, S)
        print(f"served mix tables {sizes}", flush=True)
        qs = _mix(S, tables, 11)
        used = {(k, t) for k, t, _ in qs}
        want = {kt: _expected(srv, S, *kt) for kt in used}
        rows = {}
        for k, t in used:
            col = 2 if k == S.JOIN_GROUPSUM else 0
            rows[(k, t)] = len(srv.table_column(k, t, col))
        print("served mix: 16 slots", flush=True)
        res, st = _run_checked(srv, S, qs, want, rows)
        assert st.max_concurrent > 1 and st.peak_bytes_in_flight <= 200 << 30
        print("served mix: plain loop", flush=True)
        _run_checked(srv, S, qs, want, rows, n_slots=1, transfer=True)
        print("served mix: resident command queue", flush=True)
        _, stq = _run_checked(srv, S, qs, want, rows, queue=True)
        assert stq.max_concurrent > 1
        # the timed form (no output copies) returns the same result words
        res2, _ = srv.run(qs)
        assert [list(a.r) for a in res2] == [list(b.r) for b in res]
    finally:
        srv.close()
