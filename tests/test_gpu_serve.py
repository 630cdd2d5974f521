"""Serving loop: continuous batching over device-resident tables.  Results
of every query are identical alone and in the concurrent mix; admission
respects the HBM budget; concurrency raises throughput over the plain
one-query-at-a-time loop."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _server(n_slots=8, budget=200 << 30, queue_lanes=0):
    import velodb_amd
    s = velodb_amd.Server(n_slots=n_slots, hbm_budget_bytes=budget, queue_lanes=queue_lanes)
    rng = np.random.default_rng(0)
    tables = {"f": [], "j": [], "s": []}
    for i in range(4):
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
        tables["f"].append(s.add_fact_table(100 + i, n))
        tables["j"].append(s.add_join_tables(200 + i, max(1, n // 5), n - n // 5))
        tables["s"].append(s.add_sort_table(300 + i, n))
    return s, tables


def _mix(tables, n, rate, seed):
    import velodb_amd
    rng = np.random.default_rng(seed)
    t = 0.0
    qs = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate) if rate else 0.0
        u = rng.random()
        kind, key = ((velodb_amd.Server.FILTER_SUM, "f") if u < 0.5 else
                     (velodb_amd.Server.JOIN_GROUPSUM, "j") if u < 0.8 else (velodb_amd.Server.SORT, "s"))
        qs.append((kind, int(rng.choice(tables[key])), t))
    return qs


def test_results_identical_alone_and_in_the_mix():
    s, tables = _server()
    try:
        qs = _mix(tables, 60, 0, 1)  # all arrive at once: maximum overlap
        mixed, st = s.run(qs)
        alone, _ = s.run(qs, n_slots=1)
        assert all(r.status == 0 for r in mixed) and all(r.status == 0 for r in alone)
        assert [list(r.r) for r in mixed] == [list(r.r) for r in alone]
        assert st.max_concurrent > 1
    finally:
        s.close()


def test_transfer_mode_matches_resident():
    """The reference-style path (inputs shipped from pinned host memory per
    query) computes the same results as the resident tables."""
    s, tables = _server(n_slots=4)
    try:
        qs = _mix(tables, 30, 0, 4)
        res, _ = s.run(qs)
        tr, _ = s.run(qs, n_slots=1, transfer=True)
        assert all(r.status == 0 for r in tr)
        assert [list(r.r) for r in res] == [list(r.r) for r in tr]
    finally:
        s.close()


def test_admission_respects_budget():
    s, tables = _server(n_slots=8, budget=300 << 20)
    try:
        qs = _mix(tables, 40, 0, 2)
        res, st = s.run(qs)
        ok = [r for r in res if r.status == 0]
        assert st.peak_bytes_in_flight <= 300 << 20
        # queries whose own footprint exceeds the budget are rejected, the rest run
        assert all(r.footprint_bytes > (300 << 20) for r in res if r.status != 0)
        assert len(ok) > 0
    finally:
        s.close()


def test_concurrency_raises_throughput():
    s, tables = _server(n_slots=16)
    try:
        qs = _mix(tables, 200, 0, 3)
        _, plain = s.run(qs, n_slots=1)
        _, batched = s.run(qs)
        assert batched.qps > 1.5 * plain.qps, (batched.qps, plain.qps)
    finally:
        s.close()


def test_graph_replays_equal_eager_results():
    """Resident runs replay the HIP graph captured per (slot, kind, table).
    On one slot the first run of each (kind, table) pair is eager and every
    later one a replay; runs on one and on four slots, with eager
    transfer-mode work and synchronizations between them, each give the eager
    results for every query."""
    s, tables = _server(n_slots=4)
    try:
        qs = _mix(tables, 48, 0, 7)
        pairs = len({(k, t) for k, t, _ in qs})
        eager, _ = s.run(qs, n_slots=1, transfer=True)
        assert all(r.status == 0 for r in eager)
        want = [list(r.r) for r in eager]
        got, st = s.run(qs, n_slots=1)
        assert [list(r.r) for r in got] == want
        assert st.graph_replays == len(qs) - pairs  # the replay path runs
        other = _mix(tables, 12, 0, 8)
        replays = 0
        for rnd in range(3):
            between, _ = s.run(other, n_slots=1, transfer=True)  # eager work between replays
            assert all(r.status == 0 for r in between)
            got, st = s.run(qs, n_slots=4 if rnd < 2 else 1)
            assert [list(r.r) for r in got] == want, rnd
            replays += st.graph_replays
        assert replays > 0
    finally:
        s.close()


def test_queue_mode_matches_the_graph_loop():
    """Queue mode: every admitted query runs as one program in the resident
    command queue, on its slot's lane; the results equal the graph-replay
    loop's and the eager transfer loop's, with queries in flight on several
    lanes at once."""
    s, tables = _server(n_slots=8)
    try:
        qs = _mix(tables, 60, 0, 9)
        want, _ = s.run(qs, n_slots=1, transfer=True)
        got, st = s.run(qs, queue=True)
        assert all(r.status == 0 for r in got)
        assert [list(r.r) for r in got] == [list(r.r) for r in want]
        assert st.max_concurrent > 1
        again, _ = s.run(qs, n_slots=3, queue=True)  # 3 lanes
        assert [list(r.r) for r in again] == [list(r.r) for r in want]
    finally:
        s.close()


def test_queue_mode_with_fewer_lanes_than_slots():
    """queue_lanes = 2 with 8 slots: slots share lanes (slot i on lane i mod
    2), the commands of one lane run in order, and the results are the graph
    loop's; more than 8 lanes is rejected."""
    import velodb_amd
    s, tables = _server(n_slots=8, queue_lanes=2)
    try:
        qs = _mix(tables, 40, 0, 11)
        want, _ = s.run(qs)
        got, _ = s.run(qs, queue=True)
        assert all(r.status == 0 for r in got)
        assert [list(r.r) for r in got] == [list(r.r) for r in want]
    finally:
        s.close()
    with pytest.raises(velodb_amd.VoError):
        velodb_amd.Server(n_slots=2, queue_lanes=9)
