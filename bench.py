"""Headline benchmark: oblivious bitonic sort of 2^28 (u64 key, u32 payload)
records on one MI355X, through the library's C ABI.

A step sorts one batch of 2^28 records resident in HBM: the device
generator writes each timed step's keys (SplitMix64 stream of seed = step)
and row-index payloads before the timed region, up to six distinct batches
(3.22 GB each), so the timed region holds the sorts alone; inputs never
cross PCIe.  With more steps than batches a batch is sorted again, which
costs the same: the network's work does not depend on the input order.  Prints one JSON line (rank 0) with the roofline of the dominant
kernel family (HIP events around every launch, on the library's stream) and
a CPU baseline (the oracle's single-thread bitonic network on a bounded
sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# launch families of the sort passes: local, merge, register passes of 1-5
# stages, LDS-staged passes of 4-8 stages
SORT_FAMS = (["sort_local", "sort_merge"] + [f"sort_global_s{s}" for s in range(1, 6)] +
             [f"sort_hybrid_s{s}" for s in range(4, 9)])
METRIC = "oblivious sort Mrec/s at 2^28 (%HBM peak); served QPS + p50/p99 vs plain"
REC_BYTES = 12  # u64 key + u32 payload, structure of arrays


def cpu_baseline_parallel(log2n=28, runs=3):
    """The multi-threaded restatement of obliviator's recursive parallel
    bitonic (obliviator/join/enclave/bitonic.c:72-242) on the host cores the
    box gives this GPU (at most 16 threads), on the full headline workload:
    2^log2n records from the same generator (seed 1), median of `runs`
    sorts, each on a fresh copy (the copy is not timed).  Returns the
    baseline record and the sorted (keys, payload) of the last run."""
    import numpy as np

    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    n = 1 << log2n
    k0, p0 = oracle.gen_sort_input(1, n)
    wk, wp = k0[: 1 << 20].copy(), p0[: 1 << 20].copy()
    oracle.bitonic_sort_kp_mt_inplace(wk, wp, 1 << 20, threads)  # warm-up (thread start, page faults)
    wk, wp = np.empty_like(k0), np.empty_like(p0)
    times = []
    for _ in range(runs):
        np.copyto(wk, k0)
        np.copyto(wp, p0)
        t0 = time.perf_counter()
        oracle.bitonic_sort_kp_mt_inplace(wk, wp, n, threads)
        times.append(time.perf_counter() - t0)
    ok = bool(np.all(wk[1:] >= wk[:-1]))
    rate = n / float(np.median(times)) / 1e6
    return {"value": round(rate, 3), "unit": "Mrec/s", "cores": threads, "kind": "port",
            "sample": f"oracle multi-threaded recursive bitonic (obliviator bitonic.c structure) on the full "
                      f"2^{log2n}-record workload (seed 1), {threads} threads, median of {runs} runs "
                      f"({', '.join(f'{t:.2f}' for t in times)} s); sorted: {ok}"}, wk, wp


def cpu_baseline(budget_s=12.0):
    """Oracle bitonic network (reference host path semantics, one core) on
    2^20-record samples of the same workload, repeated for ~budget_s."""
    import numpy as np

    import oracle
    n = 1 << 20
    k, p = oracle.gen_sort_input(1, n)
    oracle.bitonic_sort_kp(k, p, n, n)  # warm-up
    times = []
    t_end = time.time() + budget_s
    while time.time() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        oracle.bitonic_sort_kp(k, p, n, n)
        times.append(time.perf_counter() - t0)
        if len(times) >= 50:
            break
    med = float(np.median(times))
    rate = n / med / 1e6
    return {"value": round(rate, 3), "unit": "Mrec/s", "cores": 1, "kind": "port",
            "sample": f"oracle single-thread bitonic network (reference (k,j,i) schedule, branch-free "
                      f"cmov) on 2^20 records, median of {len(times)} runs; per-record work grows with "
                      f"log2N(log2N+1)/2, so at 2^28 the same code would run at ~{rate * 210 / 406:.2f} Mrec/s"}


def measured_traffic(family, n, profile="sort_2p28_traffic.md", n_profiled=1 << 28):
    """HBM bytes per launch of `family` from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated) taken at
    n_profiled rows, or None at other sizes."""
    path = os.path.join(ROOT, "profiles", profile)
    if n != n_profiled or not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == family:
            return int(parts[2])
    return None


def calibrate_copy(steps):
    """Plain HBM copy of 3 GiB through the library (rocprof counter calibration)."""
    import torch

    import velodb_amd
    nbytes = 3 << 30
    src = torch.empty(nbytes // 8, dtype=torch.int64, device="cuda:0")
    dst = torch.empty_like(src)
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream)
    ctx.hbm_copy(dst, src, nbytes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.hbm_copy(dst, src, nbytes)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(json.dumps({"calibration": "hbm_copy", "bytes_each_way": nbytes, "ms": round(dt * 1e3, 3),
                      "GBps_read_plus_write": round(2 * nbytes / dt / 1e9, 1)}), flush=True)
    ctx.close()


# --------------------------------------------------------------------------
# Secondary workloads (one JSON line each; the headline line stays the sort).

def _family_stats(ctx, fams):
    out = {}
    for f in fams:
        cnt, ms = ctx.kernel_stats(f)
        if cnt:
            out[f] = (cnt, ms)
    return out


def _roofline(stats, bytes_per_row, n_pad, traffic=None):
    """Roofline of the family with the largest total time; bytes_per_row maps
    a family to its algorithmic bytes per padded row and launch; traffic =
    (profile file, rows it was taken at) of the measured HBM bytes."""
    cand = {f: v for f, v in stats.items() if f in bytes_per_row}
    dom = max(cand, key=lambda f: cand[f][1])
    cnt, ms = cand[dom]
    alg = bytes_per_row[dom] * n_pad
    ach = alg / (ms / cnt * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom, n_pad, *traffic) if traffic else None, "kernel": dom,
            "avg_launch_ms": round(ms / cnt, 4), "alg_bytes_per_launch": alg}


def _line(metric, value, unit, steps, warmup, ms, dtype, config, **extra):
    d = {"metric": metric, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
         "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
         "dtype": dtype, "data": "synthetic", "config": config}
    d.update(extra)
    print(json.dumps(d), flush=True)


def bench_filter(args):
    """Q6-style filter + ORCompact of (price, disc) + exact int128 SUM over a
    fact table resident in HBM (6e6 rows / 2^23, or 3e8 rows / 2^29)."""
    import numpy as np
    import torch

    import velodb_amd
    from velodb_amd import _abi
    n_pad = 1 << (args.log2n if args.log2n != 28 else 23)
    n = 6_000_000 if n_pad == 1 << 23 else (300_000_000 if n_pad == 1 << 29 else n_pad - n_pad // 4)
    dev = torch.device("cuda:0")
    day = torch.empty(n_pad, dtype=torch.int32, device=dev)
    qty, price, disc, op, od = (torch.empty(n_pad, dtype=torch.int64, device=dev) for _ in range(5))
    out = torch.empty(4, dtype=torch.int64, device=dev)
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, flags=_abi.CFG_KERNEL_TIMING,
                             tuning=args.tuning)
    ctx.gen_fact(1, day, qty, price, disc, n)

    def step():
        ctx.filter_sum(day, qty, price, disc, n, n_pad, out, out_price=op, out_disc=od)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.kernel_stats_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    o = out.cpu().numpy().view(np.uint64)
    stats = _family_stats(ctx, ["filter_count", "compact_scan", "compact_local", "compact_hybrid", "compact_global",
                                "filter_sum"])
    # the same WHERE clause as five generic device terms (OP_PREDICATE_MASK):
    # reads day, disc, qty once each (20 B/row) and writes the mask (1 B/row)
    mask = torch.empty(n_pad, dtype=torch.uint8, device=dev)
    terms = [(day, _abi.T_INTEGER, _abi.CMP_GE, 8766), (day, _abi.T_INTEGER, _abi.CMP_LT, 9131),
             (disc, _abi.T_BIGINT, _abi.CMP_GE, 5), (disc, _abi.T_BIGINT, _abi.CMP_LE, 7),
             (qty, _abi.T_BIGINT, _abi.CMP_LT, 24)]
    ctx.predicate_mask(terms, mask, n, n_pad)
    torch.cuda.synchronize()
    ctx.kernel_stats_reset()
    for _ in range(args.steps):
        ctx.predicate_mask(terms, mask, n, n_pad)
    torch.cuda.synchronize()
    pc, pms = ctx.kernel_stats("predicate_mask")
    pred_ms = pms / max(pc, 1)
    pred = {"ms": round(pred_ms, 4), "GBps": round(21 * n_pad / (pred_ms * 1e-3) / 1e9, 1),
            "selected_equal": int(mask.sum().item()) == int(o[0])}
    # algorithmic bytes per padded row and launch: predicate reads day+qty+disc
    # and writes the mask; the local network reads the mask and both columns
    # and writes the columns; a global or LDS-staged pass reads and writes both
    bpr = {"filter_count": 4 + 8 + 8 + 1, "compact_local": 1 + 2 * 16, "compact_hybrid": 2 * 16,
           "compact_global": 2 * 16, "filter_sum": 16}
    _line("oblivious filter + compaction + int128 SUM (Mrows/s)", n / dt / 1e6, "Mrows/s", args.steps,
          args.warmup, dt * 1e3, "int64/int128",
          {"workload": f"Q6-style filter over {n} rows padded to 2^{ilog2(n_pad)}, ORCompact of (price, disc), "
                       f"exact int128 SUM(price*disc)", "n": n, "n_pad": n_pad,
           "selected": int(o[0]), "sum": str(int(o[2]) | (int(o[3]) << 64))},
          generic_predicate_mask=pred,
          roofline=_roofline(stats, bpr, n_pad, traffic=("filter_2p23_traffic.md", 1 << 23)),
          kernels_ms_per_step={f: round(v[1] / args.steps, 3) for f, v in stats.items()},
          cpu_baseline=None if args.no_cpu_baseline else cpu_filter_baseline())
    ctx.close()


def cpu_filter_baseline(budget_s=8.0):
    """The oracle's filter + ORCompact + int128 SUM restatement on one core,
    on 2^20-row samples of the same column distributions."""
    import numpy as np

    import oracle
    n, npad = 786_432, 1 << 20
    cols = oracle.gen_fact(1, n)
    oracle.filter_sum(*cols, n, npad)
    times = []
    t_end = time.time() + budget_s
    while time.time() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        oracle.filter_sum(*cols, n, npad)
        times.append(time.perf_counter() - t0)
        if len(times) >= 20:
            break
    return {"value": round(n / float(np.median(times)) / 1e6, 3), "unit": "Mrows/s", "cores": 1, "kind": "port",
            "sample": f"oracle predicate + ORCompact (obliviator oblivious_compact.c restated) + int128 SUM, "
                      f"{n} rows padded to 2^20, one thread, median of {len(times)} runs"}


def ilog2(v):
    return v.bit_length() - 1


def bench_join(args):
    """PK/FK join + GROUP BY SUM: 1.5M dimension rows (unique keys 1..1.5M,
    group 0..24), 6M fact rows (FK uniform over dimension keys, value
    100..1e7); union padded to 2^23."""
    import numpy as np
    import torch

    import velodb_amd
    from velodb_amd import _abi
    n_dim, n_fact, n_pad = 1_500_000, 6_000_000, 1 << 23
    rng = np.random.default_rng(1)
    dk = (rng.permutation(n_dim) + 1).astype(np.int32)
    dg = rng.integers(0, 25, n_dim, dtype=np.int32)
    fk = dk[rng.integers(0, n_dim, n_fact)]
    fv = rng.integers(100, 10_000_001, n_fact, dtype=np.int64)
    dev = torch.device("cuda:0")
    d_dk, d_dg, d_fk, d_fv = (torch.from_numpy(a).to(dev) for a in (dk, dg, fk, fv))
    og = torch.empty(64, dtype=torch.int32, device=dev)
    osum = torch.empty(64, dtype=torch.int64, device=dev)
    on = torch.empty(1, dtype=torch.int64, device=dev)
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, flags=_abi.CFG_KERNEL_TIMING,
                             tuning=args.tuning)

    def step(dense):
        # the group attribute's public domain is 0..24: dense GROUP BY over 25 ids
        ctx.join_groupsum(d_dk, d_dg, n_dim, d_fk, d_fv, n_fact, n_pad, og, osum, on, 25 if dense else 64,
                          dense_groups=dense)

    def timed(dense):
        for _ in range(args.warmup):
            step(dense)
        torch.cuda.synchronize()
        ctx.kernel_stats_reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(dense)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    dt_sort = timed(False)
    ok_sort = None
    ng_sort = int(on.item())
    sums_sort = osum[:ng_sort].cpu().numpy().copy()
    dt = timed(True)
    # the host check is a plain numpy group-by over the same arrays
    grp = np.zeros(n_dim + 2, np.int64)
    grp[dk] = dg
    exp = np.bincount(grp[fk], minlength=25)
    sums = np.zeros(25, np.int64)
    np.add.at(sums, grp[fk], fv)
    ng = int(on.item())
    got_g = og[:ng].cpu().numpy()
    got_s = osum[:ng].cpu().numpy()
    ok = ng == int((exp > 0).sum()) and np.array_equal(got_s, sums[got_g])
    ok_sort = ng_sort == ng and np.array_equal(sums_sort, got_s)
    fams = ["join_build", "join_fill_tiles", "join_fill_rows", "join_fill_agg", "group_dense_out",
            "compact_count", "compact_local", "compact_hybrid", "compact_global"] + SORT_FAMS
    stats = _family_stats(ctx, fams)
    # the stats are the dense run's: its union sort moves (u64 key, u32
    # payload) records, 2 x 12 B per row and pass
    bpr = {f: 24 for f in fams if f.startswith("sort_")}
    _line("oblivious PK/FK join + GROUP BY SUM (Mrows/s)", (n_dim + n_fact) / dt / 1e6, "Mrows/s", args.steps,
          args.warmup, dt * 1e3, "int64",
          {"workload": "PK/FK join of 1.5M dimension x 6M fact rows (union padded to 2^23) + exact int64 "
                       "SUM per group (25 groups)", "n_dim": n_dim, "n_fact": n_fact, "n_pad": n_pad,
           "groups": ng, "group_by": "dense oblivious aggregation over the 25-id group domain",
           "matches_numpy_groupby": bool(ok)},
          sort_groupby_ms=round(dt_sort * 1e3, 3), sort_groupby_equal=bool(ok_sort),
          roofline=_roofline(stats, bpr, n_pad, traffic=("join_2p23_traffic.md", 1 << 23)),
          kernels_ms_per_step={f: round(v[1] / args.steps, 3) for f, v in stats.items()},
          cpu_baseline=None if args.no_cpu_baseline else cpu_join_baseline())
    ctx.close()


def cpu_join_baseline(budget_s=8.0):
    """The oracle's join + GROUP BY restatement (union bitonic sort, fill,
    ORCompact, group sort, segmented sum, ORCompact) on one core, on 1/8 of
    the join workload (187,500 x 750,000 rows, union padded to 2^20)."""
    import numpy as np

    import oracle
    nd, nf, npad = 187_500, 750_000, 1 << 20
    dk, dg, fk, fv = oracle.gen_join(1, nd, nf)
    oracle.join_groupsum(dk, dg, fk, fv, npad)
    times = []
    t_end = time.time() + budget_s
    while time.time() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        oracle.join_groupsum(dk, dg, fk, fv, npad)
        times.append(time.perf_counter() - t0)
        if len(times) >= 20:
            break
    return {"value": round((nd + nf) / float(np.median(times)) / 1e6, 3), "unit": "Mrows/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle PK/FK join + sort-based GROUP BY (obliviator scalable_oblivious_join.c and "
                      f"tpch_q3 group-by restated), {nd} x {nf} rows padded to 2^20, one thread, median of "
                      f"{len(times)} runs"}


def bench_join_pairs(args):
    """Oblivious equi-join into row-index pairs (COUNT + WRITE, the
    SortMergeJoinOperator path) on the join workload's key columns: 1.5M
    unique dimension keys x 6M foreign keys -> 6M pairs; union and output
    both padded to 2^23."""
    import numpy as np
    import torch

    import velodb_amd
    from velodb_amd import _abi
    n_dim, n_fact = 1_500_000, 6_000_000
    rng = np.random.default_rng(1)
    dk = (rng.permutation(n_dim) + 1).astype(np.int32)
    fk = dk[rng.integers(0, n_dim, n_fact)]
    dev = torch.device("cuda:0")
    d_dk, d_fk = torch.from_numpy(dk).to(dev), torch.from_numpy(fk).to(dev)
    cap = 1 << 23
    ol = torch.empty(cap, dtype=torch.int64, device=dev)
    orr = torch.empty(cap, dtype=torch.int64, device=dev)
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, flags=_abi.CFG_KERNEL_TIMING,
                             tuning=args.tuning)

    def step():
        n = ctx.join_pairs_count(d_dk, n_dim, d_fk, n_fact, _abi.T_INTEGER)
        ctx.join_pairs_write(d_dk, n_dim, d_fk, n_fact, _abi.T_INTEGER, ol, orr, cap)
        return n

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.kernel_stats_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # the output is the join: equal keys, each fact row once, ordered
    l, r = ol[:n], orr[:n]
    ok = n == n_fact and bool(torch.all(d_dk[l] == d_fk[r])) and int(torch.unique(r).numel()) == n_fact
    ok = ok and bool(torch.all(d_dk[l][1:] >= d_dk[l][:-1]))
    fams = ["join_pairs_build"] + SORT_FAMS + ["join_pairs_bwd_tiles",
            "join_pairs_bwd_carry", "join_pairs_bwd_rows", "join_pairs_fwd_carry", "join_pairs_fwd_rows",
            "compact_count", "compact_scan", "compact_local", "compact_hybrid", "compact_global",
            "join_pairs_dist_init_w1", "join_pairs_dist_init_w2", "join_pairs_dist_level_w1",
            "join_pairs_dist_level_w2", "join_pairs_dist_tile_w1", "join_pairs_dist_tile_w2",
            "join_pairs_fill_carry", "join_pairs_fill_left", "join_pairs_fill_right", "join_pairs_out"]
    stats = _family_stats(ctx, fams)
    # algorithmic bytes per padded row and launch: a distribution pass (three
    # levels) reads and writes every slot once (8 or 16 B each way; the rows
    # before a slot that it also reads are the same data, read by other
    # threads); the scans read the sorted union (16 B) and write their outputs
    bpr = {"join_pairs_dist_level_w1": 16, "join_pairs_dist_level_w2": 32, "join_pairs_bwd_rows": 20,
           "join_pairs_fwd_rows": 46}
    n_pad = 1 << 23
    _line("oblivious equi-join into row pairs (Mrows/s)", (n_dim + n_fact) / dt / 1e6, "Mrows/s", args.steps,
          args.warmup, dt * 1e3, "int32 keys / int64 row ids",
          {"workload": "equi-join of 1.5M unique keys x 6M foreign keys into 6M (left, right) row pairs "
                       "(union and output padded to 2^23), COUNT + WRITE", "n_left": n_dim, "n_right": n_fact,
           "n_out": n, "output_equals_join": bool(ok)},
          roofline=_roofline(stats, bpr, n_pad),
          kernels_ms_per_step={f: round(v[1] / args.steps, 3) for f, v in stats.items()})
    ctx.close()


def bench_transfer(args):
    """Host <-> device wire format: 1 GiB of int64 rows between pageable host
    memory and a device column, through vo_transfer_* (pinned ring, copy
    threads, DMA overlap), against the reference's StagedTransfer method
    (one 32 MiB pinned chunk: memcpy, DMA, synchronize; staged_transfer.hpp:
    101-139) and the HIP runtime's own pageable copy."""
    import numpy as np
    import torch

    import velodb_amd
    nbytes = 1 << 30
    host = np.random.default_rng(3).integers(0, 1 << 62, nbytes // 8, dtype=np.int64)
    back = np.empty_like(host)
    dev = torch.empty(nbytes // 8, dtype=torch.int64, device="cuda:0")
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream)
    chunk = 32 << 20
    pinned = torch.empty(chunk // 8, dtype=torch.int64, pin_memory=True)
    pin_np = pinned.numpy()
    th = torch.from_numpy(host)

    def ours_h2d():
        ctx.transfer_to_device(dev, host)

    def ours_d2h():
        ctx.transfer_to_host(back, dev)

    def ref_h2d():
        for off in range(0, host.size, chunk // 8):
            n = min(chunk // 8, host.size - off)
            pin_np[:n] = host[off:off + n]
            dev[off:off + n].copy_(pinned[:n], non_blocking=True)
            torch.cuda.current_stream().synchronize()

    def runtime_h2d():
        dev.copy_(th)
        torch.cuda.synchronize()

    def timed(fn):
        fn()
        ts = []
        for _ in range(max(2, args.steps)):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2]

    t_ours, t_ref, t_rt, t_down = timed(ours_h2d), timed(ref_h2d), timed(runtime_h2d), timed(ours_d2h)
    ours_h2d()
    ok = bool(np.array_equal(dev[:1 << 20].cpu().numpy(), host[:1 << 20]))
    ours_d2h()
    ok = ok and bool(np.array_equal(back, host))
    gbs = lambda t: round(nbytes / t / 1e9, 2)
    _line("pageable host -> device column, padded staged transfer (GB/s)", nbytes / t_ours / 1e9, "GB/s",
          args.steps, 1, t_ours * 1e3, "bytes",
          {"workload": "1 GiB of int64 rows, pageable host memory <-> device, padded to the column capacity",
           "bytes": nbytes, "round_trip_equal": ok},
          to_host_gbs=gbs(t_down), reference_staged_method_gbs=gbs(t_ref), hip_runtime_pageable_gbs=gbs(t_rt))
    ctx.close()


def bench_gather(args):
    """Oblivious gather (the device form of the host oblivious
    materialization): one int64 column of a 2^22-row table at 2^23 record
    ids with repeats."""
    import torch

    import velodb_amd
    from velodb_amd import _abi
    nt, n = 1 << 22, 1 << 23
    g = torch.Generator(device="cuda:0").manual_seed(5)
    table = torch.randint(-(1 << 62), 1 << 62, (nt,), device="cuda:0", generator=g, dtype=torch.int64)
    ids = torch.randint(0, nt, (n,), device="cuda:0", generator=g, dtype=torch.int64)
    ids[: n // 8] = 7
    out = torch.empty(n, dtype=torch.int64, device="cuda:0")
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, flags=_abi.CFG_KERNEL_TIMING,
                             tuning=args.tuning)

    def step():
        ctx.gather_oblivious(table, nt, ids, n, out, 8)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.kernel_stats_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ok = bool(torch.all(out == table[ids]))
    fams = ["gather_obl_build"] + SORT_FAMS + ["gather_obl_tiles", "gather_obl_carry", "gather_obl_rows",
            "compact_count", "compact_scan", "compact_local", "compact_hybrid", "compact_global", "gather_obl_order",
            "gather_obl_out"]
    stats = _family_stats(ctx, fams)
    # the same ids over a 4-byte (int32) column: 12-byte union records
    t32 = (table & 0x7FFFFFFF).to(torch.int32)
    o32 = torch.empty(n, dtype=torch.int32, device="cuda:0")
    ctx.gather_oblivious(t32, nt, ids, n, o32, 4)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ctx.gather_oblivious(t32, nt, ids, n, o32, 4)
    torch.cuda.synchronize()
    dt32 = (time.perf_counter() - t1) / args.steps
    ok32 = bool(torch.all(o32 == t32[ids]))
    _line("oblivious gather by record id (Mrows/s)", n / dt / 1e6, "Mrows/s", args.steps, args.warmup, dt * 1e3,
          "int64",
          {"workload": "one int64 column of a 2^22-row table gathered at 2^23 record ids (1/8 of them one id); "
                       "union of 2^22 + 2^23 rows padded to 2^24", "n_table": nt, "n": n, "equals_indexing": ok},
          int32_column={"ms": round(dt32 * 1e3, 3), "equals_indexing": ok32},
          kernels_ms_per_step={f: round(v[1] / args.steps, 3) for f, v in stats.items()})
    ctx.close()


def bench_order_by(args):
    """SortOperator path (sort_operator.cpp:44-123): OP_SORT of (INTEGER ASC,
    DOUBLE DESC) over 12M rows padded to 2^24 with int64 row indices, then
    OP_PERMUTE of two BIGINT payload columns by those indices."""
    import numpy as np
    import torch

    import velodb_amd
    from velodb_amd import _abi
    n, n_pad = 12_000_000, 1 << 24
    g = torch.Generator(device="cuda:0").manual_seed(9)
    a0 = torch.randint(-1000, 1000, (n_pad,), device="cuda:0", generator=g, dtype=torch.int32)
    b0 = torch.randn(n_pad, device="cuda:0", generator=g, dtype=torch.float64)
    p1 = torch.arange(n_pad, device="cuda:0", dtype=torch.int64)
    p2 = p1 * 3
    a, b = a0.clone(), b0.clone()
    idx = torch.empty(n_pad, dtype=torch.int64, device="cuda:0")
    o1, o2 = torch.empty_like(p1), torch.empty_like(p2)
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, flags=_abi.CFG_KERNEL_TIMING,
                             tuning=args.tuning)

    def step():
        a.copy_(a0)
        b.copy_(b0)
        ctx.sort([a, b], [_abi.T_INTEGER, _abi.T_DOUBLE], [True, False], idx, n, n_pad)
        ctx.permute(o1, p1, idx, n, _abi.T_BIGINT)
        ctx.permute(o2, p2, idx, n, _abi.T_BIGINT)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.kernel_stats_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # checked against torch: the first column ascending, the second descending
    # within equal first keys, and the permuted payload follows the indices
    i = idx[:n]
    ka, kb = a0[i], b0[i]
    ok = bool(torch.all(ka[1:] >= ka[:-1]))
    same = ka[1:] == ka[:-1]
    ok = ok and bool(torch.all(~same | (kb[1:] <= kb[:-1])))
    ok = ok and bool(torch.all(a[:n] == ka)) and bool(torch.all(o1[:n] == i)) and bool(torch.all(o2[:n] == i * 3))
    fams = ["sort_pack", "sort_unpack", "permute"] + SORT_FAMS
    stats = _family_stats(ctx, fams)
    _line("ORDER BY two keys + permute of two columns (Mrows/s)", n / dt / 1e6, "Mrows/s", args.steps,
          args.warmup, dt * 1e3, "int32/f64 keys, int64 indices",
          {"workload": "OP_SORT of (INTEGER ASC, DOUBLE DESC) over 12M rows padded to 2^24 with int64 row "
                       "indices, then OP_PERMUTE of two BIGINT columns (the SortOperator path)", "n": n,
           "n_pad": n_pad, "order_equals_torch_check": ok},
          kernels_ms_per_step={f: round(v[1] / args.steps, 3) for f, v in stats.items()})
    ctx.close()


def bench_shuffle(args):
    """Oblivious shuffle of a position map (ObliviousTable::generatePermutation
    + compose, src/data/oblivious_table.cpp:240-255): Philox4x32-10 tag per
    row and a bitonic sort that carries the map entries, the form the table
    layer runs (no separate permutation output: 12-byte (tag, entry)
    records, sorted in place).  2^28 rows by default.  The map starts as the
    identity, so the output is the permutation itself and is checked as one.
    The 16-byte form that also writes the permutation is timed beside it."""
    import torch

    import velodb_amd
    from velodb_amd import _abi
    n_pad = 1 << args.log2n
    dev = torch.device("cuda:0")
    ident = torch.arange(n_pad, device=dev, dtype=torch.int64).to(torch.int32)
    pay = torch.empty_like(ident)
    perm = torch.empty(n_pad, dtype=torch.int32, device=dev)
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, flags=_abi.CFG_KERNEL_TIMING,
                             tuning=args.tuning)

    def timed(seed0, with_perm):
        t_ms = 0.0
        for i in range(args.steps):
            pay.copy_(ident)  # outside the events: only the library's launches are timed
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            ctx.shuffle(pay, n_pad, seed0 + i, perm_out=perm if with_perm else None)
            e1.record()
            torch.cuda.synchronize()
            t_ms += e0.elapsed_time(e1)
        return t_ms / 1e3 / args.steps

    for i in range(args.warmup):
        pay.copy_(ident)
        ctx.shuffle(pay, n_pad, 77 + i)
    torch.cuda.synchronize()
    ctx.kernel_stats_reset()
    dt = timed(77 + args.warmup, False)
    fams = ["shuffle_tags", "shuffle_unpack"] + SORT_FAMS
    stats = _family_stats(ctx, fams)
    # the map after the last timed shuffle is a permutation of the rows
    pl = pay.long()
    ok = bool(torch.equal(torch.sort(pl).values, torch.arange(n_pad, device=dev)))
    moved = float((pl != torch.arange(n_pad, device=dev)).float().mean())
    del pl
    # the 16-byte form with the permutation written out, and its agreement
    dt_perm = timed(1000, True)
    same = bool(torch.equal(pay, perm))  # identity map: both outputs are the permutation
    # per padded row and launch: a sort pass reads and writes the 12-byte
    # (tag, entry) record; the tag kernel writes the 8-byte key
    bpr = {f: 24 for f in fams if f.startswith("sort_")}
    bpr["shuffle_tags"] = 8
    _line("oblivious shuffle of a position map (Mrows/s)", n_pad / dt / 1e6, "Mrows/s", args.steps, args.warmup,
          dt * 1e3, "u64 tags, u32 map entries",
          {"workload": f"Philox4x32-10 tag shuffle of a 2^{args.log2n}-entry position map (u32), in place",
           "n_pad": n_pad, "permutation_check": ok, "fraction_moved": round(moved, 6)},
          with_perm_out={"ms": round(dt_perm * 1e3, 3), "record_bytes": 16, "map_equals_perm_out": same},
          roofline=_roofline(stats, bpr, n_pad, traffic=("shuffle_2p28_traffic.md", 1 << 28)),
          kernels_ms_per_step={f: round(v[1] / args.steps, 3) for f, v in stats.items()})
    ctx.close()


def bench_batch(args):
    """4096 independent sorts, sizes log-uniform in [2^10, 2^20]
    (numpy default_rng(0)), each padded to its power of two; one batched call
    vs one call per instance."""
    import numpy as np
    import torch

    import velodb_amd
    rng = np.random.default_rng(0)
    sizes = np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), 4096)).astype(np.int64)
    items, off = [], 0
    for n in sizes:
        p = 1 << int(np.ceil(np.log2(n)))
        off = (off + p - 1) // p * p
        items.append((off, int(n), p))
        off += p
    total = off
    padded = sum(p for _, _, p in items)
    dev = torch.device("cuda:0")
    keys = torch.empty(total, dtype=torch.int64, device=dev)
    pay = torch.empty(total, dtype=torch.int32, device=dev)
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, tuning=args.tuning)
    if args.autotune:
        ctx.autotune_sort(10, 20, rows_per_class=1 << 24)

    def run(fn, reps):
        ts = []
        for _ in range(reps):
            ctx.gen_sort_input(7, keys, pay, total)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    def batched():
        ctx.sort_batch(keys, pay, items)

    def per_instance():
        for (o, n, p) in items:
            ctx.sort_kp_async(keys[o:], pay[o:], n, p)

    run(batched, max(1, args.warmup))
    tb = run(batched, args.steps)
    kb, pb = keys.clone(), pay.clone()
    torch.cuda.synchronize()
    tp = run(per_instance, max(1, min(args.steps, 3)))
    same = bool(torch.equal(kb, keys) and torch.equal(pb, pay))
    # roofline over every sort pass of one batched call: each instance of
    # padded size p moves 24 B x p per pass, and a class runs the pass plan of
    # its size, so the call's algorithmic bytes are sum(24 p passes(p)).  The
    # size classes run on several streams forked from the context stream and
    # joined back into it, so the call is timed by HIP events on that stream
    # around the whole call; per-launch events (a second context) would add up
    # overlapping kernels.
    from velodb_amd import _abi
    plan = {}
    for _, _, p in items:
        if p not in plan:
            plan[p] = ctx.sort_plan(p)
    alg = sum(2 * 12 * p * plan[p] for _, _, p in items)
    ctx.gen_sort_input(7, keys, pay, total)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    ctx.sort_batch(keys, pay, items)
    e1.record()
    torch.cuda.synchronize()
    call_ms = e0.elapsed_time(e1)
    tctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, flags=_abi.CFG_KERNEL_TIMING,
                              tuning=args.tuning)
    ctx.gen_sort_input(7, keys, pay, total)
    tctx.kernel_stats_reset()
    tctx.sort_batch(keys, pay, items)
    torch.cuda.synchronize()
    fams = ["sort_tiny"] + SORT_FAMS
    st = _family_stats(tctx, fams)
    k_ms = sum(v[1] for v in st.values())
    tctx.close()
    ach = alg / (call_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "every sort pass of one batched call (local, merge, global, hybrid), timed as the whole call",
            "call_ms": round(call_ms, 3), "launch_ms_summed_over_streams": round(k_ms, 3),
            "alg_bytes_per_call": alg, "passes_by_class": {str(p): plan[p] for p in sorted(plan)},
            "tuning": "default"}
    resident = _batch_resident(args, ctx, keys, pay, total, items, kb, pb, tp)
    compact = _compact_batch(args)
    _line("batched oblivious sorts (Mrec/s over padded records)", padded / tb / 1e6, "Mrec/s", args.steps,
          args.warmup, tb * 1e3, "u64",
          {"workload": "4096 sorts, sizes log-uniform in [2^10, 2^20] (numpy default_rng(0)), u64 keys + u32 "
                       "payload, padded per instance, one call bucketed by size class",
           "instances": len(items), "raw_records": int(sizes.sum()), "padded_records": padded,
           "autotuned": bool(args.autotune)},
          per_instance_ms=round(tp * 1e3, 2), speedup_vs_per_instance=round(tp / tb, 2),
          batched_equals_per_instance=same, roofline=roof, resident_one_launch=resident, compact_batch=compact)
    ctx.close()


def _batch_resident(args, ctx, keys, pay, total, items, kb, pb, tp):
    """The same batch as one command of the resident queue kernel (one
    persistent launch; its size classes share the grid between barriers):
    submit -> completion, inputs generated before each run."""
    import ctypes as C

    import numpy as np
    import torch

    import velodb_amd
    from velodb_amd import _abi
    rctx = velodb_amd.Context(flags=_abi.CFG_PERSISTENT, queue_capacity=64)
    arr = (_abi.BatchItem * len(items))(*[_abi.BatchItem(o, n, p) for (o, n, p) in items])
    cmd = _abi.Command()
    cmd.opcode = _abi.OP_SORT_BATCH
    cmd.args.sort_batch = _abi.SortBatchArgs(keys.data_ptr(), pay.data_ptr(), C.cast(arr, C.c_void_p), len(items))
    ts = []
    for _ in range(max(1, args.warmup) + args.steps):
        ctx.gen_sort_input(7, keys, pay, total)
        torch.cuda.synchronize()  # before the resident kernel holds the CUs
        rctx.queue_start()
        t0 = time.perf_counter()
        rctx.wait(rctx.submit(cmd), 120000)
        ts.append(time.perf_counter() - t0)
        rctx.queue_stop()
    torch.cuda.synchronize()
    tr = float(np.median(ts[max(1, args.warmup):]))
    same = bool(torch.equal(kb, keys) and torch.equal(pb, pay))
    rctx.close()
    return {"ms": round(tr * 1e3, 2), "Mrec_s": round(sum(p for _, _, p in items) / tr / 1e6, 1),
            "speedup_vs_per_instance": round(tp / tr, 2), "equals_launch_mode_batch": same,
            "note": "one OP_SORT_BATCH command of the resident queue (512 blocks of 256 threads, 2^12-record "
                    "tiles); submit to completion, including the host recording of its program"}


def _compact_batch(args):
    """1024 independent ORCompact compactions, sizes log-uniform in
    [2^12, 2^20] (numpy default_rng(1)), two columns (u64, u32) and a u8 mask
    (30% selected, torch generator seed 5): one OP_COMPACT_BATCH call vs one
    vo_compact per item, and the same batch as one resident command."""
    import ctypes as C

    import numpy as np
    import torch

    import velodb_amd
    from velodb_amd import _abi
    rng = np.random.default_rng(1)
    sizes = np.exp(rng.uniform(np.log(2 ** 12), np.log(2 ** 20), 1024)).astype(np.int64)
    items, off = [], 0
    for n in sizes:
        p = 1 << int(np.ceil(np.log2(max(n, 4096))))
        off = (off + p - 1) // p * p
        items.append((off, int(n), p))
        off += p
    total = off
    padded = sum(p for _, _, p in items)
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(5)
    c0_in = torch.randint(-2**62, 2**62, (total,), device=dev, generator=g)
    c1_in = torch.randint(-2**31, 2**31 - 1, (total,), device=dev, dtype=torch.int32, generator=g)
    mask = (torch.rand(total, device=dev, generator=g) < 0.3).to(torch.uint8)
    c0, c1 = c0_in.clone(), c1_in.clone()
    counts = torch.zeros(len(items), dtype=torch.int64, device=dev)
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream)

    def reset():
        c0.copy_(c0_in)
        c1.copy_(c1_in)
        torch.cuda.synchronize()

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            reset()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    a, arr = velodb_amd.Context.compact_batch_args(mask, [c0, c1], [8, 4], items, counts)
    batched = lambda: ctx.check(ctx.lib.vo_compact_batch(ctx.h, C.byref(a)))  # noqa: E731
    timed(batched, max(1, args.warmup))
    tb = timed(batched, args.steps)
    r0, r1, rc = c0.clone(), c1.clone(), counts.clone()
    one = torch.zeros(len(items), dtype=torch.int64, device=dev)
    cmds = []
    for k, (o, n, p) in enumerate(items):
        cmd = _abi.Command()
        cmd.opcode = _abi.OP_COMPACT
        ca = cmd.args.compact
        ca.mask = mask.data_ptr() + o
        ca.cols[0], ca.cols[1] = c0.data_ptr() + 8 * o, c1.data_ptr() + 4 * o
        ca.col_bytes[0], ca.col_bytes[1] = 8, 4
        ca.n_cols, ca.n, ca.n_pad = 2, n, p
        ca.out_count = one.data_ptr() + 8 * k
        cmds.append(cmd)

    def per_item():
        for cmd in cmds:
            ctx.submit(cmd)
        ctx.sync()

    tp = timed(per_item, max(1, min(args.steps, 3)))
    same = bool(torch.equal(r0, c0) and torch.equal(r1, c1) and torch.equal(rc, one))
    # the same batch as one resident command
    rctx = velodb_amd.Context(flags=_abi.CFG_PERSISTENT, queue_capacity=64)
    rcmd = _abi.Command()
    rcmd.opcode = _abi.OP_COMPACT_BATCH
    rcmd.args.compact_batch = a
    ts = []
    for _ in range(max(1, args.warmup) + args.steps):
        reset()
        rctx.queue_start()
        t0 = time.perf_counter()
        rctx.wait(rctx.submit(rcmd), 120000)
        ts.append(time.perf_counter() - t0)
        rctx.queue_stop()
    torch.cuda.synchronize()
    tr = float(np.median(ts[max(1, args.warmup):]))
    same_r = bool(torch.equal(r0, c0) and torch.equal(r1, c1) and torch.equal(rc, counts))
    rctx.close()
    ctx.close()
    # algorithmic bytes: the count pass reads the mask; the local pass reads
    # the mask and reads + writes both columns; every further pass of the
    # network (levels 12 .. log2 p - 1 in groups of at most 8) reads + writes
    # both columns: 12 B per row each way
    def passes(p):
        lv = int(np.log2(p)) - 12
        return 1 + (lv + 7) // 8
    alg = sum(p * (1 + 1 + 24 * passes(p)) for _, _, p in items)
    return {"workload": "1024 compactions, sizes log-uniform in [2^12, 2^20] (numpy default_rng(1)), columns "
                        "u64 + u32, u8 mask with 30% selected",
            "padded_rows": padded, "ms": round(tb * 1e3, 2), "Mrows_s": round(padded / tb / 1e6, 1),
            "alg_GBps": round(alg / tb / 1e9, 1), "per_item_ms": round(tp * 1e3, 2),
            "speedup_vs_per_item": round(tp / tb, 2), "batched_equals_per_item": same,
            "resident_one_launch_ms": round(tr * 1e3, 2), "resident_equals_launch_mode": same_r}


def bench_queue(args):
    """Command round trip (submit -> wait returns) through the resident queue
    kernel vs launch mode: NOP commands and 4096-record sorts."""
    import numpy as np
    import torch

    import velodb_amd
    from velodb_amd import _abi
    dev = torch.device("cuda:0")
    n = 4096
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    pay = torch.empty(n, dtype=torch.int32, device=dev)
    reps = max(50, args.steps * 40)
    res = {}
    for mode, flags in (("launch", 0), ("persistent", _abi.CFG_PERSISTENT)):
        ctx = velodb_amd.Context(flags=flags)
        ctx.gen_sort_input(3, keys, pay, n)
        ctx.sync()
        ctx.queue_start()
        for kind in ("nop", "sort4096"):
            cmd = _abi.Command()
            if kind == "nop":
                cmd.opcode = _abi.OP_NOP
            else:
                cmd.opcode = _abi.OP_SORT_KP
                cmd.args.sort_kp = _abi.SortKpArgs(keys.data_ptr(), pay.data_ptr(), n, n)
            for _ in range(10):
                ctx.wait(ctx.submit(cmd))
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                ctx.wait(ctx.submit(cmd))
                ts.append(time.perf_counter() - t0)
            res[f"{mode}_{kind}_us"] = {"p50": round(float(np.median(ts)) * 1e6, 1),
                                        "p99": round(float(np.percentile(ts, 99)) * 1e6, 1)}
        ctx.queue_stop()
        ctx.close()
    p50 = res["persistent_sort4096_us"]["p50"]
    d = {"metric": "command round-trip latency, 4096-record sort (us, p50)", "value": p50, "unit": "us",
         "n_gpus": 1, "steps": reps, "warmup": 10, "ms_per_step": round(p50 / 1e3, 4), "higher_is_better": False,
         "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
         "config": {"workload": "submit + wait of one command through the resident queue kernel vs launch mode"},
         "latency": res}
    print(json.dumps(d), flush=True)


def served_workload(slots=16, rates=(0, 800), n_tables=12, check_queries=200, log=None, queue_lanes=0):
    """The served mix (50% filter+SUM, 30% join+group-by, 20% sort) over
    n_tables tables of each kind with row counts log-uniform in [1e4, 1e7],
    200 GB admission budget.  For each offered load (0 = every query arrives
    at once) the same query list runs three ways: served (`slots` concurrent
    slots over the resident tables), plain (one query at a time, inputs
    shipped from pinned host memory per query: the reference engine's path)
    plain_resident (one at a time over the resident tables) and served_queue
    (every query one program in the resident command queue, slot i on lane
    i mod lanes, `queue_lanes` lanes or min(slots, 8)).  The first
    `check_queries` queries of the first load then run served again with
    their whole results copied out, each checked against the oracle's closed
    form of its table."""
    import numpy as np

    import oracle
    import velodb_amd
    S = velodb_amd.Server
    srv = S(n_slots=slots, hbm_budget_bytes=200 << 30, queue_lanes=queue_lanes)
    rng = np.random.default_rng(0)
    tables = {"f": [], "j": [], "s": []}
    for i in range(n_tables):
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(1e7))))
        tables["f"].append(srv.add_fact_table(100 + i, n))
        tables["j"].append(srv.add_join_tables(200 + i, max(1, n // 5), n - n // 5))
        tables["s"].append(srv.add_sort_table(300 + i, n))

    def mix(count, rate, seed):
        r = np.random.default_rng(seed)
        t, qs = 0.0, []
        for _ in range(count):
            t += r.exponential(1.0 / rate) if rate else 0.0
            u = r.random()
            kind, key = ((S.FILTER_SUM, "f") if u < 0.5 else (S.JOIN_GROUPSUM, "j") if u < 0.8 else (S.SORT, "s"))
            qs.append((kind, int(r.choice(tables[key])), t))
        return qs

    names = {S.FILTER_SUM: "filter_sum", S.JOIN_GROUPSUM: "join_groupsum", S.SORT: "sort"}

    def summ(st, res=None, qs=None):
        d = {"qps": round(st.qps, 1), "p50_ms": round(st.p50_s * 1e3, 2), "p99_ms": round(st.p99_s * 1e3, 2),
             "max_concurrent": st.max_concurrent, "peak_GB": round(st.peak_bytes_in_flight / 1e9, 2),
             "graph_replays": st.graph_replays}
        if res is not None:
            # mean time from admission to completion, per query kind
            per = {}
            for r, q in zip(res, qs):
                per.setdefault(names[q[0]], []).append(r.end_s - r.admit_s)
            d["service_ms_by_kind"] = {k: round(1e3 * sum(v) / len(v), 3) for k, v in per.items()}
        return d

    srv.run(mix(64, 0, 99))  # warm-up
    srv.run(mix(16, 0, 98), n_slots=1, transfer=True)
    srv.run(mix(64, 0, 97), queue=True)
    out = {"slots": slots, "tables_per_kind": n_tables, "loads": []}
    first_qs = None
    all_equal = True
    for rate in rates:
        count = 600 if rate == 0 else max(200, rate * 2)
        qs = mix(count, rate, 7 + rate)
        first_qs = first_qs or qs
        rp, sp = srv.run(qs, n_slots=1, transfer=True)
        rr, sr = srv.run(qs, n_slots=1)
        rb, sb = srv.run(qs)
        rq, sq = srv.run(qs, queue=True)
        same = [list(a.r) for a in rp] == [list(b.r) for b in rb] == [list(c.r) for c in rr] == \
            [list(d.r) for d in rq]
        all_equal = all_equal and same
        ld = {"offered_qps": rate if rate else "all at t=0", "queries": count,
              "served": summ(sb, rb, qs), "plain": summ(sp, rp, qs), "plain_resident": summ(sr, rr, qs),
              "served_queue": summ(sq, rq, qs),
              "throughput_ratio": round(sb.qps / max(sp.qps, 1e-9), 2),
              "throughput_ratio_queue": round(sq.qps / max(sp.qps, 1e-9), 2),
              "throughput_ratio_vs_plain_resident": round(sb.qps / max(sr.qps, 1e-9), 2),
              "p99_ratio_plain_over_served": round(sp.p99_s / max(sb.p99_s, 1e-12), 2),
              "results_equal": same}
        out["loads"].append(ld)
        if log:
            log(ld)

    # oracle check of whole results, served concurrently
    qs = first_qs[:check_queries]
    want = {}
    for k, t, _ in qs:
        if (k, t) in want:
            continue
        cols = [srv.table_column(k, t, c) for c in range(2 if k == S.SORT else 4)]
        if k == S.FILTER_SUM:
            want[(k, t)] = (len(cols[0]),) + oracle.filter_closed_form(*cols)
        elif k == S.JOIN_GROUPSUM:
            want[(k, t)] = (len(cols[2]),) + oracle.join_groupsum_closed_form(*cols, max_groups=32)
        else:
            want[(k, t)] = (len(cols[0]),) + oracle.sort_closed_form(*cols)
    outs = [np.zeros(srv.query_out_bytes(k, t), np.uint8) for k, t, _ in qs]
    res, _ = srv.run(qs, outputs=outs)
    ok = all(r.status == 0 for r in res)
    for (k, t, _), o in zip(qs, outs):
        w = want[(k, t)]
        n = w[0]
        if k == S.FILTER_SUM:
            m, sm, pr, di = w[1:]
            h = o[:32].view(np.uint64)
            c = o[32:].view(np.int64)
            ok = ok and int(h[0]) == m and (int(h[2]) | (int(h[3]) << 64)) == sm
            ok = ok and np.array_equal(c[:m], pr) and np.array_equal(c[n:n + m], di)
        elif k == S.JOIN_GROUPSUM:
            g, sm = w[1:]
            G = int(o[768:776].view(np.uint64)[0])
            ok = ok and G == len(g) and np.array_equal(o[:256].view(np.int32)[:G], g)
            ok = ok and np.array_equal(o[256:768].view(np.int64)[:G], sm)
        else:
            kk, pp = w[1:]
            ok = ok and np.array_equal(o[:8 * n].view(np.uint64), kk) and np.array_equal(o[8 * n:].view(np.uint32), pp)
    out["results_equal"] = all_equal
    out["oracle_checked_queries"] = len(qs)
    out["results_equal_oracle"] = bool(ok)
    srv.close()
    return out


def bench_serve(args):
    """One line per offered load (the served loop vs the plain loops)."""
    rates = tuple(int(x) for x in args.rates.split(","))

    def log(ld):
        sb = ld["served"]
        _line(f"served queries/s (continuous batching, {args.slots} slots)", sb["qps"], "QPS", ld["queries"], 64,
              1e3 / max(sb["qps"], 1e-9), "int64/int128",
              {"workload": "served mix 50% filter+SUM / 30% join+group-by / 20% sort, tables log-uniform "
                           "in [1e4, 1e7] rows, 200 GB admission budget",
               "offered_qps": ld["offered_qps"], "queries": ld["queries"]},
              **{k: v for k, v in ld.items() if k not in ("offered_qps", "queries")})

    res = served_workload(args.slots, rates, log=log, queue_lanes=args.queue_lanes)
    print(json.dumps({"served_oracle_check": {k: res[k] for k in ("oracle_checked_queries",
                                                                  "results_equal_oracle", "results_equal")}}),
          flush=True)
    if not (res["results_equal"] and res["results_equal_oracle"]):
        sys.exit(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-served", action="store_true", help="sort line without the served sub-object")
    ap.add_argument("--tuning", type=int, default=0)
    ap.add_argument("--workload", choices=["sort", "filter", "join", "join_pairs", "batch", "serve", "queue",
                                           "transfer", "gather", "order_by", "shuffle"],
                    default="sort",
                    help="sort = the headline line; the others print their own line")
    ap.add_argument("--autotune", action="store_true", help="batch: run the autotuner first")
    ap.add_argument("--slots", type=int, default=16, help="serve: concurrent query slots")
    ap.add_argument("--queue-lanes", type=int, default=0,
                    help="serve: lanes of the resident queue in queue mode (0: min(slots, 8))")
    ap.add_argument("--rates", type=str, default="0,50,200,800",
                    help="serve: offered loads in QPS (0 = every query arrives at once)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run the plain 3 GiB HBM copy kernel instead (traffic-counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(max(1, args.steps))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the oblivious sort is not sharded across GPUs "
              "(no cross-GPU exchange stages are built); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload != "sort":
        {"filter": bench_filter, "join": bench_join, "join_pairs": bench_join_pairs, "batch": bench_batch,
         "serve": bench_serve, "queue": bench_queue, "transfer": bench_transfer,
         "gather": bench_gather, "order_by": bench_order_by, "shuffle": bench_shuffle}[args.workload](args)
        return

    import torch

    import velodb_amd
    from velodb_amd import _abi

    n = 1 << args.log2n
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    ctx = velodb_amd.Context(stream=stream.cuda_stream, flags=_abi.CFG_KERNEL_TIMING, tuning=args.tuning)
    n_batches = max(1, min(args.steps, 6))
    batches = [(torch.empty(n, dtype=torch.int64, device=dev), torch.empty(n, dtype=torch.int32, device=dev))
               for _ in range(n_batches)]

    for i in range(args.warmup):  # untimed, on the first batch
        keys, pay = batches[0]
        ctx.gen_sort_input(1000 + i, keys, pay, n)
        ctx.sort_kp_async(keys, pay, n, n)
    for b, (keys, pay) in enumerate(batches):  # the timed steps' inputs, resident before timing
        ctx.gen_sort_input(1000 + args.warmup + b, keys, pay, n)
    torch.cuda.synchronize()
    ctx.kernel_stats_reset()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        keys, pay = batches[i % n_batches]
        ctx.sort_kp_async(keys, pay, n, n)
    torch.cuda.synchronize()
    t1 = time.perf_counter()

    total_s = t1 - t0
    ms_per_step = total_s * 1e3 / args.steps
    value = n * args.steps / total_s / 1e6

    # per-family kernel time inside the timed region
    fams = {}
    # one family per kernel instantiation, so each maps to one rocprof kernel name
    for f in (["sort_local", "sort_merge", "gen_sort"] + [f"sort_global_s{s}" for s in range(1, 6)] +
              [f"sort_hybrid_s{s}" for s in range(4, 9)]):
        cnt, ms = ctx.kernel_stats(f)
        if cnt:
            fams[f] = (cnt, ms)
    sort_fams = {f: v for f, v in fams.items() if f.startswith("sort_")}
    dom = max(sort_fams, key=lambda f: sort_fams[f][1])
    cnt, ms = sort_fams[dom]
    avg_ms = ms / cnt
    alg_bytes = 2 * REC_BYTES * n  # every record read and written once per pass
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    passes = ctx.sort_plan(n)
    sort_ms = sum(v[1] for v in sort_fams.values()) / args.steps
    per_pass = {f: round(alg_bytes / (v[1] / v[0] * 1e-3) / 1e9, 1) for f, v in sort_fams.items()}

    # parity of every timed batch, on the device: ascending by (key, payload)
    # as unsigned integers, the payloads a permutation of the row indices,
    # and every key the input key of its payload's row (regenerated)
    ok = True
    ref_k = torch.empty(n, dtype=torch.int64, device=dev)
    ref_p = torch.empty(n, dtype=torch.int32, device=dev)
    top = -(2 ** 63)
    for b, (keys, pay) in enumerate(batches):
        ku = keys ^ top
        pl = pay.long()
        asc = (ku[1:] > ku[:-1]) | ((ku[1:] == ku[:-1]) & (pl[1:] > pl[:-1]))
        ok = ok and bool(torch.all(asc).item())
        ok = ok and bool(torch.equal(torch.sort(pay)[0], torch.arange(n, dtype=torch.int32, device=dev)))
        ctx.gen_sort_input(1000 + args.warmup + b, ref_k, ref_p, n)
        ok = ok and bool(torch.equal(ref_k[pl], keys))
        del ku, pl, asc
    del ref_k, ref_p

    # the GPU against the CPU baseline's own output: seed 1 sorted on both
    cpu_equal = None
    cpu_par = None
    if not args.no_cpu_baseline:
        import numpy as np
        keys, pay = batches[0]
        ctx.gen_sort_input(1, keys, pay, n)
        ctx.sort_kp_async(keys, pay, n, n)
        torch.cuda.synchronize()
        gk = keys.cpu().numpy().view(np.uint64)
        gp = pay.cpu().numpy().view(np.uint32)
        del batches
        torch.cuda.empty_cache()
        cpu_par, ck, cp = cpu_baseline_parallel(args.log2n)
        cpu_equal = bool(np.array_equal(gk, ck) and np.array_equal(gp, cp))
        del gk, gp, ck, cp
    else:
        del batches

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrec/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (SplitMix64 keys from seed = step, payload = row index, generated on device "
                "before the timed region)",
        "config": {"workload": f"oblivious bitonic sort of 2^{args.log2n} records (u64 key, u32 payload), "
                               f"ascending by key, ties by payload; {n_batches} input batches resident in HBM",
                   "n": n, "record_bytes": REC_BYTES, "layout": "SoA", "hbm_passes": passes,
                   "sort_ms_per_step": round(sort_ms, 3), "output_sorted_permutation_keys_travel": ok,
                   "bit_exact_vs_cpu_baseline_seed1": cpu_equal},
        "pct_hbm_peak": round(100 * achieved / HBM_PEAK_GBS, 1),
        "per_pass_GBps": per_pass,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, n), "kernel": dom,
                     "avg_launch_ms": round(avg_ms, 4), "alg_bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
        line["cpu_baseline_parallel"] = cpu_par
    ctx.close()
    served_ok = True
    if not args.no_served:
        # the served half of the metric: saturation and 800 QPS, oracle-checked
        sv = served_workload(args.slots, (0, 800))
        line["served"] = sv
        served_ok = sv["results_equal"] and sv["results_equal_oracle"]
    print(json.dumps(line), flush=True)
    if not ok or cpu_equal is False or not served_ok:
        print("bench.py: an output check failed (see the line above)", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
