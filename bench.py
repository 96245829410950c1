"""Benchmark: end-engine block compute on one MI355X.

Workload (the largest one-GPU share of the end engine): a box of k^4 blocks (8x8x8x8
nodes each) of the slice-24 section (3,3)(3,3)(3,3)(3,3), every open dimension computed.
The box's child blocks (standard child section (4,3)(3,3)(3,3)(3,3): every block along the
child dimension times the box cross-section) are resident in HBM as a dense slab filled with
seeded data (known ~ Bernoulli(0.9) bits, win ~ Bernoulli(0.5) & known).  One step = one
pass of the end engine over the box: mode 0 (default) launches the fused block kernel once
(every node gets all four dimensions and is written once); mode 1 launches four line passes,
one per dimension, OR-merging into the box.  Inputs are on the device before timing starts.

Prints one JSON line.  Metric: superpositions/s = box output nodes per second (the
reference's output nodes/second, pentago/mpi/toplevel.cc:83-94).

--copy-probe BYTES runs one plain device copy of BYTES bytes instead (a kernel of known
traffic, used to scale the FETCH_SIZE / WRITE_SIZE counters).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SECTION = ((3, 3), (3, 3), (3, 3), (3, 3))
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_box(P, hip, k, a0, mode=0):
    """Allocate the box output slab and the child slab (seeded) and build the plan."""
    shape = P.section_shape(SECTION)
    assert all(8 * (a0 + k) <= s for s in shape), "box must use full blocks"
    info0 = P.line_describe(P.make_line(SECTION, 0, (a0, a0, a0)))
    child = info0.standard_child.counts
    cshape = P.section_shape(child)
    cd = info0.child_dimension
    cd_len = (cshape[cd] + 7) // 8
    n_out_blocks = k ** 4
    n_child_blocks = cd_len * k ** 3
    out = hip.Buffer(64 * 4096 * n_out_blocks)
    kids = hip.Buffer(64 * 4096 * n_child_blocks)
    P.fill_seeded(kids.ptr, 4096 * n_child_blocks, 0x5EED0000 + sum(c * 10 ** i for i, c in enumerate(sum(child, ()))))
    out.zero()
    out_slab = P.Slab.of(SECTION, [a0] * 4, [a0 + k] * 4, out.ptr)
    lo = [0 if a == cd else a0 for a in range(4)]
    hi = [cd_len if a == cd else a0 + k for a in range(4)]
    kid_slab = P.Slab.of(child, lo, hi, kids.ptr)
    plan = P.Plan.slab(out_slab, [kid_slab], mode)
    # every dimension's lines read the whole child slab once (the box is a cube)
    child_nodes = cshape[cd] * (8 * k) ** 3
    return out, kids, plan, child_nodes, child, n_child_blocks


def cpu_baseline(budget_s=15.0):
    """The oracle (CPU port of compute_microline) on full block lines of the same section,
    threaded one microline per job; superpositions/s = node-line rate / open dimensions."""
    from oracle import oracle as O
    from tests.seeded import seeded_nodes
    threads = min(os.cpu_count() or 1, 16)
    lines, ins, outs = [], [], []
    nodes = 0
    for d in range(4):
        for base in [(5, 7, 9), (11, 3, 2)]:
            line = O.make_line(SECTION, d, base)
            info = O.line_describe(line)
            lines.append(line)
            ins.append(seeded_nodes(int(info.input_nodes), 99 + d))
            outs.append(np.zeros(8 * info.output_nodes, dtype=np.uint64))
            nodes += info.output_nodes
    # calibrate with one line, then run as many full rounds as fit the budget
    t0 = time.time()
    O.compute_lines_threaded(lines[:1], ins[:1], outs[:1], threads)
    one = time.time() - t0
    rounds = max(1, int(budget_s / max(one * len(lines), 1e-3)))
    t0 = time.time()
    for _ in range(rounds):
        O.compute_lines_threaded(lines, ins, outs, threads)
    dt = time.time() - t0
    rate = rounds * nodes / dt          # node-line contributions per second
    return {"value": rate / 4, "unit": "superpositions/s", "cores": threads, "kind": "port",
            "sample": "%d full block lines of section (3,3)^4 (2 per dimension, %d node contributions), "
                      "%d rounds in %.1f s; rate / 4 open dimensions" % (len(lines), nodes, rounds, dt)}


def copy_probe(nbytes):
    """One device-to-device copy of nbytes through torch's elementwise copy kernel."""
    import torch
    n = nbytes // 4
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    assert bool((y[-4:] == 1).all())
    print(json.dumps({"copy_probe_bytes": 4 * n, "read": 4 * n, "written": 4 * n}), flush=True)


def mid_bench(P, hip, n_boards=4096, seed=2):
    """Batched midsolve: n seeded boards with 18..30 stones in one work-queue launch
    (autotuned tile and workgroups) against one plain pe_midsolve_internal call per board;
    every batched result is compared with the plain one.  Also the 26-stone single board."""
    from pentago_amd.workloads import board_spots, seeded_boards
    out = {}
    # single board, 13 + 13 stones (plain per-board launch)
    one = seeded_boards(1, 26, 26, 26)[0]
    P.midsolve_internal(one)
    lat = []
    for _ in range(20):
        t0 = time.perf_counter()
        P.midsolve_internal(one)
        lat.append(1e3 * (time.perf_counter() - t0))
    out["single_26_stones_ms"] = float(np.median(lat))
    boards = seeded_boards(n_boards, seed, 18, 30)
    sample = boards[:512]
    t0 = time.time()
    tuned = P.mid_autotune(sample)
    out["autotune"] = dict(tuned, seconds=time.time() - t0, boards=len(sample))
    log("midsolve autotune: %s" % out["autotune"])
    t0 = time.time()
    plan = P.MidPlan(boards)
    info = plan.info()
    log("midsolve plan: %s, built in %.1f s" % (info, time.time() - t0))
    stream = hip.Stream()
    plan.run(stream.ptr)                       # warm-up
    res = plan.results(stream.ptr)
    ev0, ev1 = hip.Event(), hip.Event()
    reps = 2
    ms = []
    for _ in range(reps):
        ev0.record(stream)
        plan.run(stream.ptr)
        ev1.record(stream)
        ev1.synchronize()
        ms.append(ev0.elapsed_ms(ev1))
    res2 = plan.results(stream.ptr)
    plan.close()
    batched_s = min(ms) / 1e3
    log("midsolve batched: %s ms" % ms)
    # plain: one synchronous launch sequence per board
    plain = np.zeros_like(res)
    t0 = time.perf_counter()
    for i, b in enumerate(boards):
        P.midsolve_internal_raw(b, plain[i])
    plain_s = time.perf_counter() - t0
    log("midsolve plain: %.2f s" % plain_s)
    spots = np.array([board_spots(b) for b in boards])
    mism = int(np.sum(np.any(plain != res, axis=(1, 2)))) + int(np.sum(np.any(res2 != res, axis=(1, 2))))
    work = 0
    from math import comb
    for sp in spots:
        sp = int(sp)
        work += 2 * 16 * sum(comb(sp, n // 2) * comb(sp - n // 2, n - n // 2) * (sp - n + 1) for n in range(sp + 1))
    # small boards alone (24..30 stones, 6..12 empty cells): launch overhead dominates the
    # plain path, the batched kernel solves them one workgroup per board
    small = seeded_boards(n_boards, seed + 1, 24, 30)
    plan = P.MidPlan(small)
    plan.run(stream.ptr)
    sres = plan.results(stream.ptr)
    ev0.record(stream)
    plan.run(stream.ptr)
    ev1.record(stream)
    ev1.synchronize()
    small_ms = ev0.elapsed_ms(ev1)
    plan.close()
    splain = np.zeros_like(sres)
    t0 = time.perf_counter()
    for i, b in enumerate(small):
        P.midsolve_internal_raw(b, splain[i])
    small_plain_s = time.perf_counter() - t0
    out["small_boards"] = {"workload": "%d seeded boards, 24..30 stones (6..12 empty cells)" % n_boards,
                           "batched_boards_per_s": n_boards / (small_ms / 1e3), "batched_ms": small_ms,
                           "plain_boards_per_s": n_boards / small_plain_s,
                           "batched_over_plain": small_plain_s / (small_ms / 1e3),
                           "mismatches": int(np.sum(np.any(splain != sres, axis=(1, 2))))}
    log("midsolve small boards: %s" % out["small_boards"])
    out.update({
        "config": {"workload": "%d seeded boards, 18..30 stones (6..18 empty cells)" % n_boards,
                   "spots_histogram": {int(k): int(v) for k, v in zip(*np.unique(spots, return_counts=True))}},
        "batched_boards_per_s": n_boards / batched_s, "batched_ms": min(ms), "batched_ms_runs": ms,
        "plain_boards_per_s": n_boards / plain_s, "plain_s": plain_s,
        "batched_over_plain": plain_s / batched_s,
        "mismatches": mism, "plan": info,
        "algorithmic_GBps": work / batched_s / 1e9,
    })
    return out


def serve_bench(P, rates=(200, 1000, 5000), seed=3, seconds=2.0):
    """Serving loop under Poisson arrivals: p50/p99 latency and solves/s per rate; every
    answer compared with a work-queue batch solve of the same boards."""
    import threading
    from pentago_amd.workloads import seeded_boards
    out = {}
    for rate in rates:
        n = int(min(4000, max(400, rate * seconds)))
        boards = seeded_boards(n, seed + rate, 18, 34)
        rng = np.random.default_rng(seed * 7 + rate)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
        server = P.MidServer()
        tickets = [None] * n
        t_start = time.perf_counter()

        def feed():
            for i, b in enumerate(boards):
                delay = arrivals[i] - (time.perf_counter() - t_start)
                if delay > 0:
                    time.sleep(delay)
                tickets[i] = server.submit(b)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        got, lat = [], []
        for t in tickets:
            r, l = server.wait(t)
            got.append(r)
            lat.append(l)
        wall = time.perf_counter() - t_start
        stats = server.stats()
        server.close()
        plan = P.MidPlan(boards)
        plan.run()
        ref = P.results_to_lists(boards, plan.results())
        plan.close()
        mism = sum(1 for a, b in zip(got, ref) if a != b)
        lat = np.array(lat)
        out[str(rate)] = {"requests": n, "p50_ms": float(np.percentile(lat, 50)),
                          "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                          "solves_per_s": n / wall, "offered_per_s": n / float(arrivals[-1]),
                          "batches": stats["batches"], "peak_workspace_GB": stats["peak_workspace"] / 1e9,
                          "mismatches": mism}
        log("serving %d req/s: %s" % (rate, out[str(rate)]))
    return out


def mid_cpu_baseline(n=48, seed=2):
    """The oracle's midsolve (one board per thread) on a sample of the batched workload."""
    from oracle import oracle as O
    from pentago_amd.workloads import seeded_boards
    threads = min(os.cpu_count() or 1, 16)
    boards = seeded_boards(4096, seed, 18, 30)[:n]
    roots = [O.High((O.C.c_uint64 * 2)(b.side[0], b.side[1]), b.ply) for b in boards]
    t0 = time.time()
    O.midsolve_batch(roots, threads)
    dt = time.time() - t0
    return {"value": n / dt, "unit": "boards/s", "cores": threads, "kind": "port",
            "sample": "first %d boards of the batched workload (18..30 stones), one board per thread, %.1f s"
                      % (n, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--k", type=int, default=0, help="box edge in blocks (0: largest fitting ~200 GB)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", type=int, default=0, help="0: fused block kernel, 1: four line passes")
    ap.add_argument("--copy-probe", type=int, default=0, help="run one plain copy of this many bytes instead")
    ap.add_argument("--no-mid", action="store_true", help="skip the batched midsolve measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import pentago_amd as P
    from pentago_amd import hip
    P.init()
    free, total = hip.mem_info()
    shape = P.section_shape(SECTION)
    k = args.k
    if not k:
        budget = min(200e9, 0.8 * free)
        k = 1
        while True:
            kk = k + 1
            need = 64 * 4096 * (kk ** 4 + 40 * kk ** 3)
            if need > budget or 8 * kk > shape[0] - 8:
                break
            k = kk
    a0 = 0
    log("box k=%d (%d output blocks), free HBM %.1f GB" % (k, k ** 4, free / 1e9))
    t0 = time.time()
    out, kids, plan, child_nodes, child, n_child_blocks = build_box(P, hip, k, a0, args.mode)
    hip.synchronize()
    log("setup %.1f s: %d lines per pass, child slab %d blocks (%.1f GB), box %.1f GB" % (
        time.time() - t0, k ** 3, n_child_blocks,
        64 * 4096 * n_child_blocks / 1e9, 64 * 4096 * k ** 4 / 1e9))

    stream = hip.Stream()
    n_out = (8 * k) ** 4
    assert plan.nodes == n_out
    for _ in range(args.warmup):
        plan.run(stream.ptr)
    stream.synchronize()
    hip.synchronize()
    ev0, ev1 = hip.Event(), hip.Event()
    step_ms = []
    t_start = time.time()
    for step in range(args.steps):
        ev0.record(stream)
        plan.run(stream.ptr)
        ev1.record(stream)
        ev1.synchronize()
        step_ms.append(ev0.elapsed_ms(ev1))
        log("step %d: %.1f ms" % (step, step_ms[-1]))
    stream.synchronize()
    hip.synchronize()
    wall = time.time() - t_start
    ms_per_step = sum(step_ms) / args.steps
    value = n_out / (ms_per_step / 1e3)
    launches = 1 if args.mode == 0 else 4
    avg_launch_ms = ms_per_step / launches
    # algorithmic bytes per step: box output once + every dimension's child nodes once
    alg_step = 64 * (n_out + 4 * child_nodes)
    achieved = (alg_step / launches) / (avg_launch_ms / 1e3) / 1e9
    kernel = "fused_blocks_kernel" if args.mode == 0 else "end_lines_kernel"
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "%s_box%d_traffic.json" % (kernel, k))
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    result = {
        "metric": "superpositions/s",
        "value": value,
        "unit": "superpositions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded child blocks: known ~ Bernoulli(0.9), win ~ Bernoulli(0.5) & known)",
        "config": {"workload": "end-engine box of %d^4 blocks of section 24-(3,3)^4, child slab of section 25-%s, "
                               "all 4 dimensions" % (k, "".join("%d%d" % c for c in child)),
                   "box_blocks": k ** 4, "output_nodes": n_out, "child_nodes_per_dimension": child_nodes, "mode": "fused" if args.mode == 0 else "line passes",
                   "lines_per_pass": k ** 3},
        "hbm_GBps_algorithmic": alg_step / (ms_per_step / 1e3) / 1e9,
        "step_ms": step_ms,
        "wall_s": wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "kernel": kernel, "bytes_per_launch": alg_step / launches,
                     "avg_launch_ms": avg_launch_ms},
    }
    plan.close()
    del out, kids
    if not args.no_mid:
        result["midsolve"] = mid_bench(P, hip)
    if not args.no_serve:
        result["serving"] = serve_bench(P)
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(args.cpu_budget)
        except Exception as e:   # the baseline is reported, never required
            result["cpu_baseline"] = {"error": str(e)}
        if not args.no_mid:
            try:
                result["midsolve"]["cpu_baseline"] = mid_cpu_baseline()
            except Exception as e:
                result["midsolve"]["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
