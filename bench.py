"""Benchmark: end-engine block compute on one MI355X.

Workload (the largest one-GPU share of the end engine): a box of k^4 blocks (8x8x8x8
nodes each) of the slice-24 section (3,3)(3,3)(3,3)(3,3), every open dimension computed.
The box's child blocks (standard child section (4,3)(3,3)(3,3)(3,3): every block along the
child dimension times the box cross-section) are resident in HBM as a dense slab filled with
seeded data (known ~ Bernoulli(0.9) bits, win ~ Bernoulli(0.5) & known).  One step = the four
dimension passes: each pass launches every block line crossing the box (restricted to the
box along its dimension), reading child blocks from the slab and OR-merging into the box
output.  Inputs are on the device before timing starts.

Prints one JSON line.  Metric: superpositions/s = box output nodes per second (the
reference's output nodes/second, pentago/mpi/toplevel.cc:83-94).
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


def build_box(P, hip, k, a0):
    """Allocate the box output and the child slab, and the four per-dimension plans."""
    shape = P.section_shape(SECTION)
    assert all(8 * (a0 + k) <= s for s in shape), "box must use full blocks"
    info0 = P.line_describe(P.make_line(SECTION, 0, (a0, a0, a0)))
    child = info0.standard_child.counts
    cshape = P.section_shape(child)
    cd_len = (cshape[info0.child_dimension] + 7) // 8
    # child slab: blocks (c_cd in [0, cd_len)) x box range on the three other child axes
    slab_axes = [range(cd_len) if ax == info0.child_dimension else range(a0, a0 + k) for ax in range(4)]
    slab_dims = [len(r) for r in slab_axes]
    n_out_blocks = k ** 4
    n_child_blocks = int(np.prod(slab_dims))
    out = hip.Buffer(64 * 4096 * n_out_blocks)
    kids = hip.Buffer(64 * 4096 * n_child_blocks)
    P.fill_seeded(kids.ptr, 4096 * n_child_blocks, 0x5EED0000 + sum(c * 10 ** i for i, c in enumerate(sum(child, ()))))
    out.zero()

    def child_offset(cb):
        idx = 0
        for ax in range(4):
            r = slab_axes[ax]
            assert r.start <= cb[ax] < r.stop, (cb, slab_axes)
            idx = idx * slab_dims[ax] + (cb[ax] - r.start)
        return 4096 * idx

    def out_offset(b):
        idx = 0
        for ax in range(4):
            idx = idx * k + (b[ax] - a0)
        return 4096 * idx

    plans = []
    child_nodes = []
    for d in range(4):
        jobs = []
        for b1 in range(a0, a0 + k):
            for b2 in range(a0, a0 + k):
                for b3 in range(a0, a0 + k):
                    base = (b1, b2, b3)
                    line = P.make_line(SECTION, d, base)
                    info = P.line_describe(line)
                    assert info.standard_child.counts == child
                    cd = info.child_dimension
                    in_off = []
                    for kk in range(info.input_blocks):
                        cb = list(info.first_child_block)
                        cb[cd] = kk
                        in_off.append(child_offset(cb))
                    full = [0] * 4
                    for ax, v in zip([ax for ax in range(4) if ax != d], base):
                        full[ax] = v
                    out_off = []
                    for kk in range(info.length):
                        full[d] = kk
                        out_off.append(out_offset(full) if a0 <= kk < a0 + k else 0)
                    jobs.append(dict(line=line, input=kids.ptr, output=out.ptr, in_offsets=in_off,
                                     out_offsets=out_off, i_range=(8 * a0, 8 * (a0 + k)), accumulate=d > 0))
        plans.append(P.Plan(jobs))
        child_nodes.append(cshape[info0.child_dimension] * (8 * k) ** 3)
    return out, kids, plans, child_nodes, child, n_child_blocks


def cpu_baseline(budget_s=15.0):
    """The oracle (CPU port of compute_microline) on full block lines of the same section,
    threaded one microline per job; superpositions/s = node-line rate / open dimensions."""
    from oracle import oracle as O
    from tests.seeded import seeded_nodes
    threads = min(os.cpu_count() or 1, 128)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--k", type=int, default=0, help="box edge in blocks (0: largest fitting ~200 GB)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

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
    out, kids, plans, child_nodes, child, n_child_blocks = build_box(P, hip, k, a0)
    hip.synchronize()
    log("setup %.1f s: %d lines per pass, child slab %d blocks (%.1f GB), box %.1f GB" % (
        time.time() - t0, k ** 3, n_child_blocks,
        64 * 4096 * n_child_blocks / 1e9, 64 * 4096 * k ** 4 / 1e9))

    stream = hip.Stream()
    n_out = (8 * k) ** 4
    ev = [hip.Event() for _ in range(5)]
    for _ in range(args.warmup):
        for p in plans:
            p.run(stream.ptr)
    stream.synchronize()
    hip.synchronize()
    launch_ms = [0.0] * 4
    t_start = time.time()
    ev_all0, ev_all1 = hip.Event(), hip.Event()
    ev_all0.record(stream)
    for step in range(args.steps):
        ev[0].record(stream)
        for d, p in enumerate(plans):
            p.run(stream.ptr)
            ev[d + 1].record(stream)
        ev[4].synchronize()    # per-pass kernel durations of this step, on the launch stream
        for d in range(4):
            launch_ms[d] += ev[d].elapsed_ms(ev[d + 1])
        log("step %d: %.1f ms" % (step, ev[0].elapsed_ms(ev[4])))
    ev_all1.record(stream)
    stream.synchronize()
    hip.synchronize()
    wall = time.time() - t_start
    total_ms = ev_all0.elapsed_ms(ev_all1)
    ms_per_step = total_ms / args.steps
    value = n_out / (ms_per_step / 1e3)
    avg_launch_ms = sum(launch_ms) / (4 * args.steps)
    # algorithmic bytes per launch (one dimension pass): box output once + that pass's child nodes once
    alg_bytes = [64 * (n_out + c) for c in child_nodes]
    achieved = (sum(alg_bytes) / 4) / (avg_launch_ms / 1e3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "end_lines_kernel_box_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            t = json.load(f)
        if t.get("k") == k:
            traffic = t.get("bytes_per_launch")
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
                               "4 dimension passes" % (k, "".join("%d%d" % c for c in child)),
                   "box_blocks": k ** 4, "output_nodes": n_out, "child_nodes_per_pass": child_nodes[0],
                   "lines_per_pass": k ** 3},
        "hbm_GBps_algorithmic": (sum(alg_bytes)) / (ms_per_step / 1e3) / 1e9,
        "pass_ms": [x / args.steps for x in launch_ms],
        "wall_s": wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "kernel": "end_lines_kernel", "bytes_per_launch": sum(alg_bytes) / 4,
                     "avg_launch_ms": avg_launch_ms},
    }
    for p in plans:
        p.close()
    del out, kids
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(args.cpu_budget)
        except Exception as e:   # the baseline is reported, never required
            result["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
