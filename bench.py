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
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(args.cpu_budget)
        except Exception as e:   # the baseline is reported, never required
            result["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
