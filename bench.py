"""Benchmark: end-engine block compute on one MI355X (plus batched midsolve and serving).

Headline workload (the largest one-GPU share of the end engine): a box of k^4 blocks
(8x8x8x8 nodes each) of the slice-24 section (3,3)(3,3)(3,3)(3,3), mixed in the same
launch with a box of k2^4 blocks (default 8^4) of the second slice-24 section
(4,2)(2,4)(3,3)(3,3) (standard form (3,3)(2,4)(3,3)(4,2)), every open dimension computed.
The child blocks the boxes' lines read (slice-25 child sections, every block along the
child dimension times the box cross-section) are resident in HBM as dense slabs filled with
seeded data (known ~ Bernoulli(0.9) bits, win ~ Bernoulli(0.5) & known).  One step = one
pass of the end engine over both boxes: mode 0 (default) launches the fused block kernel
once (every node gets all its dimensions and is written once); mode 1 runs four line
passes per box, OR-merging into the output.  Inputs are on the device before timing starts.

Prints one JSON line.  Metric: superpositions/s = output nodes per second (the reference's
output nodes/second, pentago/mpi/toplevel.cc:83-94).  The same line carries the whole
slice-30 section, whole slices 35, 34, 33 and 32 on real data (computed from nothing, counts
checked against the production run's), the batched midsolve (boards/s against plain
per-board launches) and serving (p50/p99 latency at Poisson loads) measurements unless
--no-section30 / --no-slices / --no-mid / --no-serve.

--copy-probe BYTES runs one plain device copy of BYTES bytes instead (a kernel of known
traffic, used to scale the FETCH_SIZE / WRITE_SIZE counters); --gather-probe runs the fused
kernel's access pattern alone (hashed 64-byte gathers, no arithmetic) for its ceiling.
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
# standard form of the second section (4,2)(2,4)(3,3)(3,3), mixed into the same launch
SECTION2 = ((3, 3), (2, 4), (3, 3), (4, 2))
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


def child_slabs_for(P, counts, lo, hi):
    """The child slabs (standard child section, block lo, block hi) that the lines through
    the output box [lo, hi) read, and the algorithmic child nodes of those lines (each child
    line counted once per line)."""
    shape = P.section_shape(counts)
    needs = []
    child_nodes = 0
    for d in range(4):
        if sum(counts[d]) >= 9:
            continue
        info = P.line_describe(P.make_line(counts, d, tuple(lo[a] for a in range(4) if a != d)))
        child = info.standard_child.counts
        cshape = P.section_shape(child)
        cd = info.child_dimension
        perm = list(info.permutation)
        clo = [0 if c == cd else lo[perm[c]] for c in range(4)]
        chi = [(cshape[c] + 7) // 8 if c == cd else hi[perm[c]] for c in range(4)]
        key = (child, tuple(clo), tuple(chi))
        if key not in needs:
            needs.append(key)
        cross = 1
        for a in range(4):
            if a != d:
                cross *= min(shape[a], 8 * hi[a]) - 8 * lo[a]
        child_nodes += cshape[cd] * cross
    return needs, child_nodes


def boxes_bytes(P, specs):
    total = 0
    seen = set()
    for counts, lo, hi in specs:
        total += 64 * 4096 * int(np.prod([h - l for l, h in zip(lo, hi)]))
        for key in child_slabs_for(P, counts, lo, hi)[0]:
            if key not in seen:
                seen.add(key)
                total += 64 * 4096 * int(np.prod([h - l for l, h in zip(key[1], key[2])]))
    return total


def build_boxes(P, hip, specs, mode=0):
    """Output slabs for each (section, lo, hi) box and the seeded child slabs they read, as
    one plan.  Returns (outputs, children, plan, output nodes, child nodes)."""
    outs, kids, kid_keys = [], [], []
    out_nodes = child_nodes = 0
    for counts, lo, hi in specs:
        shape = P.section_shape(counts)
        assert all(0 <= lo[a] < hi[a] and 8 * (hi[a] - 1) < shape[a] for a in range(4))
        nblk = int(np.prod([h - l for l, h in zip(lo, hi)]))
        buf = hip.Buffer(64 * 4096 * nblk)
        buf.zero()
        outs.append((counts, list(lo), list(hi), buf))
        needs, cn = child_slabs_for(P, counts, lo, hi)
        child_nodes += cn
        out_nodes += int(np.prod([min(shape[a], 8 * hi[a]) - 8 * lo[a] for a in range(4)]))
        for key in needs:
            if key in kid_keys:
                continue
            child, clo, chi = key
            nb = int(np.prod([h - l for l, h in zip(clo, chi)]))
            kb = hip.Buffer(64 * 4096 * nb)
            seed = 0x5EED0000 + sum(c * 10 ** i for i, c in enumerate(sum(child, ()))) + 7919 * len(kid_keys)
            P.fill_seeded(kb.ptr, 4096 * nb, seed)
            kid_keys.append(key)
            kids.append((child, list(clo), list(chi), kb))
    plan = P.Plan.slab([P.Slab.of(c, lo, hi, b.ptr) for c, lo, hi, b in outs],
                       [P.Slab.of(c, lo, hi, b.ptr) for c, lo, hi, b in kids], mode)
    assert plan.nodes == out_nodes
    return outs, kids, plan, out_nodes, child_nodes


def host_info():
    """The host the CPU baselines run on: core counts and CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"nproc": os.cpu_count(), "affinity": affinity, "cpu_model": model}


def baseline_threads():
    """Host threads for the all-core CPU figures: the cores this process may use, at most 16
    (one GPU's share of the host's cores)."""
    return max(1, min(host_info()["affinity"] or os.cpu_count() or 1, 16))


def oracle_for_timing():
    """The oracle library the CPU baselines time: the -O3 x86-64-v3 (AVX2) build of the same
    C sources where the host supports it, else the test build (-O2)."""
    from oracle import oracle as O
    fast = O.fast_lib()
    build = ("oracle C sources, gcc -O3 -march=x86-64-v3 (AVX2); portable C, not the reference's SSE build "
             "(whose compute kernel is 1.81-2.25x its own 64-bit path, paper/pentago.tex:648)") if fast else \
        "oracle C sources, gcc -O2 (host without AVX2); portable C, not the reference's SSE build"
    return O, fast, build


def timed_lines(O, L, lines, ins, outs, threads, budget_s):
    """Node-line contributions per second of compute_lines_threaded over whole rounds that
    fit the budget (at least one)."""
    nodes = sum(int(O.line_describe(l).output_nodes) for l in lines)
    t0 = time.time()
    O.compute_lines_threaded(lines, ins, outs, threads, L=L)
    one = time.time() - t0
    rounds = max(1, int(budget_s / max(one, 1e-3)) - 1)
    t0 = time.time()
    for _ in range(rounds):
        O.compute_lines_threaded(lines, ins, outs, threads, L=L)
    dt = time.time() - t0 + one
    return (rounds + 1) * nodes / dt, rounds + 1, dt


def cpu_baseline(budget_s=15.0):
    """The oracle's threaded port of compute_microline (one job per microline, as
    pentago/end/compute.cc:378-404 schedules it) on full block lines of the headline section
    (3,3)^4: all-core and single-thread superpositions/s (node-line rate / 4 open dimensions);
    plus lines of the slice-30 section ((3,4),(4,4),(4,4),(4,3)) with the whole-section time
    extrapolated from them."""
    from tests.seeded import seeded_nodes
    O, L, build = oracle_for_timing()
    threads = baseline_threads()

    def lines_of(section, bases, seed):
        lines, ins, outs = [], [], []
        for d in range(4):
            if sum(section[d]) >= 9:
                continue
            for base in bases:
                line = O.make_line(section, d, base)
                info = O.line_describe(line)
                lines.append(line)
                ins.append(seeded_nodes(int(info.input_nodes), seed + d))
                outs.append(np.zeros(8 * info.output_nodes, dtype=np.uint64))
        return lines, ins, outs

    lines, ins, outs = lines_of(SECTION, [(5, 7, 9), (11, 3, 2)], 99)
    rate, rounds, dt = timed_lines(O, L, lines, ins, outs, threads, 0.5 * budget_s)
    rate1, rounds1, dt1 = timed_lines(O, L, lines[::2], ins[::2], outs[::2], 1, 0.2 * budget_s)
    # slice-30 section lines: 2 per dimension, whole-section time extrapolated
    s30 = ((3, 4), (4, 4), (4, 4), (4, 3))
    shape = O.section_shape(s30)
    l30, i30, o30 = lines_of(s30, [(3, 9, 17), (7, 2, 13)], 123)
    r30, rounds30, dt30 = timed_lines(O, L, l30, i30, o30, threads, 0.2 * budget_s)
    section_nodes = int(np.prod(shape))
    open_dims = sum(1 for d in range(4) if sum(s30[d]) < 9)
    return {"value": rate / 4, "unit": "superpositions/s", "cores": threads, "kind": "port",
            "sample": "%d full block lines of section (3,3)^4 (2 per dimension), %d rounds in %.1f s on %d threads; "
                      "superpositions/s = node-line rate / 4 open dimensions" % (len(lines), rounds, dt, threads),
            "single_thread": {"value": rate1 / 4, "unit": "superpositions/s", "cores": 1,
                              "sample": "%d of those lines, %d rounds in %.1f s" % (len(lines[::2]), rounds1, dt1)},
            "section30": {"section": "(3,4)(4,4)(4,4)(4,3) (slice 30, 2,588,774,400 nodes)",
                          "sample": "%d full block lines (2 per dimension), %d rounds in %.1f s on %d threads" % (
                              len(l30), rounds30, dt30, threads),
                          "superpositions_per_s": r30 / open_dims,
                          "extrapolated_section_s": section_nodes * open_dims / r30,
                          "extrapolated_section_s_single_thread": section_nodes * open_dims / (r30 * rate1 / rate),
                          "note": "extrapolated from the sampled lines (the whole section is not run on the CPU); "
                                  "single-thread: the all-core rate scaled by the (3,3)^4 single/all-core ratio"},
            "host": host_info(), "build": build}


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


def gather_probe(nodes=976289792, src_nodes=1 << 31, steps=3, windows=(0, 40 * 4096)):
    """The fused kernel's memory access pattern without its arithmetic
    (pentago_amd/libpentago_probe.so, a measurement probe outside the engine): one thread per
    output node of the headline launch reads 12 hashed 64-byte nodes of a 137 GB buffer
    (lanes paired on 128-byte lines, three reads at a time) and writes one node.  Window 0
    hashes every read over the whole buffer; a window of 40 blocks of nodes (10.5 MB, a child
    line of the headline's sections) confines each block's three reads per group to one
    hashed stretch of that size, as a block's moves along one dimension read one child line.
    Its rate is the ceiling the fused kernel's own gathers meet; HIP events on the launch
    stream, best of `steps` runs after one warm-up per window."""
    import ctypes as C
    from pentago_amd import hip
    L = C.CDLL(os.path.join(ROOT, "pentago_amd", "libpentago_probe.so"))
    L.pe_probe_gather.restype = C.c_int
    L.pe_probe_gather.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.c_uint64, C.c_void_p]
    src = hip.Buffer(64 * src_nodes)
    src.zero()
    dst = hip.Buffer(64 * nodes)
    stream, e0, e1 = hip.Stream(), hip.Event(), hip.Event()
    hip.synchronize()
    res = {"nodes": nodes, "reads_per_node": 12, "src_GB": 64 * src_nodes / 1e9, "windows": {}}
    requested = 64 * 13 * nodes   # 12 reads and one write of 64 bytes per node
    for window in windows:
        def run(seed):
            hip.check(L.pe_probe_gather(src.ptr, src_nodes, dst.ptr, nodes, window, seed, stream.ptr),
                      "pe_probe_gather")

        run(1)
        stream.synchronize()
        ms = []
        for k in range(steps):
            e0.record(stream)
            run(2 + k)
            e1.record(stream)
            e1.synchronize()
            ms.append(e0.elapsed_ms(e1))
        best = min(ms)
        res["windows"][str(window)] = {"window_nodes": window or src_nodes, "ms": ms, "best_ms": best,
                                       "requested_GBps": requested / (best / 1e3) / 1e9}
        log("gather probe, window %d nodes: %s" % (window or src_nodes, res["windows"][str(window)]))
    print(json.dumps({"gather_probe": res}), flush=True)
    return res


SECTION30 = ((3, 4), (4, 4), (4, 4), (4, 3))   # standard form of (4,4)(4,3)(3,4)(4,4)


def line_kernel_bench(P, hip, steps, k=8):
    """The plain line form beside the fused kernel: the same k^4-block box of section
    (3,3)^4 computed as four line passes (end_lines_kernel, one line per workgroup row, the
    schedule_compute_line drop-in) and as one fused launch; outputs compared."""
    specs = [(SECTION, [4] * 4, [4 + k] * 4)]
    out = {"workload": "box [4,%d)^4 of section 24-(3,3)^4" % (4 + k)}
    res = {}
    for mode, name in ((1, "line_passes"), (0, "fused")):
        outs, kids, plan, n_out, child_nodes = build_boxes(P, hip, specs, mode)
        stream = hip.Stream()
        plan.run(stream.ptr)
        stream.synchronize()
        ev0, ev1 = hip.Event(), hip.Event()
        ms = []
        for _ in range(max(steps, 1)):
            ev0.record(stream)
            plan.run(stream.ptr)
            ev1.record(stream)
            ev1.synchronize()
            ms.append(ev0.elapsed_ms(ev1))
        plan.close()
        t = min(ms)
        buf = outs[0][3]
        res[name] = buf.download(np.zeros(buf.nbytes // 8, dtype=np.uint64))
        out[name] = {"ms": t, "superpositions_per_s": n_out / (t / 1e3)}
        del outs, kids
    # same seeds in both builds: identical child data, so outputs must agree bit for bit
    out["outputs_equal"] = bool(np.array_equal(res["line_passes"], res["fused"]))
    out["fused_over_line_passes"] = out["line_passes"]["ms"] / out["fused"]["ms"]
    log("line kernel vs fused: %s" % out)
    return out


def line_dropin_bench(P, hip, n_lines=256, nthreads=8):
    """The schedule_compute_line drop-in (pe_end_schedule_line) driven the way the
    reference's slice loop drives schedule_compute_line (pentago/mpi/flow.cc:465-494): whole
    block lines of section (3,3)^4 in the reference's line buffers (input block k at node
    k * (block_stride + 1)), issued from 8 host threads onto their own streams, each call
    returning at once, one wakeup per line.  Lines/s from the first call to the last wakeup,
    the host time per call, and the node-line rate."""
    import ctypes as C
    import threading
    shape = P.section_shape(SECTION)
    nb = [(s + 7) // 8 for s in shape]
    rng = np.random.default_rng(7)
    jobs = []
    node_lines = 0
    for k in range(n_lines):
        d = k % 4
        base = tuple(int(rng.integers(nb[a])) for a in range(4) if a != d)
        pl = P.make_line(SECTION, d, base)
        info = P.line_describe(pl)
        din = hip.Buffer(64 * info.input_nodes)
        P.fill_seeded(din.ptr, info.input_nodes, 77 + k)
        dout = hip.Buffer(64 * info.output_nodes)
        jobs.append((pl, din, dout))
        node_lines += info.output_nodes
    hip.synchronize()
    streams = [hip.Stream() for _ in range(nthreads)]
    lock = threading.Lock()
    state = {"woken": 0}
    all_done = threading.Event()

    def wakeup(ctx):
        with lock:
            state["woken"] += 1
            if state["woken"] == len(jobs):
                all_done.set()

    cb = P.WAKEUP(wakeup)
    # the first line of each (section, dimension) builds that pair's cached tables
    for d in range(4):
        pl, din, dout = jobs[d]
        P._check(P.lib().pe_end_schedule_line(C.byref(pl), din.ptr, dout.ptr, streams[0].ptr, P.WAKEUP(), None),
                 "pe_end_schedule_line")
    streams[0].synchronize()
    call_s = [0.0] * nthreads
    errors = []

    def issue(t):
        try:
            for k in range(t, len(jobs), nthreads):
                pl, din, dout = jobs[k]
                t0 = time.perf_counter()
                st = P.lib().pe_end_schedule_line(C.byref(pl), din.ptr, dout.ptr, streams[t].ptr, cb, k + 1)
                call_s[t] += time.perf_counter() - t0
                if st:
                    raise RuntimeError(P.lib().pe_last_error())
        except Exception as e:
            errors.append(e)

    t0 = time.perf_counter()
    threads = [threading.Thread(target=issue, args=(t,)) for t in range(nthreads)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    ok = all_done.wait(120)
    dt = time.perf_counter() - t0
    for s in streams:
        s.synchronize()
    del cb
    if errors or not ok:
        return {"error": str(errors[0]) if errors else "wakeups missing"}
    out = {"workload": "%d whole block lines of section 24-(3,3)^4 (all four dimensions, random bases), reference "
                       "line buffers, %d host threads, one stream each" % (len(jobs), nthreads),
           "lines_per_s": len(jobs) / dt, "node_lines_per_s": node_lines / dt, "seconds": dt,
           "host_us_per_call": 1e6 * sum(call_s) / len(jobs), "wakeups": state["woken"]}
    log("line drop-in: %s" % out)
    return out


def slab_index(nb, b):
    idx = 0
    for a in range(4):
        idx = idx * nb[a] + b[a]
    return idx


def pass_groups(dims, child, cbytes, room):
    """The open dimensions grouped greedily, in order, into as few passes as fit in `room`
    bytes (a pass holds its dimensions' child sections, a child shared by two dimensions
    once), ordered from the most dimensions to the fewest: the finishing pass carries the
    win patterns and is cheapest with the fewest dimensions' gathers beside them (slice 30:
    [1, 2, 3] then [0], 264 ms; the other order 282 ms).  child: {dimension: child section},
    cbytes: {child section: bytes}."""
    groups = []
    for d in dims:
        if groups and sum(cbytes[c] for c in {child[x] for x in groups[-1] + [d]}) <= room:
            groups[-1].append(d)
        else:
            groups.append([d])
    groups.sort(key=len, reverse=True)
    return groups


def section_by_passes(P, hip, counts, samples, stream, seed_base=0x5EC30000, groups=None, headroom=0.92):
    """A whole section computed in passes over subsets of its open dimensions: the output
    (every block, one 4096-node slot each) stays resident, and each pass holds only its
    dimensions' child sections (seeded on the device, not timed; a child shared by two
    dimensions is held once).  By default the open dimensions are grouped greedily, in order,
    into as few passes as fit beside the output in `headroom` of the free device memory, and
    the passes run from the most dimensions to the fewest.  A
    pass is one fused launch with only its dimensions open (pe_end_slabs_plan_create_dims):
    the first pass stores its children's OR in the output slots as a partial, middle passes
    add to it, and the last finishes it (the win patterns and the result, once per node), the
    same output as the block store's merge of per-dimension results
    (pentago/end/block_store.cc:173-229).  The sequence of passes runs three times (a
    finishing pass reads the partial, so a pass alone cannot run again): with every pass on
    the static grid, then on the queue-fed form, then each pass on the faster of the two for
    it, which is the timed `ms`.  For
    every sampled block the oracle's line through it is computed from child data regenerated
    on the host from the seed.  Returns (output buffer, block counts, passes,
    {block: expected nodes})."""
    from oracle import oracle as O
    from tests.seeded import seeded_nodes
    shape = P.section_shape(counts)
    nb = [(s + 7) // 8 for s in shape]
    n_out = int(np.prod(shape))
    out_bytes = 64 * 4096 * int(np.prod(nb))
    dims = [d for d in range(4) if sum(counts[d]) < 9]
    child = {}
    for d in dims:
        child[d] = P.line_describe(P.make_line(counts, d, tuple(0 for a in range(4) if a != d))).standard_child.counts
    cnb = {c: [(s + 7) // 8 for s in P.section_shape(c)] for c in child.values()}
    cbytes = {c: 64 * 4096 * int(np.prod(n)) for c, n in cnb.items()}
    if groups is None:
        groups = pass_groups(dims, child, cbytes, headroom * hip.mem_info()[0] - out_bytes)
    out = hip.Buffer(out_bytes)
    out_slab = P.Slab.of(counts, [0] * 4, nb, out.ptr)
    ev0, ev1 = hip.Event(), hip.Event()
    passes = []
    queue = 6   # workgroups per CU of the queue-fed form
    for rep in range(3):
        for i, group in enumerate(groups):
            # 0 one pass; else 2 store the partial, 4 add to it, 3 finish it
            mode = 0 if len(groups) == 1 else 2 if i == 0 else 3 if i == len(groups) - 1 else 4
            if rep < 2:
                form = (0, queue)[rep]
            else:
                form = 0 if passes[i]["grid_ms"] <= passes[i]["queue_ms"] else queue
            kids = {}
            for c in {child[d] for d in group}:
                kids[c] = hip.Buffer(cbytes[c])
                P.fill_seeded(kids[c].ptr, 4096 * int(np.prod(cnb[c])), seed_base + sum(v * 10 ** j for j, v in enumerate(sum(c, ()))))
            t0 = time.time()
            plan = P.Plan.slab_dims(out_slab, [P.Slab.of(c, [0] * 4, cnb[c], b.ptr) for c, b in kids.items()],
                                    sum(1 << d for d in group), accumulate=mode)
            plan.set_launch(form)
            build_s = time.time() - t0
            hip.synchronize()
            ev0.record(stream)
            plan.run(stream.ptr)
            ev1.record(stream)
            ev1.synchronize()
            ms = ev0.elapsed_ms(ev1)
            plan.close()
            if rep == 0:
                # algorithmic bytes: every child node of the pass's lines once, the output
                # slots written once (and read once more by every pass after the first)
                child_nodes = sum(int(np.prod(P.section_shape(child[d]))) for d in group)
                passes.append({"dimensions": group, "mode": mode,
                               "children": sorted("".join("%d%d" % x for x in c) for c in kids),
                               "child_GB": sum(cbytes[c] for c in kids) / 1e9, "grid_ms": ms,
                               "plan_build_s": build_s,
                               "algorithmic_bytes": 64 * (child_nodes + (2 if i else 1) * n_out)})
            elif rep == 1:
                passes[i]["queue_ms"] = ms
            else:
                passes[i]["ms"] = ms
                passes[i]["launch"] = "grid" if form == 0 else "queue, %d workgroups per CU" % form
                log("section %s, dimensions %s: %s" % ("".join("%d%d" % c for c in counts), group, passes[i]))
            kids = None
    expect = {b: None for b in samples}
    for group in groups:
        for d in group:
            c = child[d]
            seed = seed_base + sum(v * 10 ** i for i, v in enumerate(sum(c, ())))
            cshape = P.section_shape(c)
            for blk in samples:
                base = tuple(blk[a] for a in range(4) if a != d)
                ol = O.make_line(counts, d, base)
                oi = O.line_describe(ol)
                inp = np.zeros(8 * oi.input_nodes, dtype=np.uint64)
                for kk in range(oi.input_blocks):
                    cb = list(oi.first_child_block)
                    cb[oi.child_dim] = kk
                    bn = int(np.prod([min(8, cshape[a] - 8 * cb[a]) for a in range(4)]))
                    off = 8 * kk * (oi.block_stride + 1)
                    inp[off:off + 8 * bn] = seeded_nodes(bn, seed, 4096 * slab_index(cnb[c], cb))
                line_out = np.zeros(8 * oi.output_nodes, dtype=np.uint64)
                O.compute_lines_threaded([ol], [inp], [line_out], baseline_threads())
                bn = int(np.prod([min(8, shape[a] - 8 * blk[a]) for a in range(4)]))
                part = line_out[8 * oi.block_stride * blk[d]:8 * oi.block_stride * blk[d] + 8 * bn]
                expect[blk] = part if expect[blk] is None else expect[blk] | part
    return out, nb, passes, expect


def section30_bench(P, hip, steps, check_blocks=2):
    """The whole slice-30 section (4,4)(4,3)(3,4)(4,4) (standard form (3,4)(4,4)(4,4)(4,3),
    shape 318 x 160 x 160 x 318, 2,588,774,400 nodes, 640,000 blocks).  The output stays
    resident (167.8 GB of slab slots); its child sections (84 + 42 + 67 GB, dimensions 1 and 2
    share one) do not fit beside it together, so the section runs in passes over subsets of
    its dimensions (section_by_passes: dimensions 1-3 fused, then dimension 0).
    Kernel time from HIP events, plan builds apart; sampled blocks against the oracle."""
    counts = SECTION30
    shape = P.section_shape(counts)
    nb = [(s + 7) // 8 for s in shape]
    n_out = int(np.prod(shape))
    rng = np.random.default_rng(30)
    samples = [tuple(int(rng.integers(n)) for n in nb) for _ in range(check_blocks - 1)] + [tuple(n - 1 for n in nb)]
    stream = hip.Stream()
    t_wall = time.time()
    out, nb, passes, expect = section_by_passes(P, hip, counts, samples, stream)
    wall = time.time() - t_wall
    mism = 0
    for blk, e in expect.items():
        got = out.download(np.zeros(e.size, dtype=np.uint64), offset=64 * 4096 * slab_index(nb, blk))
        mism += int(not np.array_equal(got, e))
    del out
    kernel_ms = sum(p["ms"] for p in passes)
    alg = sum(p["algorithmic_bytes"] for p in passes)
    res = {"workload": "whole section 30-%s (%d nodes, %d blocks), in passes over subsets of its open "
                       "dimensions with only those dimensions' child sections resident" % ("".join("%d%d" % c for c in counts), n_out, int(np.prod(nb))),
           "output_GB": 64 * 4096 * int(np.prod(nb)) / 1e9, "passes": passes, "kernel_ms": kernel_ms,
           "superpositions_per_s": n_out / (kernel_ms / 1e3),
           "algorithmic_GBps": alg / (kernel_ms / 1e3) / 1e9,
           "roofline_frac": alg / (kernel_ms / 1e3) / 1e9 / HBM_PEAK_GBPS,
           # the same time against the fused form's bytes (each node written once, each child
           # node read once per line: 64 B x (2,588,774,400 + 3,428,...) = 385 GB)
           "fused_form_algorithmic_GBps": 64 * sum(whole_section_nodes(P, counts)) / (kernel_ms / 1e3) / 1e9,
           "wall_s_including_fills_plans_and_checks": wall,
           "checked_blocks": [list(b) for b in samples], "mismatches": mism}
    log("section 30 whole: %s" % {k: v for k, v in res.items() if k != "passes"})
    return res


def whole_section_nodes(P, counts):
    """(nodes of the section, nodes of the child sections its lines read, one per open dimension)."""
    nodes = int(np.prod(P.section_shape(counts)))
    turn = sum(map(sum, counts)) & 1
    child = 0
    for d in range(4):
        if sum(counts[d]) < 9:
            c = [list(x) for x in counts]
            c[d][turn] += 1
            child += int(np.prod(P.section_shape(tuple(map(tuple, c)))))
    return nodes, child


def slices_bench(P, hip, budget=None, slice32=False):
    """Real data: slices 35, 34 and 33 (every section the empty board reaches) computed on
    the device from nothing, slice 33 in groups of sections beside slice 34; kernel time
    from HIP events on the launch stream, plan build time apart, the count kernel timed over
    each slice (one launch per group of sections, the descriptor upload inside the timed
    region), and every section's counts compared with the production run's."""
    empty = ((0, 0),) * 4
    golden = os.path.join(ROOT, "tests", "golden")
    stream = hip.Stream()
    ev0, ev1 = hip.Event(), hip.Event()
    out, stores = {}, {}

    def timed(fn):
        ev0.record(stream)
        fn()
        ev1.record(stream)
        ev1.synchronize()
        return ev0.elapsed_ms(ev1) / 1e3

    def run_group(secs, children, arena=None):
        st = P.SliceStore(secs, arena=arena)
        slabs = [st.slab(k) for k in range(len(secs))]
        t0 = time.time()
        plan = P.Plan.slab(slabs, [children.slab(k) for k in range(len(children.sections))] if children else [], 0)
        build = time.time() - t0
        kernel = timed(lambda: plan.run(stream.ptr))
        plan.close()
        d_counts = hip.Buffer(32 * sum(s.blocks for s in slabs))
        ws = hip.Buffer(P.count_workspace_bytes(len(slabs)))
        offs = np.cumsum([0] + [s.blocks for s in slabs])
        count = timed(lambda: P.count_blocks_many(slabs, d_counts.ptr, ws.ptr, stream.ptr))
        per_block = d_counts.download(np.zeros((int(offs[-1]), 4), dtype=np.uint64))
        counts = np.array([per_block[offs[k]:offs[k + 1]].sum(axis=0) for k in range(len(slabs))], dtype=np.uint64)
        return st, counts, build, kernel, count

    from pentago_amd import slices as S
    arena, cap = None, None
    for slice_ in (35, 34, 33):
        secs = P.descendent_sections(empty, slice_)
        want = np.load(os.path.join(golden, "counts-%d.npy" % slice_), allow_pickle=False)
        want = want[np.argsort(want[:, 0], kind="stable")]
        nodes = child = 0
        for s in secs:
            n, c = whole_section_nodes(P, s)
            nodes += n
            child += c if slice_ < 35 else 0
        if slice_ > 33:
            groups = [secs]
        else:
            # one arena for every group of slices 33 and 32, allocated once before timing
            # (allocating device memory costs ~25 ms per GB here, more than computing in it)
            free, _ = hip.mem_info()
            cap = (budget or free - 16e9)
            groups, cur, used = [], [], 0
            for s in secs:
                b = P.slice_bytes([s])
                if cur and used + b > cap:
                    groups.append(cur)
                    cur, used = [], 0
                cur.append(s)
                used += b
            groups.append(cur)
            need = max(P.slice_bytes(g) for g in groups)
            if slice32:
                need = max(need, S.arena_bytes(P.descendent_sections(empty, 32), cap, True))
            t0 = time.time()
            arena = hip.Buffer(need)
            hip.synchronize()
            out["arena"] = {"bytes": need, "alloc_s": time.time() - t0}
            log("arena: %s" % out["arena"])
        rows, build = [], 0.0
        kernel = count = 0.0
        asym = 0
        for g in groups:
            st, counts, b, k, c = run_group(g, stores.get(slice_ + 1), arena if slice_ == 33 else None)
            build, kernel, count = build + b, kernel + k, count + c
            asym += int(counts[:, 3].sum())
            rows.append(P.counts_rows(st.sections, counts, slice_))
            if slice_ > 33:
                stores[slice_] = st
            else:
                st.close()
        if slice_ + 1 in stores and not (slice32 and slice_ == 33):   # no longer read
            stores.pop(slice_ + 1).close()
        rows = np.concatenate(rows)
        alg = 64 * (nodes + child)
        r = {"sections": len(secs), "groups": len(groups), "nodes": nodes, "child_nodes_read": child,
             "plan_build_s": build, "kernel_s": kernel, "superpositions_per_s": nodes / kernel,
             "algorithmic_GBps": alg / kernel / 1e9, "roofline_frac": alg / kernel / 1e9 / HBM_PEAK_GBPS,
             "count_kernel_s": count, "count_GBps": 64 * nodes / count / 1e9,
             "count_roofline_frac": 64 * nodes / count / 1e9 / HBM_PEAK_GBPS,
             "asymmetric_nodes": asym, "counts_match_production": bool(np.array_equal(rows, want))}
        log("slice %d: %s" % (slice_, r))
        out[str(slice_)] = r
    if slice32:
        # slice 32 in groups, each group's slice-33 children regenerated from slice 34, in the
        # arena slice 33 used; wall time includes plan builds, the regeneration and the counts
        secs = P.descendent_sections(empty, 32)
        want = np.load(os.path.join(golden, "counts-32.npy"), allow_pickle=False)
        want = want[np.argsort(want[:, 0], kind="stable")]
        hip.synchronize()
        t0 = time.time()
        done, counts = S.grouped_counts(secs, cap, grandchildren=stores[34], arena=arena)
        hip.synchronize()
        wall = time.time() - t0
        nodes = sum(whole_section_nodes(P, s)[0] for s in secs)
        r = {"sections": len(secs), "nodes": nodes, "wall_s": wall, "superpositions_per_s_wall": nodes / wall,
             "asymmetric_nodes": int(counts[:, 3].sum()),
             "counts_match_production": bool(np.array_equal(P.counts_rows(done, counts, 32), want))}
        log("slice 32: %s" % r)
        out["32"] = r
    for st in stores.values():
        st.close()
    return out


def mid_bench(P, hip, n_boards=4096, seed=2, tune=True, plain=True):
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
    if tune:
        # the autotuner times each candidate (one warm run) on the batch itself: the best split
        # between the small-board kernel and the large boards' launches depends on the mix
        t0 = time.time()
        tuned = P.mid_autotune(boards)
        out["autotune"] = dict(tuned, seconds=time.time() - t0, boards=len(boards))
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
    # the first boards' results, for the oracle check in the CPU baseline leg (not printed)
    out["_first_results"] = P.results_to_lists(boards[:48], res2[:48])
    batched_s = min(ms) / 1e3
    log("midsolve batched: %s ms" % ms)
    spots = np.array([board_spots(b) for b in boards])
    mism = int(np.sum(np.any(res2 != res, axis=(1, 2))))
    plain_s = None
    if plain:
        # plain: one synchronous launch sequence per board
        pres = np.zeros_like(res)
        t0 = time.perf_counter()
        for i, b in enumerate(boards):
            P.midsolve_internal_raw(b, pres[i])
        plain_s = time.perf_counter() - t0
        log("midsolve plain: %.2f s" % plain_s)
        mism += int(np.sum(np.any(pres != res, axis=(1, 2))))
    work = 0
    from math import comb
    for sp in spots:
        sp = int(sp)
        work += 2 * 16 * sum(comb(sp, n // 2) * comb(sp - n // 2, n - n // 2) * (sp - n + 1) for n in range(sp + 1))
    if plain:
        # small boards alone (24..30 stones, 6..12 empty cells): launch overhead dominates the
        # plain path, the batched kernel solves them one workgroup per board
        small = seeded_boards(n_boards, seed + 1, 24, 30)
        # the small-board kernel's workgroup size, picked on this batch (an explicit plan
        # configuration: the device default stays the mix's)
        small_ms, sres, small_threads = None, None, None
        for th in (256, 512, 1024):
            plan = P.MidPlan(small, small_spots=12, threads=th)
            plan.run(stream.ptr)
            res_th = plan.results(stream.ptr)
            ev0.record(stream)
            plan.run(stream.ptr)
            ev1.record(stream)
            ev1.synchronize()
            t = ev0.elapsed_ms(ev1)
            plan.close()
            if sres is not None:
                mism += int(np.sum(np.any(res_th != sres, axis=(1, 2))))
            if small_ms is None or t < small_ms:
                small_ms, small_threads = t, th
            sres = res_th
        splain = np.zeros_like(sres)
        t0 = time.perf_counter()
        for i, b in enumerate(small):
            P.midsolve_internal_raw(b, splain[i])
        small_plain_s = time.perf_counter() - t0
        out["small_boards"] = {"workload": "%d seeded boards, 24..30 stones (6..12 empty cells)" % n_boards,
                               "threads": small_threads,
                               "batched_boards_per_s": n_boards / (small_ms / 1e3), "batched_ms": small_ms,
                               "plain_boards_per_s": n_boards / small_plain_s,
                               "batched_over_plain": small_plain_s / (small_ms / 1e3),
                               "mismatches": int(np.sum(np.any(splain != sres, axis=(1, 2))))}
        log("midsolve small boards: %s" % out["small_boards"])
    floor_s = work / (HBM_PEAK_GBPS * 1e9)
    out.update({
        "config": {"workload": "%d seeded boards, 18..30 stones (6..18 empty cells)" % n_boards,
                   "spots_histogram": {int(k): int(v) for k, v in zip(*np.unique(spots, return_counts=True))}},
        "batched_boards_per_s": n_boards / batched_s, "batched_ms": min(ms), "batched_ms_runs": ms,
        "plain_boards_per_s": n_boards / plain_s if plain_s else None, "plain_s": plain_s,
        "batched_over_plain": plain_s / batched_s if plain_s else None,
        "mismatches": mism, "plan": info,
        "algorithmic_GBps": work / batched_s / 1e9,
        # HBM bound of the mix: its algorithmic bytes (32 B per position read or written, both
        # passes in one) at 8 TB/s; batching cannot beat plain launches by more than
        # plain time / this floor unless the kernels read less than the algorithmic bytes
        "roofline": {"bound": "hbm", "achieved": work / batched_s / 1e9, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": work / batched_s / 1e9 / HBM_PEAK_GBPS, "traffic": mid_traffic(),
                     "kernel": "batched plan (mid_level1_kernel, mid_boards_kernel, mid_prep_kernel)",
                     "bytes_per_run": work, "run_ms": batched_s * 1e3},
        "hbm_floor_s": floor_s,
        "max_batched_over_plain_at_hbm_peak": plain_s / floor_s if plain_s else None,
    })
    return out


def mid_traffic():
    """HBM bytes per run of the batched midsolve workload from the committed counter profile
    (FETCH_SIZE and WRITE_SIZE scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "mid_batched_4096boards_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_run")


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


def mid_cpu_baseline(n=48, seed=2, expect=None, batch=4096):
    """The oracle's midsolve (two passes per board, one board per thread) on a sample of the
    batched workload (all-core), single-thread on the 26-stone board of the single-board
    configuration and on 24..30-stone boards of the batch.  `expect`: the batched plan's
    results for the same first boards, compared entry by entry with the oracle's."""
    from pentago_amd.workloads import seeded_boards
    O, L, build = oracle_for_timing()
    threads = baseline_threads()

    def roots(bs):
        return [O.High((O.C.c_uint64 * 2)(b.side[0], b.side[1]), b.ply) for b in bs]

    boards = seeded_boards(batch, seed, 18, 30)[:n]
    t0 = time.time()
    got = O.midsolve_batch(roots(boards), threads, L=L)
    dt = time.time() - t0
    oracle_check = None
    if expect is not None:
        bad = sum(1 for e, g, b in zip(expect, got, boards) if list(e) != list(g[:len(e)]))
        oracle_check = {"boards": min(len(expect), n), "mismatches": bad}
    one = seeded_boards(1, 26, 26, 26)[0]
    r1 = roots([one])
    t0 = time.time()
    reps = 0
    while time.time() - t0 < 1.0:
        O.midsolve_batch(r1, 1, L=L)
        reps += 1
    dt1 = time.time() - t0
    small = [b for b in seeded_boards(256, seed, 24, 30)][:64]
    t0 = time.time()
    O.midsolve_batch(roots(small), 1, L=L)
    dts = time.time() - t0
    return {"value": n / dt, "unit": "boards/s", "cores": threads, "kind": "port",
            "sample": "first %d boards of the batched workload (18..30 stones), one board per thread, %.1f s on %d "
                      "threads" % (n, dt, threads),
            "single_26_stones": {"boards_per_s": reps / dt1, "ms": 1e3 * dt1 / reps, "cores": 1},
            "single_thread_24_30_stones": {"boards_per_s": len(small) / dts, "boards": len(small), "cores": 1},
            "batched_vs_oracle": oracle_check, "host": host_info(), "build": build}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--k", type=int, default=0, help="box edge in blocks (0: largest fitting ~200 GB)")
    ap.add_argument("--k2", type=int, default=8, help="edge of the second section's box (0: none)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", type=int, default=0, help="0: fused block kernel, 1: four line passes")
    ap.add_argument("--copy-probe", type=int, default=0, help="run one plain copy of this many bytes instead")
    ap.add_argument("--gather-probe", action="store_true",
                    help="run the gather probe (the fused kernel's access pattern alone) instead")
    ap.add_argument("--no-mid", action="store_true", help="skip the batched midsolve measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-end", action="store_true", help="skip the end-engine measurement (profiling the others)")
    ap.add_argument("--mid-boards", type=int, default=4096, help="boards in the batched midsolve measurement")
    ap.add_argument("--no-section30", dest="section30", action="store_false",
                    help="skip the whole slice-30 section measurement")
    ap.add_argument("--no-slices", dest="slices", action="store_false",
                    help="skip the whole-slice (35, 34, 33) measurement on real data")
    ap.add_argument("--no-line-kernel", dest="line_kernel", action="store_false",
                    help="skip the line-pass (schedule_compute_line form) comparison box")
    ap.add_argument("--launch", type=int, default=-1,
                    help="fused launch form: -1 autotune (default), 0 grid, n queue with n workgroups per CU")
    ap.add_argument("--no-mid-tune", dest="mid_tune", action="store_false",
                    help="skip the midsolve autotuner (built-in defaults)")
    ap.add_argument("--no-mid-plain", dest="mid_plain", action="store_false",
                    help="skip the plain per-board midsolve comparison and the small-board leg")
    ap.add_argument("--no-line-dropin", dest="line_dropin", action="store_false",
                    help="skip the line drop-in measurement (lines scheduled from host threads)")
    ap.add_argument("--no-slice32", dest="slice32", action="store_false",
                    help="skip slice 32 (computed in groups with regenerated children, ~13 s)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gather_probe:
        gather_probe()
        return

    import pentago_amd as P
    from pentago_amd import hip
    P.init()
    result = {"metric": "superpositions/s", "value": None, "unit": "superpositions/s", "n_gpus": 1}
    if not args.no_end:
        free, total = hip.mem_info()
        shape = P.section_shape(SECTION)
        # the second section's box: k2^4 interior blocks, mixed into the same launch
        k2, a2 = args.k2, 8
        specs2 = [(SECTION2, [a2] * 4, [a2 + k2] * 4)] if k2 else []
        k = args.k
        if not k:
            budget = min(200e9, 0.8 * free)
            k = 1
            while 8 * (k + 1) <= shape[0] - 8 and boxes_bytes(P, [(SECTION, [0] * 4, [k + 1] * 4)] + specs2) <= budget:
                k += 1
        specs = [(SECTION, [0] * 4, [k] * 4)] + specs2
        log("boxes: %s, free HBM %.1f GB" % (specs, free / 1e9))
        t0 = time.time()
        outs, kids, plan, n_out, child_nodes = build_boxes(P, hip, specs, args.mode)
        hip.synchronize()
        log("setup %.1f s: %d output blocks (%.1f GB), %d child slabs (%.1f GB)" % (
            time.time() - t0, sum(int(np.prod([h - l for l, h in zip(lo, hi)])) for _, lo, hi, _ in outs),
            sum(b.nbytes for *_, b in outs) / 1e9, len(kids), sum(b.nbytes for *_, b in kids) / 1e9))

        stream = hip.Stream()
        launch = None
        form = 0
        if args.mode == 0 and args.launch >= 0:
            plan.set_launch(args.launch)
            form = args.launch
            launch = {"choice": "grid" if args.launch == 0 else "queue, %d workgroups per CU" % args.launch}
        elif args.mode == 0:
            # launch form: static grid against the persistent queue-fed grid, fastest kept
            t0 = time.time()
            forms, choice = plan.autotune(stream.ptr)
            form = choice
            launch = {"choice": "grid" if choice == 0 else "queue, %d workgroups per CU" % choice,
                      "timed_ms": {("grid" if f == 0 else "queue_%d_per_cu" % f): ms for f, ms in forms.items()},
                      "seconds": time.time() - t0}
            log("launch forms: %s" % launch)
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
        # algorithmic bytes per step: box output once + every line's child nodes once
        alg_step = 64 * (n_out + child_nodes)
        achieved = (alg_step / launches) / (avg_launch_ms / 1e3) / 1e9
        # the dominant kernel: the grid form, the queue-fed form, or the line kernel
        kernel = ("fused_blocks_kernel" if form == 0 else "fused_blocks_queue_kernel") if args.mode == 0 \
            else "end_lines_kernel"
        traffic = None
        tag = "box%d" % k + ("_mixed%d" % k2 if k2 else "")
        tfile = os.path.join(ROOT, "profiles", "%s_%s_traffic.json" % (kernel, tag))
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
            "config": {"workload": "end-engine: box of %d^4 blocks of section 24-(3,3)^4%s, all open dimensions, "
                                   "child slabs of the slice-25 children" % (
                                       k, (" mixed with a box of %d^4 blocks of section 24-%s" % (
                                           k2, "".join("%d%d" % c for c in SECTION2))) if k2 else ""),
                       "boxes": [{"section": "".join("%d%d" % c for c in sec), "lo": lo, "hi": hi} for sec, lo, hi in specs],
                       "output_nodes": n_out, "child_nodes_read": child_nodes,
                       "mode": "fused" if args.mode == 0 else "line passes"},
            "launch": launch,
            "hbm_GBps_algorithmic": alg_step / (ms_per_step / 1e3) / 1e9,
            "step_ms": step_ms,
            "wall_s": wall,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                         "kernel": kernel, "bytes_per_launch": alg_step / launches,
                         "avg_launch_ms": avg_launch_ms},
        }
        plan.close()
        del outs, kids
    if not args.no_end and args.section30:
        result["section30"] = section30_bench(P, hip, args.steps)
    if not args.no_end and args.line_kernel:
        result["line_kernel"] = line_kernel_bench(P, hip, args.steps)
    if not args.no_end and args.line_dropin:
        result["line_dropin"] = line_dropin_bench(P, hip)
    if not args.no_end and args.slices:
        result["slices"] = slices_bench(P, hip, slice32=args.slice32)
    if not args.no_mid:
        result["midsolve"] = mid_bench(P, hip, args.mid_boards, tune=args.mid_tune, plain=args.mid_plain)
        first = result["midsolve"].pop("_first_results")
    if not args.no_serve:
        result["serving"] = serve_bench(P)
    if not args.no_cpu and not args.no_end:
        try:
            result["cpu_baseline"] = cpu_baseline(args.cpu_budget)
        except Exception as e:   # the baseline is reported, never required
            result["cpu_baseline"] = {"error": str(e)}
    if not args.no_cpu and not args.no_mid:
        try:
            result["midsolve"]["cpu_baseline"] = mid_cpu_baseline(expect=first, batch=args.mid_boards)
        except Exception as e:
            result["midsolve"]["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
