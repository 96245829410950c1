"""End-engine block lines on the GPU against the oracle, bit for bit.

Mirrors pentago/end/compute.cc semantics: every output word of a line must equal the
oracle's for the same input buffer (reference layout).  Real-data pins: production sparse
samples at slices 34 and 35 (pentago/high/high_test.cc:47-57) computed end to end on the
GPU.
"""
import ctypes as C

import numpy as np
import pytest

import pentago_amd as P
from pentago_amd import hip
from oracle import oracle as O
from tests.common import golden
from tests.seeded import seeded_nodes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    P.init()


# (section counts, dimension, block base selector) -- all 8 global transforms, reflections,
# partial last blocks and the slice-35 path
CASES = [
    (((3, 4), (4, 4), (4, 4), (4, 3)), 0, "last"),
    (((3, 4), (4, 4), (4, 4), (4, 3)), 2, "first"),   # reflected child (transform 6)
    (((3, 4), (4, 4), (4, 4), (4, 3)), 3, "last"),
    (((4, 4), (4, 3), (3, 4), (4, 4)), 0, "last"),    # transform 3
    (((4, 4), (4, 3), (3, 4), (4, 4)), 3, "last"),    # transform 5
    (((3, 3), (3, 3), (3, 3), (3, 3)), 1, "last"),    # transform 1
    (((3, 3), (3, 3), (3, 3), (3, 3)), 3, "last"),    # transform 2
    (((4, 2), (2, 4), (3, 3), (3, 3)), 0, "last"),    # transform 4
    (((4, 2), (2, 4), (3, 3), (3, 3)), 2, "first"),
    (((4, 5), (5, 4), (4, 4), (4, 4)), 3, "first"),   # transform 7, slice 34
    (((2, 3), (4, 4), (3, 2), (4, 4)), 1, "last"),
    (((3, 2), (2, 3), (3, 3), (4, 4)), 3, "last"),
    (((5, 4), (4, 5), (4, 4), (5, 4)), 2, "last"),    # slice 35
    (((4, 4), (4, 4), (4, 4), (5, 4)), 2, "last"),    # slice 33, transform 5
]


def make_case(counts, dim, where):
    shape = P.section_shape(counts)
    base = tuple(0 if where == "first" else (shape[a] - 1) // 8 for a in range(4) if a != dim)
    return P.make_line(counts, dim, base), O.make_line(counts, dim, base)


def line_input(info, seed):
    return seeded_nodes(int(info.input_nodes), seed) if info.input_nodes else None


@pytest.mark.parametrize("case", range(len(CASES)))
def test_line_parity(case):
    counts, dim, where = CASES[case]
    pl, ol = make_case(counts, dim, where)
    info = P.line_describe(pl)
    print("case", case, counts, dim, where, "output nodes", info.output_nodes, flush=True)
    inp = line_input(info, 1000 + case)
    gpu = P.compute_line_host(pl, inp)
    ref = np.zeros(8 * info.output_nodes, dtype=np.uint64)
    O.compute_lines_threaded([ol], [inp], [ref], 16)
    assert np.array_equal(gpu, ref)


def test_batched_lines_match_single_lines():
    """pe_end_compute_lines on mixed lines equals one launch per line."""
    picks = [CASES[i] for i in (0, 3, 5, 7, 12)]
    lines, infos, ins, outs_b, outs_s, keep = [], [], [], [], [], []
    for k, (counts, dim, where) in enumerate(picks):
        pl, _ = make_case(counts, dim, where)
        info = P.line_describe(pl)
        inp = line_input(info, 77 + k)
        din = hip.Buffer(max(64, 64 * info.input_nodes))
        if inp is not None:
            din.upload(inp)
        ob, os_ = hip.Buffer(64 * info.output_nodes), hip.Buffer(64 * info.output_nodes)
        ob.zero()
        os_.zero()
        lines.append(pl); infos.append(info); ins.append(din); outs_b.append(ob); outs_s.append(os_)
    P.compute_lines(lines, [b.ptr for b in ins], [b.ptr for b in outs_b])
    for pl, din, os_ in zip(lines, ins, outs_s):
        P.compute_lines([pl], [din.ptr], [os_.ptr])
    hip.synchronize()
    for info, ob, os_ in zip(infos, outs_b, outs_s):
        a = ob.download(np.zeros(8 * info.output_nodes, dtype=np.uint64))
        b = os_.download(np.zeros(8 * info.output_nodes, dtype=np.uint64))
        assert np.array_equal(a, b)


def test_schedule_line_calls_wakeup_once():
    counts, dim, where = CASES[12]
    pl, _ = make_case(counts, dim, where)
    info = P.line_describe(pl)
    out = hip.Buffer(64 * info.output_nodes)
    calls = []
    cb = P.schedule_compute_line(pl, None, out.ptr, lambda: calls.append(1))
    hip.synchronize()
    import time
    for _ in range(100):
        if calls:
            break
        time.sleep(0.01)
    assert calls == [1]
    del cb


HOST_FN = C.CFUNCTYPE(None, C.c_void_p)


def test_schedule_line_from_threads_is_asynchronous_and_wakes_once():
    """schedule_compute_line's contract (pentago/end/compute.cc:359-370,394-404,
    compute.h:84-86; called from pool threads, pentago/mpi/flow.cc:465-494): 8 host threads
    issue 64 lines at once on their own streams, every call returns while its stream is still
    held back by a host-side gate (so no call waited for its line), wakeup fires exactly once
    per line after its output is complete, and every output equals the oracle's."""
    import threading
    import time
    nthreads, per_thread = 8, 8
    jobs = []
    for k in range(nthreads * per_thread):
        counts, dim, where = CASES[k % len(CASES)]
        pl, ol = make_case(counts, dim, where)
        info = P.line_describe(pl)
        inp = line_input(info, 5000 + k)
        din = hip.Buffer(max(64, 64 * info.input_nodes))
        if inp is not None:
            din.upload(inp)
        dout = hip.Buffer(64 * info.output_nodes)
        dout.zero()
        jobs.append((pl, ol, info, inp, din, dout))
    streams = [hip.Stream() for _ in range(nthreads)]
    # the gate: a host function on its own stream that returns once the test opens it (or
    # after 10 s at most); every line stream waits for the event recorded behind it
    gate, gate_stream, gate_event = threading.Event(), hip.Stream(), hip.Event()
    gate_fn = HOST_FN(lambda _: gate.wait(10.0))
    hip.check(hip.hip().hipLaunchHostFunc(C.c_void_p(gate_stream.ptr), C.cast(gate_fn, C.c_void_p), None),
              "hipLaunchHostFunc")
    gate_event.record(gate_stream)
    for s in streams:
        hip.check(hip.hip().hipStreamWaitEvent(C.c_void_p(s.ptr), C.c_void_p(gate_event.ptr), 0), "hipStreamWaitEvent")
    lock = threading.Lock()
    wakes = {}

    def wakeup(ctx):
        with lock:
            wakes[ctx] = wakes.get(ctx, 0) + 1

    cb = P.WAKEUP(wakeup)
    errors = []

    def issue(t):
        try:
            for k in range(t, len(jobs), nthreads):
                pl, _, _, _, din, dout = jobs[k]
                st = P.lib().pe_end_schedule_line(C.byref(pl), din.ptr, dout.ptr, streams[t].ptr, cb, k + 1)
                assert st == 0, P.lib().pe_last_error()
        except Exception as e:  # reported by the main thread
            errors.append(e)

    try:
        threads = [threading.Thread(target=issue, args=(t,)) for t in range(nthreads)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        with lock:
            early = dict(wakes)
    finally:
        gate.set()
    assert not errors, errors
    assert early == {}, "a line completed while its stream was held back"
    for s in streams:
        s.synchronize()
    deadline = time.time() + 20
    while len(wakes) < len(jobs) and time.time() < deadline:
        time.sleep(0.01)
    time.sleep(0.2)   # a second wakeup of any line would land here
    assert sorted(wakes) == list(range(1, len(jobs) + 1))
    assert all(v == 1 for v in wakes.values()), wakes
    outs = [np.zeros(8 * j[2].output_nodes, dtype=np.uint64) for j in jobs]
    refs = [np.zeros_like(o) for o in outs]
    O.compute_lines_threaded([j[1] for j in jobs], [j[3] for j in jobs], refs, 16)
    for k, (j, out, ref) in enumerate(zip(jobs, outs, refs)):
        j[5].download(out)
        assert np.array_equal(out, ref), (k, CASES[k % len(CASES)])
    gate_stream.synchronize()
    del cb, gate_fn


def test_compute_line_accumulating_plan_jobs_slice35():
    """35-stone whole-line jobs now run on the fused kernel's no-children form: a plan of
    slice-35 lines (one of them OR-accumulating into prefilled output) equals the oracle."""
    counts = ((5, 4), (4, 5), (4, 4), (5, 4))
    shape = P.section_shape(counts)
    jobs, refs, bufs = [], [], []
    rng = np.random.default_rng(35)
    for k, base in enumerate([(0, 0, 0), tuple((shape[a] - 1) // 8 for a in range(4) if a != 2)]):
        pl, ol = P.make_line(counts, 2, base), O.make_line(counts, 2, base)
        info = P.line_describe(pl)
        ref = np.zeros(8 * info.output_nodes, dtype=np.uint64)
        O.compute_lines_threaded([ol], [None], [ref], 4)
        prefill = rng.integers(0, 2**63, size=ref.size, dtype=np.uint64) if k else np.zeros_like(ref)
        out = hip.Buffer(64 * info.output_nodes)
        out.upload(prefill)
        jobs.append({"line": pl, "input": None, "output": out.ptr, "accumulate": k})
        refs.append(ref | prefill)
        bufs.append(out)
    plan = P.Plan(jobs)
    plan.run()
    hip.synchronize()
    for out, ref in zip(bufs, refs):
        assert np.array_equal(out.download(np.zeros_like(ref)), ref)
    plan.close()


def test_seeded_fill_past_the_dispatch_work_item_limit():
    """A fill of more than 2^30 nodes (more than 2^32 (node, word) pairs, the most one dispatch
    counts) writes the last nodes too: the slice-30 section's largest child slab has 1.31e9."""
    free, _ = hip.mem_info()
    n = (1 << 30) + (1 << 24)
    if 64 * n > 0.8 * free:
        pytest.skip("needs %.0f GB of device memory" % (64 * n / 1e9))
    buf = hip.Buffer(64 * n)
    P.fill_seeded(buf.ptr, n, 4242)
    hip.synchronize()
    for first in (0, n - 3000, (1 << 30) - 1000):
        got = buf.download(np.zeros(8 * 3000, dtype=np.uint64), offset=64 * first)
        assert np.array_equal(got, seeded_nodes(3000, 4242, first)), first


def test_seeded_fill_matches_host_generator():
    n = 5000
    buf = hip.Buffer(64 * n)
    P.fill_seeded(buf.ptr, n, 12345, 987654321)
    hip.synchronize()
    got = buf.download(np.zeros(8 * n, dtype=np.uint64))
    assert np.array_equal(got, seeded_nodes(n, 12345, 987654321))


# ---------------- real data: production samples computed on the GPU ----------------

class GpuBlocks:
    """Block values of standard sections, computed with GPU lines (one per open dimension,
    OR-merged as pentago/end/block_store.cc:173-229) from recursively computed child blocks."""

    def __init__(self):
        self.lines = {}

    def line_output(self, counts, dim, base):
        key = (counts, dim, base)
        if key not in self.lines:
            pl = P.make_line(counts, dim, base)
            info = P.line_describe(pl)
            inp = None
            if info.input_nodes:
                inp = np.zeros(8 * info.input_nodes, dtype=np.uint64)
                child = info.standard_child.counts
                for k in range(info.input_blocks):
                    cb = list(info.first_child_block)
                    cb[info.child_dimension] = k
                    data = self.block(child, tuple(cb))
                    off = 8 * k * (info.block_stride + 1)
                    inp[off:off + data.size] = data
            self.lines[key] = (info, P.compute_line_host(pl, inp))
        return self.lines[key]

    def block(self, counts, block):
        shape = P.section_shape(counts)
        bshape = [min(shape[a], 8 * (block[a] + 1)) - 8 * block[a] for a in range(4)]
        acc = np.zeros(8 * int(np.prod(bshape)), dtype=np.uint64)
        for d in range(4):
            if sum(counts[d]) >= 9:
                continue
            base = tuple(block[a] for a in range(4) if a != d)
            info, out = self.line_output(counts, d, base)
            off = 8 * info.block_stride * block[d]
            acc |= out[off:off + acc.size]
        return acc


def board_counts_and_index(board):
    counts, idx = [], []
    for q in range(4):
        quad = (board >> (16 * q)) & 0xffff
        ir = O.rmin_inverse(quad)
        assert ir & 3 == 0
        idx.append(ir >> 2)
        b = w = 0
        for i in range(9):
            d = quad % 3
            quad //= 3
            b += d == 1
            w += d == 2
        counts.append((b, w))
    return tuple(counts), idx


@pytest.mark.parametrize("slice_", [35, 34])
def test_production_samples_on_gpu(slice_):
    data = golden("samples.json")[str(slice_)]
    G = GpuBlocks()
    for board, wins in zip(data["boards"], data["wins"]):
        counts, idx = board_counts_and_index(board)
        block = tuple(i // 8 for i in idx)
        shape = P.section_shape(counts)
        bshape = [min(shape[a], 8 * (block[a] + 1)) - 8 * block[a] for a in range(4)]
        local = [i % 8 for i in idx]
        pos = ((local[0] * bshape[1] + local[1]) * bshape[2] + local[2]) * bshape[3] + local[3]
        node = G.block(counts, block)[8 * pos:8 * pos + 8]
        win = sum(int(node[i]) << (64 * i) for i in range(4))
        wot = sum(int(node[4 + i]) << (64 * i) for i in range(4))
        black, white = win, (~wot) & (2**256 - 1)
        if slice_ & 1:
            black, white = white, black
        expect_b = sum(wins[i] << (64 * i) for i in range(4))
        expect_w = sum(wins[4 + i] << (64 * i) for i in range(4))
        assert (black, white) == (expect_b, expect_w), board


@pytest.mark.parametrize("case", [1, 5, 9])
def test_line_jobs_offsets_accumulate_and_ranges(case):
    """Line-job plans with scattered child and output blocks (reversed, gapped), OR-
    accumulating into a prefilled output: the whole-line job (fused kernel) and a job
    restricted to indices 1.. (line kernel) against the oracle's line."""
    counts, dim, where = CASES[case]
    pl, ol = make_case(counts, dim, where)
    info = P.line_describe(pl)
    inp = line_input(info, 70 + case)
    ref = np.zeros(8 * info.output_nodes, dtype=np.uint64)
    O.compute_lines_threaded([ol], [inp], [ref], 16)
    bs, nin, nout = info.block_stride, info.input_blocks, info.length
    # child block k at node (nin - 1 - k) * (bs + 9): evenly spaced, descending, with gaps
    in_off = [(nin - 1 - k) * (bs + 9) for k in range(nin)]
    dev_in = np.zeros(8 * (nin * (bs + 9)), dtype=np.uint64)
    for k in range(nin):
        src = inp[8 * k * (bs + 1):8 * (k * (bs + 1) + bs + 1)]
        dev_in[8 * in_off[k]:8 * in_off[k] + src.size] = src
    # output block k at node (nout - 1 - k) * (bs + 3)
    out_off = [(nout - 1 - k) * (bs + 3) for k in range(nout)]
    prefill = seeded_nodes(nout * (bs + 3), 90 + case)
    din = hip.Buffer(dev_in.nbytes)
    din.upload(dev_in)
    # the index along the line of every output node (reference layout, per block)
    oshape = list(info.output_shape)
    along = np.zeros(info.output_nodes, dtype=np.int64)
    for k in range(nout):
        bshape = list(oshape)
        bshape[dim] = min(8, oshape[dim] - 8 * k)
        idx = np.indices(bshape)[dim].reshape(-1) + 8 * k
        along[k * bs:k * bs + idx.size] = idx
    for i_range in (None, (1, oshape[dim])):
        dout = hip.Buffer(prefill.nbytes)
        dout.upload(prefill)
        job = dict(line=pl, input=din.ptr, output=dout.ptr, in_offsets=in_off, out_offsets=out_off, accumulate=1)
        if i_range:
            job["i_range"] = i_range
        plan = P.Plan([job])
        plan.run()
        hip.synchronize()
        plan.close()
        got = dout.download(np.zeros_like(prefill))
        want = prefill.copy()
        for k in range(nout):
            n = min(bs, info.output_nodes - k * bs)
            sel = np.arange(k * bs, k * bs + n)
            if i_range:
                sel = sel[along[sel] >= i_range[0]]
            for j in sel:
                o = out_off[k] + (j - k * bs)
                want[8 * o:8 * o + 8] |= ref[8 * j:8 * j + 8]
        assert np.array_equal(got, want), (case, i_range)
