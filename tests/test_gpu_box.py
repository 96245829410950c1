"""The benchmark's box workload (output and child slabs, fused block kernel or four OR-merged
line passes) against the oracle's line kernel on small boxes, plus fused slab plans on
sections with reflected children and partial edge blocks."""
import numpy as np
import pytest

import bench
import pentago_amd as P
from pentago_amd import hip
from oracle import oracle as O
from tests.seeded import seeded_nodes

pytestmark = pytest.mark.gpu


def oracle_block(counts, blk, child_block):
    """OR over the open dimensions of the oracle's line output for one block."""
    shape = O.section_shape(counts)
    bshape = [min(shape[a], 8 * (blk[a] + 1)) - 8 * blk[a] for a in range(4)]
    acc = np.zeros(8 * int(np.prod(bshape)), dtype=np.uint64)
    for d in range(4):
        if sum(counts[d]) >= 9:
            continue
        base = tuple(blk[ax] for ax in range(4) if ax != d)
        ol = O.make_line(counts, d, base)
        oi = O.line_describe(ol)
        inp = np.zeros(8 * oi.input_nodes, dtype=np.uint64)
        child = O.section_counts(oi.standard_child)
        for kk in range(oi.input_blocks):
            cb = list(oi.first_child_block)
            cb[oi.child_dim] = kk
            data = child_block(child, tuple(cb))
            off = 8 * kk * (oi.block_stride + 1)
            inp[off:off + data.size] = data
        line_out = np.zeros(8 * oi.output_nodes, dtype=np.uint64)
        O.compute_lines_threaded([ol], [inp], [line_out], 16)
        off = 8 * oi.block_stride * blk[d]
        acc |= line_out[off:off + acc.size]
    return acc


class SlabSet:
    """Host copies of child slabs: block lookup by (section, block coordinates)."""

    def __init__(self):
        self.slabs = []

    def add(self, counts, lo, hi, data):
        self.slabs.append((counts, lo, hi, data))

    def block(self, counts, cb):
        for c, lo, hi, data in self.slabs:
            if c != counts or any(not (lo[a] <= cb[a] < hi[a]) for a in range(4)):
                continue
            idx = 0
            for a in range(4):
                idx = idx * (hi[a] - lo[a]) + (cb[a] - lo[a])
            shape = O.section_shape(counts)
            bshape = [min(shape[a], 8 * (cb[a] + 1)) - 8 * cb[a] for a in range(4)]
            return data[idx, :int(np.prod(bshape))].reshape(-1)
        raise KeyError((counts, cb))


@pytest.mark.parametrize("k,a0,mode", [(1, 51, 0), (2, 7, 0), (2, 7, 1)])
def test_box_matches_oracle(k, a0, mode):
    P.init()
    out, kids, plan, child_nodes, child, n_child_blocks = bench.build_box(P, hip, k, a0, mode)
    plan.run()
    hip.synchronize()
    box = out.download(np.zeros(8 * 4096 * k ** 4, dtype=np.uint64)).reshape(k ** 4, 4096, 8)
    slab = kids.download(np.zeros(8 * 4096 * n_child_blocks, dtype=np.uint64)).reshape(n_child_blocks, 4096, 8)
    info0 = P.line_describe(P.make_line(bench.SECTION, 0, (a0, a0, a0)))
    cd = info0.child_dimension
    cd_len = (P.section_shape(child)[cd] + 7) // 8
    S = SlabSet()
    S.add(child, [0 if a == cd else a0 for a in range(4)], [cd_len if a == cd else a0 + k for a in range(4)], slab)
    rng = np.random.default_rng(k + 10 * mode)
    for _ in range(2):
        blk = tuple(a0 + int(x) for x in rng.integers(k, size=4))
        idx = 0
        for ax in range(4):
            idx = idx * k + (blk[ax] - a0)
        print("mode", mode, "block", blk, flush=True)
        assert np.array_equal(box[idx].reshape(-1), oracle_block(bench.SECTION, blk, S.block)), blk
    plan.close()


# Output boxes touching partial edge blocks, with children standardized through
# reflections (transforms 4..7), several child sections.
SLAB_CASES = [
    ((4, 2), (2, 4), (3, 3), (3, 3)),     # all dimensions reflect (transform 4)
    ((4, 4), (4, 3), (3, 4), (4, 4)),     # transforms 3 and 5
    ((4, 5), (5, 4), (4, 4), (4, 4)),     # slice 34, transform 7
]


@pytest.mark.parametrize("case", range(len(SLAB_CASES)))
def test_fused_slab_edges_and_reflections(case):
    P.init()
    counts = SLAB_CASES[case]
    shape = P.section_shape(counts)
    nb = [(s + 7) // 8 for s in shape]
    lo = [n - 1 for n in nb]                       # last (partial) block along every axis
    hi = nb
    out = hip.Buffer(64 * 4096)
    out.zero()
    S = SlabSet()
    children, keep = [], []
    seen = set()
    for d in range(4):
        if sum(counts[d]) >= 9:
            continue
        info = P.line_describe(P.make_line(counts, d, tuple(lo[a] for a in range(4) if a != d)))
        child = info.standard_child.counts
        cshape = P.section_shape(child)
        cd = info.child_dimension
        clo = list(info.first_child_block)
        chi = [c + 1 for c in clo]
        clo[cd], chi[cd] = 0, (cshape[cd] + 7) // 8
        key = (child, tuple(clo), tuple(chi))
        if key in seen:
            continue
        seen.add(key)
        nblk = int(np.prod([h - l for l, h in zip(clo, chi)]))
        host = seeded_nodes(4096 * nblk, 31 + len(keep)).reshape(nblk, 4096, 8)
        buf = hip.Buffer(host.nbytes)
        buf.upload(host)
        keep.append(buf)
        children.append(P.Slab.of(child, clo, chi, buf.ptr))
        S.add(child, clo, chi, host)
    plan = P.Plan.slab(P.Slab.of(counts, lo, hi, out.ptr), children, 0)
    plan.run()
    hip.synchronize()
    got = out.download(np.zeros(8 * 4096, dtype=np.uint64))
    expect = oracle_block(counts, tuple(lo), S.block)
    assert np.array_equal(got[:expect.size], expect)
    plan.close()


def test_slab_plan_rejects_missing_children():
    P.init()
    out = hip.Buffer(64 * 4096)
    with pytest.raises(P.PentagoError):
        P.Plan.slab(P.Slab.of(bench.SECTION, [0] * 4, [1] * 4, out.ptr), [], 0)


@pytest.mark.parametrize("mode", [0, 1])
def test_mixed_sections_in_one_plan(mode):
    """Boxes of two sections (the benchmark's mix, at 1^4 and 2^4 blocks) in one plan: every
    block of both against the oracle."""
    P.init()
    specs = [(bench.SECTION, [5] * 4, [6] * 4), (bench.SECTION2, [9, 9, 9, 9], [10, 11, 10, 11])]
    outs, kids, plan, n_out, child_nodes = bench.build_boxes(P, hip, specs, mode)
    plan.run()
    hip.synchronize()
    S = SlabSet()
    for child, lo, hi, buf in kids:
        nb = int(np.prod([h - l for l, h in zip(lo, hi)]))
        S.add(child, lo, hi, buf.download(np.zeros(8 * 4096 * nb, dtype=np.uint64)).reshape(nb, 4096, 8))
    for counts, lo, hi, buf in outs:
        nb = int(np.prod([h - l for l, h in zip(lo, hi)]))
        got = buf.download(np.zeros(8 * 4096 * nb, dtype=np.uint64)).reshape(nb, 4096, 8)
        shape = O.section_shape(counts)
        idx = 0
        for b0 in range(lo[0], hi[0]):
            for b1 in range(lo[1], hi[1]):
                for b2 in range(lo[2], hi[2]):
                    for b3 in range(lo[3], hi[3]):
                        blk = (b0, b1, b2, b3)
                        print("mode", mode, "section", counts, "block", blk, flush=True)
                        expect = oracle_block(counts, blk, S.block)
                        assert np.array_equal(got[idx].reshape(-1)[:expect.size], expect), (counts, blk)
                        idx += 1
    plan.close()


def download_outs(outs):
    res = []
    for counts, lo, hi, buf in outs:
        nb = int(np.prod([h - l for l, h in zip(lo, hi)]))
        res.append(buf.download(np.zeros(8 * 4096 * nb, dtype=np.uint64)))
    return res


def test_queue_launch_matches_grid_and_oracle():
    """The persistent queue-fed form of the fused kernel (per-XCD queues, work taken from the
    other XCDs' queues at the end) on boxes of two sections of uneven size: every workgroup
    count gives the grid form's output bit for bit, and sampled blocks equal the oracle."""
    P.init()
    specs = [(bench.SECTION, [3] * 4, [6, 6, 5, 6]), (bench.SECTION2, [9, 9, 9, 9], [10, 11, 10, 11])]
    outs, kids, plan, n_out, child_nodes = bench.build_boxes(P, hip, specs, 0)
    plan.set_launch(0)
    plan.run()
    hip.synchronize()
    grid = download_outs(outs)
    for wgs in (1, 2, 4, 6):
        for _, _, _, buf in outs:
            buf.zero()
        plan.set_launch(wgs)
        print("queue form, %d workgroups per CU" % wgs, flush=True)
        plan.run()
        plan.run()   # a second run resets the queue heads
        hip.synchronize()
        for a, b in zip(grid, download_outs(outs)):
            assert np.array_equal(a, b), wgs
    S = SlabSet()
    for child, lo, hi, buf in kids:
        nb = int(np.prod([h - l for l, h in zip(lo, hi)]))
        S.add(child, lo, hi, buf.download(np.zeros(8 * 4096 * nb, dtype=np.uint64)).reshape(nb, 4096, 8))
    got = download_outs(outs)
    for (counts, lo, hi, _), g in zip(outs, got):
        nb = int(np.prod([h - l for l, h in zip(lo, hi)]))
        g = g.reshape(nb, 4096, 8)
        for idx in (0, nb - 1):
            blk = np.unravel_index(idx, [h - l for l, h in zip(lo, hi)])
            blk = tuple(int(lo[a] + blk[a]) for a in range(4))
            expect = oracle_block(counts, blk, S.block)
            assert np.array_equal(g[idx].reshape(-1)[:expect.size], expect), (counts, blk)
    plan.close()


def test_autotune_picks_a_form_and_keeps_the_output():
    P.init()
    specs = [(bench.SECTION, [3] * 4, [5] * 4), (bench.SECTION2, [9] * 4, [10] * 4)]
    outs, kids, plan, n_out, child_nodes = bench.build_boxes(P, hip, specs, 0)
    plan.set_launch(0)
    plan.run()
    hip.synchronize()
    before = download_outs(outs)
    times, choice = plan.autotune()
    print("autotune", times, choice, flush=True)
    assert set(times) == {0, 2, 3, 4, 5, 6} and choice in times
    assert times[choice] == min(times.values())
    for _, _, _, buf in outs:
        buf.zero()
    plan.run()
    hip.synchronize()
    for a, b in zip(before, download_outs(outs)):
        assert np.array_equal(a, b)
    plan.close()
    # the choice became the device default of later plans; put the grid form back
    plan2 = bench.build_boxes(P, hip, specs[:1], 0)[2]
    assert P.lib().pe_end_plan_set_launch(plan2.handle, 0) == 0
    P.set_default_launch(0)
    plan2.close()


class DeviceSlabs:
    """Child blocks read from device slabs on demand (for boxes too large to copy to the host)."""

    def __init__(self, kids):
        self.kids = kids

    def block(self, counts, cb):
        for c, lo, hi, buf in self.kids:
            if c != counts or any(not (lo[a] <= cb[a] < hi[a]) for a in range(4)):
                continue
            idx = 0
            for a in range(4):
                idx = idx * (hi[a] - lo[a]) + (cb[a] - lo[a])
            shape = O.section_shape(counts)
            bshape = [min(shape[a], 8 * (cb[a] + 1)) - 8 * cb[a] for a in range(4)]
            n = int(np.prod(bshape))
            return buf.download(np.zeros(8 * n, dtype=np.uint64), offset=64 * 4096 * idx)
        raise KeyError((counts, cb))


def test_full_size_mixed_box_samples():
    """The benchmark's full workload (22^4 blocks of (3,3)^4 and 8^4 blocks of the second
    section, ~193 GB resident): two sampled blocks of each section against the oracle."""
    P.init()
    free, _ = hip.mem_info()
    specs = [(bench.SECTION, [0] * 4, [22] * 4), (bench.SECTION2, [8] * 4, [16] * 4)]
    need = bench.boxes_bytes(P, specs)
    if need > 0.9 * free:
        pytest.skip("needs %.0f GB of device memory, %.0f GB free" % (need / 1e9, free / 1e9))
    outs, kids, plan, n_out, child_nodes = bench.build_boxes(P, hip, specs, 0)
    plan.run()
    hip.synchronize()
    D = DeviceSlabs(kids)
    rng = np.random.default_rng(2024)
    for counts, lo, hi, buf in outs:
        for _ in range(2):
            blk = tuple(int(rng.integers(lo[a], hi[a])) for a in range(4))
            idx = 0
            for a in range(4):
                idx = idx * (hi[a] - lo[a]) + (blk[a] - lo[a])
            expect = oracle_block(counts, blk, D.block)
            got = buf.download(np.zeros(expect.size, dtype=np.uint64), offset=64 * 4096 * idx)
            print("full box", counts, "block", blk, flush=True)
            assert np.array_equal(got, expect), (counts, blk)
    plan.close()
