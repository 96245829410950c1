"""Whole sections computed in passes over subsets of their open dimensions (the form the
benchmark's slice-30 section runs in: output resident, only the pass's child sections
resident, fused launches with the other dimensions closed; the first pass stores a partial in
the slab slots, middle passes add to it and the last finishes it) against the fused slab plan
with every child resident, on every block, and against the oracle on sampled blocks
(pentago/end/compute.cc:219-371 per line, pentago/end/block_store.cc:173-229 merge)."""
import numpy as np
import pytest

import bench
import pentago_amd as P
from pentago_amd import hip

pytestmark = pytest.mark.gpu

SEED = 0x7E570000


@pytest.mark.parametrize("counts,groups", [
    (((1, 7), (1, 7), (7, 1), (7, 1)), [[0], [1], [2], [3]]),
    (((1, 7), (1, 7), (7, 1), (7, 1)), [[0, 3], [1, 2]]),
    (((3, 5), (2, 6), (5, 3), (6, 2)), [[0], [1, 2, 3]]),
    (((3, 5), (2, 6), (5, 3), (6, 2)), None),
    # quadrant 0 full (dimension 0 closed): the finishing pass's fields for it carry no moves
    (((5, 4), (1, 7), (7, 1), (2, 3)), [[1], [2], [3]]),
    (((5, 4), (1, 7), (7, 1), (2, 3)), [[3], [1, 2]]),
])
def test_section_by_passes_matches_fused_plan_and_oracle(counts, groups):
    P.init()
    shape = P.section_shape(counts)
    nb = [(s + 7) // 8 for s in shape]
    rng = np.random.default_rng(sum(shape))
    samples = [tuple(int(rng.integers(n)) for n in nb) for _ in range(2)] + [tuple(n - 1 for n in nb), (0, 0, 0, 0)]
    print("section", counts, "shape", shape, "samples", samples, flush=True)
    stream = hip.Stream()
    out, nb, passes, expect = bench.section_by_passes(P, hip, counts, samples, stream, seed_base=SEED, groups=groups)
    print("passes", passes, flush=True)
    # the fused plan over the whole section, every child section resident with the same data
    kids = []
    for d in range(4):
        if sum(counts[d]) >= 9:
            continue
        child = P.line_describe(P.make_line(counts, d, (0, 0, 0))).standard_child.counts
        if any(c == child for c, _ in kids):
            continue
        cnb = [(s + 7) // 8 for s in P.section_shape(child)]
        buf = hip.Buffer(64 * 4096 * int(np.prod(cnb)))
        P.fill_seeded(buf.ptr, 4096 * int(np.prod(cnb)), SEED + sum(c * 10 ** i for i, c in enumerate(sum(child, ()))))
        kids.append((child, (cnb, buf)))
    fused = hip.Buffer(64 * 4096 * int(np.prod(nb)))
    plan = P.Plan.slab(P.Slab.of(counts, [0] * 4, nb, fused.ptr),
                       [P.Slab.of(c, [0] * 4, cnb, buf.ptr) for c, (cnb, buf) in kids], 0)
    plan.run()
    hip.synchronize()
    plan.close()
    a = out.download(np.zeros(8 * 4096 * int(np.prod(nb)), dtype=np.uint64)).reshape(-1, 4096 * 8)
    b = fused.download(np.zeros(8 * 4096 * int(np.prod(nb)), dtype=np.uint64)).reshape(-1, 4096 * 8)
    for idx, blk in enumerate(np.ndindex(*nb)):
        m = 8 * int(np.prod([min(8, shape[x] - 8 * blk[x]) for x in range(4)]))
        assert np.array_equal(a[idx, :m], b[idx, :m]), ("passes vs fused", blk)
    for blk, e in expect.items():
        idx = bench.slab_index(nb, blk)
        assert np.array_equal(a[idx, :e.size], e), ("passes vs oracle", blk)


def test_finishing_pass_is_not_autotuned():
    """A finishing pass (accumulate 3) replaces the partial it reads by the result, so the
    launch autotuner, which runs a plan repeatedly, refuses it; the other modes are accepted."""
    P.init()
    counts = ((1, 7), (1, 7), (7, 1), (7, 1))
    nb = [(s + 7) // 8 for s in P.section_shape(counts)]
    child = P.line_describe(P.make_line(counts, 0, (0, 0, 0))).standard_child.counts
    cnb = [(s + 7) // 8 for s in P.section_shape(child)]
    out = hip.Buffer(64 * 4096 * int(np.prod(nb)))
    kid = hip.Buffer(64 * 4096 * int(np.prod(cnb)))
    P.fill_seeded(kid.ptr, 4096 * int(np.prod(cnb)), SEED)
    children = [P.Slab.of(child, [0] * 4, cnb, kid.ptr)]
    plan = P.Plan.slab_dims(P.Slab.of(counts, [0] * 4, nb, out.ptr), children, 1, accumulate=3)
    with pytest.raises(P.PentagoError):
        plan.autotune()
    plan.close()
    plan = P.Plan.slab_dims(P.Slab.of(counts, [0] * 4, nb, out.ptr), children, 1, accumulate=2)
    try:
        plan.autotune()
        hip.synchronize()
    finally:
        plan.close()
        P.set_default_launch(0)   # the autotuner's choice would carry over to later plans


@pytest.mark.parametrize("groups", [[[0], [1], [2], [3]], [[1, 2, 3], [0]]])
def test_or_merged_passes_match_fused_plan(groups):
    """accumulate 0 then 1: each pass a complete result over its dimensions, OR-merged into the
    slots (the block store's merge, pentago/end/block_store.cc:173-229), equals the fused plan
    over every dimension on every block."""
    P.init()
    counts = ((1, 7), (1, 7), (7, 1), (7, 1))
    shape = P.section_shape(counts)
    nb = [(s + 7) // 8 for s in shape]
    kids = {}
    for d in range(4):
        child = P.line_describe(P.make_line(counts, d, (0, 0, 0))).standard_child.counts
        if child not in kids:
            cnb = [(s + 7) // 8 for s in P.section_shape(child)]
            buf = hip.Buffer(64 * 4096 * int(np.prod(cnb)))
            P.fill_seeded(buf.ptr, 4096 * int(np.prod(cnb)), SEED + sum(c * 10 ** i for i, c in enumerate(sum(child, ()))))
            kids[child] = (cnb, buf)
    slabs = [P.Slab.of(c, [0] * 4, cnb, buf.ptr) for c, (cnb, buf) in kids.items()]
    nbytes = 64 * 4096 * int(np.prod(nb))
    merged, fused = hip.Buffer(nbytes), hip.Buffer(nbytes)
    for k, group in enumerate(groups):
        print("pass", group, "accumulate", int(k > 0), flush=True)
        plan = P.Plan.slab_dims(P.Slab.of(counts, [0] * 4, nb, merged.ptr), slabs, sum(1 << d for d in group),
                                accumulate=int(k > 0))
        plan.run()
        hip.synchronize()
        plan.close()
    plan = P.Plan.slab(P.Slab.of(counts, [0] * 4, nb, fused.ptr), slabs, 0)
    plan.run()
    hip.synchronize()
    plan.close()
    a = merged.download(np.zeros(nbytes // 8, dtype=np.uint64)).reshape(-1, 4096 * 8)
    b = fused.download(np.zeros(nbytes // 8, dtype=np.uint64)).reshape(-1, 4096 * 8)
    for idx, blk in enumerate(np.ndindex(*nb)):
        m = 8 * int(np.prod([min(8, shape[x] - 8 * blk[x]) for x in range(4)]))
        assert np.array_equal(a[idx, :m], b[idx, :m]), blk
