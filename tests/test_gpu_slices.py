"""Whole slices on the GPU against the production run.

Slices 35, 34, 33 and 32 (every section the empty board reaches, pentago/end/sections.cc:18-49)
are computed on the device in slice stores -- slice 35 from nothing, 34 from 35, 33 from 34
in groups that fit next to slice 34, 32 in groups whose slice-33 children are regenerated
from slice 34 -- and each section's count_block_wins total
(pentago/end/block_store.cc:330-349) is compared, converted as pentago/mpi/io.cc:573-581,
with the reference's data/counts-<slice>.npy rows (the production run's counts, committed
as fixtures).  The production sparse samples of each slice (pentago/high/high_test.cc:41-58)
are read out of the stores (pentago/end/check.cc:149-158 format).  The count kernel itself
is checked on seeded data, stabilized boards included, against the oracle's
popcounts_over_stabilizers.
"""
import os

import numpy as np
import pytest

import pentago_amd as P
from pentago_amd import hip
from pentago_amd import slices as S
from oracle import oracle as O
from tests.common import golden

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EMPTY = ((0, 0),) * 4
_stores = {}


def golden_rows(slice_):
    rows = np.load(os.path.join(GOLDEN, "counts-%d.npy" % slice_), allow_pickle=False)
    return rows[np.argsort(rows[:, 0], kind="stable")]


def board_section(board):
    counts = []
    for q in range(4):
        quad = (board >> (16 * q)) & 0xffff
        b = w = 0
        for _ in range(9):
            d = quad % 3
            quad //= 3
            b += d == 1
            w += d == 2
        counts.append((b, w))
    return tuple(counts)


def check_samples(slice_, slab_of):
    """Production samples whose section slab_of(section) returns (None: not resident)."""
    data = golden("samples.json")[str(slice_)]
    checked = 0
    full = (1 << 256) - 1
    for board, wins in zip(data["boards"], data["wins"]):
        slab = slab_of(board_section(board))
        if slab is None:
            continue
        node = P.slab_samples(slab, [board])[0]
        win = sum(int(node[i]) << (64 * i) for i in range(4))
        wot = sum(int(node[4 + i]) << (64 * i) for i in range(4))
        black, white = win, full & ~wot
        if slice_ & 1:
            black, white = white, black
        assert black == sum(wins[i] << (64 * i) for i in range(4)), board
        assert white == sum(wins[4 + i] << (64 * i) for i in range(4)), board
        checked += 1
    return checked


def store(slice_):
    if slice_ not in _stores:
        secs = P.descendent_sections(EMPTY, slice_)
        s = P.SliceStore(secs)
        s.compute(store(slice_ + 1) if slice_ < 35 else None)
        _stores[slice_] = s
    return _stores[slice_]


def release(slice_):
    s = _stores.pop(slice_, None)
    if s is not None:
        s.close()


def test_count_blocks_seeded_against_oracle():
    # (2,0) has opposite pairs (fixed by a half turn) and (0,1) the centre stone (fixed by a
    # quarter turn), so node stabilizers have 1, 2, 4, 8, 16 and 32 elements, and seeded data
    # is mostly not invariant under them (counted in column 3)
    counts = ((2, 0), (1, 1), (0, 1), (0, 1))
    shape = P.section_shape(counts)
    nb = [(s + 7) // 8 for s in shape]
    blocks = int(np.prod(nb))
    buf = hip.Buffer(64 * 4096 * blocks)
    P.fill_seeded(buf.ptr, 4096 * blocks, 77, 0)
    slab = P.Slab.of(counts, [0] * 4, nb, buf.ptr)
    out = hip.Buffer(32 * blocks)
    P.count_blocks(slab, out.ptr)
    hip.synchronize()
    got = out.download(np.zeros(4 * blocks, dtype=np.uint64)).reshape(blocks, 4)
    data = buf.download(np.zeros(8 * 4096 * blocks, dtype=np.uint64)).reshape(blocks, 4096, 8)
    bits = np.unpackbits(data.view(np.uint8), axis=2).reshape(blocks, 4096, 2, 256).sum(axis=3)
    rm = [O.rmin(*counts[q])[0] for q in range(4)]
    full = (1 << 256) - 1
    want = np.zeros((blocks, 4), dtype=np.uint64)
    shifts = set()
    for k in range(blocks):
        b = np.unravel_index(k, nb)
        bs = [min(8, shape[a] - 8 * b[a]) for a in range(4)]
        for p in range(int(np.prod(bs))):
            i = np.unravel_index(p, bs)
            board = sum(int(rm[q][8 * b[q] + i[q]]) << (16 * q) for q in range(4))
            total = O.popcounts_over_stabilizers(board, full, full)[2]
            shift = (256 // total).bit_length() - 1
            shifts.add(shift)
            w0, w1 = int(bits[k, p, 0]), int(bits[k, p, 1])
            res = O.popcounts_over_stabilizers(
                board, sum(int(data[k, p, j]) << (64 * j) for j in range(4)),
                sum(int(data[k, p, 4 + j]) << (64 * j) for j in range(4)))
            bad = ((w0 | w1) & ((1 << shift) - 1)) != 0
            assert (res is None) == bad
            want[k] += np.array([w0 >> shift, w1 >> shift, total, int(bad)], dtype=np.uint64)
    assert np.array_equal(got, want)
    assert shifts == {0, 1, 2, 3, 4, 5}
    assert 0 < int(want[:, 3].sum()) < int(np.prod(shape))
    # the synchronous per-slab sum agrees
    assert np.array_equal(P.count_slabs([slab])[0], want.sum(axis=0))


def test_count_blocks_many_matches_per_slab_launches():
    # 150 slabs (more than 64, so the kernel's slab search takes two rounds): single blocks,
    # partial boxes and whole small sections, each with its own seeded data
    rng = np.random.default_rng(5)
    secs = [((2, 0), (1, 1), (0, 1), (0, 1)), ((3, 4), (4, 4), (4, 4), (4, 3)), ((1, 2), (2, 1), (0, 0), (3, 3))]
    slabs, bufs = [], []
    for k in range(150):
        counts = secs[k % 3]
        nb = [(s + 7) // 8 for s in P.section_shape(counts)]
        lo = [int(rng.integers(0, n)) for n in nb]
        hi = [int(rng.integers(l + 1, min(n, l + 2) + 1)) for l, n in zip(lo, nb)]
        blocks = int(np.prod([h - l for l, h in zip(lo, hi)]))
        buf = hip.Buffer(64 * 4096 * blocks)
        P.fill_seeded(buf.ptr, 4096 * blocks, 100 + k, 0)
        bufs.append(buf)
        slabs.append(P.Slab.of(counts, lo, hi, buf.ptr))
    total = sum(s.blocks for s in slabs)
    offs = np.cumsum([0] + [s.blocks for s in slabs])
    one, many = hip.Buffer(32 * total), hip.Buffer(32 * total)
    ws = hip.Buffer(P.count_workspace_bytes(len(slabs)))
    for s, o in zip(slabs, offs):
        P.count_blocks(s, one.ptr + 32 * int(o))
    P.count_blocks_many(slabs, many.ptr, ws.ptr)
    hip.synchronize()
    a = one.download(np.zeros((total, 4), dtype=np.uint64))
    b = many.download(np.zeros((total, 4), dtype=np.uint64))
    assert np.array_equal(a, b)
    assert int(a[:, 2].sum()) > 0
    sums = np.array([a[offs[k]:offs[k + 1]].sum(axis=0) for k in range(len(slabs))], dtype=np.uint64)
    assert np.array_equal(P.count_slabs(slabs), sums)


def test_slice_35_matches_production_counts():
    s = store(35)
    counts = s.counts()
    assert int(counts[:, 3].sum()) == 0
    assert np.array_equal(P.counts_rows(s.sections, counts, 35), golden_rows(35))
    assert check_samples(35, lambda sec: s.slab(s.index[sec]) if sec in s.index else None) == 32


def test_slice_34_matches_production_counts():
    s = store(34)
    release(35)
    counts = s.counts()
    assert int(counts[:, 3].sum()) == 0
    assert np.array_equal(P.counts_rows(s.sections, counts, 34), golden_rows(34))
    assert check_samples(34, lambda sec: s.slab(s.index[sec]) if sec in s.index else None) == 32


def test_slice_33_matches_production_counts():
    """Slice 33 (518 GB as slabs) in groups of sections that fit beside slice 34."""
    children = store(34)
    secs = P.descendent_sections(EMPTY, 33)
    free, _ = hip.mem_info()
    budget = free - 16e9
    assert budget > 20e9, "not enough device memory beside slice 34"
    samples = [0]

    def visit(st):
        samples[0] += check_samples(33, lambda sec: st.slab(st.index[sec]) if sec in st.index else None)

    done, counts = S.grouped_counts(secs, budget, children=children, visit=visit)
    assert int(counts[:, 3].sum()) == 0
    assert np.array_equal(P.counts_rows(done, counts, 33), golden_rows(33))
    assert samples[0] == 32


def test_slice_32_matches_production_counts():
    """Slice 32 (1.4 TB dense) in groups, each group's slice-33 children regenerated from
    the resident slice 34 first."""
    grandchildren = store(34)
    secs = P.descendent_sections(EMPTY, 32)
    free, _ = hip.mem_info()
    budget = free - 16e9
    # in a caller arena (pe_slice_grouped_counts_in); one byte short is refused
    need = S.arena_bytes(secs, budget, True)
    assert 0 < need <= budget
    arena = hip.Buffer(need)
    with pytest.raises(P.PentagoError):
        S.grouped_counts(secs, budget, grandchildren=grandchildren, arena=type("A", (), {"ptr": arena.ptr, "nbytes": need - 1})())
    done, counts = S.grouped_counts(secs, budget, grandchildren=grandchildren, arena=arena)
    del arena
    release(34)
    assert int(counts[:, 3].sum()) == 0
    assert np.array_equal(P.counts_rows(done, counts, 32), golden_rows(32))


def test_plan_mixing_slice_34_and_35_outputs():
    """One slab plan with 35-stone outputs (the fused kernel's no-children form) and 34-stone
    outputs (fused kernel) equals the two computed separately: the 35-stone sections through
    line passes on the line kernel (mode 1), the 34-stone ones on a fused plan of their own.
    Every valid node of every block (partial edge blocks included) is compared."""
    s35 = P.descendent_sections(EMPTY, 35)
    s34 = P.descendent_sections(EMPTY, 34)
    small34 = sorted(s34, key=lambda s: int(np.prod(P.section_shape(s))))[:4]
    kids = sorted({c for s in small34 for c in S.child_sections(s)})
    st35 = P.SliceStore(kids)
    st35.compute()
    extra35 = [s for s in s35 if s not in set(kids)][:3]
    a34, b34 = P.SliceStore(small34), P.SliceStore(small34)
    a35, b35 = P.SliceStore(extra35), P.SliceStore(extra35)
    children = [st35.slab(k) for k in range(len(kids))]
    # separately: the 35-stone sections by line passes (mode 1)
    lined = P.Plan.slab([a35.slab(k) for k in range(len(extra35))], [], 1)
    lined.run()
    hip.synchronize()
    lined.close()
    a34.compute(st35)
    # together: one fused plan over both stores' slabs
    plan = P.Plan.slab([b35.slab(k) for k in range(len(extra35))] + [b34.slab(k) for k in range(len(small34))],
                       children, 0)
    plan.run()
    hip.synchronize()
    plan.close()
    for x, y, secs in ((a34, b34, small34), (a35, b35, extra35)):
        assert np.array_equal(x.counts(), y.counts())
        for k, sec in enumerate(secs):
            shape = P.section_shape(sec)
            nb = [(v + 7) // 8 for v in shape]
            n = int(np.prod(nb)) * 4096 * 8
            got = np.zeros(n, dtype=np.uint64)
            want = np.zeros(n, dtype=np.uint64)
            for arr, store in ((got, y), (want, x)):
                hip.check(hip.hip().hipMemcpy(hip.C.c_void_p(arr.ctypes.data), hip.C.c_void_p(store.slab(k).data),
                                              arr.nbytes, hip.D2H), "hipMemcpy D2H")
            # compare the valid prefix of every block slot
            for b in range(int(np.prod(nb))):
                blk = np.unravel_index(b, nb)
                m = 8 * int(np.prod([min(8, shape[a] - 8 * blk[a]) for a in range(4)]))
                assert np.array_equal(got[b * 4096 * 8:b * 4096 * 8 + m], want[b * 4096 * 8:b * 4096 * 8 + m]), (sec, b)
    for s in (st35, a34, b34, a35, b35):
        s.close()
