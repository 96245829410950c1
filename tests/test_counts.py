"""Counts and section enumeration, CPU side.

- popcounts_over_stabilizers (oracle) against the reference's own slow count over orbits
  (pentago/base/pentago_test.cc:203-235): for random boards with repeated quadrants and
  meaningless data, every distinct board reachable by a symmetry that keeps each quadrant's
  rotation class is counted once.
- Whole-section counts of the oracle (recursive evaluator + popcounts_over_stabilizers)
  against the production run's counts files (data/counts-<slice>.npy, rows
  (sig, black wins, white wins, total), pentago/mpi/io.cc:573-581) for the small sections of
  slices 35, 34 and 33.
- The library's section enumeration (descendent_sections, pentago/end/sections.cc:18-49)
  against the section lists of all 36 counts files.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import pentago_amd as P
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_rows(slice_):
    return np.load(os.path.join(GOLDEN, "counts-%d.npy" % slice_), allow_pickle=False)


def sections_of(rows):
    return [tuple(((int(s) >> (16 * q)) & 255, (int(s) >> (16 * q + 8)) & 255) for q in range(4)) for s in rows[:, 0]]


def random_board(rng, stones):
    cells = rng.permutation(36)[:stones]
    quads = [0, 0, 0, 0]
    for k, c in enumerate(cells):
        q, i = divmod(int(c), 9)
        quads[q] += (1 if k % 2 == 0 else 2) * 3 ** i
    return quads


def board_of(quads):
    return sum(int(q) << (16 * i) for i, q in enumerate(quads))


def test_popcounts_over_stabilizers_against_orbit_count():
    rng = np.random.default_rng(83191941)
    full = (1 << 256) - 1
    stabilized = 0
    canon = {}

    def cls(q):
        if q not in canon:
            s0, s1 = O.lib().po_unpack(q, 0), O.lib().po_unpack(q, 1)
            canon[q] = (bin(s0).count("1"), bin(s1).count("1"), O.rmin_inverse(q) >> 2)
        return canon[q]

    for step in range(40):
        quads = random_board(rng, int(rng.integers(0, 37)))
        pick = int(rng.integers(0, 256))
        quads = [quads[pick >> (2 * i) & 3] for i in range(4)]
        board = board_of(quads)
        wins = O.super_meaningless(board)
        fast = O.popcounts_over_stabilizers(board, wins, full & ~wins)
        assert fast is not None and fast[0] + fast[1] == fast[2]
        seen, slow0 = set(), 0
        base = [cls(q) for q in quads]
        for g in range(8):
            for l in range(256):
                b = O.transform_board((g, l), board)
                if all(cls((b >> (16 * q)) & 0xffff) == base[q] for q in range(4)) and b not in seen:
                    seen.add(b)
                    slow0 += O.meaningless(b)
        assert (slow0, len(seen) - slow0, len(seen)) == fast, (step, board)
        stabilized += fast[2] < 256
    assert stabilized >= 5   # the repeated quadrants make stabilized boards common


def test_popcounts_rejects_non_invariant_data():
    # the empty board is fixed by every local rotation (shift 8): one rotation's bit alone
    # is not a super evaluation of it
    assert O.popcounts_over_stabilizers(0, 1, 1) is None
    assert O.popcounts_over_stabilizers(0, (1 << 256) - 1, (1 << 256) - 1) == (1, 1, 1)


def small_sections(slice_, limit):
    rows = golden_rows(slice_)
    secs = sections_of(rows)
    sizes = [int(np.prod(O.section_shape(s))) for s in secs]
    pick = [k for k in np.argsort(sizes, kind="stable") if sizes[k] <= limit]
    return rows, secs, pick


@pytest.mark.parametrize("slice_,limit,count", [(35, 20000, 40), (34, 3000, 24), (33, 400, 6)])
def test_oracle_section_counts_match_production(slice_, limit, count):
    rows, secs, pick = small_sections(slice_, limit)
    pick = pick[:count]
    assert len(pick) == count
    O.lib().po_eval_reset()
    got = [O.count_section(secs[k]) for k in pick]
    want = rows[pick]
    assert np.array_equal(P.counts_rows([secs[k] for k in pick], got, slice_), want)
    # some picked sections have wins for both colors, so the orientation of the rows matters
    assert any(int(r[1]) and int(r[2]) for r in want) or slice_ == 35


def test_counts_rows_orientation():
    secs = [((0, 0),) * 4]
    # (wins, wins-or-ties, total) for the mover -> (black wins, white wins): even slices keep
    # the mover's wins first, odd slices swap
    assert P.counts_rows(secs, [(3, 10, 16)], 34).tolist() == [[0, 3, 6, 16]]
    assert P.counts_rows(secs, [(3, 10, 16)], 35).tolist() == [[0, 6, 3, 16]]


def test_descendent_sections_match_production_lists():
    want = json.load(open(os.path.join(GOLDEN, "section_sigs.json")))
    root = ((0, 0),) * 4
    for slice_ in range(36):
        secs = P.descendent_sections(root, slice_)
        sigs = np.sort(np.array([P.section_sig(s) for s in secs], dtype="<u8"))
        assert len(sigs) == want[str(slice_)]["n"], slice_
        assert hashlib.sha1(sigs.tobytes()).hexdigest() == want[str(slice_)]["sha1"], slice_
    # the full list at slices 33..35 in the library's order (sorted by sig)
    for slice_ in (33, 34, 35):
        secs = P.descendent_sections(root, slice_)
        assert [P.section_sig(s) for s in secs] == sorted(int(x) for x in golden_rows(slice_)[:, 0])


def test_descendent_sections_of_a_small_root():
    # the sections a 30-stone root reaches at 32 stones: two stones of alternating colour
    root = ((3, 4), (4, 4), (4, 4), (4, 3))
    got = P.descendent_sections(root, 32)
    assert got and all(sum(map(sum, s)) == 32 for s in got)
    expect = set()
    for a in range(4):
        for b in range(4):
            c = [list(x) for x in root]
            c[a][0] += 1
            c[b][1] += 1
            if all(sum(x) <= 9 for x in c):
                expect.add(P.section_sig(P.standardize(tuple(map(tuple, c)))[0]))
    assert sorted(P.section_sig(s) for s in got) == sorted(expect)


def test_slice_driver_grouping():
    from pentago_amd import slices as S
    root = ((0, 0),) * 4
    s32 = P.descendent_sections(root, 32)
    kids = set()
    for s in s32:
        kids.update(S.child_sections(s))
    assert sorted(kids) == sorted(P.descendent_sections(root, 33))
    budget = 150e9
    gs = S.plan_groups(s32, budget, True)
    assert sorted(s for g in gs for s in g) == sorted(s32) and len(gs) > 1
    kids = {s: S.child_sections(s) for s in s32}
    own = {s: P.slice_bytes([s]) for s in s32}
    kb = {c: P.slice_bytes([c]) for s in s32 for c in kids[s]}

    def need(g):
        return sum(own[s] for s in g) + sum(kb[c] for c in {c for s in g for c in kids[s]})

    later = set(s32)
    for g in gs:
        assert need(g) <= budget or len(g) == 1
        later -= set(g)
        # a group is closed only when no remaining section fits in it
        assert all(need(g + [s]) > budget for s in later)
    # sharing children: fewer child bytes regenerated than taking the sections in order
    regen = sum(need(g) - sum(own[s] for s in g) for g in gs)
    seq, cur = 0, []
    for s in s32:
        if cur and need(cur + [s]) > budget:
            seq += need(cur) - sum(own[x] for x in cur)
            cur = []
        cur.append(s)
    seq += need(cur) - sum(own[x] for x in cur)
    assert regen < seq
    # without regeneration only the sections' own slabs count
    gs = S.plan_groups(s32, budget, False)
    assert all(P.slice_bytes(g) <= budget for g in gs)
    assert all(P.slice_bytes(g + h[:1]) > budget for g, h in zip(gs, gs[1:]))


def test_slice24_enumeration_totals():
    """pentago/end/end_test.cc:110-118: all 24-stone sections hold 1,921,672,470,396 nodes in
    500,235,319 blocks of 8^4 and 95,263,785 block lines (lines only along dimensions whose
    quadrant has an empty cell)."""
    sections = P.descendent_sections(((0, 0),) * 4, 24)
    nodes = blocks = lines = 0
    for c in sections:
        shape = P.section_shape(c)
        nb = [(s + 7) // 8 for s in shape]
        nodes += int(np.prod(shape, dtype=np.int64))
        b = int(np.prod(nb, dtype=np.int64))
        blocks += b
        lines += sum(b // nb[d] for d in range(4) if sum(c[d]) < 9)
    assert (nodes, blocks, lines) == (1921672470396, 500235319, 95263785)
