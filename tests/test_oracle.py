"""Pins for the oracle (CPU restatement) against the reference's own known answers,
closed forms and brute force.  No GPU.

Reference tests mirrored: pentago/base/super_test.cc, pentago/mid/mid_test.cc,
pentago/mid/benchmark_test.cc, pentago/high/high_test.cc.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.common import golden, random_sides, random_side, SIDE_MASK

MASK64 = 2**64 - 1


def test_table_hashes():
    """precompute.cc:169-1038: every generated table hashes to the reference digest."""
    import ctypes
    for name, digest in golden("table_hashes.json").items():
        buf = ctypes.create_string_buffer(41)
        assert O.lib().po_table_hash(name.encode(), buf) == 0
        assert buf.value.decode() == digest, name


def rotate_quadrant_left(side, q, times):
    for _ in range(times):
        quad = (side >> (16 * q)) & 0x1ff
        new = 0
        for x in range(3):
            for y in range(3):
                if quad >> (3 * x + y) & 1:
                    new |= 1 << (3 * (2 - y) + x)
        side = (side & ~(0x1ff << (16 * q))) | new << (16 * q)
    return side


def test_super_wins_brute_force():
    """super_test.cc:20-69: super_wins(side)(r) == won(rotated side) for all 256 r."""
    rng = np.random.default_rng(1740291)
    for count in [5, 8, 12, 18, 24, 30]:
        for _ in range(6):
            side = random_side(rng, count / 36)
            wins = O.super_wins(side)
            for r in range(256):
                rs = side
                for q in range(4):
                    rs = rotate_quadrant_left(rs, q, (r >> (2 * q)) & 3)
                assert bool(wins >> r & 1) == O.won(rs), (side, r)


def won_python(side):
    """Independent five-in-a-row test on the 6x6 grid (cell (x,y) at 16*(2*(x//3)+y//3)+3*(x%3)+y%3)."""
    def cell(x, y):
        return side >> (16 * (2 * (x // 3) + y // 3) + 3 * (x % 3) + y % 3) & 1
    for dx, dy in ((1, 0), (0, 1), (1, 1), (1, -1)):
        for x in range(6):
            for y in range(6):
                if all(0 <= x + i * dx < 6 and 0 <= y + i * dy < 6 and cell(x + i * dx, y + i * dy) for i in range(5)):
                    return True
    return False


def test_won_matches_grid_definition():
    rng = np.random.default_rng(5)
    for _ in range(3000):
        side = random_side(rng, rng.uniform(0.2, 0.8))
        assert O.won(side) == won_python(side)


def test_rmax_closed_form():
    """super_test.cc:71-87: rmax(s)(r) = OR over the 8 single rotations b of s(r+b)."""
    rng = np.random.default_rng(1740291)
    for _ in range(200):
        s = int.from_bytes(rng.bytes(32), "little")
        rs = O.rmax(s)
        for r in range(256):
            c = [(r >> (2 * q)) & 3 for q in range(4)]
            o = 0
            for q in range(4):
                for d in (1, -1):
                    cc = list(c)
                    cc[q] = (cc[q] + d) & 3
                    o |= s >> (cc[0] + 4 * cc[1] + 16 * cc[2] + 64 * cc[3]) & 1
            assert (rs >> r & 1) == o


def test_group_axioms():
    """super_test.cc:98-150 (sampled): identity, inverse, associativity."""
    rng = np.random.default_rng(3)
    syms = [(int(rng.integers(8)), int(rng.integers(256))) for _ in range(60)]
    e = (0, 0)
    for g in syms:
        assert O.sym_mul(g, e) == g and O.sym_mul(e, g) == g
        assert O.sym_mul(g, O.sym_inverse(g)) == e
    for a in syms[:15]:
        for b in syms[:15]:
            for c in syms[:6]:
                assert O.sym_mul(O.sym_mul(a, b), c) == O.sym_mul(a, O.sym_mul(b, c))


def test_action_consistency():
    """super_test.cc:152-164: transform_side is a group action."""
    rng = np.random.default_rng(875431)
    for _ in range(500):
        s0 = (int(rng.integers(8)), int(rng.integers(256)))
        s1 = (int(rng.integers(8)), int(rng.integers(256)))
        side = random_side(rng)
        assert O.transform_side(s0, O.transform_side(s1, side)) == O.transform_side(O.sym_mul(s0, s1), side)


def test_transform_super_definition():
    """super_test.cc:187-201: transform_super(s, C) = {g z y' g' : z in C}."""
    rng = np.random.default_rng(72831)
    for _ in range(300):
        g, l = int(rng.integers(8)), int(rng.integers(256))
        C = int.from_bytes(rng.bytes(32), "little") & int.from_bytes(rng.bytes(32), "little")
        sg, sl = (g, 0), (0, l)
        right = O.sym_mul(O.sym_inverse(sl), O.sym_inverse(sg))
        correct = 0
        for r in range(256):
            if C >> r & 1:
                correct |= 1 << O.sym_mul(O.sym_mul(sg, (0, r)), right)[1]
        assert O.transform_super((g, l), C) == correct


def test_meaningless_golden_and_equivariance():
    """super_test.cc:223-240 golden values, then super_test.cc:203-211 equivariance."""
    for board, m in golden("meaningless.json"):
        assert O.meaningless(board) == m
    rng = np.random.default_rng(99)
    for _ in range(12):
        s = (int(rng.integers(8)), int(rng.integers(256)))
        s0, s1 = random_sides(rng, 7, 7)
        board = O.pack_board(s0, s1)
        assert O.super_meaningless(O.transform_board(s, board)) == O.transform_super(s, O.super_meaningless(board))


def test_halfsuper_relations():
    """mid_test.cc:174-249: split/merge roundtrip, rmax halves, halfsuper_wins halves."""
    rng = np.random.default_rng(667731)
    for step in range(300):
        if step < 256:
            s = 1 << step
        else:
            s = (2**256 - 1)
            for _ in range(4):
                s &= int.from_bytes(rng.bytes(32), "little")
        h0, h1 = O.split(s)
        assert O.merge(h0, h1) == s
        assert O.merge(O.half_rmax(h1), O.half_rmax(h0)) == O.rmax(s)
    for _ in range(2000):
        side = random_side(rng)
        assert O.merge(O.halfsuper_wins(side, 0), O.halfsuper_wins(side, 1)) == O.super_wins(side)


def test_rmax_limit_is_parity():
    """mid_test.cc:253-272: iterating rmax from a singleton reaches the parity set."""
    s, seen = 1, []
    while s not in seen:
        seen.append(s)
        s = O.rmax(s)
    for r in range(256):
        assert (s >> r & 1) == (sum((r >> (2 * q)) & 3 for q in range(4)) & 1)


def test_subsets():
    """mid_test.cc:274-325: subsets in large-entry-major order; index and mask round trips."""
    from itertools import combinations
    for n in range(19):
        for k in range(4):
            sets = sorted(combinations(range(n), k), key=lambda c: tuple(reversed(c)))
            assert len(sets) == O.lib().po_choose(n, k)
            for i, c in enumerate(sets):
                packed = sum(v << (5 * j) for j, v in enumerate(c))
                assert O.lib().po_subset(n, k, i) == packed
                assert O.lib().po_subset_index(n, k, packed) == i
                assert O.lib().po_subset_mask(n, k, i) == sum(1 << v for v in c)


def test_bottleneck_golden():
    """mid_test.cc:129-134."""
    g = golden("bottleneck.json")
    for sl, size in zip(g["slices"], g["sizes"]):
        assert O.lib().po_midsolve_workspace_size(sl) == size


def test_high_board_conversions():
    """high_test.cc:12-28."""
    rng = np.random.default_rng(5554)
    for stones in range(0, 37, 3):
        b, w = (stones + 1) // 2, stones // 2
        s0, s1 = random_sides(rng, b, w)
        board = O.pack_board(s0, s1)
        for middle in (0, 1):
            h = O.high_from_board(board, middle)
            assert O.lib().po_high_board(h) == board
            raw = O.lib().po_high_raw(h)
            assert raw >> 63 == middle
            assert O.lib().po_high_board(O.high_from_board(raw & (2**63 - 1), raw >> 63)) == board


def test_midsolve_kat():
    """benchmark_test.cc:39-58: all 163 values of the 18-stone known answer."""
    kat = golden("mid_kat.json")
    root = O.high_from_board(kat["board"], 0)
    values = O.midsolve(root)
    assert len(values) == 1 + 18 + 8 * 18
    expect = {b | (m << 63): v for b, m, v in kat["values"]}
    assert values == expect


def brute_mid_values(root):
    """Brute-force values of the boards midsolve reports, from game rules alone."""
    out = {}
    L = O.lib()

    def value(h):
        v = O.C.c_int()
        if L.po_high_done_and_value(h, O.C.byref(v)):
            return v.value
        turn = (h.ply >> 1) & 1
        if not (h.ply & 1):
            me, them = h.side[turn], h.side[1 - turn]
            return O.brute_value(me, them)
        best = -1
        for q in range(4):
            for d in (1, -1):
                best = max(best, -value(L.po_high_rotate(h, q, d)))
        return best

    for raw in O.midsolve(root):
        h = O.high_from_board(raw & (2**63 - 1), raw >> 63)
        out[raw] = value(h)
    return out


def test_midsolve_brute_force():
    """mid_test.cc:95-126 pattern: midsolve values equal an independent game search."""
    rng = np.random.default_rng(2223)
    for stones in (31, 32, 33, 34, 35):
        for trial in range(3):
            for middle in (0, 1):
                while True:
                    s0, s1 = random_sides(rng, (stones + 1) // 2, stones // 2)
                    if not O.won(s0) and not O.won(s1):
                        break
                root = O.high_from_board(O.pack_board(s0, s1), middle)
                got = O.midsolve(root)
                assert got == brute_mid_values(root)


def convert_node(node0, node1, turn):
    """(we-win, we-win-or-tie) -> (black-wins, white-wins), pentago/end/check.cc:153-156."""
    a, b = node0, (~node1) & (2**256 - 1)
    return (b, a) if turn else (a, b)


def sample_index(board):
    idx = []
    for q in range(4):
        ir = O.rmin_inverse((board >> (16 * q)) & 0xffff)
        assert ir & 3 == 0
        idx.append(ir >> 2)
    return idx


def board_section(board):
    counts = []
    for q in range(4):
        quad = (board >> (16 * q)) & 0xffff
        b = w = 0
        for i in range(9):
            d = quad % 3
            quad //= 3
            b += d == 1
            w += d == 2
        counts.append((b, w))
    return tuple(counts)


@pytest.mark.parametrize("slice_,count", [(35, 32), (34, 32), (33, 32)])
def test_end_line_production_samples(slice_, count):
    """high_test.cc:41-58 samples from the production run, checked against the oracle's
    line kernel through the recursive evaluator (format conversion as check.cc:149-158)."""
    data = golden("samples.json")[str(slice_)]
    O.lib().po_eval_reset()
    for board, wins in list(zip(data["boards"], data["wins"]))[:count]:
        counts = board_section(board)
        std, g = O.standardize(counts)
        assert std == counts and g == 0
        idx = sample_index(board)
        n0, n1 = O.eval_node(counts, idx)
        bw, ww = convert_node(n0, n1, slice_ & 1)
        expect_b = sum(wins[i] << (64 * i) for i in range(4))
        expect_w = sum(wins[4 + i] << (64 * i) for i in range(4))
        assert (bw, ww) == (expect_b, expect_w), board
    O.lib().po_eval_reset()


def test_end_line_brute_force_slice34():
    """Whole nodes of a 34-stone section against brute-force minimax."""
    rng = np.random.default_rng(11)
    counts = ((4, 4), (4, 5), (4, 4), (5, 4))    # 34 stones, black to move, empties in q0 and q2
    std, _ = O.standardize(counts)
    shape = O.section_shape(std)
    turn = sum(b + w for b, w in std) & 1
    O.lib().po_eval_reset()
    for _ in range(6):
        idx = [int(rng.integers(s)) for s in shape]
        quads = [int(O.rmin(*std[q])[0][idx[q]]) for q in range(4)]
        board = sum(quads[q] << (16 * q) for q in range(4))
        me, them = O.unpack_board(board, turn), O.unpack_board(board, 1 - turn)
        assert O.eval_node(std, idx) == O.brute_super(me, them)
    O.lib().po_eval_reset()


def result_boards(s0, s1):
    """The boards of midsolve_internal's results (pentago/mid/mid_test.cc:32-44): the root,
    then the root with the next stone on each empty cell in increasing bit order."""
    count = bin(s0 | s1).count("1")
    boards = [(s0, s1)]
    empty = 0x01ff01ff01ff01ff & ~(s0 | s1)
    for bit in range(64):
        if empty >> bit & 1:
            after = [s0, s1]
            after[count & 1] |= 1 << bit
            boards.append(tuple(after))
    return boards


@pytest.mark.parametrize("stones,boards", [(35, 4), (34, 4), (33, 4), (32, 4), (31, 2), (30, 1)])
def test_midsolve_internal_all_bits_brute_force(stones, boards):
    """mid_test.cc:45-93: every one of the 128 bits of every (win, notlose) result of
    midsolve_internal equals a brute-force game search of its board under that rotation;
    the halfsuper of parity p sits at the rotations merge() gives parity p, the root at the
    board's own parity (middle) and its children at the other."""
    rng = np.random.default_rng(5554 + stones)
    for trial in range(boards):
        for middle in (0, 1):
            while True:
                s0, s1 = random_sides(rng, (stones + 1) // 2, stones // 2)
                if not O.won(s0) and not O.won(s1):
                    break
            root = O.high_from_board(O.pack_board(s0, s1), middle)
            results = O.midsolve_internal(root)[:37 - stones]
            rboards = result_boards(s0, s1)
            assert len(results) == len(rboards) == 37 - stones
            for i, ((b0, b1), (win, notlose)) in enumerate(zip(rboards, results)):
                assert win & ~notlose == 0, (stones, middle, i)
                turn = bin(b0 | b1).count("1") & 1
                me, them = (b1, b0) if turn else (b0, b1)
                rparity = (i > 0) != bool(middle)
                expand = (lambda h: O.merge(0, h)) if rparity else (lambda h: O.merge(h, 0))
                known = expand(2**128 - 1)
                correct_win, correct_notlose = O.brute_super(me, them)
                assert (correct_win ^ expand(win)) & known == 0, (stones, middle, i, hex(b0), hex(b1))
                assert (correct_notlose ^ expand(notlose)) & known == 0, (stones, middle, i, hex(b0), hex(b1))
