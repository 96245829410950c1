"""Host-side checks of batched midsolve plans (no GPU): how a batch is split between the
small-board kernel and the large-board launches, how much device memory it needs, and that
tight budgets shrink the plan or are refused."""
import pytest

import pentago_amd as P
from pentago_amd.workloads import board_spots, seeded_boards


def work_bytes(spots):
    """32-byte workspace entries: the larger of two adjacent levels, summed (bottleneck)."""
    from math import comb
    worst, prev = 0, 1
    for n in range(spots + 1):
        nxt = comb(spots, n + 1) * comb(n + 1, (n + 1) // 2) if n + 1 <= spots else 0
        worst = max(worst, prev + nxt)
        prev = nxt
    return (32 * worst + 255) & ~255


def table_bytes(spots):
    """Bytes of one board's tables (set0_info, wins1, cs1ps, subset masks), 256-aligned."""
    from math import comb
    up = lambda x: (x + 255) & ~255
    set0 = sum(comb(spots, n // 2) for n in range(spots + 1))
    wins1 = sum(comb(spots, n - n // 2) for n in range(spots + 1))
    s1p = [comb(spots - n // 2, n - n // 2) for n in range(spots + 1)]
    cs1ps = sum(c * (spots - n // 2) for n, c in enumerate(s1p))
    return up(256 * set0) + up(32 * wins1) + up(2 * cs1ps) + up(4 * sum(s1p))


def expected_bytes(boards, small_spots, workgroups, streams):
    """Device memory of a plan: descriptors, small tables, one workspace per workgroup, one
    tables+workspace slot per large-board stream (descriptor part bounded from above)."""
    small = [board_spots(b) for b in boards if board_spots(b) <= small_spots]
    big = [board_spots(b) for b in boards if board_spots(b) > small_spots]
    wgs = min(workgroups, len(small))
    fixed = sum(table_bytes(s) for s in small)
    slots = wgs * (max(work_bytes(s) for s in small) if small else 0)
    bigs = (max(streams, 1) if big else 0) * (max(table_bytes(s) + work_bytes(s) for s in big) if big else 0)
    return fixed + slots + bigs


def test_plan_memory_matches_the_layout():
    boards = seeded_boards(512, 7, 18, 30)
    launches, ws = P.mid_schedule_info(boards, int(200e9), small_spots=14, workgroups=1024, threads=512, streams=2)
    est = expected_bytes(boards, 14, 1024, 2)
    assert est <= ws <= est + 4096 * len(boards) + (1 << 20)      # + descriptors
    big = [b for b in boards if board_spots(b) > 14]
    assert launches == 3 + sum(2 + board_spots(b) for b in big)


@pytest.mark.parametrize("small_spots", [10, 14, 18])
def test_split_by_size(small_spots):
    boards = seeded_boards(200, 11, 18, 34)
    launches, ws = P.mid_schedule_info(boards, int(200e9), small_spots=small_spots, workgroups=256, streams=1)
    big = [b for b in boards if board_spots(b) > small_spots]
    small = [b for b in boards if board_spots(b) <= small_spots]
    assert launches == (3 if small else 0) + sum(2 + board_spots(b) for b in big)


def test_tight_budget_shrinks_the_plan():
    """A budget just above the minimum (one workspace slot, one large-board slot) is
    accepted with fewer workgroups and streams; anything below it is refused."""
    boards = seeded_boards(200, 12, 20, 32)
    small = [b for b in boards if board_spots(b) <= 14]
    minimum = expected_bytes(boards, 14, 1, 1)
    launches, ws = P.mid_schedule_info(boards, minimum + (2 << 20), small_spots=14, workgroups=1024, streams=4)
    assert ws <= minimum + (2 << 20)
    with pytest.raises(P.PentagoError):
        P.mid_schedule_info(boards, minimum - work_bytes(max(board_spots(b) for b in small)) // 2, small_spots=14,
                            workgroups=1024, streams=4)


def test_schedule_rejects_incomplete_config():
    boards = seeded_boards(2, 3, 26, 26)
    with pytest.raises(P.PentagoError):
        P.mid_schedule_info(boards, 0)


def test_seeded_boards_follow_the_rules():
    boards = seeded_boards(300, 5, 18, 34)
    for b in boards:
        s0, s1 = int(b.side[0]), int(b.side[1])
        n0, n1 = bin(s0).count("1"), bin(s1).count("1")
        assert s0 & s1 == 0
        assert n0 - n1 in (0, 1)
        assert b.ply == 2 * (n0 + n1)
        assert not P.high_done(b)[0]
        assert 18 <= n0 + n1 <= 34
    assert seeded_boards(5, 9, 20, 30)[3].side[0] == seeded_boards(5, 9, 20, 30)[3].side[0]


def test_high_done_matches_oracle():
    """pe_high_done against the oracle's five-in-a-row test on boards of every density."""
    import numpy as np
    from oracle import oracle as O
    from tests.common import random_sides
    rng = np.random.default_rng(12)
    hits = 0
    for _ in range(400):
        stones = int(rng.integers(8, 36))
        s0, s1 = random_sides(rng, (stones + 1) // 2, stones // 2)
        done, value = P.high_done(P.high_board(s0, s1, 2 * stones))
        w0, w1 = O.won(s0), O.won(s1)
        assert done == (w0 or w1)
        if done:
            hits += 1
            turn = stones & 1
            assert value == (0 if w0 and w1 else (1 if (w0 if turn == 0 else w1) else -1))
    assert hits > 20
