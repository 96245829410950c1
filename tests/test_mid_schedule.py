"""Host-side checks of the midsolve work-queue schedule (no GPU): the task order respects
every dependency, every step's tasks are present, boards sharing a workspace slot are
chained through their gates and the workspace stays within the budget.  The library
checks these invariants itself while building the order; these tests drive it over the
batched and serving board mixes and over tight budgets."""
import pytest

import pentago_amd as P
from pentago_amd.workloads import board_spots, seeded_boards


def board_bytes_upper(spots):
    # 32-byte workspace entries dominate; generous bound used only to size budgets
    from math import comb
    worst, prev = 0, 1
    for n in range(spots + 1):
        nxt = comb(spots, n + 1) * comb(n + 1, (n + 1) // 2) if n + 1 <= spots else 0
        worst = max(worst, prev + nxt)
        prev = nxt
    return 32 * worst


def test_schedule_batched_mix():
    boards = seeded_boards(512, 7, 18, 30)
    tasks, ws = P.mid_schedule_info(boards, int(200e9), 4096, 2048)
    assert tasks > len(boards) * 3
    assert 0 < ws <= int(200e9)


@pytest.mark.parametrize("budget_boards", [1, 2, 5])
def test_schedule_tight_budget_reuses_slots(budget_boards):
    """A budget of a few of the largest boards forces slot reuse (gate chains)."""
    boards = seeded_boards(200, 11, 22, 30)
    big = max(board_spots(b) for b in boards)
    budget = int(1.3 * budget_boards * board_bytes_upper(big)) + (64 << 20)
    tasks, ws = P.mid_schedule_info(boards, budget, 1024, 512)
    assert ws <= budget


def test_schedule_rejects_board_over_budget():
    boards = seeded_boards(4, 3, 18, 18)
    with pytest.raises(P.PentagoError):
        P.mid_schedule_info(boards, 1 << 20, 4096, 256)


def test_schedule_rejects_incomplete_config():
    boards = seeded_boards(2, 3, 26, 26)
    with pytest.raises(P.PentagoError):
        P.mid_schedule_info(boards, 0, 4096, 256)


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
