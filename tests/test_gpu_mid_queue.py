"""The persistent work-queue midsolve (pe_mid_plan_*, pe_midsolve_batch mode 1) against
the level-synchronous batch and the oracle: mixed sizes, tight workspace budgets that force
slot reuse, graph replays, the autotuner and the bounded-wait error path."""
import numpy as np
import pytest

import pentago_amd as P
from oracle import oracle as O
from pentago_amd.workloads import board_spots, seeded_boards
from tests.common import random_sides

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    P.init()


def mixed_boards(seed, n, lo, hi):
    """Seeded boards plus boards in the middle of a move (rotation pending)."""
    rng = np.random.default_rng(seed)
    out = []
    for k, stones in enumerate(rng.integers(lo, hi + 1, size=n)):
        stones = int(stones)
        while True:
            s0, s1 = random_sides(rng, (stones + 1) // 2, stones // 2)
            b = P.high_board(s0, s1, 2 * stones - (k & 1))
            if not P.high_done(b)[0]:
                out.append(b)
                break
    return out


def to_oracle(h):
    return O.High((O.C.c_uint64 * 2)(h.side[0], h.side[1]), h.ply)


def solve_queue(boards, **kw):
    plan = P.MidPlan(boards, **kw)
    print("plan", plan.info(), flush=True)
    plan.run()
    res = plan.results()
    plan.close()
    return P.results_to_lists(boards, res)


def test_queue_small_boards_match_oracle():
    boards = mixed_boards(3, 24, 28, 34)
    got = solve_queue(boards)
    for b, r in zip(boards, got):
        assert r == O.midsolve_internal(to_oracle(b))[:board_spots(b) + 1]


def test_queue_matches_level_batch_mixed_sizes():
    boards = mixed_boards(4, 80, 21, 33) + seeded_boards(40, 8, 18, 30)
    assert solve_queue(boards) == P.midsolve_batch(boards, mode=0)


def test_batch_mode1_matches_mode0():
    boards = seeded_boards(64, 9, 22, 32)
    assert P.midsolve_batch(boards, mode=1) == P.midsolve_batch(boards, mode=0)


@pytest.mark.parametrize("slots", [1, 3])
def test_queue_tight_budget_reuses_workspace(slots):
    """A budget of one or three of the largest boards: boards wait for slots (gate chains)."""
    boards = mixed_boards(5, 60, 22, 30)
    big = max(board_spots(b) for b in boards)
    from tests.test_mid_schedule import board_bytes_upper
    budget = int(1.3 * slots * board_bytes_upper(big)) + (64 << 20)
    got = solve_queue(boards, workspace_budget=budget, tile=512, workgroups=128)
    assert got == P.midsolve_batch(boards, mode=0)


def test_graph_replays_match_eager():
    boards = mixed_boards(6, 40, 23, 32)
    expect = P.midsolve_batch(boards, mode=0)
    plan = P.MidPlan(boards, use_graph=True)
    for rep in range(3):
        plan.run()
        assert P.results_to_lists(boards, plan.results()) == expect, rep
        # eager work between replays
        assert P.midsolve_batch(boards[:4], mode=0) == expect[:4]
    plan.close()


def test_autotune_picks_a_candidate_and_becomes_default():
    boards = seeded_boards(48, 10, 22, 30)
    best = P.mid_autotune(boards)
    assert best["tile"] in (1024, 4096, 16384) and best["workgroups"] > 0
    plan = P.MidPlan(boards)
    info = plan.info()
    assert info["tile"] == best["tile"] and info["workgroups"] == best["workgroups"]
    plan.run()
    assert P.results_to_lists(boards, plan.results()) == P.midsolve_batch(boards, mode=0)
    plan.close()


def test_wait_limit_reports_error_and_queue_recovers():
    """A 1 us bound on every dependency wait: the kernel reports the failed wait, skips the
    remaining work and still exits; the next plan runs normally."""
    boards = seeded_boards(32, 12, 20, 26)
    plan = P.MidPlan(boards, wait_limit_us=1, workgroups=512, tile=256)
    plan.run()
    with pytest.raises(P.PentagoError):
        plan.results()
    plan.close()
    assert solve_queue(boards[:8]) == P.midsolve_batch(boards[:8], mode=0)
