"""Batched midsolve plans (pe_mid_plan_*, pe_midsolve_batch mode 1) against the plain
per-board solve and the oracle: mixed sizes, every split between the small-board queue
kernel and the large-board launches, tight budgets, graph replays and the autotuner."""
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


@pytest.mark.parametrize("small_spots,threads,streams", [(4, 256, 1), (10, 1024, 2), (14, 512, -1), (18, 256, 3)])
def test_every_split_matches_plain(small_spots, threads, streams):
    boards = mixed_boards(7, 50, 22, 32)
    got = solve_queue(boards, small_spots=small_spots, threads=threads, streams=streams)
    assert got == [P.midsolve_internal(b) for b in boards]


def test_tight_budget_shrinks_plan():
    """A budget of one small-board workspace and one large-board slot (plus tables)."""
    boards = mixed_boards(5, 60, 20, 30)
    from tests.test_mid_schedule import expected_bytes
    budget = expected_bytes(boards, 14, 1, 1) + (2 << 20)
    plan = P.MidPlan(boards, workspace_budget=budget, small_spots=14, streams=4)
    info = plan.info()
    print("plan", info, flush=True)
    assert info["workspace_bytes"] <= budget and info["streams"] <= 1
    plan.run()
    got = P.results_to_lists(boards, plan.results())
    plan.close()
    assert got == P.midsolve_batch(boards, mode=0)


def test_graph_replays_match_eager():
    boards = mixed_boards(6, 40, 20, 32)
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
    assert best["small_spots"] in (12, 14, 16) and best["threads"] in (256, 512, 1024)
    plan = P.MidPlan(boards)
    info = plan.info()
    assert (info["small_spots"], info["threads"]) == (best["small_spots"], best["threads"])
    plan.run()
    assert P.results_to_lists(boards, plan.results()) == P.midsolve_batch(boards, mode=0)
    plan.close()
