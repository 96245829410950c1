"""Midsolve at 18-21 stones (15-18 empty cells) on the GPU against the oracle, bit for bit.

These are the boards whose levels run as full-device launches (mid_level_kernel on the plain
path, mid_level1_kernel in batched plans, spread over the XCDs by s1p bands) or, up to the
small-board limit, one workgroup each in the persistent queue kernel.  Every result entry
(all 128 bits of win and notlose, root and children) is compared with the oracle's
midsolve_internal (pentago/mid/midengine.cc:93-125), which the CPU suite pins bit for bit
against brute-force search (pentago/mid/mid_test.cc:45-93 pattern).  The 163-value known
answer (pentago/mid/benchmark_test.cc:39-58) also runs through a batched plan.
"""
import numpy as np
import pytest

import pentago_amd as P
from oracle import oracle as O
from tests.common import golden, random_sides

pytestmark = pytest.mark.gpu

STONES = (18, 19, 20, 21)


def deep_boards():
    rng = np.random.default_rng(1821)
    out = []
    for stones in STONES:
        for middle in (0, 1):
            while True:
                s0, s1 = random_sides(rng, (stones + 1) // 2, stones // 2)
                if not O.won(s0) and not O.won(s1):
                    break
            out.append(P.high_board(s0, s1, 2 * stones - middle))
    return out


def to_oracle(h):
    return O.High((O.C.c_uint64 * 2)(h.side[0], h.side[1]), h.ply)


@pytest.fixture(scope="module")
def cases():
    P.init()
    boards = deep_boards()
    print("oracle midsolve of %d boards (18-21 stones) on 8 threads" % len(boards), flush=True)
    ref = O.midsolve_batch([to_oracle(h) for h in boards], 8)
    expect = [r[:37 - ((h.ply + 1) >> 1)] for h, r in zip(boards, ref)]
    return boards, expect


def stones_of(h):
    return (h.ply + 1) >> 1


def test_plain_matches_oracle(cases):
    boards, expect = cases
    for h, e in zip(boards, expect):
        print("plain", stones_of(h), "stones, ply", h.ply, flush=True)
        assert P.midsolve_internal(h) == e, (hex(h.side[0]), hex(h.side[1]), h.ply)


def run_plan(boards, **kw):
    plan = P.MidPlan(boards, **kw)
    info = plan.info()
    print("plan", info, flush=True)
    plan.run()
    got = P.results_to_lists(boards, plan.results())
    plan.close()
    return got, info


def test_large_board_plan_matches_oracle(cases):
    """small_spots 4: every board takes the level-by-level large-board path."""
    boards, expect = cases
    got, info = run_plan(boards, small_spots=4, streams=2)
    assert info["large_boards"] == len(boards) and info["small_boards"] == 0
    assert got == expect


def test_small_board_queue_matches_oracle(cases):
    """small_spots 16: the 20- and 21-stone boards (16 and 15 empty cells) are solved whole
    by single workgroups of the persistent queue kernel, the rest level by level."""
    boards, expect = cases
    got, info = run_plan(boards, small_spots=16, threads=1024, streams=1)
    assert info["small_boards"] == 4 and info["large_boards"] == 4
    assert got == expect


def test_graph_replays_match_oracle(cases):
    boards, expect = cases
    plan = P.MidPlan(boards, use_graph=True)
    print("plan", plan.info(), flush=True)
    for rep in range(2):
        plan.run()
        assert P.results_to_lists(boards, plan.results()) == expect, rep
    plan.close()


def test_known_answer_through_plan():
    """benchmark_test.cc:39-58: the 163 values of the 18-stone board from a batched plan's
    results (large-board path and queue kernel alike), traversed."""
    P.init()
    kat = golden("mid_kat.json")
    expect = {b | (m << 63): v for b, m, v in kat["values"]}
    h = P.high_from_board(kat["board"], 0)
    for small_spots in (4, 18):
        plan = P.MidPlan([h], small_spots=small_spots, threads=1024)
        print("plan", plan.info(), flush=True)
        plan.run()
        res = plan.results()
        plan.close()
        assert P.midsolve_traverse(h, res[0]) == expect, small_spots
