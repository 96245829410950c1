"""Midsolve on the GPU: known answers, game paths and parity with the oracle.

Mirrors pentago/mid/benchmark_test.cc (163-value known answer), web/client/src/unit.js
(game paths) and pentago/mid/mid_test.cc:45-93 (internal results).
"""
import numpy as np
import pytest

import pentago_amd as P
from oracle import oracle as O
from tests.common import golden, random_sides

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    P.init()


def test_midsolve_known_answer():
    kat = golden("mid_kat.json")
    values = P.midsolve(P.high_from_board(kat["board"], 0))
    assert len(values) == 1 + 18 + 8 * 18
    assert values == {b | (m << 63): v for b, m, v in kat["values"]}


def parse(name):
    middle = name.endswith("m")
    return int(name.rstrip("m")), middle


def test_game_paths():
    """web/client/src/unit.js:171-205: every value the game files record for boards the
    midsolves of the path's deep boards produce."""
    checked = 0
    for game in golden("games.json"):
        values = {}
        for k, v in game["values"].items():
            b, m = parse(k)
            values[b | (int(m) << 63)] = v
        for name in game["path"]:
            b, m = parse(name)
            h = P.high_from_board(b, m)
            if (h.ply + 1) >> 1 < 18:
                continue
            for raw, v in P.midsolve(h).items():
                if raw in values:
                    assert values[raw] == v, (name, raw)
                    checked += 1
    assert checked > 100


def random_board(rng, stones, middle):
    while True:
        s0, s1 = random_sides(rng, (stones + 1) // 2, stones // 2)
        if not O.won(s0) and not O.won(s1):
            return P.high_board(s0, s1, 2 * stones - middle)


def to_oracle(h):
    return O.High((O.C.c_uint64 * 2)(h.side[0], h.side[1]), h.ply)


@pytest.mark.parametrize("stones", list(range(22, 36)))
def test_internal_matches_oracle(stones):
    rng = np.random.default_rng(5554 + stones)
    for middle in (0, 1):
        h = random_board(rng, stones, middle)
        print("stones", stones, "middle", middle, hex(h.side[0]), hex(h.side[1]), flush=True)
        assert P.midsolve_internal(h) == O.midsolve_internal(to_oracle(h))[:37 - stones]


def test_batch_matches_single_boards():
    rng = np.random.default_rng(42)
    boards = [random_board(rng, int(s), int(rng.integers(2))) for s in rng.integers(20, 36, size=40)]
    batch = P.midsolve_batch(boards, mode=0)
    for h, got in zip(boards, batch):
        assert got == P.midsolve_internal(h)
