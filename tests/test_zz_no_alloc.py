"""The allocation-free midsolve entry point (pe_midsolve_internal_noalloc) against the
allocating one (pe_midsolve_internal), called through the C ABI, bit for bit.

Calls in a row reuse the library's fixed device scratch array: a large board (15 empty
cells, the largest that fits), then smaller ones, then the large one again; a board that
does not fit is refused with PE_ERR_CAPACITY and leaves the next call working."""
import ctypes as C

import numpy as np
import pytest

import pentago_amd as P
from pentago_amd.workloads import seeded_board

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    P.init()


def solve_noalloc(board):
    res = (P.HalfSupers * 19)()
    st = P.lib().pe_midsolve_internal_noalloc(C.byref(board), res, None)
    return st, res


def as_lists(board, res):
    spots = 36 - ((board.ply + 1) >> 1)
    return [P._halves(res[i]) for i in range(spots + 1)], [P._halves(res[i]) for i in range(spots + 1, 19)]


def test_noalloc_matches_allocating_entry_point():
    rng = np.random.default_rng(717)
    boards = [seeded_board(rng, s) for s in (21, 27, 24, 33, 21)]
    for b in boards:
        stones = (b.ply + 1) >> 1
        print("noalloc", stones, "stones", hex(b.side[0]), hex(b.side[1]), flush=True)
        st, res = solve_noalloc(b)
        assert st == 0, P.lib().pe_last_error()
        got, tail = as_lists(b, res)
        assert got == P.midsolve_internal(b), stones
        assert all(x == (0, 0) for x in tail)


def test_noalloc_middle_boards():
    rng = np.random.default_rng(718)
    for stones in (22, 30):
        b = seeded_board(rng, stones)
        b.ply -= 1   # the same stones, rotation pending
        st, res = solve_noalloc(b)
        assert st == 0, P.lib().pe_last_error()
        assert as_lists(b, res)[0] == P.midsolve_internal(b)


def test_noalloc_refuses_boards_beyond_its_scratch():
    rng = np.random.default_rng(719)
    big = seeded_board(rng, 20)   # 16 empty cells: more than the fixed array holds
    st, _ = solve_noalloc(big)
    assert st == 4   # PE_ERR_CAPACITY
    assert b"scratch" in P.lib().pe_last_error()
    small = seeded_board(rng, 25)
    st, res = solve_noalloc(small)
    assert st == 0, P.lib().pe_last_error()
    assert as_lists(small, res)[0] == P.midsolve_internal(small)
