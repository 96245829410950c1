"""Seeded synthetic midsolve boards for the batched and serving workloads.

A board has `stones` stones, black = white or white + 1 (black moves when the counts are
equal), placed uniformly on the 36 cells; boards that already hold a five in a row are
redrawn.  Deterministic in the seed (numpy PCG64).
"""
import numpy as np

from . import high_board, high_done

CELLS = [16 * q + i for q in range(4) for i in range(9)]


def seeded_board(rng, stones):
    while True:
        cells = rng.permutation(36)[:stones]
        black = (stones + 1) // 2
        s0 = s1 = 0
        for k, c in enumerate(cells):
            if k < black:
                s0 |= 1 << CELLS[int(c)]
            else:
                s1 |= 1 << CELLS[int(c)]
        b = high_board(s0, s1, 2 * stones)
        if not high_done(b)[0]:
            return b


def seeded_boards(n, seed, min_stones, max_stones):
    """n boards with stone counts uniform in [min_stones, max_stones]."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(min_stones, max_stones + 1, size=n)
    return [seeded_board(rng, int(s)) for s in counts]


def board_spots(board):
    return 36 - ((board.ply + 1) >> 1)
