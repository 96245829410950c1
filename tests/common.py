"""Shared helpers for the test suite: seeded boards and golden fixtures."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIDE_MASK = 0x01ff01ff01ff01ff
CELLS = [16 * q + i for q in range(4) for i in range(9)]


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def random_sides(rng, black, white):
    """Two disjoint sides with the given stone counts, placed uniformly."""
    cells = rng.permutation(36)[:black + white]
    s0 = s1 = 0
    for k, c in enumerate(cells):
        bit = 1 << CELLS[int(c)]
        if k < black:
            s0 |= bit
        else:
            s1 |= bit
    return s0, s1


def random_side(rng, p=0.5):
    s = 0
    for c in CELLS:
        if rng.random() < p:
            s |= 1 << c
    return s
