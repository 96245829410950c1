"""Supertensor files, CPU side: the production run's slice 0-5 files (data/slice-<n>.pentago,
xz-compressed, interleave filter) read by the host reader, unfiltered by the oracle's
restatement of pentago/data/filter.h, and counted per section with popcounts_over_stabilizers:
the totals equal the production counts files (data/counts-<n>.npy rows (sig, black wins,
white wins, total)).  Counting the file's (black, white) supers directly needs no turn
convention, so this pins the container format, the index and the filter on real data.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from pentago_amd import supertensor as ST
from pentago_amd import section_sig

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def words_to_int(w):
    return sum(int(w[i]) << (64 * i) for i in range(4))


@pytest.mark.parametrize("slice_", range(6))
def test_production_files_count_like_the_production_run(slice_):
    readers = ST.open_supertensors(os.path.join(GOLDEN, "slice-%d.pentago" % slice_))
    want = np.load(os.path.join(GOLDEN, "counts-%d.npy" % slice_), allow_pickle=False)
    want = {int(r[0]): tuple(int(x) for x in r[1:]) for r in want}
    assert sorted(section_sig(r.header.section) for r in readers) == sorted(want)
    for r in readers:
        h = r.header
        assert h.stones == slice_ and h.filter == 1 and h.block_size == 8
        rm = [O.rmin(*h.section[q])[0] for q in range(4)]
        got = np.zeros(3, dtype=np.int64)
        for i in range(int(np.prod(h.blocks))):
            block = np.unravel_index(i, h.blocks)
            raw = r.read_raw(block)
            nodes = O.uninterleave(raw)
            # the filter is a bijection on real data: filtering again gives the file's words
            assert np.array_equal(O.interleave(nodes), raw)
            bs = h.block_shape(block)
            for p in range(len(nodes)):
                idx = np.unravel_index(p, bs)
                board = sum(int(rm[q][8 * block[q] + idx[q]]) << (16 * q) for q in range(4))
                c = O.popcounts_over_stabilizers(board, words_to_int(nodes[p, :4]), words_to_int(nodes[p, 4:]))
                assert c is not None
                got += c
        assert tuple(got) == want[section_sig(h.section)], h.section


def test_filter_round_trip_and_bit_positions():
    rng = np.random.default_rng(5)
    x = rng.integers(0, 2**63, size=(17, 8), dtype=np.uint64) * 2 + rng.integers(0, 2, size=(17, 8), dtype=np.uint64)
    y = O.interleave(x)
    assert np.array_equal(O.uninterleave(y), x)
    # a single black bit (word k, bit i) lands at word 2k + i // 32, bit 2 (i % 32); white's
    # one position higher
    for k, i in [(0, 0), (1, 31), (2, 32), (3, 63)]:
        z = np.zeros((1, 8), dtype=np.uint64)
        z[0, k] = np.uint64(1) << np.uint64(i)
        z[0, 4 + k] = np.uint64(1) << np.uint64(i)
        out = O.interleave(z)
        expect = np.zeros((1, 8), dtype=np.uint64)
        expect[0, 2 * k + i // 32] = np.uint64(3) << np.uint64(2 * (i % 32))
        assert np.array_equal(out, expect), (k, i)


def test_header_round_trip_and_rejections():
    h = ST.Header(((2, 1), (0, 0), (1, 1), (0, 1)), 8, 1, valid=True, index=(72, 40, 123))
    g = ST.Header.unpack(h.pack())
    assert (g.section, g.shape, g.blocks, g.index, g.filter) == (h.section, h.shape, h.blocks, h.index, 1)
    bad = bytearray(h.pack())
    bad[0] ^= 1
    with pytest.raises(Exception):
        ST.Header.unpack(bytes(bad))
    h.valid = False
    with pytest.raises(Exception):
        ST.Header.unpack(h.pack())
