"""Host-side logic of the benchmark (no GPU): how a section too large for the device with all
of its children is split into passes over subsets of its open dimensions."""
import bench

# the slice-30 section 30-(3,4)(4,4)(4,4)(4,3): child sections of dimensions 0 / 1 and 2
# (shared) / 3, in bytes of slab slots
CHILD = {0: "A", 1: "B", 2: "B", 3: "C"}
BYTES = {"A": 83.9e9, "B": 41.9e9, "C": 67.1e9}


def test_slice30_passes_on_one_device():
    # 0.92 of 307.5 GB free, less the 167.8 GB of output slots
    assert bench.pass_groups([0, 1, 2, 3], CHILD, BYTES, 0.92 * 307.5e9 - 167.8e9) == [[1, 2, 3], [0]]


def test_everything_fits_in_one_pass():
    assert bench.pass_groups([0, 1, 2, 3], CHILD, BYTES, 1e12) == [[0, 1, 2, 3]]


def test_tight_room_and_ordering():
    groups = bench.pass_groups([0, 1, 2, 3], CHILD, BYTES, 50e9)
    # a shared child is held once ([1, 2] fits), a dimension alone always makes a pass, and
    # the passes run from the most dimensions to the fewest
    assert groups == [[1, 2], [0], [3]]
    assert sorted(d for g in groups for d in g) == [0, 1, 2, 3]
    for g in groups:
        assert len(g) == 1 or sum(BYTES[c] for c in {CHILD[d] for d in g}) <= 50e9


def test_closed_dimensions_are_skipped():
    assert bench.pass_groups([1, 3], CHILD, BYTES, 1e12) == [[1, 3]]
