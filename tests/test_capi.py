"""The C-ABI library loads, exports every entry point the header declares, and its host
side (tables, sections, line index state) agrees with the oracle.  No GPU calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pentago_amd as P
from oracle import oracle as O
from tests.common import golden
from tests.seeded import seeded_nodes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pentago_amd.h")).read()
    return sorted(set(re.findall(r"^(?:pe_status|int64_t|int|const char\*)\s+(pe_[a-z_0-9]+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol():
    lib = C.CDLL(P.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name


def test_product_tables_match_reference_digests():
    for name, digest in golden("table_hashes.json").items():
        if name == "superstandardize_table":
            continue   # superstandardize is not on this path
        assert P.table_digest(name) == digest, name


SECTIONS = [((3, 4), (4, 4), (4, 4), (4, 3)), ((4, 4), (4, 3), (3, 4), (4, 4)), ((3, 3), (3, 3), (3, 3), (3, 3)),
            ((4, 2), (2, 4), (3, 3), (3, 3)), ((5, 4), (4, 5), (4, 4), (5, 4)), ((4, 5), (5, 4), (4, 4), (4, 4)),
            ((2, 3), (4, 4), (3, 2), (4, 4)), ((4, 4), (4, 4), (4, 4), (5, 4)), ((3, 2), (2, 3), (3, 3), (4, 4))]


def test_sections_match_oracle():
    for c in SECTIONS:
        assert P.section_shape(c) == O.section_shape(c)
        assert P.standardize(c) == O.standardize(c)


def test_line_index_state_matches_oracle():
    for c in SECTIONS:
        shape = P.section_shape(c)
        for d in range(4):
            if sum(c[d]) >= 9:
                continue
            for base in [(0, 0, 0), tuple((shape[a] - 1) // 8 for a in range(4) if a != d)]:
                pi = P.line_describe(P.make_line(c, d, base))
                oi = O.line_describe(O.make_line(c, d, base))
                for f in ("input_shape", "output_shape", "permutation", "first_child_block"):
                    assert list(getattr(pi, f)) == list(getattr(oi, f)), (c, d, f)
                for f in ("block_stride", "input_blocks", "length", "input_nodes", "output_nodes", "section_transform"):
                    assert getattr(pi, f) == getattr(oi, f), (c, d, f)
                assert pi.child_dimension == oi.child_dim
                assert pi.standard_child.counts == O.section_counts(oi.standard_child)


def test_invalid_lines_are_rejected():
    line = P.make_line(((3, 4), (4, 4), (4, 4), (4, 3)), 0, (0, 0, 0))
    line.length = 7
    with pytest.raises(P.PentagoError):
        P.line_describe(line)
    line = P.make_line(((3, 4), (4, 4), (4, 4), (4, 3)), 0, (30, 0, 0))
    with pytest.raises(P.PentagoError):
        P.line_describe(line)
    with pytest.raises(P.PentagoError):
        P.section_shape(((9, 1), (0, 0), (0, 0), (0, 0)))


def test_midsolve_workspace_sizes():
    g = golden("bottleneck.json")
    for sl, size in zip(g["slices"], g["sizes"]):
        assert P.lib().pe_midsolve_workspace_size(sl) == size


def test_seeded_generator_statistics():
    x = seeded_nodes(4096, 7).reshape(-1, 2, 4)
    win, known = x[:, 0], x[:, 1]
    assert not np.any(win & ~known)
    bits = lambda a: np.unpackbits(a.view(np.uint8)).mean()
    assert abs(bits(known) - 58982 / 65536) < 0.003
    assert abs(bits(win) - 58982 / 65536 / 2) < 0.003


def test_count_many_arguments_checked_on_the_host():
    # descriptor workspace grows with the slab count; missing buffers are rejected before any
    # device call
    assert 0 < P.count_workspace_bytes(1) <= P.count_workspace_bytes(2) < P.count_workspace_bytes(100)
    slab = P.Slab.of(((3, 4), (4, 4), (4, 4), (4, 3)), [0] * 4, [1] * 4, 0)
    with pytest.raises(P.PentagoError):
        P.count_blocks_many([slab, slab], 1 << 20, None)
    with pytest.raises(P.PentagoError):
        P.count_blocks_many([slab], None, 1 << 20)


def test_slab_dims_arguments_checked_on_the_host():
    # the dimension mask, the accumulate mode (0-4) and the stone count are rejected before
    # any device call; no slab memory is touched
    slab = P.Slab.of(((3, 4), (4, 4), (4, 4), (4, 3)), [0] * 4, [1] * 4, 1 << 20)
    for dims, acc in ((0, 0), (0x10, 0), (1, -1), (1, 5)):
        with pytest.raises(P.PentagoError):
            P.Plan.slab_dims(slab, [], dims, accumulate=acc)
    full = P.Slab.of(((4, 4), (4, 4), (4, 4), (4, 3)), [0] * 4, [1] * 4, 1 << 20)
    with pytest.raises(P.PentagoError):
        P.Plan.slab_dims(full, [], 1, accumulate=3)


def test_gather_probe_library_checks_its_arguments():
    # the measurement probe is its own library (the engine does not export it); bad shapes
    # are refused before any launch: odd or tiny buffers, windows outside the buffer, no nodes
    lib = C.CDLL(os.path.join(os.path.dirname(P.LIB_PATH), "libpentago_probe.so"))
    assert not hasattr(C.CDLL(P.LIB_PATH), "pe_probe_gather")
    f = lib.pe_probe_gather
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.c_uint64, C.c_void_p]
    fake = 1 << 20   # never dereferenced: every call below is refused first
    for src_nodes, n, window in ((3, 1, 0), (0, 1, 0), (8, 0, 0), (8, 1, 10), (8, 1, 3), (8, 1 << 32, 0)):
        assert f(fake, src_nodes, fake, n, window, 1, None) != 0, (src_nodes, n, window)
    assert f(None, 8, fake, 1, 0, 1, None) != 0
