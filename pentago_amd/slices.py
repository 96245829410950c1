"""One-GPU slice driver over slice stores (the slice loop of pentago/mpi/toplevel.cc:263-421
without MPI, compression or files).

Slices are computed from 35 down.  A slice that fits beside the next one is kept whole
(SliceStore); one that does not is computed in groups of sections by the native loop
pe_slice_grouped_counts, each group counted (and sampled) and freed before the next.  Two
slices below the last resident one, a group's child sections are regenerated from the
resident slice first (they are needed only while the group runs).
"""
import ctypes as C

import numpy as np

from . import Section, Slab, _check, lib, standardize


def child_sections(section):
    """Standardized child sections of a section: one stone of the player to move in each
    quadrant with room (pentago/base/section.cc:55-60, pentago/end/sections.cc:33-37)."""
    turn = sum(map(sum, section)) & 1
    out = set()
    for q in range(4):
        if sum(section[q]) < 9:
            c = [list(x) for x in section]
            c[q][turn] += 1
            out.add(standardize(tuple(map(tuple, c)))[0])
    return sorted(out)


def plan_groups(sections, budget, regenerate):
    """The groups grouped_counts computes (host only): list of section lists, in the order
    they are computed (with regeneration, sections sharing child sections go together)."""
    arr = (Section * len(sections))(*[Section.of(c) for c in sections])
    order = (C.c_int * len(sections))()
    ends = (C.c_int * len(sections))()
    n = C.c_int()
    _check(lib().pe_slice_plan_groups(len(sections), arr, int(budget), int(bool(regenerate)), order, ends,
                                      C.byref(n)), "pe_slice_plan_groups")
    out, begin = [], 0
    for g in range(n.value):
        out.append([sections[order[k]] for k in range(begin, ends[g])])
        begin = ends[g]
    return out


def arena_bytes(sections, budget, regenerate):
    """Device bytes grouped_counts needs for its largest group (host only)."""
    arr = (Section * len(sections))(*[Section.of(c) for c in sections])
    b = int(lib().pe_slice_arena_bytes(len(sections), arr, int(budget), int(bool(regenerate))))
    if b < 0:
        raise ValueError("invalid sections or budget")
    return b


class GroupView:
    """A computed group inside grouped_counts (valid during the visit only)."""

    def __init__(self, handle):
        self.handle = handle
        n = lib().pe_slice_sections(handle)
        self.slabs = []
        for k in range(n):
            slab = Slab()
            _check(lib().pe_slice_slab(handle, k, C.byref(slab)), "pe_slice_slab")
            self.slabs.append(slab)
        self.sections = [s.section.counts for s in self.slabs]
        self.index = {s: k for k, s in enumerate(self.sections)}

    def slab(self, k):
        return self.slabs[k]


VISIT = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)


def grouped_counts(sections, budget, children=None, grandchildren=None, visit=None, stream=None, arena=None):
    """Counts of every section of one slice, computed in groups within `budget` bytes by the
    native slice loop (pe_slice_grouped_counts).

    children: the whole store of the next slice, or None with grandchildren: the whole store
    of the slice after it, from which each group's child sections are regenerated.
    visit(group) sees each computed group (a GroupView) before it is freed.  arena: device
    memory (an object with .ptr and .nbytes, at least arena_bytes(...)) to compute in, reused
    across calls; None allocates one for the call.  Returns (sections in order, (n, 4)
    counts)."""
    if children is None and grandchildren is None:
        raise ValueError("need the next slice's store or the one after it")
    sections = [tuple(tuple(x) for x in s) for s in sections]
    arr = (Section * len(sections))(*[Section.of(c) for c in sections])
    counts = np.zeros((len(sections), 4), dtype=np.uint64)
    errors = []

    def cb(ctx, handle):
        try:
            visit(GroupView(handle))
        except Exception as e:   # re-raised after the native loop returns
            errors.append(e)

    fn = VISIT(cb) if visit else VISIT()
    ng = C.c_int()
    ch = children.handle if children else None
    gch = grandchildren.handle if grandchildren else None
    if arena is None:
        _check(lib().pe_slice_grouped_counts(len(sections), arr, int(budget), ch, gch, counts.ctypes.data,
                                             fn, None, C.byref(ng), stream), "pe_slice_grouped_counts")
    else:
        _check(lib().pe_slice_grouped_counts_in(len(sections), arr, int(budget), ch, gch, counts.ctypes.data,
                                                fn, None, C.byref(ng), arena.ptr, int(arena.nbytes), stream),
               "pe_slice_grouped_counts_in")
    if errors:
        raise errors[0]
    return sections, counts
