"""One-GPU slice driver over slice stores (the slice loop of pentago/mpi/toplevel.cc:263-421
without MPI, compression or files).

Slices are computed from 35 down.  A slice that fits beside the next one is kept whole
(SliceStore); one that does not is computed in groups of sections, each group counted (and
sampled) and freed before the next.  Two slices below the last resident one, a group's child
sections are regenerated from the resident slice first (they are needed only while the
group runs).
"""
import numpy as np

from . import SliceStore, slice_bytes, standardize


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


def groups(sections, budget, kids_of=None):
    """Split sections (in order) into groups whose slabs fit `budget` bytes, counting also
    the slabs of the union of kids_of(section) over the group when given."""
    size = {}

    def nbytes(sec):
        if sec not in size:
            size[sec] = slice_bytes([sec])
        return size[sec]

    out, cur, kids, used = [], [], set(), 0
    for s in sections:
        new = set(kids_of(s)) - kids if kids_of else set()
        need = nbytes(s) + sum(nbytes(k) for k in new)
        if cur and used + need > budget:
            out.append(cur)
            cur, kids, used = [], set(), 0
            new = set(kids_of(s)) if kids_of else set()
            need = nbytes(s) + sum(nbytes(k) for k in new)
        cur.append(s)
        kids |= new
        used += need
    if cur:
        out.append(cur)
    return out


def grouped_counts(sections, budget, children=None, grandchildren=None, visit=None, stream=None):
    """Counts of every section, computed in groups that fit `budget` bytes.

    children: the whole store of the next slice, or None with grandchildren: the whole store
    of the slice after it, from which each group's child sections are regenerated.
    visit(store) is called with each computed group before it is freed.  Returns
    (sections in order, (n, 4) counts)."""
    if children is None and grandchildren is None:
        raise ValueError("need the next slice's store or the one after it")
    regen = children is None
    done, counts = [], []
    for g in groups(sections, budget, child_sections if regen else None):
        kids = children
        if regen:
            kids = SliceStore(sorted({c for s in g for c in child_sections(s)}))
            kids.compute(grandchildren, stream)
        st = SliceStore(g)
        try:
            st.compute(kids, stream)
            counts.append(st.counts())
            done.extend(st.sections)
            if visit:
                visit(st)
        finally:
            st.close()
            if regen:
                kids.close()
    return done, np.concatenate(counts) if counts else np.zeros((0, 4), dtype=np.uint64)
