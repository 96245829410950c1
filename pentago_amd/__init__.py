"""pentago_amd: MI355X engine for pentago's retrograde superscore path.

Python view of the C ABI in include/pentago_amd.h (libpentago_amd.so, HIP for gfx950).
Names follow the reference (section_t, line_t, schedule_compute_line, midsolve,
midsolve_internal, high_board_t).  There is no CPU fallback: every compute call runs the
HIP kernels, and loading fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PENTAGO_AMD_LIB selects another build of the library (same C-ABI), e.g. for A/B timing.
LIB_PATH = os.environ.get("PENTAGO_AMD_LIB") or os.path.join(HERE, "libpentago_amd.so")


class PentagoError(RuntimeError):
    pass


class Section(C.Structure):
    """section_t: per-quadrant (black, white) counts (pentago/base/section.h:32-33)."""
    _fields_ = [("c%d%d" % (q, k), C.c_uint8) for q in range(4) for k in range(2)]

    @classmethod
    def of(cls, counts):
        s = cls()
        for q in range(4):
            setattr(s, "c%d0" % q, counts[q][0])
            setattr(s, "c%d1" % q, counts[q][1])
        return s

    @property
    def counts(self):
        return tuple((getattr(self, "c%d0" % q), getattr(self, "c%d1" % q)) for q in range(4))


class Line(C.Structure):
    """line_t: one block line (pentago/end/line.h:15-35)."""
    _fields_ = [("section", Section), ("dimension", C.c_uint8), ("length", C.c_uint8),
                ("block_base", C.c_uint8 * 3)]


class LineInfo(C.Structure):
    _fields_ = [("input_shape", C.c_int32 * 4), ("output_shape", C.c_int32 * 4),
                ("block_stride", C.c_int32), ("input_blocks", C.c_int32), ("length", C.c_int32),
                ("dimension", C.c_int32), ("child_dimension", C.c_int32),
                ("input_nodes", C.c_int64), ("output_nodes", C.c_int64),
                ("standard_child", Section), ("section_transform", C.c_int32),
                ("permutation", C.c_int32 * 4), ("first_child_block", C.c_int32 * 4)]


class LineJob(C.Structure):
    """pe_line_job: a line plus where its child blocks and output blocks live."""
    _fields_ = [("line", Line), ("input", C.c_void_p), ("output", C.c_void_p),
                ("input_block_offsets", C.POINTER(C.c_int64)), ("output_block_offsets", C.POINTER(C.c_int64)),
                ("i_begin", C.c_int32), ("i_end", C.c_int32), ("accumulate", C.c_int32), ("pad", C.c_int32)]


class Slab(C.Structure):
    """pe_slab: a dense box of blocks [lo, hi) of one section in device memory."""
    _fields_ = [("section", Section), ("lo", C.c_int32 * 4), ("hi", C.c_int32 * 4), ("data", C.c_void_p)]

    @classmethod
    def of(cls, counts, lo, hi, ptr):
        s = cls()
        s.section = Section.of(counts)
        for a in range(4):
            s.lo[a], s.hi[a] = lo[a], hi[a]
        s.data = ptr
        return s

    @property
    def blocks(self):
        n = 1
        for a in range(4):
            n *= self.hi[a] - self.lo[a]
        return n


class HighBoard(C.Structure):
    """high_board_s (pentago/high/board_c.h:9-12)."""
    _fields_ = [("side", C.c_uint64 * 2), ("ply", C.c_uint32)]


class HalfSupers(C.Structure):
    """halfsupers_t (pentago/mid/halfsuper_c.h:16-18)."""
    _fields_ = [("win", C.c_uint64 * 2), ("notlose", C.c_uint64 * 2)]


class MidConfig(C.Structure):
    """pe_mid_config: zero fields take the device defaults."""
    _fields_ = [("workspace_budget", C.c_int64), ("small_spots", C.c_int32), ("workgroups", C.c_int32),
                ("threads", C.c_int32), ("streams", C.c_int32), ("use_graph", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {"workspace_budget": int(self.workspace_budget), "small_spots": int(self.small_spots),
                "workgroups": int(self.workgroups), "threads": int(self.threads), "streams": int(self.streams),
                "use_graph": self.use_graph > 0}


class MidPlanStats(C.Structure):
    _fields_ = [("small_boards", C.c_int64), ("large_boards", C.c_int64), ("launches", C.c_int64),
                ("workspace_bytes", C.c_int64), ("workgroups", C.c_int32), ("streams", C.c_int32)]


class MidServerConfig(C.Structure):
    _fields_ = [("workspace_budget", C.c_int64), ("streams", C.c_int32), ("max_batch", C.c_int32),
                ("small_spots", C.c_int32), ("workgroups", C.c_int32)]


class MidServerInfo(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("submitted", "completed", "batches", "peak_workspace", "pending",
                                         "in_flight")]


class MidValue(C.Structure):
    _fields_ = [("raw", C.c_uint64), ("value", C.c_int32), ("pad", C.c_int32)]


class EndTuning(C.Structure):
    _fields_ = [("queue_wgs_per_cu", C.c_int32), ("candidates", C.c_int32), ("forms", C.c_int32 * 8),
                ("ms", C.c_float * 8)]


WAKEUP = C.CFUNCTYPE(None, C.c_void_p)

_lib = None


def lib():
    """Load libpentago_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PentagoError("libpentago_amd.so is not built: run __graft_entry__.build() "
                               "(or make -C pentago_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        P, V = C.POINTER, C.c_void_p
        sig = {
            "pe_init": (C.c_int, []),
            "pe_last_error": (C.c_char_p, []),
            "pe_table_digest": (C.c_int, [C.c_char_p, C.c_char_p]),
            "pe_section_shape": (C.c_int, [P(Section), P(C.c_int32)]),
            "pe_section_standardize": (C.c_int, [P(Section), P(Section), P(C.c_int32)]),
            "pe_line_describe": (C.c_int, [P(Line), P(LineInfo)]),
            "pe_end_compute_line": (C.c_int, [P(Line), V, V, V]),
            "pe_end_schedule_line": (C.c_int, [P(Line), V, V, V, WAKEUP, V]),
            "pe_end_compute_lines": (C.c_int, [C.c_int, P(Line), P(V), P(V), V]),
            "pe_end_compute_line_host": (C.c_int, [P(Line), V, V]),
            "pe_end_plan_create": (C.c_int, [C.c_int, P(LineJob), P(V)]),
            "pe_end_plan_run": (C.c_int, [V, V]),
            "pe_end_slab_plan_create": (C.c_int, [P(Slab), C.c_int, P(Slab), C.c_int, P(V)]),
            "pe_end_slabs_plan_create": (C.c_int, [C.c_int, P(Slab), C.c_int, P(Slab), C.c_int, P(V)]),
            "pe_end_slabs_plan_create_dims": (C.c_int, [C.c_int, P(Slab), C.c_int, P(Slab), C.c_uint32, C.c_int,
                                                        P(V)]),
            "pe_end_plan_nodes": (C.c_int64, [V]),
            "pe_end_plan_destroy": (C.c_int, [V]),
            "pe_end_plan_set_launch": (C.c_int, [V, C.c_int]),
            "pe_end_set_default_launch": (C.c_int, [C.c_int]),
            "pe_end_plan_autotune": (C.c_int, [V, V, P(EndTuning)]),
            "pe_fill_seeded": (C.c_int, [V, C.c_int64, C.c_uint64, C.c_int64, V]),
            "pe_end_count_blocks": (C.c_int, [P(Slab), V, V]),
            "pe_end_count_slabs": (C.c_int, [C.c_int, P(Slab), V]),
            "pe_end_count_blocks_many": (C.c_int, [C.c_int, P(Slab), V, V, V]),
            "pe_end_count_workspace_bytes": (C.c_int64, [C.c_int]),
            "pe_end_slab_samples": (C.c_int, [P(Slab), C.c_int, P(C.c_uint64), V]),
            "pe_supertensor_decode": (C.c_int, [C.c_int, C.c_int, V, C.c_int64, V, V]),
            "pe_supertensor_encode": (C.c_int, [C.c_int, C.c_int, V, C.c_int64, V, V]),
            "pe_descendent_sections": (C.c_int, [P(Section), C.c_int, C.c_int, P(Section), P(C.c_int)]),
            "pe_slice_create": (C.c_int, [C.c_int, P(Section), P(V)]),
            "pe_slice_create_in": (C.c_int, [C.c_int, P(Section), V, C.c_int64, P(V)]),
            "pe_slice_bytes": (C.c_int64, [C.c_int, P(Section)]),
            "pe_slice_compute": (C.c_int, [V, V, V]),
            "pe_slice_slab": (C.c_int, [V, C.c_int, P(Slab)]),
            "pe_slice_sections": (C.c_int, [V]),
            "pe_slice_counts": (C.c_int, [V, V]),
            "pe_slice_destroy": (C.c_int, [V]),
            "pe_slice_plan_groups": (C.c_int, [C.c_int, P(Section), C.c_int64, C.c_int, P(C.c_int), P(C.c_int),
                                               P(C.c_int)]),
            "pe_slice_grouped_counts": (C.c_int, [C.c_int, P(Section), C.c_int64, V, V, V,
                                                  C.CFUNCTYPE(None, V, V), V, P(C.c_int), V]),
            "pe_slice_grouped_counts_in": (C.c_int, [C.c_int, P(Section), C.c_int64, V, V, V,
                                                     C.CFUNCTYPE(None, V, V), V, P(C.c_int), V, C.c_int64, V]),
            "pe_slice_arena_bytes": (C.c_int64, [C.c_int, P(Section), C.c_int64, C.c_int]),
            "pe_midsolve_workspace_size": (C.c_int64, [C.c_int]),
            "pe_midsolve_internal": (C.c_int, [P(HighBoard), P(HalfSupers)]),
            "pe_midsolve": (C.c_int, [P(HighBoard), P(MidValue), P(C.c_int)]),
            "pe_midsolve_batch": (C.c_int, [C.c_int, P(HighBoard), V, C.c_int, V]),
            "pe_midsolve_traverse": (C.c_int, [P(HighBoard), P(HalfSupers), P(MidValue), P(C.c_int)]),
            "pe_midsolve_internal_noalloc": (C.c_int, [P(HighBoard), P(HalfSupers), V]),
            "pe_midsolve_noalloc_scratch_bytes": (C.c_int64, []),
            "pe_high_done": (C.c_int, [P(HighBoard), P(C.c_int), P(C.c_int)]),
            "pe_mid_server_create": (C.c_int, [P(MidServerConfig), P(V)]),
            "pe_mid_server_submit": (C.c_int, [V, P(HighBoard), P(C.c_int64)]),
            "pe_mid_server_poll": (C.c_int, [V, C.c_int64, V, P(C.c_int), P(C.c_double)]),
            "pe_mid_server_wait": (C.c_int, [V, C.c_int64, V, P(C.c_double)]),
            "pe_mid_server_stats": (C.c_int, [V, P(MidServerInfo)]),
            "pe_mid_server_destroy": (C.c_int, [V]),
            "pe_mid_plan_create": (C.c_int, [C.c_int, P(HighBoard), P(MidConfig), P(V)]),
            "pe_mid_plan_run": (C.c_int, [V, V]),
            "pe_mid_plan_results": (C.c_int, [V, V, V]),
            "pe_mid_plan_info": (C.c_int, [V, P(MidPlanStats), P(MidConfig)]),
            "pe_mid_plan_destroy": (C.c_int, [V]),
            "pe_mid_schedule_info": (C.c_int, [C.c_int, P(HighBoard), P(MidConfig), P(C.c_int64), P(C.c_int64)]),
            "pe_mid_autotune": (C.c_int, [C.c_int, P(HighBoard), C.c_int64, P(MidConfig)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status, what):
    if status != 0:
        raise PentagoError("%s: %s (status %d)" % (what, lib().pe_last_error().decode(), status))


def init():
    """Build and self-check the tables and upload them to the current device."""
    _check(lib().pe_init(), "pe_init")


def table_digest(name):
    out = C.create_string_buffer(41)
    _check(lib().pe_table_digest(name.encode(), out), "pe_table_digest")
    return out.value.decode()


# ---------------- end engine ----------------

def section_shape(counts):
    """section_t::shape(): rotation-minimal quadrant counts per dimension."""
    shape = (C.c_int32 * 4)()
    _check(lib().pe_section_shape(C.byref(Section.of(counts)), shape), "pe_section_shape")
    return tuple(shape)


def standardize(counts):
    """section_t::standardize<8>(): (standard counts, symmetry)."""
    out = Section()
    g = C.c_int32()
    _check(lib().pe_section_standardize(C.byref(Section.of(counts)), C.byref(out), C.byref(g)),
           "pe_section_standardize")
    return out.counts, g.value


def make_line(counts, dimension, block_base):
    """line_t for section `counts`, a dimension and the other three block indices."""
    line = Line()
    line.section = Section.of(counts)
    line.dimension = dimension
    line.length = (section_shape(counts)[dimension] + 7) // 8
    for i in range(3):
        line.block_base[i] = block_base[i]
    return line


def line_describe(line):
    info = LineInfo()
    _check(lib().pe_line_describe(C.byref(line), C.byref(info)), "pe_line_describe")
    return info


def compute_line_host(line, inp):
    """Synchronous: input (uint64, 8 words per node, reference layout) -> output array."""
    info = line_describe(line)
    out = np.zeros(8 * info.output_nodes, dtype=np.uint64)
    ip = inp.ctypes.data if inp is not None and inp.size else None
    _check(lib().pe_end_compute_line_host(C.byref(line), ip, out.ctypes.data), "pe_end_compute_line_host")
    return out


def compute_lines(lines, d_inputs, d_outputs, stream=None):
    """Asynchronous batch launch on device pointers."""
    n = len(lines)
    arr = (Line * n)(*lines)
    ins = (C.c_void_p * n)(*d_inputs)
    outs = (C.c_void_p * n)(*d_outputs)
    _check(lib().pe_end_compute_lines(n, arr, ins, outs, stream), "pe_end_compute_lines")


def schedule_compute_line(line, d_input, d_output, wakeup, stream=None):
    """schedule_compute_line (pentago/end/compute.h:123): asynchronous, calls wakeup() once
    the output is complete.  Keep the returned callback object alive until then."""
    cb = WAKEUP(lambda ctx: wakeup())
    _check(lib().pe_end_schedule_line(C.byref(line), d_input, d_output, stream, cb, None), "pe_end_schedule_line")
    return cb


def set_default_launch(queue_wgs_per_cu):
    """pe_end_set_default_launch: launch form of the current device's later end plans."""
    _check(lib().pe_end_set_default_launch(int(queue_wgs_per_cu)), "pe_end_set_default_launch")


class Plan:
    """A device-resident set of line jobs (pe_end_plan), launched with run()."""

    @classmethod
    def slab(cls, output, children, mode=0):
        """Plan computing every block of the output slab (or list of output slabs) from
        child slabs (mode 0 fused, mode 1 four line passes)."""
        self = cls.__new__(cls)
        outs = list(output) if isinstance(output, (list, tuple)) else [output]
        n = len(children)
        arr = (Slab * max(n, 1))(*children)
        oarr = (Slab * len(outs))(*outs)
        h = C.c_void_p()
        _check(lib().pe_end_slabs_plan_create(len(outs), oarr, n, arr, mode, C.byref(h)), "pe_end_slabs_plan_create")
        self.handle = h
        self.nodes = lib().pe_end_plan_nodes(h)
        return self

    @classmethod
    def slab_dims(cls, output, children, dims, accumulate=0):
        """Fused plan over the output slab(s) computing only the open dimensions in the bit
        mask `dims` (pe_end_slabs_plan_create_dims).  accumulate: 0 store, 1 OR into the
        output, 2 store a partial (children only), 4 add to the partial, 3 finish from it."""
        self = cls.__new__(cls)
        outs = list(output) if isinstance(output, (list, tuple)) else [output]
        n = len(children)
        arr = (Slab * max(n, 1))(*children)
        oarr = (Slab * len(outs))(*outs)
        h = C.c_void_p()
        _check(lib().pe_end_slabs_plan_create_dims(len(outs), oarr, n, arr, int(dims), int(accumulate),
                                                   C.byref(h)), "pe_end_slabs_plan_create_dims")
        self.handle = h
        self.nodes = lib().pe_end_plan_nodes(h)
        return self

    def __init__(self, jobs):
        """jobs: list of dicts with keys line, input, output and optionally in_offsets,
        out_offsets (sequences of node offsets), i_range (begin, end), accumulate."""
        n = len(jobs)
        arr = (LineJob * n)()
        keep = []
        for k, j in enumerate(jobs):
            J = arr[k]
            J.line = j["line"]
            J.input = j.get("input")
            J.output = j["output"]
            for key, field in (("in_offsets", "input_block_offsets"), ("out_offsets", "output_block_offsets")):
                if j.get(key) is not None:
                    a = np.ascontiguousarray(j[key], dtype=np.int64)
                    keep.append(a)
                    setattr(J, field, a.ctypes.data_as(C.POINTER(C.c_int64)))
            if j.get("i_range"):
                J.i_begin, J.i_end = j["i_range"]
            J.accumulate = int(bool(j.get("accumulate", 0)))
        h = C.c_void_p()
        _check(lib().pe_end_plan_create(n, arr, C.byref(h)), "pe_end_plan_create")
        self.handle = h
        self.nodes = lib().pe_end_plan_nodes(h)

    def run(self, stream=None):
        _check(lib().pe_end_plan_run(self.handle, stream), "pe_end_plan_run")

    def set_launch(self, queue_wgs_per_cu):
        """0: static grid; n > 0: persistent grid of n workgroups per CU on device queues."""
        _check(lib().pe_end_plan_set_launch(self.handle, int(queue_wgs_per_cu)), "pe_end_plan_set_launch")

    def autotune(self, stream=None):
        """Time the launch forms on this plan and keep the fastest: {form: ms}, choice."""
        t = EndTuning()
        _check(lib().pe_end_plan_autotune(self.handle, stream, C.byref(t)), "pe_end_plan_autotune")
        return {int(t.forms[i]): float(t.ms[i]) for i in range(t.candidates)}, int(t.queue_wgs_per_cu)

    def close(self):
        if self.handle:
            lib().pe_end_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fill_seeded(d_nodes, nodes, seed, first_node=0, stream=None):
    _check(lib().pe_fill_seeded(d_nodes, nodes, seed, first_node, stream), "pe_fill_seeded")


def count_slabs(slabs):
    """count_block_wins summed per slab: (n, 4) uint64 array of (wins, wins-or-ties, total,
    nodes not invariant under their stabilizer) for the player to move."""
    slabs = list(slabs)
    out = np.zeros((len(slabs), 4), dtype=np.uint64)
    arr = (Slab * max(len(slabs), 1))(*slabs)
    _check(lib().pe_end_count_slabs(len(slabs), arr, out.ctypes.data), "pe_end_count_slabs")
    return out


def count_blocks(slab, d_counts, stream=None):
    """Per-block counts of one slab into device memory (4 uint64 per block), asynchronous."""
    _check(lib().pe_end_count_blocks(C.byref(slab), d_counts, stream), "pe_end_count_blocks")


def count_blocks_many(slabs, d_counts, d_workspace, stream=None):
    """Per-block counts of many slabs in one launch into device memory (4 uint64 per block,
    slab after slab); d_workspace holds count_workspace_bytes(len(slabs)) bytes."""
    slabs = list(slabs)
    arr = (Slab * max(len(slabs), 1))(*slabs)
    _check(lib().pe_end_count_blocks_many(len(slabs), arr, d_counts, d_workspace, stream),
           "pe_end_count_blocks_many")


def count_workspace_bytes(n):
    return int(lib().pe_end_count_workspace_bytes(int(n)))


def slab_samples(slab, boards):
    """Nodes of rotation-standardized boards inside a slab: (n, 8) uint64 (win words, then
    win-or-tie words, for the player to move)."""
    b = np.ascontiguousarray(boards, dtype=np.uint64)
    out = np.zeros((len(b), 8), dtype=np.uint64)
    _check(lib().pe_end_slab_samples(C.byref(slab), len(b), b.ctypes.data_as(C.POINTER(C.c_uint64)),
                                      out.ctypes.data), "pe_end_slab_samples")
    return out


def descendent_sections(root, slice_):
    """Standardized sections of a slice that root depends on, sorted by sig
    (pentago/end/sections.cc:18-49)."""
    r = Section.of(root)
    n = C.c_int()
    _check(lib().pe_descendent_sections(C.byref(r), slice_, 0, None, C.byref(n)), "pe_descendent_sections")
    arr = (Section * max(n.value, 1))()
    _check(lib().pe_descendent_sections(C.byref(r), slice_, n.value, arr, C.byref(n)), "pe_descendent_sections")
    return [arr[k].counts for k in range(n.value)]


def section_sig(counts):
    """section_t::sig() (pentago/base/section.h:40-48)."""
    return sum(int(counts[q][0]) << (16 * q) | int(counts[q][1]) << (16 * q + 8) for q in range(4))


def counts_rows(sections, counts, slice_):
    """Rows (sig, black wins, white wins, total) as the reference writes counts-<slice>.npy
    (pentago/mpi/io.cc:573-581): wins are the mover's, losses = total - wins-or-ties, and
    odd slices swap them so wins are black's."""
    rows = np.zeros((len(sections), 4), dtype=np.uint64)
    for k, sec in enumerate(sections):
        wins, losses = int(counts[k][0]), int(counts[k][2]) - int(counts[k][1])
        if slice_ & 1:
            wins, losses = losses, wins
        rows[k] = (section_sig(sec), wins, losses, int(counts[k][2]))
    return rows


def slice_bytes(sections):
    arr = (Section * max(len(sections), 1))(*[Section.of(c) for c in sections])
    return lib().pe_slice_bytes(len(sections), arr)


class SliceStore:
    """Whole sections of one slice resident on the device (pe_slice): compute() fills them
    from the next slice's store, counts() gives count_block_wins per section.  With `arena`
    (an object with .ptr and .nbytes that outlives the store) the slabs are carved from it
    instead of allocated."""

    def __init__(self, sections, arena=None):
        self.sections = [tuple(tuple(c) for c in s) for s in sections]
        arr = (Section * len(self.sections))(*[Section.of(c) for c in self.sections])
        h = C.c_void_p()
        if arena is None:
            _check(lib().pe_slice_create(len(self.sections), arr, C.byref(h)), "pe_slice_create")
        else:
            _check(lib().pe_slice_create_in(len(self.sections), arr, arena.ptr, int(arena.nbytes), C.byref(h)),
                   "pe_slice_create_in")
        self.handle = h
        self.slice = sum(sum(c) for c in self.sections[0])
        self.index = {s: k for k, s in enumerate(self.sections)}

    def compute(self, children=None, stream=None):
        _check(lib().pe_slice_compute(self.handle, children.handle if children else None, stream), "pe_slice_compute")

    def slab(self, k):
        out = Slab()
        _check(lib().pe_slice_slab(self.handle, k, C.byref(out)), "pe_slice_slab")
        return out

    def counts(self):
        out = np.zeros((len(self.sections), 4), dtype=np.uint64)
        _check(lib().pe_slice_counts(self.handle, out.ctypes.data), "pe_slice_counts")
        return out

    def close(self):
        if self.handle:
            lib().pe_slice_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------- midsolve ----------------

def high_board(side0, side1, ply):
    b = HighBoard()
    b.side[0], b.side[1], b.ply = side0, side1, ply
    return b


def high_from_board(board, middle):
    """high_board_t::from_board (pentago/high/board.cc:208-221)."""
    s0 = s1 = 0
    for q in range(4):
        quad = (board >> (16 * q)) & 0xffff
        for i in range(9):
            d = quad % 3
            quad //= 3
            if d == 1:
                s0 |= 1 << (16 * q + i)
            elif d == 2:
                s1 |= 1 << (16 * q + i)
    return high_board(s0, s1, 2 * bin(s0 | s1).count("1") - int(bool(middle)))


def high_done(board):
    """high_board_t::done_and_value (pentago/high/board.h:56): (done, value)."""
    done, value = C.c_int(), C.c_int()
    _check(lib().pe_high_done(C.byref(board), C.byref(done), C.byref(value)), "pe_high_done")
    return bool(done.value), int(value.value)


def _halves(r):
    return (int(r.win[0]) | int(r.win[1]) << 64, int(r.notlose[0]) | int(r.notlose[1]) << 64)


def midsolve_internal(board):
    """midsolve_internal (pentago/mid/midengine.h:40): list of (win, notlose) halfsupers."""
    res = (HalfSupers * 19)()
    _check(lib().pe_midsolve_internal(C.byref(board), res), "pe_midsolve_internal")
    spots = 36 - ((board.ply + 1) >> 1)
    return [_halves(res[i]) for i in range(spots + 1)]


def midsolve_internal_noalloc(board, stream=None):
    """pe_midsolve_internal_noalloc: midsolve_internal without any device allocation (boards
    with at least 21 stones; the library's fixed device scratch array)."""
    res = (HalfSupers * 19)()
    _check(lib().pe_midsolve_internal_noalloc(C.byref(board), res, stream), "pe_midsolve_internal_noalloc")
    spots = 36 - ((board.ply + 1) >> 1)
    return [_halves(res[i]) for i in range(spots + 1)]


def midsolve_internal_raw(board, out):
    """midsolve_internal into a (19, 4) uint64 numpy row (win lo/hi, notlose lo/hi)."""
    _check(lib().pe_midsolve_internal(C.byref(board), out.ctypes.data_as(C.POINTER(HalfSupers))),
           "pe_midsolve_internal")


def midsolve(board):
    """midsolve (pentago/mid/midengine.h:45): {raw board: value}."""
    vals = (MidValue * 163)()
    n = C.c_int()
    _check(lib().pe_midsolve(C.byref(board), vals, C.byref(n)), "pe_midsolve")
    return {int(vals[i].raw): int(vals[i].value) for i in range(n.value)}


def midsolve_traverse(board, results):
    """midsolve_traverse (pentago/mid/midengine.cc:127-158) over precomputed results (a
    (19, 4) uint64 row as MidPlan.results() returns): {raw board: value}."""
    import numpy as np
    row = np.ascontiguousarray(results, dtype=np.uint64).reshape(19, 4)
    vals = (MidValue * 163)()
    n = C.c_int()
    _check(lib().pe_midsolve_traverse(C.byref(board), row.ctypes.data_as(C.POINTER(HalfSupers)), vals, C.byref(n)),
           "pe_midsolve_traverse")
    return {int(vals[i].raw): int(vals[i].value) for i in range(n.value)}


def midsolve_batch(boards, mode=0, stream=None):
    """pe_midsolve_batch: mode 0 level-synchronous launches, 1 persistent work queue."""
    n = len(boards)
    arr = (HighBoard * n)(*boards)
    res = ((HalfSupers * 19) * n)()
    _check(lib().pe_midsolve_batch(n, arr, C.cast(res, C.c_void_p), mode, stream), "pe_midsolve_batch")
    out = []
    for i in range(n):
        spots = 36 - ((boards[i].ply + 1) >> 1)
        out.append([_halves(res[i][k]) for k in range(spots + 1)])
    return out


def _spots(board):
    return 36 - ((board.ply + 1) >> 1)


def results_to_lists(boards, res):
    """(n, 19, 4) uint64 results -> per board the list of (win, notlose) halfsuper ints."""
    out = []
    for i, b in enumerate(boards):
        r = res[i]
        out.append([(int(r[k, 0]) | int(r[k, 1]) << 64, int(r[k, 2]) | int(r[k, 3]) << 64)
                    for k in range(_spots(b) + 1)])
    return out


class MidPlan:
    """A batch of boards solved by a batched plan (pe_mid_plan_*): small boards one per
    workgroup through a persistent device-queue kernel, large boards level by level.

    Build once, run any number of times; results() waits for the last run and returns an
    (n, 19, 4) uint64 array (win lo, win hi, notlose lo, notlose hi per entry)."""

    def __init__(self, boards, workspace_budget=0, small_spots=0, workgroups=0, threads=0, streams=0,
                 use_graph=None):
        self.boards = list(boards)
        n = len(self.boards)
        arr = (HighBoard * max(n, 1))(*self.boards)
        graph = 0 if use_graph is None else (1 if use_graph else -1)   # None: the device default
        cfg = MidConfig(workspace_budget, small_spots, workgroups, threads, streams, graph, 0)
        self.ptr = C.c_void_p()
        _check(lib().pe_mid_plan_create(n, arr, C.byref(cfg), C.byref(self.ptr)), "pe_mid_plan_create")

    def run(self, stream=None):
        _check(lib().pe_mid_plan_run(self.ptr, stream), "pe_mid_plan_run")

    def results(self, stream=None):
        import numpy as np
        res = np.zeros((max(len(self.boards), 1), 19, 4), dtype=np.uint64)
        _check(lib().pe_mid_plan_results(self.ptr, res.ctypes.data_as(C.c_void_p), stream), "pe_mid_plan_results")
        return res[:len(self.boards)]

    def info(self):
        st, cfg = MidPlanStats(), MidConfig()
        _check(lib().pe_mid_plan_info(self.ptr, C.byref(st), C.byref(cfg)), "pe_mid_plan_info")
        d = cfg.as_dict()
        d.update({n: int(getattr(st, n)) for n, _ in MidPlanStats._fields_})
        return d

    def close(self):
        if self.ptr:
            lib().pe_mid_plan_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mid_schedule_info(boards, workspace_budget, small_spots=14, workgroups=1024, threads=512, streams=2):
    """pe_mid_schedule_info: (launches per run, device bytes) of a plan, host only."""
    n = len(boards)
    arr = (HighBoard * max(n, 1))(*boards)
    cfg = MidConfig(workspace_budget, small_spots, workgroups, threads, streams, 0, 0)
    launches, ws = C.c_int64(), C.c_int64()
    _check(lib().pe_mid_schedule_info(n, arr, C.byref(cfg), C.byref(launches), C.byref(ws)), "pe_mid_schedule_info")
    return int(launches.value), int(ws.value)


def mid_autotune(boards, workspace_budget=0):
    """pe_mid_autotune: time candidate configurations (small/large split, threads per
    small-board workgroup, graph replay), the faster of two runs of these boards each; the best becomes
    the device default.  Returns the chosen configuration."""
    n = len(boards)
    arr = (HighBoard * n)(*boards)
    best = MidConfig()
    _check(lib().pe_mid_autotune(n, arr, workspace_budget, C.byref(best)), "pe_mid_autotune")
    return best.as_dict()


class MidServer:
    """Serving loop (pe_mid_server_*): submit boards as they arrive, collect results.

    submit() returns a ticket; wait(ticket) / poll(ticket) return (results, latency_ms) with
    results the list of (win, notlose) halfsuper ints as midsolve_internal gives."""

    def __init__(self, workspace_budget=0, streams=0, max_batch=0, small_spots=0, workgroups=0):
        cfg = MidServerConfig(workspace_budget, streams, max_batch, small_spots, workgroups)
        self.ptr = C.c_void_p()
        self.spots = {}
        _check(lib().pe_mid_server_create(C.byref(cfg), C.byref(self.ptr)), "pe_mid_server_create")

    def submit(self, board):
        t = C.c_int64()
        _check(lib().pe_mid_server_submit(self.ptr, C.byref(board), C.byref(t)), "pe_mid_server_submit")
        self.spots[t.value] = _spots(board)
        return t.value

    def _out(self, ticket, res, lat):
        sp = self.spots.pop(ticket)
        return [_halves(res[i]) for i in range(sp + 1)], lat.value

    def wait(self, ticket):
        res = (HalfSupers * 19)()
        lat = C.c_double()
        _check(lib().pe_mid_server_wait(self.ptr, ticket, res, C.byref(lat)), "pe_mid_server_wait")
        return self._out(ticket, res, lat)

    def poll(self, ticket):
        res = (HalfSupers * 19)()
        lat = C.c_double()
        done = C.c_int()
        _check(lib().pe_mid_server_poll(self.ptr, ticket, res, C.byref(done), C.byref(lat)), "pe_mid_server_poll")
        return self._out(ticket, res, lat) if done.value else None

    def stats(self):
        info = MidServerInfo()
        _check(lib().pe_mid_server_stats(self.ptr, C.byref(info)), "pe_mid_server_stats")
        return {n: int(getattr(info, n)) for n, _ in MidServerInfo._fields_}

    def close(self):
        if self.ptr:
            lib().pe_mid_server_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
