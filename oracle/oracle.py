"""ctypes view of the oracle (plain-C restatement of the reference path).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product (pentago_amd/) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ORACLE_LIB", os.path.join(HERE, "liboracle.so"))


class Super(C.Structure):
    _fields_ = [("w", C.c_uint64 * 4)]


class Half(C.Structure):
    _fields_ = [("w", C.c_uint64 * 2)]


class Halves(C.Structure):
    _fields_ = [("win", Half), ("notlose", Half)]


class Sym(C.Structure):
    _fields_ = [("glob", C.c_uint8), ("local", C.c_uint8)]


class Section(C.Structure):
    # Eight scalar fields (not a nested array) so that libffi classifies the 8-byte struct
    # correctly when it is passed by value.
    _fields_ = [("c%d%d" % (q, k), C.c_uint8) for q in range(4) for k in range(2)]


class Line(C.Structure):
    _fields_ = [("section", Section), ("dimension", C.c_uint8), ("length", C.c_uint8),
                ("block_base", C.c_uint8 * 3)]


class High(C.Structure):
    _fields_ = [("side", C.c_uint64 * 2), ("ply", C.c_uint32)]


class LineInfo(C.Structure):
    _fields_ = [("input_shape", C.c_int * 4), ("output_shape", C.c_int * 4),
                ("block_stride", C.c_int), ("input_blocks", C.c_int), ("length", C.c_int),
                ("dim", C.c_int), ("child_dim", C.c_int),
                ("input_nodes", C.c_int64), ("output_nodes", C.c_int64),
                ("standard_child", Section), ("section_transform", C.c_int),
                ("permutation", C.c_int * 4), ("first_child_block", C.c_int * 4)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None
_fast = None
# The same sources built with -O3 -march=x86-64-v3 (AVX2) for the CPU baseline timings.
FAST_PATH = os.path.join(HERE, "liboracle_fast.so")


def _load(path):
    L = C.CDLL(path)
    P = C.POINTER
    sigs = {
        "po_init": (None, []),
        "po_table_hash": (C.c_int, [C.c_char_p, C.c_char_p]),
        "po_rmax": (Super, [Super]),
        "po_super_wins": (Super, [C.c_uint64]),
        "po_won": (C.c_int, [C.c_uint64]),
        "po_sym_mul": (Sym, [Sym, Sym]),
        "po_sym_inverse": (Sym, [Sym]),
        "po_transform_super": (Super, [Sym, Super]),
        "po_transform_side": (C.c_uint64, [Sym, C.c_uint64]),
        "po_transform_board": (C.c_uint64, [Sym, C.c_uint64]),
        "po_meaningless": (C.c_int, [C.c_uint64, C.c_uint64]),
        "po_super_meaningless": (Super, [C.c_uint64, C.c_uint64]),
        "po_pack_board": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "po_unpack_board": (C.c_uint64, [C.c_uint64, C.c_int]),
        "po_rmin": (C.c_int, [C.c_int, C.c_int, P(P(C.c_uint16)), P(C.c_int)]),
        "po_rmin_inverse": (C.c_uint16, [C.c_uint16]),
        "po_section_shape": (C.c_int, [Section, P(C.c_int)]),
        "po_section_standardize": (C.c_int, [Section, P(Section)]),
        "po_line_describe": (C.c_int, [P(Line), P(LineInfo)]),
        "po_compute_line": (C.c_int, [P(Line), C.c_void_p, C.c_void_p]),
        "po_compute_microlines": (C.c_int, [P(Line), C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
        "po_compute_lines_threaded": (C.c_int, [C.c_int, P(Line), P(C.c_void_p), P(C.c_void_p), C.c_int]),
        "po_eval_reset": (None, []),
        "po_eval_node": (C.c_int, [Section, P(C.c_int), P(Super)]),
        "po_popcounts_over_stabilizers": (C.c_int, [C.c_uint64, P(Super), P(C.c_uint16)]),
        "po_section_counts": (C.c_int, [Section, P(C.c_uint64)]),
        "po_interleave": (None, [C.c_int64, C.c_void_p]),
        "po_uninterleave": (None, [C.c_int64, C.c_void_p]),
        "po_brute_value": (C.c_int, [C.c_uint64, C.c_uint64]),
        "po_brute_super": (None, [C.c_uint64, C.c_uint64, P(Super)]),
        "po_half_rmax": (Half, [Half]),
        "po_split": (None, [Super, P(Half)]),
        "po_merge": (Super, [Half, Half]),
        "po_halfsuper_wins": (Half, [C.c_uint64, C.c_int]),
        "po_choose": (C.c_int, [C.c_int, C.c_int]),
        "po_subset": (C.c_uint64, [C.c_int, C.c_int, C.c_int]),
        "po_subset_mask": (C.c_uint32, [C.c_int, C.c_int, C.c_int]),
        "po_subset_index": (C.c_int, [C.c_int, C.c_int, C.c_uint64]),
        "po_bottleneck": (C.c_int, [C.c_int]),
        "po_midsolve_workspace_size": (C.c_int, [C.c_int]),
        "po_midsolve_internal": (C.c_int, [High, C.c_void_p, P(Halves)]),
        "po_midsolve": (C.c_int, [High, C.c_void_p, P(C.c_uint64), P(C.c_int)]),
        "po_midsolve_traverse": (C.c_int, [High, P(Halves), P(C.c_uint64), P(C.c_int)]),
        "po_midsolve_batch": (C.c_int, [C.c_int, P(High), C.c_void_p, C.c_int]),
        "po_high_from_board": (High, [C.c_uint64, C.c_int]),
        "po_high_board": (C.c_uint64, [High]),
        "po_high_raw": (C.c_uint64, [High]),
        "po_high_place": (High, [High, C.c_int]),
        "po_high_rotate": (High, [High, C.c_int, C.c_int]),
        "po_high_done_and_value": (C.c_int, [High, P(C.c_int)]),
        "po_high_count": (C.c_int, [High]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L.po_init()
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


def fast_lib():
    """The optimized build (for timing), or None where it is missing or the host lacks AVX2."""
    global _fast
    if _fast is None:
        flags = ""
        try:
            with open("/proc/cpuinfo") as f:
                flags = next((l for l in f if l.startswith("flags")), "")
        except OSError:
            pass
        _fast = _load(FAST_PATH) if os.path.exists(FAST_PATH) and "avx2" in flags.split() else False
    return _fast or None


# ---------------- conversions ----------------

def super_from_int(x):
    s = Super()
    for i in range(4):
        s.w[i] = (x >> (64 * i)) & (2**64 - 1)
    return s


def super_to_int(s):
    return sum(int(s.w[i]) << (64 * i) for i in range(4))


def half_to_int(h):
    return int(h.w[0]) | int(h.w[1]) << 64


def half_from_int(x):
    h = Half()
    h.w[0] = x & (2**64 - 1)
    h.w[1] = x >> 64
    return h


def section(counts):
    s = Section()
    for q in range(4):
        setattr(s, "c%d0" % q, counts[q][0])
        setattr(s, "c%d1" % q, counts[q][1])
    return s


def section_counts(s):
    return tuple((getattr(s, "c%d0" % q), getattr(s, "c%d1" % q)) for q in range(4))


def make_line(counts, dim, block_base, length=None):
    sec = section(counts)
    shape = section_shape(counts)
    line = Line()
    line.section = sec
    line.dimension = dim
    line.length = (shape[dim] + 7) // 8 if length is None else length
    for i in range(3):
        line.block_base[i] = block_base[i]
    return line


# ---------------- wrappers ----------------

def rmax(x):
    return super_to_int(lib().po_rmax(super_from_int(x)))


def super_wins(side):
    return super_to_int(lib().po_super_wins(side))


def won(side):
    return bool(lib().po_won(side))


def sym(g, l):
    return Sym(g, l)


def sym_mul(a, b):
    r = lib().po_sym_mul(Sym(*a), Sym(*b))
    return (r.glob, r.local)


def sym_inverse(a):
    r = lib().po_sym_inverse(Sym(*a))
    return (r.glob, r.local)


def transform_super(s, x):
    return super_to_int(lib().po_transform_super(Sym(*s), super_from_int(x)))


def transform_side(s, side):
    return lib().po_transform_side(Sym(*s), side)


def transform_board(s, board):
    return lib().po_transform_board(Sym(*s), board)


def pack_board(s0, s1):
    return lib().po_pack_board(s0, s1)


def unpack_board(b, s):
    return lib().po_unpack_board(b, s)


def meaningless(board, salt=0):
    return lib().po_meaningless(board, salt)


def super_meaningless(board, salt=0):
    return super_to_int(lib().po_super_meaningless(board, salt))


def rmin(b, w):
    p = C.POINTER(C.c_uint16)()
    moved = C.c_int()
    n = lib().po_rmin(b, w, C.byref(p), C.byref(moved))
    return np.array([p[i] for i in range(n)], dtype=np.uint16), moved.value


def rmin_inverse(q):
    return lib().po_rmin_inverse(q)


def section_shape(counts):
    shape = (C.c_int * 4)()
    assert lib().po_section_shape(section(counts), shape) == 0
    return tuple(shape)


def standardize(counts):
    out = Section()
    g = lib().po_section_standardize(section(counts), C.byref(out))
    return section_counts(out), g


def line_describe(line):
    info = LineInfo()
    r = lib().po_line_describe(C.byref(line), C.byref(info))
    if r:
        raise ValueError("invalid line (%d)" % r)
    return info


def compute_line(line, inp):
    """inp: uint64 array of 8*input_nodes words (node = 2 supers).  Returns output words."""
    info = line_describe(line)
    out = np.zeros(8 * info.output_nodes, dtype=np.uint64)
    ip = inp.ctypes.data if inp is not None and inp.size else None
    assert lib().po_compute_line(C.byref(line), ip, out.ctypes.data) == 0
    return out


def compute_microlines(line, inp, lo, hi, out=None):
    info = line_describe(line)
    if out is None:
        out = np.zeros(8 * info.output_nodes, dtype=np.uint64)
    ip = inp.ctypes.data if inp is not None and inp.size else None
    assert lib().po_compute_microlines(C.byref(line), ip, out.ctypes.data, lo, hi) == 0
    return out


def compute_lines_threaded(lines, inputs, outputs, threads, L=None):
    n = len(lines)
    arr = (Line * n)(*lines)
    ins = (C.c_void_p * n)(*[(i.ctypes.data if i is not None and i.size else None) for i in inputs])
    outs = (C.c_void_p * n)(*[o.ctypes.data for o in outputs])
    assert (L or lib()).po_compute_lines_threaded(n, arr, ins, outs, threads) == 0


def eval_node(counts, index):
    out = (Super * 2)()
    idx = (C.c_int * 4)(*index)
    assert lib().po_eval_node(section(counts), idx, out) == 0
    return super_to_int(out[0]), super_to_int(out[1])


def popcounts_over_stabilizers(board, wins, notlose):
    """(wins, wins-or-ties, total) over distinct rotated positions; None where the reference
    asserts (counts not multiples of the stabilizer size)."""
    v = (Super * 2)(super_from_int(wins), super_from_int(notlose))
    out = (C.c_uint16 * 3)()
    r = lib().po_popcounts_over_stabilizers(board, v, out)
    return None if r else tuple(out)


def count_section(counts):
    """(wins, wins-or-ties, total) summed over every node of a standard section."""
    out = (C.c_uint64 * 3)()
    r = lib().po_section_counts(section(counts), out)
    assert r == 0, r
    return tuple(out)


def interleave(nodes):
    """(n, 8) uint64 (black, white) nodes -> filter-1 file words (a new array)."""
    a = np.ascontiguousarray(nodes, dtype=np.uint64).copy()
    lib().po_interleave(len(a), a.ctypes.data)
    return a


def uninterleave(words):
    a = np.ascontiguousarray(words, dtype=np.uint64).copy()
    lib().po_uninterleave(len(a), a.ctypes.data)
    return a


def brute_super(me, them):
    out = (Super * 2)()
    lib().po_brute_super(me, them, out)
    return super_to_int(out[0]), super_to_int(out[1])


def brute_value(me, them):
    return lib().po_brute_value(me, them)


def half_rmax(h):
    return half_to_int(lib().po_half_rmax(half_from_int(h)))


def split(s):
    out = (Half * 2)()
    lib().po_split(super_from_int(s), out)
    return half_to_int(out[0]), half_to_int(out[1])


def merge(even, odd):
    return super_to_int(lib().po_merge(half_from_int(even), half_from_int(odd)))


def halfsuper_wins(side, parity):
    return half_to_int(lib().po_halfsuper_wins(side, parity))


def high_from_board(board, middle):
    return lib().po_high_from_board(board, int(middle))


def midsolve_internal(root):
    """root: High.  Returns list of 19 (win, notlose) pairs as ints."""
    spots = 36 - lib().po_high_count(root)
    ws = np.zeros(2 * max(1, lib().po_bottleneck(spots)), dtype=np.uint64)
    res = (Halves * 19)()
    assert lib().po_midsolve_internal(root, ws.ctypes.data, res) == 0
    return [(half_to_int(r.win), half_to_int(r.notlose)) for r in res]


def midsolve(root):
    spots = 36 - lib().po_high_count(root)
    ws = np.zeros(2 * max(1, lib().po_bottleneck(spots)), dtype=np.uint64)
    raws = (C.c_uint64 * 163)()
    vals = (C.c_int * 163)()
    n = lib().po_midsolve(root, ws.ctypes.data, raws, vals)
    assert n >= 0
    return {int(raws[i]): int(vals[i]) for i in range(n)}


def midsolve_traverse(root, supers):
    """supers: list of (win, notlose) ints for the first len(supers) entries."""
    arr = (Halves * 19)()
    for i, (w, n) in enumerate(supers):
        arr[i].win = half_from_int(w)
        arr[i].notlose = half_from_int(n)
    raws = (C.c_uint64 * 163)()
    vals = (C.c_int * 163)()
    n = lib().po_midsolve_traverse(root, arr, raws, vals)
    return [(int(raws[i]), int(vals[i])) for i in range(n)]


def midsolve_batch(roots, threads, L=None):
    n = len(roots)
    arr = (High * n)(*roots)
    res = ((Halves * 19) * n)()
    assert (L or lib()).po_midsolve_batch(n, arr, C.cast(res, C.c_void_p), threads) == 0
    return [[(half_to_int(r.win), half_to_int(r.notlose)) for r in res[i]] for i in range(n)]
