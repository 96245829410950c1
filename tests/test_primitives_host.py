"""Host build of the device primitives (pentago_amd/csrc/super.h) against the oracle.

The same header is compiled for gfx950 into the product; here g++ compiles it for the CPU so
the bit manipulation is checked before it runs on the GPU.  No GPU needed."""
import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from tests.common import random_side

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def hc():
    d = tempfile.mkdtemp()
    so = os.path.join(d, "hostcheck.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-o", so,
                           os.path.join(HERE, "hostcheck", "hostcheck.cc")])
    return C.CDLL(so)


def arr(x, n):
    return (C.c_uint64 * n)(*[(x >> (64 * i)) & (2**64 - 1) for i in range(n)])


def val(a):
    return sum(int(a[i]) << (64 * i) for i in range(len(a)))


def test_rmax(hc):
    rng = np.random.default_rng(1)
    for _ in range(300):
        s = int.from_bytes(rng.bytes(32), "little")
        out = (C.c_uint64 * 4)()
        hc.hc_rmax(arr(s, 4), out)
        assert val(out) == O.rmax(s)


def test_transform_all_symmetries(hc):
    rng = np.random.default_rng(2)
    for g in range(8):
        for _ in range(60):
            y = int(rng.integers(256))
            s = int.from_bytes(rng.bytes(32), "little")
            out = (C.c_uint64 * 4)()
            hc.hc_transform(g, y, arr(s, 4), out)
            assert val(out) == O.transform_super((g, y), s), (g, y)


def test_shift_one_coordinate(hc):
    """shift_coord_var(X, q, k)(z) = X(z + k e_q), the identity transform with y = k << 2q."""
    rng = np.random.default_rng(5)
    for q in range(4):
        for k in range(4):
            for _ in range(20):
                s = int.from_bytes(rng.bytes(32), "little")
                out = (C.c_uint64 * 4)()
                hc.hc_shift_var(q, k, arr(s, 4), out)
                assert val(out) == O.transform_super((0, k << (2 * q)), s), (q, k)


@pytest.mark.parametrize("g", [1, 2, 3])
def test_line_transforms_of_unreflected_lines(hc, g):
    """The generated straight-line transforms equal transform_super((g, y)) of the oracle for
    the three rotations an unreflected line uses, y = r (1,1,1,1) with r = -g mod 4
    (pentago/end/compute.cc:69), on every single bit (a permutation is fixed by where each
    bit goes) and on random supers."""
    y = 0x55 * ((4 - g) & 3)
    for j in range(256):
        out = (C.c_uint64 * 4)()
        hc.hc_line_transform(g, arr(1 << j, 4), out)
        assert val(out) == O.transform_super((g, y), 1 << j), (g, j)
    rng = np.random.default_rng(6 + g)
    for _ in range(50):
        s = int.from_bytes(rng.bytes(32), "little")
        out = (C.c_uint64 * 4)()
        hc.hc_line_transform(g, arr(s, 4), out)
        assert val(out) == O.transform_super((g, y), s)


def test_super_wins(hc):
    f = O.lib().po_superwin_info
    f.restype = C.POINTER(C.c_uint64)
    tab = f()
    rng = np.random.default_rng(3)
    for _ in range(2000):
        side = random_side(rng, rng.uniform(0.1, 0.9))
        out = (C.c_uint64 * 4)()
        hc.hc_super_wins(C.c_uint64(side), tab, out)
        assert val(out) == O.super_wins(side)


def test_split_and_half_rmax(hc):
    rng = np.random.default_rng(4)
    for _ in range(500):
        s = int.from_bytes(rng.bytes(32), "little")
        h = O.split(s)
        for p in (0, 1):
            out = (C.c_uint64 * 2)()
            hc.hc_split(arr(s, 4), p, out)
            assert val(out) == h[p]
        x = int.from_bytes(rng.bytes(16), "little")
        out = (C.c_uint64 * 2)()
        hc.hc_hrmax(arr(x, 2), out)
        assert val(out) == O.half_rmax(x)


def test_generated_transforms_header_is_current():
    """pentago_amd/csrc/transforms.h is the output of gen_transforms.py (whose plans are
    simulated on all 256 bit positions before they are written)."""
    csrc = os.path.join(HERE, "..", "pentago_amd", "csrc")
    out = subprocess.check_output([sys.executable, os.path.join(csrc, "gen_transforms.py")], text=True)
    with open(os.path.join(csrc, "transforms.h")) as f:
        assert f.read() == out
