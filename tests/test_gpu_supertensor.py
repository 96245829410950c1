"""Supertensor files on the GPU: the production run's slice 0-5 files decoded into device
slabs (pe_supertensor_decode) and counted by the count kernel equal the production counts;
the device encode/decode kernels agree with the oracle's interleave filter and the
(black, white) orientation of pentago/end/check.cc:149-158; computed sections written as a
multiple-section file (zlib and xz) read back bit for bit.
"""
import os
import tempfile

import numpy as np
import pytest

import pentago_amd as P
from pentago_amd import hip
from pentago_amd import supertensor as ST
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FULL = np.uint64(0xFFFFFFFFFFFFFFFF)


def whole_slab(counts):
    shape = P.section_shape(counts)
    nb = [(s + 7) // 8 for s in shape]
    buf = hip.Buffer(64 * 4096 * int(np.prod(nb)))
    return P.Slab.of(counts, [0] * 4, nb, buf.ptr), buf


def slab_nodes(slab, counts):
    """Valid nodes of every block of a whole slab, (blocks, nodes of the block) concatenated."""
    shape = P.section_shape(counts)
    nb = [slab.hi[a] - slab.lo[a] for a in range(4)]
    out = []
    for k in range(int(np.prod(nb))):
        b = np.unravel_index(k, nb)
        n = int(np.prod([min(8, shape[a] - 8 * b[a]) for a in range(4)]))
        a = np.zeros((n, 8), dtype=np.uint64)
        hip.check(hip.hip().hipMemcpy(hip.C.c_void_p(a.ctypes.data), hip.C.c_void_p(slab.data + 64 * 4096 * k),
                                      a.nbytes, hip.D2H), "hipMemcpy D2H")
        out.append(a)
    return np.concatenate(out)


@pytest.mark.parametrize("slice_", range(6))
def test_production_files_decode_and_count_on_device(slice_):
    readers = ST.open_supertensors(os.path.join(GOLDEN, "slice-%d.pentago" % slice_))
    want = np.load(os.path.join(GOLDEN, "counts-%d.npy" % slice_), allow_pickle=False)
    want = want[np.argsort(want[:, 0], kind="stable")]
    slabs, keep, secs = [], [], []
    for r in readers:
        slab, buf = whole_slab(r.header.section)
        r.load(slab)
        slabs.append(slab)
        keep.append(buf)
        secs.append(r.header.section)
    hip.synchronize()
    counts = P.count_slabs(slabs)
    assert int(counts[:, 3].sum()) == 0
    rows = P.counts_rows(secs, counts, slice_)
    rows = rows[np.argsort(rows[:, 0], kind="stable")]
    assert np.array_equal(rows, want)


@pytest.mark.parametrize("turn", [0, 1])
def test_encode_decode_match_oracle_filter(turn):
    rng = np.random.default_rng(11 + turn)
    n = 3000
    nodes = rng.integers(0, 2**63, size=(n, 8), dtype=np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, size=(n, 8), dtype=np.uint64)
    buf = hip.Buffer(64 * n)
    buf.upload(nodes)
    out = hip.Buffer(64 * n)
    P._check(P.lib().pe_supertensor_encode(1, turn, buf.ptr, n, out.ptr, None), "encode")
    hip.synchronize()
    got = out.download(np.zeros((n, 8), dtype=np.uint64))
    win, wot = nodes[:, :4], nodes[:, 4:]
    black, white = (~wot & FULL, win) if turn else (win, ~wot & FULL)
    assert np.array_equal(got, O.interleave(np.concatenate([black, white], axis=1)))
    # decode in place restores the nodes; filter 0 is the plain (black, white) pair
    P._check(P.lib().pe_supertensor_decode(1, turn, out.ptr, n, out.ptr, None), "decode")
    hip.synchronize()
    assert np.array_equal(out.download(np.zeros((n, 8), dtype=np.uint64)), nodes)
    P._check(P.lib().pe_supertensor_encode(0, turn, buf.ptr, n, out.ptr, None), "encode")
    hip.synchronize()
    assert np.array_equal(out.download(np.zeros((n, 8), dtype=np.uint64)), np.concatenate([black, white], axis=1))


@pytest.mark.parametrize("level", [6, 26])
def test_write_read_round_trip(level):
    # a few slice-35 sections computed on the device, one with partial edge blocks
    secs = P.descendent_sections(((0, 0),) * 4, 35)
    secs = sorted(secs, key=lambda s: int(np.prod(P.section_shape(s))))
    pick = secs[:3] + [secs[len(secs) // 2]]
    assert any(x % 8 for x in P.section_shape(pick[-1]))
    st = P.SliceStore(pick)
    st.compute()
    slabs = [st.slab(k) for k in range(len(pick))]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "slice-35-part.pentago")
        ST.write_supertensors(path, slabs, level=level)
        readers = ST.open_supertensors(path)
        assert [r.header.section for r in readers] == [tuple(s) for s in st.sections]
        for r, slab, sec in zip(readers, slabs, st.sections):
            orig = slab_nodes(slab, sec)
            # host view: the file's (black, white) supers are the nodes turned round
            raw = np.concatenate([r.read_raw(np.unravel_index(i, r.header.blocks))
                                  for i in range(int(np.prod(r.header.blocks)))])
            bw = O.uninterleave(raw)
            assert np.array_equal(bw[:, :4], orig[:, 4:] ^ FULL)      # slice 35: white to move
            assert np.array_equal(bw[:, 4:], orig[:, :4])
            # device view: loading into a fresh slab restores the nodes
            fresh, buf = whole_slab(sec)
            r.load(fresh)
            hip.synchronize()
            assert np.array_equal(slab_nodes(fresh, sec), orig)
    st.close()
