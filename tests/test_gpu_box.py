"""The benchmark's box workload (child blocks read from a section slab, box-restricted index
ranges, four OR-merged dimension passes) against the oracle on a small box."""
import numpy as np
import pytest

import bench
import pentago_amd as P
from pentago_amd import hip
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,a0", [(1, 51), (2, 7)])
def test_box_matches_oracle(k, a0):
    P.init()
    out, kids, plans, child_nodes, child, n_child_blocks = bench.build_box(P, hip, k, a0)
    for p in plans:
        p.run()
    hip.synchronize()
    box = out.download(np.zeros(8 * 4096 * k ** 4, dtype=np.uint64)).reshape(k ** 4, 4096, 8)
    slab = kids.download(np.zeros(8 * 4096 * n_child_blocks, dtype=np.uint64)).reshape(n_child_blocks, 4096, 8)
    info0 = P.line_describe(P.make_line(bench.SECTION, 0, (a0, a0, a0)))
    cd0 = info0.child_dimension
    cshape = P.section_shape(child)
    cd_len = (cshape[cd0] + 7) // 8
    slab_dims = [cd_len if ax == cd0 else k for ax in range(4)]
    slab_lo = [0 if ax == cd0 else a0 for ax in range(4)]

    def child_block(cb):
        idx = 0
        for ax in range(4):
            idx = idx * slab_dims[ax] + (cb[ax] - slab_lo[ax])
        bshape = [min(cshape[ax], 8 * (cb[ax] + 1)) - 8 * cb[ax] for ax in range(4)]
        return slab[idx, :int(np.prod(bshape))].reshape(-1)

    expect = np.zeros_like(box)
    rng = np.random.default_rng(k)
    blocks = [tuple(a0 + int(x) for x in rng.integers(k, size=4)) for _ in range(2)]
    for blk in blocks:
        acc = np.zeros(8 * 4096, dtype=np.uint64)
        for d in range(4):
            base = tuple(blk[ax] for ax in range(4) if ax != d)
            ol = O.make_line(bench.SECTION, d, base)
            oi = O.line_describe(ol)
            inp = np.zeros(8 * oi.input_nodes, dtype=np.uint64)
            for kk in range(oi.input_blocks):
                cb = list(oi.first_child_block)
                cb[oi.child_dim] = kk
                data = child_block(cb)
                off = 8 * kk * (oi.block_stride + 1)
                inp[off:off + data.size] = data
            line_out = np.zeros(8 * oi.output_nodes, dtype=np.uint64)
            O.compute_lines_threaded([ol], [inp], [line_out], 16)
            off = 8 * oi.block_stride * blk[d]
            acc |= line_out[off:off + acc.size]
        idx = 0
        for ax in range(4):
            idx = idx * k + (blk[ax] - a0)
        assert np.array_equal(box[idx].reshape(-1), acc), blk
