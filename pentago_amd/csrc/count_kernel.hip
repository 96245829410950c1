// Win counts of computed blocks: pentago/end/block_store.cc:330-349 (count_block_wins), the
// sum over a block's nodes of popcounts_over_stabilizers (pentago/base/count.cc:254-262).
//
// A node's board is the rotation-minimal quadrant tuple its index names; its local
// stabilizer (the rotations of single quadrants that fix the board) has 2^shift elements,
// shift = the sum over quadrants of 2 (fixed by a quarter turn), 1 (by a half turn) or 0.
// Every super evaluation is constant on stabilizer cosets, so the node's distinct rotated
// positions number 256 >> shift and its wins (wins-or-ties) are popcount >> shift.
//
// One workgroup per block reads the block's nodes once (64 bytes each, coalesced: thread t
// takes nodes t, t+256, ...), reduces in registers, across the wave and through LDS, and
// writes four u64 per block: wins, wins-or-ties, total, and the number of nodes whose
// popcounts are not multiples of 2^shift (the reference asserts there are none; a nonzero
// value means the data is not a super evaluation of its boards).  HBM-bound: 64 B per node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "device.h"
#include "end.h"
#include "tables.h"

namespace pe {
namespace {

constexpr int kCountThreads = 256;

struct CountArgs {
  const uint64_t* data;      // slab base: block k's slot at node 4096 k
  uint64_t* counts;          // 4 u64 per block of the box
  int32_t lo[4], ext[4];     // box origin and extent, in blocks
  int32_t rmin_off[4], rmin_size[4];
  int64_t first;             // index of the slab's first block among all blocks of the launch
};

__device__ inline uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block blk of slab A, by one workgroup.
__device__ inline void count_block(const CountArgs& A, int64_t blk, const uint8_t* __restrict__ stab) {
  int b[4], n[4];
  int64_t r = blk;
  for (int a = 3; a >= 0; a--) {
    b[a] = A.lo[a] + int(r % A.ext[a]);
    r /= A.ext[a];
  }
  for (int a = 0; a < 4; a++) n[a] = min(8, A.rmin_size[a] - 8 * b[a]);
  const int nodes = n[0] * n[1] * n[2] * n[3];
  const ulonglong2* base = reinterpret_cast<const ulonglong2*>(A.data + blk * (4096 * 8));
  const uint8_t* s0 = stab + A.rmin_off[0] + 8 * b[0];
  const uint8_t* s1 = stab + A.rmin_off[1] + 8 * b[1];
  const uint8_t* s2 = stab + A.rmin_off[2] + 8 * b[2];
  const uint8_t* s3 = stab + A.rmin_off[3] + 8 * b[3];
  uint32_t c0 = 0, c1 = 0, c2 = 0, bad = 0;
  for (int p = threadIdx.x; p < nodes; p += kCountThreads) {
    const ulonglong2 x0 = base[4 * p], x1 = base[4 * p + 1], x2 = base[4 * p + 2], x3 = base[4 * p + 3];
    int t = p;
    const int i3 = t % n[3]; t /= n[3];
    const int i2 = t % n[2]; t /= n[2];
    const int i1 = t % n[1];
    const int i0 = t / n[1];
    const int shift = s0[i0] + s1[i1] + s2[i2] + s3[i3];
    const uint32_t w0 = __popcll(x0.x) + __popcll(x0.y) + __popcll(x1.x) + __popcll(x1.y);
    const uint32_t w1 = __popcll(x2.x) + __popcll(x2.y) + __popcll(x3.x) + __popcll(x3.y);
    bad += ((w0 | w1) & ((1u << shift) - 1)) != 0;
    c0 += w0 >> shift;
    c1 += w1 >> shift;
    c2 += 256u >> shift;
  }
  c0 = wave_sum(c0);
  c1 = wave_sum(c1);
  c2 = wave_sum(c2);
  bad = wave_sum(bad);
  __shared__ uint32_t part[kCountThreads / 64][4];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[wave][0] = c0;
    part[wave][1] = c1;
    part[wave][2] = c2;
    part[wave][3] = bad;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    uint64_t s = 0;
    for (int w = 0; w < kCountThreads / 64; w++) s += part[w][threadIdx.x];
    A.counts[4 * blk + threadIdx.x] = s;
  }
}

// One slab, its descriptor passed by value: workgroup k counts block k.
__global__ __launch_bounds__(kCountThreads) void count_blocks_kernel(CountArgs A, const uint8_t* __restrict__ stab) {
  count_block(A, blockIdx.x, stab);
}

// Many slabs in one launch: slab s owns workgroups [first_s, first_{s+1}) (first ascending,
// first_0 = 0).  Each wave finds its slab by a 64-way search: the lanes test 64 evenly spaced
// descriptors of the remaining range at once, so 4096 slabs take two dependent loads.
__global__ __launch_bounds__(kCountThreads) void count_slabs_kernel(const CountArgs* __restrict__ slabs, int nslabs,
                                                                     const uint8_t* __restrict__ stab) {
  const int64_t wg = blockIdx.x;
  const int lane = threadIdx.x & 63;
  int lo = 0, len = nslabs;   // the slab is in [lo, lo + len) and first_lo <= wg
  while (len > 1) {
    const int step = (len + 63) / 64;
    const int idx = lo + lane * step;
    const bool le = idx < lo + len && slabs[idx].first <= wg;
    const int c = __popcll(__ballot(le));   // a prefix of the lanes, lane 0 among them
    const int end = lo + len;
    lo += (c - 1) * step;
    len = min(step, end - lo);
  }
  const CountArgs A = slabs[lo];
  count_block(A, wg - A.first, stab);
}

}  // namespace

int64_t slab_blocks(const pe_slab& slab) {
  int64_t n = 1;
  for (int a = 0; a < 4; a++) n *= std::max(0, slab.hi[a] - slab.lo[a]);
  return n;
}

// Launch descriptor of one slab (counts land at d_counts, 4 u64 per block).
static pe_status count_args(const pe_slab& slab, uint64_t* d_counts, int64_t first, CountArgs* A, std::string* err) {
  Section s;
  std::memcpy(s.c, slab.section.counts, 8);
  int shape[4];
  if (!s.valid_counts() || !s.shape(shape)) { *err = "invalid slab section"; return PE_ERR_INVALID; }
  if (!slab.data || !d_counts) { *err = "null slab data or counts"; return PE_ERR_INVALID; }
  A->data = static_cast<const uint64_t*>(slab.data);
  A->counts = d_counts;
  A->first = first;
  const auto& Tb = tables();
  for (int a = 0; a < 4; a++) {
    if (slab.lo[a] < 0 || slab.hi[a] <= slab.lo[a] || 8 * (slab.hi[a] - 1) >= shape[a]) {
      *err = "slab box outside its section";
      return PE_ERR_INVALID;
    }
    A->lo[a] = slab.lo[a];
    A->ext[a] = slab.hi[a] - slab.lo[a];
    const int bucket = rmin_bucket(s.c[a][0], s.c[a][1]);
    A->rmin_off[a] = Tb.rmin_offsets[bucket];
    A->rmin_size[a] = shape[a];
  }
  return PE_OK;
}

// Per-block counts of n slabs in one launch: slab k's blocks land at d_counts[k] (4 u64 per
// block, C-order over its box).  With desc null (one slab only) the descriptor is passed by
// value and the call is asynchronous; otherwise the descriptors go up into `desc`, device
// memory of count_desc_bytes(n) bytes, by a copy that waits for the stream first.
pe_status count_blocks_launch(int n, const pe_slab* slabs, uint64_t* const* d_counts, void* desc, hipStream_t stream,
                              std::string* err) {
  DevTables T;
  if (device_tables(&T, err) != PE_OK) return PE_ERR_DEVICE;
  std::vector<CountArgs> args(size_t(std::max(n, 0)));
  int64_t total = 0;
  for (int k = 0; k < n; k++) {
    const pe_status st = count_args(slabs[k], d_counts[k], total, &args[size_t(k)], err);
    if (st != PE_OK) return st;
    total += slab_blocks(slabs[k]);
  }
  if (!total) return PE_OK;
  if (total > 0x7fffffff) { *err = "too many blocks in one count launch"; return PE_ERR_CAPACITY; }
  if (n == 1 && !desc) {
    hipLaunchKernelGGL(count_blocks_kernel, dim3(unsigned(total)), dim3(kCountThreads), 0, stream, args[0], T.rmin_stab);
  } else {
    if (!desc) { *err = "no descriptor memory for a multi-slab count"; return PE_ERR_INVALID; }
    // Synchronous: the descriptors live on this function's stack frame.
    if (hipStreamSynchronize(stream) != hipSuccess ||
        hipMemcpy(desc, args.data(), sizeof(CountArgs) * args.size(), hipMemcpyHostToDevice) != hipSuccess) {
      *err = "count descriptor upload failed";
      return PE_ERR_DEVICE;
    }
    hipLaunchKernelGGL(count_slabs_kernel, dim3(unsigned(total)), dim3(kCountThreads), 0, stream,
                       static_cast<const CountArgs*>(desc), n, T.rmin_stab);
  }
  if (hipGetLastError() != hipSuccess) { *err = "count kernel launch failed"; return PE_ERR_DEVICE; }
  return PE_OK;
}

size_t count_desc_bytes(int n) { return sizeof(CountArgs) * size_t(std::max(n, 1)); }

}  // namespace pe
