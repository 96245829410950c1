// Measurement probe, built as its own library (libpentago_probe.so), not part of the engine:
// the device's rate for the fused block kernel's memory access pattern without its
// arithmetic.  Thread i reads twelve 64-byte nodes of `src` at hashed positions, three at a
// time as the kernel loads a dimension's moves (lanes 2k and 2k+1 read the two nodes of one
// 128-byte line, as the kernel's neighbouring lanes mostly do), XORs them and writes one
// 64-byte node to dst[i].  Positions are hashed over the whole buffer (window 0: apart from
// the line pairs no read is served by the L2, and nearly every read is a new page), or,
// like the kernel's four child lines per block, each group of three reads of the threads of
// one 4096-node block falls in its own window of `window` nodes at a hashed place.
#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

constexpr int kReads = 12, kGroup = 3;

__device__ inline uint64_t mix64(uint64_t x) {   // splitmix64 finalizer (Steele et al. 2014)
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

__device__ inline uint4 xor4(uint4 a, uint4 b) { return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w); }

// pairs: 128-byte lines per window; bases: places a window can start at (in lines)
__global__ __launch_bounds__(256) void gather_probe_kernel(const uint4* __restrict__ src, uint64_t pairs,
                                                           uint64_t bases, uint4* __restrict__ dst, int64_t n,
                                                           uint64_t seed) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  uint4 acc[4] = {};
  const uint64_t pair = uint64_t(i) >> 1, block = uint64_t(i) >> 12;
  for (int g = 0; g < kReads; g += kGroup) {
    const uint64_t base = __umul64hi(mix64(~seed ^ (block * kReads + uint64_t(g))), bases);
    uint4 v[kGroup][4];
#pragma unroll
    for (int r = 0; r < kGroup; r++) {
      // the same line for both lanes of a pair; the node within it by the lane's parity
      const uint64_t line = base + __umul64hi(mix64(seed + pair * kReads + uint64_t(g + r)), pairs);
      const uint4* p = src + 4 * (2 * line + (uint64_t(i) & 1));
#pragma unroll
      for (int w = 0; w < 4; w++) v[r][w] = p[w];
    }
#pragma unroll
    for (int r = 0; r < kGroup; r++)
#pragma unroll
      for (int w = 0; w < 4; w++) acc[w] = xor4(acc[w], v[r][w]);
  }
  uint4* o = dst + 4 * i;
#pragma unroll
  for (int w = 0; w < 4; w++) o[w] = acc[w];
}

}  // namespace

// src: src_nodes 64-byte nodes (even, >= 2); window: 0 (the whole buffer) or an even node
// count in [2, src_nodes]; dst: n nodes, 1 <= n < 2^32.  Returns the hipError_t of the launch.
// Every read stays in src: a window starts at line b < bases = (src_nodes - window) / 2 + 1
// and spans window / 2 lines.
extern "C" int pe_probe_gather(const void* src, int64_t src_nodes, void* dst, int64_t n, int64_t window,
                               uint64_t seed, void* stream) {
  if (!window) window = src_nodes;
  if (!src || !dst || src_nodes < 2 || (src_nodes & 1) || window < 2 || (window & 1) || window > src_nodes ||
      n < 1 || n >= (int64_t(1) << 32))
    return int(hipErrorInvalidValue);
  const unsigned blocks = unsigned((n + 255) / 256);
  hipLaunchKernelGGL(gather_probe_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const uint4*>(src), uint64_t(window) >> 1, uint64_t(src_nodes - window) / 2 + 1,
                     static_cast<uint4*>(dst), n, seed);
  return int(hipGetLastError());
}
