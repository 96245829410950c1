// Midsolve kernels for gfx950.
//
// Per board: one table-building pass (per-s0 tables, the other side's wins, child index
// remaps; pentago/mid/internal.h:122-259) and then one sweep per level n = spots..0 in
// which one thread evaluates one position (s0, s1p) exactly as internal.h:358-408 does,
// carrying (win, notlose) together.  Consecutive threads share s0 and walk s1p, so the
// children they OR (rows child_s0s[i] of the previous level) are read as nearly contiguous
// runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mid_device.h"
#include "mid_queue.h"

namespace pe {

namespace {

__device__ __forceinline__ int find_board(const int64_t* prefix, int nb, int64_t t) {
  int lo = 0, hi = nb;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

// The binomial table (4 KB) is staged in LDS first: the subset unranking and the offset
// loops below read it in chains of dependent lookups.
__global__ __launch_bounds__(256) void mid_prep_kernel(const MidBoardDev* __restrict__ boards,
                                                       const int64_t* __restrict__ prefix, int nb,
                                                       int64_t total, DevTables T) {
  __shared__ int32_t sC[32 * 32];
  for (int i = threadIdx.x; i < 32 * 32; i += blockDim.x) sC[i] = T.choose[i];
  __syncthreads();
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int b = find_board(prefix, nb, t);
  mid_prep_entry(boards[b], t - prefix[b], T, sC);
}

// Level launches of one board are spread over the 8 XCDs (workgroup i runs on XCD i % 8) by
// bands of s1p: XCD x sweeps every s0 for bands x, x + 8, ... in turn.  A position's
// children sit in the rows of the next level at columns that follow its s1p, so an XCD's L2
// then holds its band of each row instead of every XCD touching every column.
__device__ __forceinline__ bool banded_position(uint32_t wg, uint32_t tid, uint32_t band, uint32_t nsets1p,
                                                uint32_t nsets0, int64_t* pos) {
  const int64_t q = int64_t(wg >> 3) * 256 + tid;
  const int64_t per_band = int64_t(nsets0) * band;
  const int64_t bi = q / per_band;
  const int64_t r = q - bi * per_band;
  const int64_t s0 = r / band;
  const int64_t s1p = (int64_t(wg & 7) + 8 * bi) * band + (r - s0 * band);
  if (s1p >= nsets1p) return false;   // a band past the last, or the last band's tail
  *pos = s0 * nsets1p + s1p;
  return true;
}

// step: level n = spots_b - step for every board with spots_b >= step
// band > 0 (one board only): banded order, see banded_position
__global__ __launch_bounds__(256) void mid_level_kernel(const MidBoardDev* __restrict__ boards,
                                                        const int64_t* __restrict__ prefix, int nb,
                                                        int64_t total, int step, LevelGrid g, DevTables T) {
  if (g.band) {
    int64_t pos;
    if (!banded_position(blockIdx.x, threadIdx.x, g.band, g.nsets1p, g.nsets0, &pos)) return;
    mid_position(boards[0], boards[0].I.spots - step, pos, boards[0].work, T);
    return;
  }
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int b = find_board(prefix, nb, t);
  const MidBoardDev& B = boards[b];
  mid_position(B, B.I.spots - step, t - prefix[b], B.work, T);
}

__global__ __launch_bounds__(256) void mid_level1_kernel(const MidBoardDev* __restrict__ board, int n,
                                                         int64_t positions, LevelGrid g, DevTables T) {
  int64_t t;
  if (g.band) {
    if (!banded_position(blockIdx.x, threadIdx.x, g.band, g.nsets1p, g.nsets0, &t)) return;
  } else {
    t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= positions) return;
  }
  mid_position(*board, n, t, board->work, T);
}

}  // namespace

LevelGrid level_grid(const MidInfo& I, int n) {
  const int64_t positions = I.positions[n];
  const uint32_t n1 = uint32_t(I.sets1p_off[n + 1] - I.sets1p_off[n]);
  LevelGrid g;
  g.blocks = unsigned((positions + 255) / 256);
  g.band = 0;
  g.nsets1p = n1;
  g.nsets0 = n1 ? uint32_t(positions / n1) : 0;
  const uint32_t band = std::min<uint32_t>(kLevelBand, (n1 + 7) / 8);
  if (kLevelBand == 0 || band < 16 || positions < (int64_t(1) << 16)) return g;   // small levels: linear order
  const uint32_t nbands = (n1 + band - 1) / band;
  const int64_t per_xcd = int64_t((nbands + 7) / 8) * g.nsets0 * band;
  g.band = band;
  g.blocks = unsigned(8 * ((per_xcd + 255) / 256));
  return g;
}

// Allocation of one board's device scratch inside a caller-provided arena.
static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

size_t mid_table_bytes(const MidInfo& I) {
  const MidSizes s = mid_sizes(I);
  return align256(sizeof(Set0Info) * s.set0) + align256(sizeof(Wins1) * s.wins1) +
         align256(sizeof(uint16_t) * s.cs1ps) + align256(sizeof(uint32_t) * s.set1p);
}
size_t mid_work_bytes(const MidInfo& I) { return align256(32 * size_t(mid_sizes(I).work)); }
size_t mid_board_bytes(const MidInfo& I) {
  return mid_table_bytes(I) + mid_work_bytes(I) + align256(sizeof(pe_halfsupers) * 19);
}

void mid_carve_tables(MidBoardDev* B, uint8_t* base) {
  const MidSizes s = mid_sizes(B->I);
  size_t off = 0;
  B->set0 = reinterpret_cast<Set0Info*>(base + off); off += align256(sizeof(Set0Info) * s.set0);
  B->wins1 = reinterpret_cast<Wins1*>(base + off); off += align256(sizeof(Wins1) * s.wins1);
  B->cs1ps = reinterpret_cast<uint16_t*>(base + off); off += align256(sizeof(uint16_t) * s.cs1ps);
  B->set1p_bits = reinterpret_cast<uint32_t*>(base + off);
}

void mid_carve(MidBoardDev* B, uint8_t* base) {
  mid_carve_tables(B, base);
  const size_t t = mid_table_bytes(B->I);
  B->work = reinterpret_cast<uint64_t*>(base + t);
  B->results = reinterpret_cast<pe_halfsupers*>(base + t + mid_work_bytes(B->I));
}

pe_status mid_launch_level(const MidBoardDev* board, const MidInfo& info, int n, int64_t positions,
                           const DevTables& T, hipStream_t stream) {
  if (positions <= 0) return PE_OK;
  const LevelGrid g = level_grid(info, n);
  hipLaunchKernelGGL(mid_level1_kernel, dim3(g.blocks), dim3(256), 0, stream, board, n, positions, g, T);
  PE_HIP(hipGetLastError(), "mid_level1_kernel launch");
  return PE_OK;
}

pe_status mid_launch_prep(const MidBoardDev* boards, const int64_t* prefix, int nb, int64_t total,
                          const DevTables& T, hipStream_t stream) {
  if (total <= 0) return PE_OK;
  hipLaunchKernelGGL(mid_prep_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, stream, boards, prefix, nb,
                     total, T);
  PE_HIP(hipGetLastError(), "mid_prep_kernel launch");
  return PE_OK;
}

// Level-synchronous solve of many boards whose scratch lives in `arena` (device memory of
// at least the sum of mid_board_bytes).  Results land in each board's scratch; the caller
// copies them out (offsets returned in result_offsets).
pe_status mid_solve_levels(const std::vector<MidInfo>& infos, uint8_t* arena, std::vector<size_t>* result_offsets,
                           hipStream_t stream, DeviceBuffer* desc_cache) {
  DevTables T;
  std::string err;
  if (device_tables(&T, &err) != PE_OK) { set_error(err); return PE_ERR_DEVICE; }
  const int nb = int(infos.size());
  if (!nb) return PE_OK;
  std::vector<MidBoardDev> boards(nb);
  result_offsets->resize(nb);
  size_t off = 0;
  int max_spots = 0;
  for (int b = 0; b < nb; b++) {
    boards[b].I = infos[b];
    mid_carve(&boards[b], arena + off);
    (*result_offsets)[b] = size_t(reinterpret_cast<uint8_t*>(boards[b].results) - arena);
    off += mid_board_bytes(infos[b]);
    max_spots = std::max(max_spots, infos[b].spots);
  }
  // prefix sums: prep entries, then positions per step
  const int steps = max_spots + 1;
  std::vector<int64_t> prefix(size_t(nb + 1) * (steps + 1), 0);
  for (int b = 0; b < nb; b++) {
    const MidSizes s = mid_sizes(infos[b]);
    prefix[b + 1] = prefix[b] + s.set0 + s.wins1 + s.set1p;
  }
  for (int t = 0; t < steps; t++) {
    int64_t* p = prefix.data() + size_t(nb + 1) * (t + 1);
    for (int b = 0; b < nb; b++) {
      const int n = infos[b].spots - t;
      p[b + 1] = p[b] + (n >= 0 ? infos[b].positions[n] : 0);
    }
  }
  const size_t desc_bytes = sizeof(MidBoardDev) * nb;
  const size_t bytes = desc_bytes + sizeof(int64_t) * prefix.size();
  std::vector<uint8_t> blob(bytes);
  std::memcpy(blob.data(), boards.data(), desc_bytes);
  std::memcpy(blob.data() + desc_bytes, prefix.data(), sizeof(int64_t) * prefix.size());
  void* d = nullptr;
  if (desc_cache) {
    PE_HIP(desc_cache->reserve(bytes), "hipMalloc(mid descriptors)");
    d = desc_cache->ptr;
  } else {
    PE_HIP(hipMallocAsync(&d, bytes, stream), "hipMallocAsync(mid descriptors)");
  }
  PE_HIP(hipMemcpyAsync(d, blob.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync(mid descriptors)");
  const auto* dboards = static_cast<const MidBoardDev*>(d);
  const auto* dprefix = reinterpret_cast<const int64_t*>(static_cast<uint8_t*>(d) + desc_bytes);
  {
    const int64_t total = prefix[nb];
    if (total > 0)
      hipLaunchKernelGGL(mid_prep_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, stream, dboards,
                         dprefix, nb, total, T);
    PE_HIP(hipGetLastError(), "mid_prep_kernel launch");
  }
  for (int t = 0; t < steps; t++) {
    const int64_t* p = prefix.data() + size_t(nb + 1) * (t + 1);
    const int64_t total = p[nb];
    if (total <= 0) continue;
    LevelGrid g;
    g.blocks = unsigned((total + 255) / 256);
    g.band = g.nsets1p = g.nsets0 = 0;
    if (nb == 1) g = level_grid(infos[0], infos[0].spots - t);
    hipLaunchKernelGGL(mid_level_kernel, dim3(g.blocks), dim3(256), 0, stream, dboards,
                       dprefix + size_t(nb + 1) * (t + 1), nb, total, t, g, T);
    PE_HIP(hipGetLastError(), "mid_level_kernel launch");
  }
  if (!desc_cache) PE_HIP(hipFreeAsync(d, stream), "hipFreeAsync(mid descriptors)");
  return PE_OK;
}

// ---- allocation-free single-board solve ----
// For callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc); the reference has no counterpart (its
// caller passes a workspace it owns).  The launch descriptors, the board's tables, its
// workspace and its results live in `scratch`, a fixed array that every device the library
// runs on holds until the process ends.  Host code never uses the array's name as an
// address: each device's address comes from hipGetSymbolAddress, once per device.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu[kScratchDevices];      // one solve at a time per device in this process
uint8_t* g_scratch_ptr[kScratchDevices];       // device address of `scratch` on each device

// The descriptors of one board's launches: the board record, then the prefix sums the prep
// and level kernels read (mid_solve_levels' layout for one board: [0, entries] per launch).
struct NoAllocDesc {
  MidBoardDev board;
  int64_t prefix[2 * 21];
};

}  // namespace

int64_t mid_noalloc_scratch_bytes() { return int64_t(kScratchBytes); }

pe_status mid_solve_noalloc(const pe_high_board& hb, pe_halfsupers* results, hipStream_t stream) {
  MidInfo I;
  if (!make_mid_info(hb, &I)) { set_error("not a valid board with at least 18 stones"); return PE_ERR_INVALID; }
  DevTables T;
  if (!device_tables_ready(&T)) {
    set_error("the device's tables are not uploaded: call pe_init on this device first (this entry point allocates nothing)");
    return PE_ERR_INVALID;
  }
  const size_t desc_bytes = align256(sizeof(NoAllocDesc));
  const size_t need = desc_bytes + mid_board_bytes(I);
  if (need > kScratchBytes) {
    set_error("board with " + std::to_string(I.spots) + " empty cells needs " + std::to_string(need) +
              " bytes of device scratch; the allocation-free solve has " + std::to_string(kScratchBytes) +
              " (boards with at most 15 empty cells, i.e. at least 21 stones)");
    return PE_ERR_CAPACITY;
  }
  int dev = 0;
  PE_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kScratchDevices) { set_error("device index out of range"); return PE_ERR_DEVICE; }
  std::lock_guard<std::mutex> lock(g_scratch_mu[dev]);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<uint8_t*>(p);
  }
  uint8_t* const base = g_scratch_ptr[dev];
  // every table, workspace and result entry the kernels read is written by them first
  // (prep, then level n = spots .. 0), so nothing in the array needs clearing between calls
  NoAllocDesc d;
  std::memset(&d, 0, sizeof d);
  d.board.I = I;
  mid_carve(&d.board, base + desc_bytes);
  const MidSizes sz = mid_sizes(I);
  const int steps = I.spots + 1;
  d.prefix[1] = sz.set0 + sz.wins1 + sz.set1p;
  for (int t = 0; t < steps; t++) d.prefix[2 * (t + 1) + 1] = I.positions[I.spots - t];
  // every way out waits for the work queued on the stream: the descriptors are read from
  // this frame, and the next call (on any stream) reuses the array
  struct Drain {
    hipStream_t s;
    ~Drain() { (void)hipStreamSynchronize(s); }
  } drain{stream};
  PE_HIP(hipMemcpyAsync(base, &d, sizeof d, hipMemcpyHostToDevice, stream), "hipMemcpyAsync(descriptors)");
  const auto* dboard = reinterpret_cast<const MidBoardDev*>(base);
  const auto* dprefix = reinterpret_cast<const int64_t*>(base + offsetof(NoAllocDesc, prefix));
  if (d.prefix[1] > 0) {
    hipLaunchKernelGGL(mid_prep_kernel, dim3(unsigned((d.prefix[1] + 255) / 256)), dim3(256), 0, stream, dboard,
                       dprefix, 1, d.prefix[1], T);
    PE_HIP(hipGetLastError(), "mid_prep_kernel launch");
  }
  for (int t = 0; t < steps; t++) {
    const int64_t total = d.prefix[2 * (t + 1) + 1];
    if (total <= 0) continue;
    const LevelGrid g = level_grid(I, I.spots - t);
    hipLaunchKernelGGL(mid_level_kernel, dim3(g.blocks), dim3(256), 0, stream, dboard, dprefix + 2 * (t + 1), 1, total,
                       t, g, T);
    PE_HIP(hipGetLastError(), "mid_level_kernel launch");
  }
  PE_HIP(hipMemcpyAsync(results, d.board.results, sizeof(pe_halfsupers) * 19, hipMemcpyDeviceToHost, stream),
         "hipMemcpyAsync(results)");
  PE_HIP(hipStreamSynchronize(stream), "allocation-free midsolve");
  for (int i = I.spots + 1; i < 19; i++) std::memset(&results[i], 0, sizeof(pe_halfsupers));
  return PE_OK;
}

}  // namespace pe
