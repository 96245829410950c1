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
#include <cstring>
#include <vector>

#include "device.h"
#include "mid.h"
#include "super.h"

namespace pe {

struct MidBoardDev {
  MidInfo I;
  Set0Info* set0;
  Wins1* wins1;
  uint16_t* cs1ps;
  uint64_t* work;           // 4 u64 per entry: win[2], notlose[2]
  pe_halfsupers* results;   // 19 entries
};

namespace {

__device__ __forceinline__ int dchoose(const int32_t* C, int n, int k) {
  return (n < 0 || k < 0 || k > n) ? 0 : C[n * 32 + k];
}

// k-subset of [0,n) with the given index, 5 bits per entry, large-entry-major order
// (pentago/mid/subsets.h:33-48)
__device__ uint64_t dsubset(const int32_t* C, int n, int k, int index) {
  uint64_t set = 0;
  for (int i = k - 1; i >= 0; i--) {
    int j = i, base = 0;
    for (; j < n; j++) {
      const int next = base + dchoose(C, j, i);
      if (index < next) break;
      base = next;
    }
    set |= uint64_t(j) << (5 * i);
    index -= base;
  }
  return set;
}

__device__ __forceinline__ uint64_t dside(const MidInfo& I, int k, uint64_t set) {
  uint64_t s = 0;
  for (int i = 0; i < k; i++) s |= uint64_t(1) << I.empty[set >> (5 * i) & 0x1f];
  return s;
}

__device__ __forceinline__ int level_of(const int32_t* off, int spots, int64_t idx) {
  int n = 0;
  while (n < spots && off[n + 1] <= idx) n++;
  return n;
}

// One table entry: index over [set0 | wins1 | cs1ps] of one board.
__device__ void mid_prep_entry(const MidBoardDev& B, int64_t idx, const DevTables& T) {
  const MidInfo& I = B.I;
  const int spots = I.spots;
  const int32_t* C = T.choose;
  const int64_t n_set0 = I.sets0_off[spots + 1], n_wins1 = I.wins1_off[spots + 1];
  if (idx < n_set0) {
    // set0_info (internal.h:174-259)
    const int n = level_of(I.sets0_off, spots, idx);
    const int s0 = int(idx - I.sets0_off[n]);
    const int k0 = n >> 1, k1 = n - k0;
    const int parity = (n + I.ply) & 1;
    const uint64_t set0 = dsubset(C, spots, k0, s0);
    const uint64_t side0 = I.side[(spots + n) & 1] | dside(I, k0, set0);
    Set0Info out;
    const H2 w0 = halfsuper_wins(side0, parity, T.superwin);
    out.wins0[0] = w0.w[0]; out.wins0[1] = w0.w[1];
    const uint64_t free_ = 0x01ff01ff01ff01ffull & ~side0;
    for (int i = 0, k = 0; i < spots; i++)
      if (free_ >> I.empty[i] & 1) out.empty1[k++] = uint8_t(i);
    for (int i = 0; i < spots - k0; i++) {
      const int j = out.empty1[i] - i;
      uint16_t c = uint16_t(dchoose(C, out.empty1[i], j + 1));
      for (int a = 0; a < k0; a++) c = uint16_t(c + dchoose(C, int(set0 >> (5 * a) & 0x1f), a + (a >= j) + 1));
      out.child_s0s[i] = c;
    }
    for (int a = 0; a < k1; a++)
      for (int q = 0; q < spots - k0; q++) {
        uint16_t offset = uint16_t(a ? 0 : s0);
        for (int i = out.empty1[q] - q; i < k0; i++) {
          const int v = int(set0 >> (5 * i) & 0x1f);
          if (v > a) offset = uint16_t(offset - dchoose(C, v - a - 1, i));
        }
        out.offset0[a * (spots - k0) + q] = offset;
      }
    B.set0[idx] = out;
  } else if (idx < n_set0 + n_wins1) {
    // wins1 (internal.h:122-131)
    const int64_t e = idx - n_set0;
    const int n = level_of(I.wins1_off, spots, e);
    const int s1 = int(e - I.wins1_off[n]);
    const int k0 = n >> 1, k1 = n - k0;
    (void)k0;
    const int parity = (n + I.ply) & 1;
    const uint64_t side1 = I.side[(spots + n + 1) & 1] | dside(I, k1, dsubset(C, spots, k1, s1));
    const S4 w = super_wins(side1, T.superwin);
    const H2 after = split_half(w, parity), before = split_half(w, !parity);
    Wins1 out;
    out.after[0] = after.w[0]; out.after[1] = after.w[1];
    out.before[0] = before.w[0]; out.before[1] = before.w[1];
    B.wins1[e] = out;
  } else {
    // cs1ps (internal.h:133-148): remap s1p when a stone of the mover lands at empty1[j]
    const int64_t e = idx - n_set0 - n_wins1;
    const int n = level_of(I.cs1ps_off, spots, e);
    const int index = int(e - I.cs1ps_off[n]);
    const int k0 = n >> 1, k1 = n - k0;
    const int s1p = index / (spots - k0);
    const int j = index - s1p * (spots - k0);
    const uint64_t set1p = dsubset(C, spots - k0, k1, s1p);
    uint16_t c = uint16_t(s1p);
    for (int a = 0; a < k1; a++) {
      const int v = int(set1p >> (5 * a) & 0x1f);
      if (j < v) c = uint16_t(c - dchoose(C, v - 1, a));
    }
    B.cs1ps[e] = c;
  }
}

// One position (internal.h:358-408), both aggressive passes at once.
__device__ void mid_position(const MidBoardDev& B, int n, int64_t pos, const DevTables& T) {
  const MidInfo& I = B.I;
  const int spots = I.spots;
  const int k0 = n >> 1, k1 = n - k0;
  const int32_t* C = T.choose;
  const int nsets1p = dchoose(C, spots - k0, k1);
  const int s0 = int(pos / nsets1p);
  const int s1p = int(pos - int64_t(s0) * nsets1p);
  const Set0Info& I0 = B.set0[I.sets0_off[n] + s0];
  const uint64_t set1p = dsubset(C, spots - k0, k1, s1p);
  uint32_t filled1p = 0;
  uint16_t s1 = 0, s0p = 0;
  for (int i = 0; i < k1; i++) {
    const int q = int(set1p >> (5 * i) & 0x1f);
    filled1p |= 1u << q;
    s1 = uint16_t(s1 + dchoose(C, I0.empty1[q], i + 1));
    s0p = uint16_t(s0p + I0.offset0[i * (spots - k0) + q]);
  }
  uint64_t win0 = 0, win1 = 0, nl0 = 0, nl1 = 0;
  if (n == spots) nl0 = nl1 = ~uint64_t(0);
  const MidGrab gin = I.spaces[n + 1];
  const uint16_t* cs = B.cs1ps + I.cs1ps_off[n] + s1p * (spots - k0);
  uint32_t unmoved = ~filled1p;
  for (int m = 0; m < spots - n; m++) {
    const int i = __builtin_ctz(unmoved);
    unmoved &= unmoved - 1;
    const uint64_t* c = B.work + 4 * (int64_t(gin.lo) + int64_t(I0.child_s0s[i]) * gin.ny + cs[i]);
    const ulonglong2 a = reinterpret_cast<const ulonglong2*>(c)[0];
    const ulonglong2 b = reinterpret_cast<const ulonglong2*>(c)[1];
    win0 |= a.x; win1 |= a.y; nl0 |= b.x; nl1 |= b.y;
  }
  const Wins1 w1 = B.wins1[I.wins1_off[n] + s1];
  const uint64_t ip0 = ~(I0.wins0[0] | w1.after[0]), ip1 = ~(I0.wins0[1] | w1.after[1]);
  win0 = (ip0 & win0) | (I0.wins0[0] & ~w1.after[0]);
  win1 = (ip1 & win1) | (I0.wins0[1] & ~w1.after[1]);
  nl0 = (ip0 & nl0) | I0.wins0[0];
  nl1 = (ip1 & nl1) | I0.wins0[1];
  if (n <= 1) {
    pe_halfsupers& r = B.results[n + s1p];
    r.win[0] = win0; r.win[1] = win1; r.notlose[0] = nl0; r.notlose[1] = nl1;
  }
  H2 nw; nw.w[0] = ~nl0; nw.w[1] = ~nl1;     // ~(win, notlose) = (~notlose, ~win)
  H2 nn; nn.w[0] = ~win0; nn.w[1] = ~win1;
  const H2 sw = hrmax(nw), sn = hrmax(nn);
  const MidGrab gout = I.spaces[n];
  uint64_t* o = B.work + 4 * (int64_t(gout.lo) + int64_t(s1) * gout.ny + s0p);
  reinterpret_cast<ulonglong2*>(o)[0] = make_ulonglong2(sw.w[0] | w1.before[0], sw.w[1] | w1.before[1]);
  reinterpret_cast<ulonglong2*>(o)[1] = make_ulonglong2(sn.w[0] | w1.before[0], sn.w[1] | w1.before[1]);
}

__device__ __forceinline__ int find_board(const int64_t* prefix, int nb, int64_t t) {
  int lo = 0, hi = nb;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void mid_prep_kernel(const MidBoardDev* __restrict__ boards,
                                                       const int64_t* __restrict__ prefix, int nb,
                                                       int64_t total, DevTables T) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int b = find_board(prefix, nb, t);
  mid_prep_entry(boards[b], t - prefix[b], T);
}

// step: level n = spots_b - step for every board with spots_b >= step
__global__ __launch_bounds__(256) void mid_level_kernel(const MidBoardDev* __restrict__ boards,
                                                        const int64_t* __restrict__ prefix, int nb,
                                                        int64_t total, int step, DevTables T) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int b = find_board(prefix, nb, t);
  const MidBoardDev& B = boards[b];
  mid_position(B, B.I.spots - step, t - prefix[b], T);
}

}  // namespace

// Allocation of one board's device scratch inside a caller-provided arena.
static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

size_t mid_board_bytes(const MidInfo& I) {
  const MidSizes s = mid_sizes(I);
  return align256(sizeof(Set0Info) * s.set0) + align256(sizeof(Wins1) * s.wins1) +
         align256(sizeof(uint16_t) * s.cs1ps) + align256(32 * s.work) + align256(sizeof(pe_halfsupers) * 19);
}

static void carve(MidBoardDev* B, uint8_t* base) {
  const MidSizes s = mid_sizes(B->I);
  size_t off = 0;
  B->set0 = reinterpret_cast<Set0Info*>(base + off); off += align256(sizeof(Set0Info) * s.set0);
  B->wins1 = reinterpret_cast<Wins1*>(base + off); off += align256(sizeof(Wins1) * s.wins1);
  B->cs1ps = reinterpret_cast<uint16_t*>(base + off); off += align256(sizeof(uint16_t) * s.cs1ps);
  B->work = reinterpret_cast<uint64_t*>(base + off); off += align256(32 * s.work);
  B->results = reinterpret_cast<pe_halfsupers*>(base + off);
}

// Level-synchronous solve of many boards whose scratch lives in `arena` (device memory of
// at least the sum of mid_board_bytes).  Results land in each board's scratch; the caller
// copies them out (offsets returned in result_offsets).
pe_status mid_solve_levels(const std::vector<MidInfo>& infos, uint8_t* arena, std::vector<size_t>* result_offsets,
                           hipStream_t stream) {
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
    carve(&boards[b], arena + off);
    (*result_offsets)[b] = size_t(reinterpret_cast<uint8_t*>(boards[b].results) - arena);
    off += mid_board_bytes(infos[b]);
    max_spots = std::max(max_spots, infos[b].spots);
  }
  // prefix sums: prep entries, then positions per step
  const int steps = max_spots + 1;
  std::vector<int64_t> prefix(size_t(nb + 1) * (steps + 1), 0);
  for (int b = 0; b < nb; b++) {
    const MidSizes s = mid_sizes(infos[b]);
    prefix[b + 1] = prefix[b] + s.set0 + s.wins1 + s.cs1ps;
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
  PE_HIP(hipMallocAsync(&d, bytes, stream), "hipMallocAsync(mid descriptors)");
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
    hipLaunchKernelGGL(mid_level_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, stream, dboards,
                       dprefix + size_t(nb + 1) * (t + 1), nb, total, t, T);
    PE_HIP(hipGetLastError(), "mid_level_kernel launch");
  }
  PE_HIP(hipFreeAsync(d, stream), "hipFreeAsync(mid descriptors)");
  return PE_OK;
}

}  // namespace pe
