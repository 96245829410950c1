// End-engine block line kernel for gfx950.
//
// One thread computes one output node.  A workgroup of 256 threads takes one line, one
// index i along the line dimension and 256 consecutive cross-section positions, so the
// quadrant of the line dimension (rmin[d][i]) and everything derived from it -- the moves,
// the child index j, its local rotation -- are uniform across the wave, while the other
// three quadrants vary across lanes.  Consecutive lanes read consecutive child nodes
// (64 bytes each), so each move's gather is a contiguous run per wave.
//
// Per node (pentago/end/compute.cc:308-356):
//   wins0 = super_wins(side0), wins1 = super_wins(side1)
//   wins  = OR_m super_wins(side0 + move_m) | rmax(OR_m T_m(~child_m[1]))
//   nl    = rmax(OR_m T_m(~child_m[0]))
//   out   = ((wins & ~imm) | (wins0 & ~wins1), (nl & ~imm) | wins0),  imm = wins0 | wins1
// using rmax(A) | rmax(B) = rmax(A | B) and T_m = rearrange_g o shift_{y_m} with the
// global part g uniform per line, so the rearrangement and rmax run once per node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "device.h"
#include "end.h"
#include "end_device.h"
#include "super.h"

namespace pe {

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads)
void end_lines_kernel(const LineDev* __restrict__ lines, const uint8_t* __restrict__ refl,
                      const int64_t* __restrict__ prefix, int nlines, DevTables T) {
  // Find this workgroup's line: prefix[l] <= blockIdx.x < prefix[l+1]
  const int64_t task = blockIdx.x;
  int lo = 0, hi = nlines;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= task) lo = mid; else hi = mid;
  }
  const LineDev& L = lines[lo];
  const int chunks = (L.cross_count + kThreads - 1) / kThreads;
  const int64_t local = task - prefix[lo];
  const int i = L.i_begin + int(local / chunks);
  const int p = int(local % chunks) * kThreads + threadIdx.x;
  if (p >= L.cross_count) return;

  const int dim = L.dim, cd = L.child_dim;
  // cross-section coordinates (C-order over the three other dimensions)
  int full_base[4];
  {
    int rem = p;
    const int c2 = rem % L.cross_shape[2]; rem /= L.cross_shape[2];
    const int c1 = rem % L.cross_shape[1]; rem /= L.cross_shape[1];
    const int c[3] = {rem, c1, c2};
    for (int a = 0, k = 0; a < 4; a++) full_base[a] = a == dim ? 0 : c[k++];
  }
  const int turn = L.turn;
  uint64_t base_side0 = 0, base_side1 = 0;
  for (int a = 0; a < 4; a++) {
    if (a == dim) continue;
    const uint32_t sides = T.rmin_sides[L.rmin_off[a] + 8 * L.first_block[a] + full_base[a]];
    const uint64_t s0 = sides & 0x1ff, s1 = sides >> 16;
    base_side0 |= (turn ? s1 : s0) << (16 * a);
    base_side1 |= (turn ? s0 : s1) << (16 * a);
  }

  // output index (compute.cc:244-251, 267-269)
  int obs[4];
  for (int a = 0; a < 4; a++) obs[a] = L.out_shape[a];
  obs[dim] = 8;
  const int length = L.length;
  const int last_blk = i >> 3;
  if (last_blk == length - 1) obs[dim] = (L.rmin_size[dim] & 7) ? (L.rmin_size[dim] & 7) : 8;
  int ost[4]; strides4(obs, ost);
  const int64_t oblock = L.out_off ? L.out_off[last_blk] : int64_t(L.block_stride) * last_blk;
  const int64_t oidx = oblock + int64_t(dot4(ost, full_base)) + int64_t(ost[dim]) * (i & 7);

  // node sides
  const int qi = L.rmin_off[dim] + i;
  const uint16_t quad = T.rmin_flat[qi];
  const uint32_t qs = T.rmin_sides[qi];
  const uint32_t sq0 = turn ? qs >> 16 : qs & 0x1ff, sq1 = turn ? qs & 0x1ff : qs >> 16;
  const uint64_t side0 = base_side0 | uint64_t(sq0) << (16 * dim);
  const uint64_t side1 = base_side1 | uint64_t(sq1) << (16 * dim);
  const S4 wins0 = super_wins(side0, T.superwin);
  const S4 wins1 = super_wins(side1, T.superwin);
  S4 wins = s4_zero(), nl = s4_zero();
  const uint32_t moves = 511 & ~(sq0 | sq1);

  if (L.slice35) {
    // compute.cc:329-332, with rmax applied once to the OR of the per-move terms
    S4 a1 = s4_zero(), a0 = s4_zero();
    for (uint32_t m = moves; m; m &= m - 1) {
      const uint64_t move = uint64_t(m & (0u - m)) << (16 * dim);
      const S4 nw = super_wins(side0 | move, T.superwin);
      wins = s4_or(wins, nw);
      a1 = s4_or(a1, s4_andnot(nw, wins1));
      a0 = s4_or(a0, s4_or(nw, s4_not(wins1)));
    }
    wins = s4_or(wins, rmax(a1));
    nl = rmax(a0);
  } else {
    // child base indices (compute.cc:253-271)
    int ibs[4], fcb[4], rcb[4];
    for (int a = 0; a < 4; a++) { ibs[a] = obs[L.perm[a]]; fcb[a] = full_base[L.perm[a]]; }
    ibs[cd] = 8;
    for (int a = 0; a < 4; a++) {
      const int fcn = 8 * L.first_child_block[a] + fcb[a];
      rcb[a] = fcb[a] ^ int(a != cd && fcn < L.reflection_moves[a]);
    }
    int ist[4]; strides4(ibs, ist);
    const int input_stride = ist[cd];
    const int input_base = dot4(ist, rcb);
    ibs[cd] = (L.in_dim_size & 7) ? (L.in_dim_size & 7) : 8;
    int ist2[4]; strides4(ibs, ist2);
    const int last_input_base = dot4(ist2, rcb);
    const int line_moves = L.reflection_moves[cd];
    // local part of base_transform (compute.cc:274-281)
    const uint8_t* syms = refl + L.refl_base;
    int local_sym = 0;
    if (L.rs_off[4] > 0)
      for (int a = 0; a < 4; a++)
        if (a != cd) local_sym |= syms[L.rs_off[a] + fcb[a]];
    const int g = L.inv_global;
    const int base_local = local_add(L.inv_local, local_sym);
    const uint64_t* in = L.input;
    S4 acc1 = s4_zero(), acc0 = s4_zero();
    for (uint32_t m = moves; m; m &= m - 1) {
      const uint32_t mv = m & (0u - m);
      wins = s4_or(wins, super_wins(side0 | uint64_t(mv) << (16 * dim), T.superwin));
      const int new_q = quad + (1 + turn) * T.pack[mv];
      const int ir = T.rmin_inverse[new_q];
      const int j = ir >> 2;
      const int y = local_add(local_add(T.commute[g * 256 + ((ir & 3) << (2 * dim))], base_local),
                              syms[L.rs_off[cd] + j]);
      const int64_t cblock = L.in_off ? L.in_off[j >> 3] : int64_t(L.block_stride + 1) * (j >> 3);
      const int64_t cidx = cblock + int64_t((j >> 3) == L.child_length - 1 ? last_input_base : input_base) +
                           int64_t(input_stride) * ((j ^ int(j < line_moves)) & 7);
      const uint64_t* c = in + 8 * cidx;
      acc0 = s4_or(acc0, shift_local(s4_not(ld(c)), y));
      acc1 = s4_or(acc1, shift_local(s4_not(ld(c + 4)), y));
    }
    wins = s4_or(wins, rmax(rearrange(acc1, g)));
    nl = rmax(rearrange(acc0, g));
  }
  const S4 imm = s4_or(wins0, wins1);
  uint64_t* o = L.output + 8 * oidx;
  S4 r0 = s4_or(s4_andnot(wins, imm), s4_andnot(wins0, wins1));
  S4 r1 = s4_or(s4_andnot(nl, imm), wins0);
  if (L.accumulate) {   // OR-merge with earlier dimensions (pentago/end/block_store.cc:173-229)
    r0 = s4_or(r0, ld(o));
    r1 = s4_or(r1, ld(o + 4));
  }
  st(o, r0);
  st(o + 4, r1);
}

// Seeded child data: known ~ Bernoulli(58982/65536) per bit, win ~ Bernoulli(1/2) & known.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// One thread per (node, word), grid-stride: a dispatch counts at most 2^32 work-items, and a
// large slab has more (node, word) pairs than that.
__global__ void fill_seeded_kernel(uint64_t* __restrict__ nodes, int64_t count, uint64_t seed, int64_t first) {
  const uint64_t key = mix64(seed);
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < count * 4; t += stride) {
    const int64_t node = t >> 2;
    const int w = int(t & 3);
    const uint64_t ctr = (uint64_t(first + node) * 4 + uint64_t(w)) * 32;
    // p = 0.1110011001100110 in binary, consumed from the last digit to the first
    const uint32_t digits = 0xE666;   // bit 15 = first binary digit
    uint64_t known = 0;
    for (int k = 0; k < 16; k++) {
      const uint64_t r = mix64(key ^ (ctr + uint64_t(k)));
      known = (digits >> k & 1) ? (known | r) : (known & r);
    }
    const uint64_t win = mix64(key ^ (ctr + 16)) & known;
    nodes[8 * node + w] = win;
    nodes[8 * node + 4 + w] = known;
  }
}

}  // namespace

// Pack line descriptors, the workgroup prefix, reflection symmetries and block offset tables
// into one device allocation.  Pointers inside the blob are patched to device addresses.
pe_status plan_create(const std::vector<LineHost>& lines, EndPlan* plan) {
  DevTables T;
  std::string err;
  if (device_tables(&T, &err) != PE_OK) { set_error(err); return PE_ERR_DEVICE; }
  const int n = int(lines.size());
  std::vector<LineDev> devs(n);
  std::vector<int64_t> prefix(n + 1, 0);
  std::vector<uint8_t> refl;
  std::vector<int64_t> offs;
  std::vector<int64_t> in_at(n, -1), out_at(n, -1);
  int64_t out_nodes = 0;
  for (int k = 0; k < n; k++) {
    devs[k] = lines[k].dev;
    devs[k].refl_base = int32_t(refl.size());
    refl.insert(refl.end(), lines[k].refl_syms.begin(), lines[k].refl_syms.end());
    if (!lines[k].in_off.empty()) { in_at[k] = int64_t(offs.size()); offs.insert(offs.end(), lines[k].in_off.begin(), lines[k].in_off.end()); }
    if (!lines[k].out_off.empty()) { out_at[k] = int64_t(offs.size()); offs.insert(offs.end(), lines[k].out_off.begin(), lines[k].out_off.end()); }
    const int chunks = (devs[k].cross_count + kThreads - 1) / kThreads;
    const int span = devs[k].i_end - devs[k].i_begin;
    prefix[k + 1] = prefix[k] + int64_t(span) * chunks;
    out_nodes += int64_t(span) * devs[k].cross_count;
  }
  if (prefix[n] > 0x7fffffff) { set_error("too many line tasks in one plan"); return PE_ERR_CAPACITY; }
  refl.push_back(0);
  const size_t off_prefix = sizeof(LineDev) * n;
  const size_t off_offs = (off_prefix + sizeof(int64_t) * (n + 1) + 255) & ~size_t(255);
  const size_t off_refl = off_offs + sizeof(int64_t) * offs.size();
  const size_t bytes = off_refl + refl.size();
  void* d = nullptr;
  PE_HIP(hipMalloc(&d, bytes), "hipMalloc(line plan)");
  auto* base = static_cast<uint8_t*>(d);
  const int64_t* doffs = reinterpret_cast<const int64_t*>(base + off_offs);
  for (int k = 0; k < n; k++) {
    devs[k].in_off = in_at[k] >= 0 ? doffs + in_at[k] : nullptr;
    devs[k].out_off = out_at[k] >= 0 ? doffs + out_at[k] : nullptr;
  }
  std::vector<uint8_t> blob(bytes, 0);
  std::memcpy(blob.data(), devs.data(), off_prefix);
  std::memcpy(blob.data() + off_prefix, prefix.data(), sizeof(int64_t) * (n + 1));
  if (!offs.empty()) std::memcpy(blob.data() + off_offs, offs.data(), sizeof(int64_t) * offs.size());
  std::memcpy(blob.data() + off_refl, refl.data(), refl.size());
  if (hipMemcpy(d, blob.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    set_error("hipMemcpy(line plan) failed");
    return PE_ERR_DEVICE;
  }
  plan->blob = d;
  plan->bytes = bytes;
  plan->nlines = n;
  plan->tasks = prefix[n];
  plan->output_nodes = out_nodes;
  plan->off_prefix = off_prefix;
  plan->off_refl = off_refl;
  PE_HIP(hipGetDevice(&plan->device), "hipGetDevice");
  return PE_OK;
}

pe_status plan_run(const EndPlan& plan, hipStream_t stream) {
  if (!plan.nlines || !plan.tasks) return PE_OK;
  DevTables T;
  std::string err;
  if (device_tables(&T, &err) != PE_OK) { set_error(err); return PE_ERR_DEVICE; }
  auto* base = static_cast<uint8_t*>(plan.blob);
  hipLaunchKernelGGL(end_lines_kernel, dim3(unsigned(plan.tasks)), dim3(kThreads), 0, stream,
                     reinterpret_cast<const LineDev*>(base), base + plan.off_refl,
                     reinterpret_cast<const int64_t*>(base + plan.off_prefix), plan.nlines, T);
  PE_HIP(hipGetLastError(), "end_lines_kernel launch");
  return PE_OK;
}

void plan_destroy(EndPlan* plan) {
  if (plan->blob) (void)hipFree(plan->blob);
  plan->blob = nullptr;
  plan->nlines = 0;
}

pe_status launch_fill_seeded(void* d_nodes, int64_t nodes, uint64_t seed, int64_t first, hipStream_t stream) {
  if (nodes <= 0) return PE_OK;
  const int64_t threads = nodes * 4;
  const int64_t blocks = std::min<int64_t>((threads + 255) / 256, int64_t(1) << 20);   // <= 2^28 work-items
  hipLaunchKernelGGL(fill_seeded_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream,
                     static_cast<uint64_t*>(d_nodes), nodes, seed, first);
  PE_HIP(hipGetLastError(), "fill_seeded_kernel launch");
  return PE_OK;
}

}  // namespace pe
