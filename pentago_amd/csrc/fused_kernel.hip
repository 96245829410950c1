// Fused end-engine block kernel for gfx950: every node of an output block gets all of its
// open dimensions in one pass and is written once.
//
// For a node with sides (side0 = mover, side1) the reference computes, per open dimension d
// (pentago/end/compute.cc:308-356, one line per dimension) and then ORs the four results
// (pentago/end/block_store.cc:173-229):
//   r_d = ((wins_d & ~imm) | (wins0 & ~wins1), (nl_d & ~imm) | wins0),  imm = wins0 | wins1
//   wins_d = OR_{m in d} super_wins(side0 + m) | OR_m rmax(T_m(~child_m[1]))
//   nl_d   = OR_m rmax(T_m(~child_m[0]))
// Exact regrouping used here:
//   * OR_d r_d = ((OR_d wins_d) & ~imm) | (wins0 & ~wins1), likewise for nl: wins0, wins1
//     and the final combination happen once per node.
//   * super_wins(side) = OR over eight pattern groups G of reduce_G(AND_{q in G} field_q,G),
//     and a move in quadrant d changes only d's field.  Groups without d contribute bits of
//     wins0, which ~imm removes anyway, so OR_m super_wins(side0 + m) & ~imm equals
//     OR_G reduce_G(OR_{d in G} Fm_d,G & AND_{q in G, q != d} F0_q,G) & ~imm, with
//     Fm = OR over the moves of the moved quadrant's field (AND distributes over OR).
//     The device's field pool holds F0, F1 and Fm per (quadrant, rotation-minimal index).
//   * rmax(A) | rmax(B) = rmax(A | B) and T_m = R_g o shift_{y_m}: the children are shifted
//     and ORed per dimension, rearranged once per dimension, and rmax runs once per node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <tuple>
#include <vector>

#include "device.h"
#include "end.h"
#include "end_device.h"
#include "super.h"
#include "tables.h"
#include "transforms.h"

namespace pe {

pe_status plan_create(const std::vector<LineHost>& lines, EndPlan* plan);

namespace {

enum { FH = 0, FV = 1, FLO = 2, FHI = 3, FAS = 4 };

// Field words of one node: word k of field f of table t (0: F0, 1: F1, 2: Fm), quadrant q.
// Quadrants 0 and 1 are read from global memory (the same entry for the whole wave),
// quadrants 2 and 3 from the block's copy of its eight entries in LDS.
struct NodeFields {
  const uint64_t* glob[2];
  const uint64_t* lds[2];
  __device__ __forceinline__ uint64_t w(int q, int k, int t, int f) const {
    return q < 2 ? glob[q][15 * k + 5 * t + f] : lds[q - 2][15 * k + 5 * t + f];
  }
};

// Word k of a pattern-group term for the three quantities: x[0] mover's side, x[1] other
// side, x[2] the mover's side with one quadrant's field replaced by its OR over that
// quadrant's moves (all moves at once).
__device__ __forceinline__ void g2(const NodeFields& F, int k, int a, int fa, int b, int fb, uint64_t x[3]) {
  const uint64_t a0 = F.w(a, k, 0, fa), b0 = F.w(b, k, 0, fb);
  x[0] = a0 & b0;
  x[1] = F.w(a, k, 1, fa) & F.w(b, k, 1, fb);
  x[2] = (F.w(a, k, 2, fa) & b0) | (a0 & F.w(b, k, 2, fb));
}
// The same terms with three-input logic operations (at least one of the quadrants is lane-
// varying; products of two wave-uniform quadrants stay scalar).  g2_acc ORs the term into
// acc; g2_term / g3_term return it.
__device__ __forceinline__ void g2_acc(const NodeFields& F, int k, int a, int fa, int b, int fb, uint64_t acc[3]) {
  const uint64_t a0 = F.w(a, k, 0, fa), b0 = F.w(b, k, 0, fb);
  acc[0] = lop3<kOrAnd>(acc[0], a0, b0);
  acc[1] = lop3<kOrAnd>(acc[1], F.w(a, k, 1, fa), F.w(b, k, 1, fb));
  acc[2] = lop3<kOrAnd>(lop3<kOrAnd>(acc[2], F.w(a, k, 2, fa), b0), a0, F.w(b, k, 2, fb));
}
__device__ __forceinline__ void g2_term(const NodeFields& F, int k, int a, int fa, int b, int fb, uint64_t x[3]) {
  const uint64_t a0 = F.w(a, k, 0, fa), b0 = F.w(b, k, 0, fb);
  x[0] = a0 & b0;
  x[1] = F.w(a, k, 1, fa) & F.w(b, k, 1, fb);
  x[2] = lop3<kOrAnd>(F.w(a, k, 2, fa) & b0, a0, F.w(b, k, 2, fb));
}
// a2 b0 c0 | a0 b2 c0 | a0 b0 c2, with a (and b when b < 2) wave-uniform
__device__ __forceinline__ void g3_term(const NodeFields& F, int k, int a, int fa, int b, int fb, int c, int fc,
                                        uint64_t x[3]) {
  const uint64_t a0 = F.w(a, k, 0, fa), b0 = F.w(b, k, 0, fb), c0 = F.w(c, k, 0, fc);
  if (b < 2) {
    const uint64_t ab0 = a0 & b0, ab2 = (F.w(a, k, 2, fa) & b0) | (a0 & F.w(b, k, 2, fb));
    x[0] = ab0 & c0;
    x[1] = (F.w(a, k, 1, fa) & F.w(b, k, 1, fb)) & F.w(c, k, 1, fc);
    x[2] = lop3<kOrAnd>(ab2 & c0, ab0, F.w(c, k, 2, fc));
  } else {
    const uint64_t bc0 = b0 & c0;
    x[0] = a0 & bc0;
    x[1] = lop3<kAndAnd>(F.w(a, k, 1, fa), F.w(b, k, 1, fb), F.w(c, k, 1, fc));
    const uint64_t inner = lop3<kOrAnd>(F.w(b, k, 2, fb) & c0, b0, F.w(c, k, 2, fc));
    x[2] = lop3<kOrAnd>(F.w(a, k, 2, fa) & bc0, a0, inner);
  }
}

// Compiler-only memory fence: keeps the field loads of one word from being hoisted above
// the previous word's work, which would otherwise keep ~200 loaded values live.
#define PE_FENCE() asm volatile("" ::: "memory")

constexpr int kThreads = 256;
constexpr int kParts = 4096 / kThreads;
constexpr int kChunk = 1 << 19;   // blocks per grid launch: 2^31 work-items
// Occupancy: each node's running OR over its dimensions (and then its output words) lives in
// LDS (16 KB per workgroup, 24 KB with the field entries), so a wave needs 66 VGPRs and six
// waves share a SIMD (the LDS limit); children are loaded one move at a time.  Measured on
// the headline launch against keeping them in registers (123 VGPRs, four waves, moves loaded
// three at a time): 157.4 vs 168.4 ms; two moves at a time 162.5.  The launch bound asks for
// at least five waves' worth of registers.
constexpr int kWaves = 5;

// One output block of 4096 nodes, computed by 16 workgroups (`part` = which 256 nodes).
// S35: blocks of 35-stone sections, whose one move fills the board (no child data).
// B.line[d] indexes `lines` for every open dimension d (-1: closed or not computed).
template <bool S35>
__device__ __forceinline__ void fused_block(const BlockDev& B, const SecDev& S, const LineDev* __restrict__ lines,
                                            const uint64_t* __restrict__ fields,
                                            const uint16_t* __restrict__ movetab, const uint8_t* __restrict__ refl,
                                            int part) {
  // wave w of the block holds (loc0, loc1) = (w / 8, w % 8) and its 64 lanes (loc2, loc3):
  // quadrants 0 and 1 are the same for the whole wave, so their table reads are scalar
  // loads and their partial products scalar operations.  Lanes outside a partial edge
  // block idle.
  const int wave = part * (kThreads / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // the block record's fields as plain int arrays (indexed by the dimension loop below)
  const int bb[4] = {B.b[0], B.b[1], B.b[2], B.b[3]};
  const int bline[4] = {B.line[0], B.line[1], B.line[2], B.line[3]};
  const int bs0 = min(8, S.rmin_size[0] - 8 * bb[0]), bs1 = min(8, S.rmin_size[1] - 8 * bb[1]);
  const int bs2 = min(8, S.rmin_size[2] - 8 * bb[2]), bs3 = min(8, S.rmin_size[3] - 8 * bb[3]);
  // B.accumulate: 0 store the result, 1 OR it into the output; partial passes (sections run
  // in passes over subsets of their dimensions): 2 store the children's OR only, 4 add to
  // the partial in the output slots, 3 finish from it (the result, stored)
  const int mode = B.accumulate;
  const bool partial_out = !S35 && (mode == 2 || mode == 4);
  // the block's field entries of quadrants 2 and 3 (eight each) staged in LDS (the win
  // patterns need them; a partial pass does not)
  __shared__ uint64_t sf[2][8][60];
  if (!partial_out) {   // uniform over the workgroup (one block)
    for (int i = threadIdx.x; i < 2 * 8 * 60; i += kThreads) {
      const int q = 2 + i / 480, r = (i / 60) % 8, w = i % 60;
      const int bsq = q == 2 ? bs2 : bs3, bq = q == 2 ? bb[2] : bb[3];
      const int64_t fo = q == 2 ? S.field_off[2] : S.field_off[3];
      sf[q - 2][r][w] = r < bsq ? fields[4 * fo + 60 * int64_t(8 * bq + r) + w] : 0;
    }
    __syncthreads();
  }
  int loc[4] = {wave >> 3, wave & 7, lane >> 3, lane & 7};
  if (loc[0] >= bs0 || loc[1] >= bs1 || loc[2] >= bs2 || loc[3] >= bs3) return;
  const int p = ((loc[0] * bs1 + loc[1]) * bs2 + loc[2]) * bs3 + loc[3];

  // ---- children: per open dimension, AND the moves' children shifted along the child
  //      dimension (OR of complements = complement of AND), then apply the rest of the
  //      local rotation and the rearrangement ----
  // the node's running OR over the dimensions, then its output, in LDS: its own 8 words
  // ([word][thread], so a wave's accesses are consecutive 8-byte slots), never shared
  __shared__ uint64_t sacc[S35 ? 1 : 8][kThreads];
  uint64_t* o = B.out + 8 * p;
  if (!S35) {
    if (mode >= 3) {   // the partial left in the output slots by an earlier pass
      const S4 a = ld(o), b = ld(o + 4);
      for (int w = 0; w < 4; w++) {
        sacc[w][threadIdx.x] = a.w[w];
        sacc[4 + w][threadIdx.x] = b.w[w];
      }
    } else {
      for (int w = 0; w < 8; w++) sacc[w][threadIdx.x] = 0;
    }
  }
#pragma unroll
  for (int d = 0; d < 4; d++) {
    if (S35 || bline[d] < 0) continue;
    const LineDev& L = lines[bline[d]];
    const int cd = L.child_dim;
    int base = 0, base_last = 0, local_sym = 0;
    const uint8_t* syms = refl + L.refl_base;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      if (a == d) continue;
      const int x = loc[a] ^ int(L.ax_fcn[a] + loc[a] < L.ax_moves[a]);
      base += L.ax_stride[a] * x;
      base_last += L.ax_stride_last[a] * x;
      if (L.reflected) local_sym |= syms[L.ax_rs[a] + loc[a]];
    }
    const uint16_t* row = movetab + S.move_off[d] + 16 * (8 * bb[d] + loc[d]);
    const int nm = S.nmoves[d];
    const int last = L.child_length - 1;
    const uint64_t* in = L.input + 8 * L.in_base;
    S4 acc0, acc1;
    for (int i = 0; i < 4; i++) acc0.w[i] = acc1.w[i] = ~uint64_t(0);
    // one move at a time: its child's four loads, then its shift
    for (int m = 0; m < nm; m++) {
      const int e = row[m];
      const int kk = e & 63;
      const int64_t cidx = int64_t(kk) * L.in_kstride + (kk == last ? base_last : base) +
                           int64_t(L.input_stride) * ((e >> 6) & 7);
      const uint64_t* c = in + 8 * cidx;
      const S4 c0 = ld(c), c1 = ld(c + 4);
      acc0 = s4_and(acc0, shift_coord_var(c0, cd, e >> 9));
      acc1 = s4_and(acc1, shift_coord_var(c1, cd, e >> 9));
    }
    // child frame -> parent frame: X(commute[g][x] + y).  Unreflected lines use one of four
    // fixed symmetries (straight-line code); reflected ones have a per-node y.
    S4 x0 = s4_not(acc0), x1 = s4_not(acc1);
    const int ft = L.fast_t;
    if (ft == 1) {
      x0 = line_transform1(x0);
      x1 = line_transform1(x1);
    } else if (ft == 2) {
      x0 = line_transform2(x0);
      x1 = line_transform2(x1);
    } else if (ft == 3) {
      x0 = line_transform3(x0);
      x1 = line_transform3(x1);
    } else if (ft < 0) {
      const int y = local_add(L.inv_local, local_sym);
      x0 = rearrange(shift_local(x0, y), L.inv_global);
      x1 = rearrange(shift_local(x1, y), L.inv_global);
    }
#pragma unroll
    for (int w = 0; w < 4; w++) sacc[w][threadIdx.x] |= x0.w[w];
#pragma unroll
    for (int w = 0; w < 4; w++) sacc[4 + w][threadIdx.x] |= x1.w[w];
  }
  if (partial_out) {   // rmax(A) | rmax(B) = rmax(A | B): the finishing pass applies it once
    S4 a, b;
    for (int w = 0; w < 4; w++) {
      a.w[w] = sacc[w][threadIdx.x];
      b.w[w] = sacc[4 + w][threadIdx.x];
    }
    st(o, a);
    st(o + 4, b);
    return;
  }
  if (!S35) {   // rmax in place: words 0-3 from the children's [0] halves, 4-7 from [1]
#pragma unroll
    for (int h = 0; h < 2; h++) {
      S4 t;
      for (int w = 0; w < 4; w++) t.w[w] = sacc[4 * h + w][threadIdx.x];
      t = rmax(t);
      for (int w = 0; w < 4; w++) sacc[4 * h + w][threadIdx.x] = t.w[w];
    }
  }

  // ---- wins0, wins1 and the wins of every move (pentago/base/superscore.cc:38-96 pattern
  //      groups), one 64-bit word at a time: every reduction except the one over r3 stays
  //      inside a word, and the three groups reduced over r3 become ORs across words ----
  NodeFields F;
#pragma unroll
  for (int a = 0; a < 2; a++)
    F.glob[a] = fields + 4 * int64_t(S.field_off[a]) + 60 * int64_t(8 * bb[a] + loc[a]);
  F.lds[0] = sf[0][loc[2]];
  F.lds[1] = sf[1][loc[3]];
  uint64_t A[3] = {0, 0, 0}, Bv[3] = {0, 0, 0}, Cv[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint64_t x[3];
    g2_acc(F, k, 0, FH, 2, FH, A);
    PE_FENCE();
    g2(F, k, 0, FV, 1, FV, x);   // both quadrants wave-uniform: scalar
    for (int t = 0; t < 3; t++) Bv[t] |= x[t];
    PE_FENCE();
    g3_term(F, k, 1, FLO, 0, FAS, 2, FLO, x);
    for (int t = 0; t < 3; t++) Cv[t] |= x[t];
    PE_FENCE();
  }
  S4 o0, o1, wv, w0v, w1v;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint64_t w[3], x[3], y[3];
    g2_term(F, k, 2, FV, 3, FV, x);
    g3_term(F, k, 0, FLO, 2, FAS, 3, FLO, y);
    for (int t = 0; t < 3; t++) w[t] = red1(lop3<kOr3>(A[t], red0(x[t]), y[t])) | Cv[t];
    PE_FENCE();
    g2_term(F, k, 1, FH, 3, FH, x);
    g3_term(F, k, 0, FHI, 1, FAS, 3, FHI, y);
    for (int t = 0; t < 3; t++) w[t] |= red2(lop3<kOr3>(Bv[t], red0(x[t]), y[t]));
    PE_FENCE();
    g3_term(F, k, 1, FHI, 3, FAS, 2, FHI, x);
    for (int t = 0; t < 3; t++) w[t] |= red0(x[t]);
    if (S35) {
      // OR over the moves of super_wins(side0 + move) is w[2] | w[0] exactly: groups through
      // the moved quadrant are in w[2], the rest are groups of side0 (a move only adds wins)
      wv.w[k] = w[2] | w[0];
      w0v.w[k] = w[0];
      w1v.w[k] = w[1];
    } else {
      // the output words replace the child words they are made from
      // ((w2 | r1) & ~(w0 | w1)) | (w0 & ~w1) and (r0 & ~(w0 | w1)) | w0 as functions of
      // (u, w0, w1): truth tables 0x54 and 0xDC
      const uint64_t r0k = sacc[k][threadIdx.x], r1k = sacc[4 + k][threadIdx.x];
      sacc[4 + k][threadIdx.x] = lop3<0x54>(w[2] | r1k, w[0], w[1]);
      sacc[k][threadIdx.x] = lop3<0xDC>(r0k, w[0], w[1]);
    }
    PE_FENCE();
  }
  if (S35) {   // compute.cc:329-332: after the board fills, a position nobody has won is a tie
    const S4 win = s4_or(wv, rmax(s4_andnot(wv, w1v)));
    const S4 nl = rmax(s4_or(wv, s4_not(w1v)));
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint64_t imm = w0v.w[k] | w1v.w[k];
      o0.w[k] = (win.w[k] & ~imm) | (w0v.w[k] & ~w1v.w[k]);
      o1.w[k] = (nl.w[k] & ~imm) | w0v.w[k];
    }
  }
  if (!S35)
#pragma unroll
    for (int k = 0; k < 4; k++) {
      o0.w[k] = sacc[4 + k][threadIdx.x];
      o1.w[k] = sacc[k][threadIdx.x];
    }
  if (mode == 1) {   // OR-merge with other dimensions' lines (block_store.cc:173-229)
    o0 = s4_or(o0, ld(o));
    o1 = s4_or(o1, ld(o + 4));
  }
  st(o, o0);
  st(o + 4, o1);
}

template <bool S35>
__global__ __launch_bounds__(kThreads, kWaves)
void fused_blocks_kernel(const BlockDev* __restrict__ blocks, const LineDev* __restrict__ lines,
                         const SecDev* __restrict__ secs, const uint64_t* __restrict__ fields,
                         const uint16_t* __restrict__ movetab, const uint8_t* __restrict__ refl,
                         int nblocks, int blocks_per_xcd) {
  // `blocks` and `nblocks` describe this launch's chunk of the plan's blocks
  // Workgroups are dispatched to the 8 XCDs round-robin: workgroup i runs on XCD i % 8.
  // XCD x takes the contiguous run of blocks [x * blocks_per_xcd, ...), so the blocks that
  // share child lines (neighbours along the last axes) and the two halves of a 128-byte line
  // meet in the same L2.
  const int xcd = blockIdx.x & 7, r = blockIdx.x >> 3;
  const int blk = xcd * blocks_per_xcd + r / kParts, part = r % kParts;
  if (r / kParts >= blocks_per_xcd || blk >= nblocks) return;
  const BlockDev& B = blocks[blk];
  fused_block<S35>(B, secs[B.sec], lines, fields, movetab, refl, part);
}

// The same blocks fed from device work queues to a persistent grid: heads[x] counts the
// tasks (block * 16 + part, from the start of XCD x's run of blocks) taken from XCD x's
// run.  A workgroup takes tasks from the run of its own XCD (workgroup i runs on XCD
// i % 8) and, once that is exhausted, from the other runs in turn; every pop past a run's
// end fails, so each workgroup stops after at most eight failed pops.
template <bool S35>
__global__ __launch_bounds__(kThreads, kWaves)
void fused_blocks_queue_kernel(const BlockDev* __restrict__ blocks, const LineDev* __restrict__ lines,
                               const SecDev* __restrict__ secs, const uint64_t* __restrict__ fields,
                               const uint16_t* __restrict__ movetab, const uint8_t* __restrict__ refl,
                               int nblocks, int blocks_per_xcd, unsigned int* __restrict__ heads) {
  __shared__ int s_task;
  const int home = int(blockIdx.x) & 7;
  for (int x = 0; x < 8;) {
    const int xcd = (home + x) & 7;
    const int run = max(0, min(blocks_per_xcd, nblocks - xcd * blocks_per_xcd));
    if (threadIdx.x == 0) s_task = int(atomicAdd(&heads[xcd], 1u));
    __syncthreads();
    const int t = s_task;
    __syncthreads();   // every thread has read the task before the next pop overwrites it
    if (t >= run * kParts) {
      x++;
      continue;
    }
    const BlockDev& B = blocks[xcd * blocks_per_xcd + t / kParts];
    fused_block<S35>(B, secs[B.sec], lines, fields, movetab, refl, t % kParts);
    __syncthreads();   // the block's staged fields are no longer read
  }
}

__global__ void fused_queue_reset_kernel(unsigned int* heads) {
  if (threadIdx.x < 8) heads[threadIdx.x] = 0;
}

// One block line passed by value (no device descriptors): the line and its section record.
struct LineLaunch {
  LineDev line;
  SecDev sec;
};

// Workgroup i computes part i % 16 of block i / 16 of the line; output block k at node
// k * block_stride, child block k at node k * (block_stride + 1) (the reference's buffers).
template <bool S35>
__global__ __launch_bounds__(kThreads, kWaves)
void fused_line_kernel(const LineLaunch a, const uint64_t* __restrict__ fields,
                       const uint16_t* __restrict__ movetab, const uint8_t* __restrict__ refl) {
  const int k = int(blockIdx.x) / kParts, part = int(blockIdx.x) % kParts;
  const int dim = a.line.dim;
  BlockDev B;
  B.sec = 0;
#pragma unroll
  for (int x = 0; x < 4; x++) {
    B.b[x] = x == dim ? k : a.line.first_block[x];
    B.line[x] = x == dim ? 0 : -1;
  }
  B.accumulate = a.line.accumulate;
  B.out = a.line.output + 8 * int64_t(k) * a.line.block_stride;
  fused_block<S35>(B, a.sec, &a.line, fields, movetab, refl, part);
}

int64_t slab_index(const pe_slab& s, const int b[4]) {
  int64_t idx = 0;
  for (int a = 0; a < 4; a++) {
    if (b[a] < s.lo[a] || b[a] >= s.hi[a]) return -1;
    idx = idx * (s.hi[a] - s.lo[a]) + (b[a] - s.lo[a]);
  }
  return idx;
}

bool same_section(const pe_section& a, const pe_section& b) { return !std::memcmp(a.counts, b.counts, 8); }

}  // namespace

// Build the lines through an output slab (one per open dimension and cross-section block)
// with child addressing into the child slabs.
static pe_status slab_lines(const pe_slab& out, int nch, const pe_slab* ch, std::vector<LineHost>* lines,
                            std::map<std::tuple<int, int, int, int>, int>* index, std::string* err,
                            uint32_t dims = 0xf) {
  Section s;
  std::memcpy(s.c, out.section.counts, 8);
  int shape[4];
  if (!s.valid_counts() || !s.shape(shape)) { *err = "invalid output section"; return PE_ERR_INVALID; }
  if (s.sum() > 35) { *err = "slab plans need sections with at most 35 stones"; return PE_ERR_INVALID; }
  for (int a = 0; a < 4; a++)
    if (out.lo[a] < 0 || out.hi[a] <= out.lo[a] || 8 * (out.hi[a] - 1) >= shape[a]) { *err = "output box outside the section"; return PE_ERR_INVALID; }
  // child slabs by section
  std::unordered_map<uint64_t, std::vector<int>> by_section;
  for (int c = 0; c < nch; c++) {
    Section cs;
    std::memcpy(cs.c, ch[c].section.counts, 8);
    by_section[cs.sig()].push_back(c);
  }
  for (int d = 0; d < 4; d++) {
    if (s.c[d][0] + s.c[d][1] >= 9 || !(dims >> d & 1)) continue;
    int o[3], k = 0;
    for (int a = 0; a < 4; a++) if (a != d) o[k++] = a;
    for (int x = out.lo[o[0]]; x < out.hi[o[0]]; x++)
      for (int y = out.lo[o[1]]; y < out.hi[o[1]]; y++)
        for (int z = out.lo[o[2]]; z < out.hi[o[2]]; z++) {
          pe_line L;
          L.section = out.section;
          L.dimension = uint8_t(d);
          L.length = uint8_t((shape[d] + 7) / 8);
          L.block_base[0] = uint8_t(x); L.block_base[1] = uint8_t(y); L.block_base[2] = uint8_t(z);
          LineHost h;
          const pe_status st = make_line(L, &h, err);
          if (st != PE_OK) return st;
          if (h.dev.slice35) {   // the game ends after the move: no child data
            (*index)[std::make_tuple(d, x, y, z)] = int(lines->size());
            lines->push_back(std::move(h));
            continue;
          }
          const int cd = h.info.child_dimension;
          const pe_slab* found = nullptr;
          Section want;
          std::memcpy(want.c, h.info.standard_child.counts, 8);
          const auto it = by_section.find(want.sig());
          const std::vector<int> none;
          for (int c : it == by_section.end() ? none : it->second) {
            if (found) break;
            int cb[4];
            for (int a = 0; a < 4; a++) cb[a] = h.info.first_child_block[a];
            bool ok = true;
            for (int kk = 0; kk < h.info.input_blocks && ok; kk++) { cb[cd] = kk; ok = slab_index(ch[c], cb) >= 0; }
            if (ok) found = &ch[c];
          }
          if (!found) { *err = "no child slab covers the child blocks of a line"; return PE_ERR_INVALID; }
          h.dev.input = static_cast<const uint64_t*>(found->data);
          int cb[4];
          for (int a = 0; a < 4; a++) cb[a] = h.info.first_child_block[a];
          for (int kk = 0; kk < h.info.input_blocks; kk++) { cb[cd] = kk; h.in_off.push_back(4096 * slab_index(*found, cb)); }
          (*index)[std::make_tuple(d, x, y, z)] = int(lines->size());
          lines->push_back(std::move(h));
        }
  }
  return PE_OK;
}

// Pack blocks, lines (with their child offsets and reflection symmetries), sections and
// move tables into one device allocation (field tables come from the device's field pool).
static pe_status pack_fused_plan(const std::vector<BlockDev>& blocks, const std::vector<LineHost>& lines,
                                 const std::vector<SecDev>& secs, const std::vector<uint16_t>& movetab,
                                 int64_t nodes, FusedPlan* plan, std::string* err) {
  if (blocks.size() > size_t(0x7fffffff) / kParts) { *err = "fused plan too large"; return PE_ERR_CAPACITY; }
  // blob: blocks | lines | sections | moves | offsets | symmetries
  std::vector<LineDev> devs(lines.size());
  std::vector<int64_t> offs;
  std::vector<uint8_t> refl;
  std::vector<int64_t> at(lines.size());
  for (size_t k = 0; k < lines.size(); k++) {
    devs[k] = lines[k].dev;
    // child blocks along the child dimension sit at affine offsets in their slab
    const std::vector<int64_t>& io = lines[k].in_off;   // empty for 35-stone lines (no child data)
    devs[k].in_base = io.empty() ? 0 : io[0];
    devs[k].in_kstride = io.size() > 1 ? io[1] - io[0] : 0;
    for (size_t kk = 0; kk < io.size(); kk++)
      if (io[kk] != devs[k].in_base + int64_t(kk) * devs[k].in_kstride) {
        *err = "internal: child blocks of a line are not evenly spaced";
        return PE_ERR_INVALID;
      }
    devs[k].refl_base = int32_t(refl.size());
    refl.insert(refl.end(), lines[k].refl_syms.begin(), lines[k].refl_syms.end());
    at[k] = int64_t(offs.size());
    offs.insert(offs.end(), lines[k].in_off.begin(), lines[k].in_off.end());
  }
  refl.push_back(0);
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t o_lines = up(sizeof(BlockDev) * blocks.size());
  const size_t o_secs = o_lines + up(sizeof(LineDev) * devs.size());
  const size_t o_moves = o_secs + up(sizeof(SecDev) * secs.size());
  const size_t o_offs = o_moves + up(sizeof(uint16_t) * std::max<size_t>(movetab.size(), 1));
  const size_t o_refl = o_offs + up(sizeof(int64_t) * offs.size());
  const size_t o_heads = o_refl + up(refl.size());
  const size_t bytes = o_heads + 256;
  void* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) { *err = "hipMalloc(fused plan) failed"; return PE_ERR_DEVICE; }
  auto* base = static_cast<uint8_t*>(d);
  for (size_t k = 0; k < devs.size(); k++) devs[k].in_off = reinterpret_cast<const int64_t*>(base + o_offs) + at[k];
  std::vector<uint8_t> blob(bytes, 0);
  std::memcpy(blob.data(), blocks.data(), sizeof(BlockDev) * blocks.size());
  std::memcpy(blob.data() + o_lines, devs.data(), sizeof(LineDev) * devs.size());
  std::memcpy(blob.data() + o_secs, secs.data(), sizeof(SecDev) * secs.size());
  if (!movetab.empty()) std::memcpy(blob.data() + o_moves, movetab.data(), sizeof(uint16_t) * movetab.size());
  std::memcpy(blob.data() + o_offs, offs.data(), sizeof(int64_t) * offs.size());
  std::memcpy(blob.data() + o_refl, refl.data(), refl.size());
  // uploaded on a stream of its own that does not wait for other work, so that a plan can be
  // built while another plan's kernels run
  hipStream_t ups = nullptr;
  const bool copied = hipStreamCreateWithFlags(&ups, hipStreamNonBlocking) == hipSuccess &&
                      hipMemcpyAsync(d, blob.data(), bytes, hipMemcpyHostToDevice, ups) == hipSuccess &&
                      hipStreamSynchronize(ups) == hipSuccess;
  if (ups) (void)hipStreamDestroy(ups);
  if (!copied) {
    (void)hipFree(d);
    *err = "hipMemcpy(fused plan) failed";
    return PE_ERR_DEVICE;
  }
  plan->blob = d;
  plan->nblocks = int(blocks.size());
  plan->off_lines = o_lines;
  plan->off_secs = o_secs;
  plan->off_moves = o_moves;
  plan->off_refl = o_refl;
  plan->off_heads = o_heads;
  plan->output_nodes = nodes;
  return PE_OK;
}

// One output slab's share of a fused plan, built independently of the others: its lines,
// section record, field and move tables, and blocks (section 0, lines counted from 0).
struct SlabBuild {
  std::vector<LineHost> lines;
  SecDev sec;
  std::vector<uint16_t> movetab;
  std::vector<BlockDev> blocks;
  int64_t nodes = 0;
  pe_status status = PE_OK;
  std::string err;
};

static void build_slab_body(const pe_slab& out, int nch, const pe_slab* ch, uint32_t dims, int accumulate,
                            SlabBuild* r) {
  const auto& Tb = tables();
  if (!out.data) { r->err = "output slab without memory"; r->status = PE_ERR_INVALID; return; }
  std::map<std::tuple<int, int, int, int>, int> index;
  r->status = slab_lines(out, nch, ch, &r->lines, &index, &r->err, dims);
  if (r->status != PE_OK) return;
  Section s;
  std::memcpy(s.c, out.section.counts, 8);
  if (s.sum() > 35) { r->err = "the fused kernel needs sections with at most 35 stones"; r->status = PE_ERR_INVALID; return; }
  SecDev& sec = r->sec;
  std::memset(&sec, 0, sizeof sec);
  sec.turn = s.sum() & 1;
  r->status = section_move_tables(s, &r->movetab, sec.move_off, sec.nmoves, &r->err);
  if (r->status != PE_OK) return;
  for (int a = 0; a < 4; a++) {
    const int b = rmin_bucket(s.c[a][0], s.c[a][1]);
    sec.rmin_off[a] = Tb.rmin_offsets[b];
    sec.rmin_size[a] = Tb.rmin_offsets[b + 1] - Tb.rmin_offsets[b];
    // a quadrant's moves count only when its dimension is computed, except in a finishing
    // pass (accumulate 3), which evaluates the win patterns of every move for the whole node
    sec.field_off[a] = field_pool_off(a, sec.turn, accumulate == 3 || ((dims >> a) & 1), s.c[a][0], s.c[a][1]);
    if (!(dims >> a & 1)) { sec.move_off[a] = -1; sec.nmoves[a] = 0; }
  }
  // blocks in C order over the box, except that the last computed dimension runs innermost
  // (with every dimension open, the plain C order): consecutive blocks of an XCD's run then
  // lie along that dimension's lines, whose child reads the L2 can share
  int inner = 3;
  for (int d = 0; d < 4; d++)
    if (s.c[d][0] + s.c[d][1] < 9 && (dims >> d & 1)) inner = d;
  int ord[4], nord = 0;
  for (int a = 0; a < 4; a++)
    if (a != inner) ord[nord++] = a;
  ord[3] = inner;
  int64_t total = 1;
  for (int a = 0; a < 4; a++) total *= std::max(0, out.hi[a] - out.lo[a]);
  int b[4];
  for (int64_t t = 0; t < total; t++) {
    int64_t rest = t;
    for (int j = 3; j >= 0; j--) {
      const int a = ord[j], e = out.hi[a] - out.lo[a];
      b[a] = out.lo[a] + int(rest % e);
      rest /= e;
    }
    BlockDev B;
    B.sec = 0;
    B.accumulate = accumulate;
    int n = 1;
    for (int a = 0; a < 4; a++) { B.b[a] = b[a]; n *= std::min(8, sec.rmin_size[a] - 8 * b[a]); }
    for (int d = 0; d < 4; d++) {
      B.line[d] = -1;
      if (s.c[d][0] + s.c[d][1] >= 9 || !(dims >> d & 1)) continue;
      int o[3], kk = 0;
      for (int a = 0; a < 4; a++) if (a != d) o[kk++] = b[a];
      B.line[d] = index.at(std::make_tuple(d, o[0], o[1], o[2]));
    }
    B.out = static_cast<uint64_t*>(out.data) + 8 * 4096 * slab_index(out, b);
    r->blocks.push_back(B);
    r->nodes += n;
  }
}

// Worker-thread entry: no exception leaves the thread (a host allocation failure or a bad
// index becomes the slab's error status).
static void build_slab(const pe_slab& out, int nch, const pe_slab* ch, uint32_t dims, int accumulate, SlabBuild* r) {
  try {
    build_slab_body(out, nch, ch, dims, accumulate, r);
  } catch (const std::bad_alloc&) {
    r->status = PE_ERR_CAPACITY;
    r->err = "host memory exhausted while building a fused plan";
  } catch (const std::exception& e) {
    r->status = PE_ERR_INVALID;
    r->err = std::string("fused plan: ") + e.what();
  }
}

pe_status fused_plan_create(int nout, const pe_slab* outs, int nch, const pe_slab* ch, FusedPlan* plan,
                            std::string* err, uint32_t dims, int accumulate) {
  DevTables T;
  if (device_tables(&T, err) != PE_OK) return PE_ERR_DEVICE;
  // one plan holds either sections with at most 34 stones or 35-stone sections only
  bool any35 = false, all35 = true;
  for (int k = 0; k < nout; k++) {
    int sum = 0;
    for (int q = 0; q < 4; q++) sum += outs[k].section.counts[q][0] + outs[k].section.counts[q][1];
    any35 |= sum == 35;
    all35 &= sum == 35;
  }
  if (any35 && !all35) { *err = "a fused plan mixes 35-stone sections with others"; return PE_ERR_INVALID; }
  plan->s35 = any35;
  // the slabs' shares are built on host threads (a slice has thousands of sections), then
  // concatenated in slab order, so the plan does not depend on the thread count
  std::vector<SlabBuild> builds(size_t(std::max(nout, 0)));
  const int nthreads = std::max(1, std::min({nout, int(std::thread::hardware_concurrency()), 16}));
  std::atomic<int> next(0);
  auto work = [&]() {
    for (int k; (k = next.fetch_add(1)) < nout;) build_slab(outs[k], nch, ch, dims, accumulate, &builds[size_t(k)]);
  };
  std::vector<std::thread> pool;
  struct Join {   // joins the started threads on every way out, a failed thread start included
    std::vector<std::thread>& p;
    ~Join() {
      for (auto& th : p)
        if (th.joinable()) th.join();
    }
  } join{pool};
  for (int t = 1; t < nthreads; t++) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  std::vector<LineHost> lines;
  std::vector<SecDev> secs;
  std::vector<uint16_t> movetab;
  std::vector<BlockDev> blocks;
  int64_t nodes = 0;
  for (SlabBuild& r : builds) {
    if (r.status != PE_OK) { *err = r.err; return r.status; }
    SecDev sec = r.sec;
    for (int a = 0; a < 4; a++)
      if (sec.move_off[a] >= 0) sec.move_off[a] += int32_t(movetab.size());
    movetab.insert(movetab.end(), r.movetab.begin(), r.movetab.end());
    const int line0 = int(lines.size());
    for (auto& h : r.lines) lines.push_back(std::move(h));
    for (BlockDev B : r.blocks) {
      B.sec = int32_t(secs.size());
      for (int d = 0; d < 4; d++)
        if (B.line[d] >= 0) B.line[d] += line0;
      blocks.push_back(B);
    }
    nodes += r.nodes;
    secs.push_back(sec);
    std::vector<BlockDev>().swap(r.blocks);
  }
  return pack_fused_plan(blocks, lines, secs, movetab, nodes, plan, err);
}

bool fused_line_eligible(const LineHost& h) {
  if (h.dev.i_begin != 0 || h.dev.i_end != h.dev.rmin_size[h.dev.dim]) return false;
  const std::vector<int64_t>& io = h.in_off;
  for (size_t k = 2; k < io.size(); k++)
    if (io[k] - io[k - 1] != io[1] - io[0]) return false;
  return true;
}

// Each job's blocks become fused blocks with only the job's dimension open: the section's
// fields from the pool with the move fields (Fm) of the other quadrants zero, and only that
// dimension's move table, so a block computes exactly the line's contribution
// (compute.cc:219-371 for one dimension); accumulate jobs OR it into the output.  The jobs of
// one plan are all 35-stone lines (no child data) or all have at most 34 stones.
pe_status fused_line_plan_create(const std::vector<LineHost>& jobs, FusedPlan* plan, std::string* err) {
  DevTables T;
  if (device_tables(&T, err) != PE_OK) return PE_ERR_DEVICE;
  const auto& Tb = tables();
  std::vector<LineHost> lines;
  std::vector<SecDev> secs;
  std::vector<uint16_t> movetab;
  std::vector<BlockDev> blocks;
  std::map<std::pair<uint64_t, int>, int> sec_index;   // (section sig, dimension) -> SecDev
  int64_t nodes = 0;
  plan->s35 = !jobs.empty() && jobs[0].dev.slice35;
  for (const LineHost& job : jobs) {
    if (!fused_line_eligible(job)) { *err = "internal: line job not eligible for the fused kernel"; return PE_ERR_INVALID; }
    if (bool(job.dev.slice35) != plan->s35) { *err = "internal: fused line plan mixes 35-stone lines with others"; return PE_ERR_INVALID; }
    const int d = job.dev.dim;
    Section own;   // the line's own section
    std::memcpy(own.c, job.section.counts, 8);
    const auto key = std::make_pair(own.sig(), d);
    auto it = sec_index.find(key);
    if (it == sec_index.end()) {
      SecDev sec;
      std::memset(&sec, 0, sizeof sec);
      sec.turn = own.sum() & 1;
      std::vector<uint16_t> mt;
      pe_status st = section_move_tables(own, &mt, sec.move_off, sec.nmoves, err, d);
      if (st != PE_OK) return st;
      for (int a = 0; a < 4; a++) {
        // only d's moves are this line's
        sec.field_off[a] = field_pool_off(a, sec.turn, a == d, own.c[a][0], own.c[a][1]);
        if (a != d) { sec.move_off[a] = -1; sec.nmoves[a] = 0; }
        else if (sec.move_off[a] >= 0) sec.move_off[a] += int32_t(movetab.size());
        const int b = rmin_bucket(own.c[a][0], own.c[a][1]);
        sec.rmin_off[a] = Tb.rmin_offsets[b];
        sec.rmin_size[a] = Tb.rmin_offsets[b + 1] - Tb.rmin_offsets[b];
      }
      movetab.insert(movetab.end(), mt.begin(), mt.end());
      it = sec_index.emplace(key, int(secs.size())).first;
      secs.push_back(sec);
    }
    const SecDev& sec = secs[size_t(it->second)];
    LineHost h = job;
    if (h.in_off.empty())
      for (int k = 0; k < h.info.input_blocks; k++) h.in_off.push_back(int64_t(k) * (h.dev.block_stride + 1));
    const int line_index = int(lines.size());
    for (int k = 0; k < h.dev.length; k++) {
      BlockDev B;
      B.sec = it->second;
      int n = 1;
      for (int a = 0; a < 4; a++) {
        B.b[a] = a == d ? k : h.dev.first_block[a];
        n *= std::min(8, sec.rmin_size[a] - 8 * B.b[a]);
        B.line[a] = a == d ? line_index : -1;
      }
      B.accumulate = h.dev.accumulate;
      B.out = h.dev.output + 8 * (h.out_off.empty() ? int64_t(k) * h.dev.block_stride : h.out_off[size_t(k)]);
      blocks.push_back(B);
      nodes += n;
    }
    lines.push_back(std::move(h));
  }
  return pack_fused_plan(blocks, lines, secs, movetab, nodes, plan, err);
}

pe_status fused_plan_run(const FusedPlan& plan, hipStream_t stream) {
  if (!plan.nblocks) return PE_OK;
  DevTables T;
  std::string err;
  if (device_tables(&T, &err) != PE_OK) { set_error(err); return PE_ERR_DEVICE; }
  auto* base = static_cast<uint8_t*>(plan.blob);
  const int per_xcd = (plan.nblocks + 7) / 8;
  const auto* blocks = reinterpret_cast<const BlockDev*>(base);
  const auto* lines = reinterpret_cast<const LineDev*>(base + plan.off_lines);
  const auto* secs = reinterpret_cast<const SecDev*>(base + plan.off_secs);
  const auto* moves = reinterpret_cast<const uint16_t*>(base + plan.off_moves);
  const uint8_t* refl = base + plan.off_refl;
  if (plan.queue_wgs_per_cu > 0) {
    // persistent form: a whole number of workgroups per XCD, queue heads cleared by a kernel
    // on the same stream (no memset node, so the sequence may be captured in a graph)
    int dev = 0, cus = 0;
    PE_HIP(hipGetDevice(&dev), "hipGetDevice");
    PE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute(CUs)");
    const int64_t want = int64_t(cus) * plan.queue_wgs_per_cu;
    const unsigned grid = unsigned(std::max<int64_t>(8, std::min<int64_t>(want, int64_t(per_xcd) * kParts * 8)) / 8 * 8);
    auto* heads = reinterpret_cast<unsigned int*>(base + plan.off_heads);
    hipLaunchKernelGGL(fused_queue_reset_kernel, dim3(1), dim3(64), 0, stream, heads);
    hipLaunchKernelGGL(plan.s35 ? fused_blocks_queue_kernel<true> : fused_blocks_queue_kernel<false>, dim3(grid),
                       dim3(kThreads), 0, stream, blocks, lines, secs, T.fields, moves, refl, plan.nblocks, per_xcd,
                       heads);
    PE_HIP(hipGetLastError(), "fused_blocks_queue_kernel launch");
    return PE_OK;
  }
  // a dispatch counts at most 2^32 work-items (4096 per block): plans with more blocks run
  // as consecutive launches of at most kChunk blocks, each spread over the XCDs as above
  for (int b0 = 0; b0 < plan.nblocks; b0 += kChunk) {
    const int n = std::min(kChunk, plan.nblocks - b0);
    const int px = (n + 7) / 8;
    hipLaunchKernelGGL(plan.s35 ? fused_blocks_kernel<true> : fused_blocks_kernel<false>,
                       dim3(unsigned(int64_t(px) * kParts * 8)), dim3(kThreads), 0, stream, blocks + b0, lines, secs,
                       T.fields, moves, refl, n, px);
    PE_HIP(hipGetLastError(), "fused_blocks_kernel launch");
  }
  return PE_OK;
}

void fused_plan_destroy(FusedPlan* plan) {
  if (plan->blob) (void)hipFree(plan->blob);
  plan->blob = nullptr;
  plan->nblocks = 0;
}

// Mode 1: four line passes over the slab, later passes OR-accumulating.  The lines of pass p
// (the slab's p-th open dimension) are appended to (*passes)[p], so that the passes of many
// slabs can share one launch each.
pe_status slab_line_passes(const pe_slab& out, int nch, const pe_slab* ch, std::vector<std::vector<LineHost>>* passes,
                           int64_t* nodes_out, std::string* err) {
  std::vector<LineHost> lines;
  std::map<std::tuple<int, int, int, int>, int> index;
  pe_status st = slab_lines(out, nch, ch, &lines, &index, err);
  if (st != PE_OK) return st;
  Section s;
  std::memcpy(s.c, out.section.counts, 8);
  int shape[4];
  s.shape(shape);
  int64_t nodes = 1;
  for (int a = 0; a < 4; a++) nodes *= std::min(shape[a], 8 * out.hi[a]) - 8 * out.lo[a];
  *nodes_out = nodes;
  size_t pass = 0;
  for (int d = 0; d < 4; d++) {
    std::vector<LineHost> group;
    for (auto& h : lines) {
      if (h.dev.dim != d) continue;
      LineHost g = h;
      g.dev.output = static_cast<uint64_t*>(out.data);
      g.dev.i_begin = 8 * out.lo[d];
      g.dev.i_end = std::min(shape[d], 8 * out.hi[d]);
      g.dev.accumulate = pass > 0;
      int b[4];
      for (int a = 0; a < 4; a++) b[a] = a == d ? 0 : g.dev.first_block[a];
      for (int kk = 0; kk < g.dev.length; kk++) {
        b[d] = kk;
        const int64_t idx = slab_index(out, b);
        g.out_off.push_back(idx < 0 ? 0 : 4096 * idx);
      }
      group.push_back(std::move(g));
    }
    if (group.empty()) continue;
    if (passes->size() <= pass) passes->resize(pass + 1);
    for (auto& g : group) (*passes)[pass].push_back(std::move(g));
    pass++;
  }
  return PE_OK;
}

// ---- single block lines launched by value (the asynchronous schedule_compute_line form) ----

namespace {

// Per (device, section, dimension): the section record of a one-dimension fused block, that
// dimension's move rows and, for a child standardized through a reflection, the reflection
// symmetry of every rotation-minimal quadrant of each child axis (compute.cc:96-122; a
// line's all_reflection_symmetries are the entries of its child blocks), in device memory.
// Built on first use and kept until the process ends (32 bytes per rotation-minimal index
// of the line's quadrant plus at most 1,680 bytes of symmetries: a few KB per section and
// dimension); the entries are immutable, so launches in flight may read them while other
// threads add entries.
struct LineTabs {
  SecDev sec;
  const uint16_t* moves = nullptr;
  const uint8_t* refl = nullptr;
  int32_t refl_off[4] = {0, 0, 0, 0};   // child axis c's symmetries start at refl + refl_off[c]
};
struct LineCache {
  std::mutex mu;
  std::map<std::pair<uint64_t, int>, LineTabs> tabs;
  uint8_t* chunk = nullptr;      // bump allocation of move rows
  size_t left = 0;
  hipStream_t upload = nullptr;  // non-blocking: uploads do not wait for the caller's work
};
constexpr int kMaxLineDevices = 64;
LineCache g_line_cache[kMaxLineDevices];

pe_status line_tabs(const LineHost& h, LineTabs* out, std::string* err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxLineDevices) { *err = "no HIP device"; return PE_ERR_DEVICE; }
  LineCache& C = g_line_cache[dev];
  Section own;
  std::memcpy(own.c, h.section.counts, 8);
  const int d = h.dev.dim;
  const auto key = std::make_pair(own.sig(), d);
  std::lock_guard<std::mutex> lock(C.mu);
  const auto it = C.tabs.find(key);
  if (it != C.tabs.end()) { *out = it->second; return PE_OK; }
  const auto& Tb = tables();
  LineTabs t;
  SecDev& sec = t.sec;
  std::memset(&sec, 0, sizeof sec);
  sec.turn = own.sum() & 1;
  std::vector<uint16_t> mt;
  const pe_status st = section_move_tables(own, &mt, sec.move_off, sec.nmoves, err, d);
  if (st != PE_OK) return st;
  const int32_t rows_off = sec.move_off[d];
  for (int a = 0; a < 4; a++) {
    sec.field_off[a] = field_pool_off(a, sec.turn, a == d, own.c[a][0], own.c[a][1]);
    const int b = rmin_bucket(own.c[a][0], own.c[a][1]);
    sec.rmin_off[a] = Tb.rmin_offsets[b];
    sec.rmin_size[a] = Tb.rmin_offsets[b + 1] - Tb.rmin_offsets[b];
    if (a != d) { sec.move_off[a] = -1; sec.nmoves[a] = 0; }
  }
  // reflection symmetries per child axis (zero unless the child is reflected)
  std::vector<uint8_t> refl;
  if (rows_off >= 0 && (h.info.section_transform & 4)) {
    for (int c = 0; c < 4; c++) {
      const int b = rmin_bucket(h.info.standard_child.counts[c][0], h.info.standard_child.counts[c][1]);
      t.refl_off[c] = int32_t(refl.size());
      for (int i = Tb.rmin_offsets[b]; i < Tb.rmin_offsets[b + 1]; i++) {
        const uint16_t q = Tb.rmin_flat[i];
        const uint16_t r = uint16_t(Tb.pack[Tb.refl[Tb.unpack[q][0]]] + 2 * Tb.pack[Tb.refl[Tb.unpack[q][1]]]);
        refl.push_back(uint8_t((Tb.rmin_inverse[r] & 3) << (2 * c)));
      }
    }
  }
  if (rows_off >= 0) {   // lines with child data: upload d's move rows and the symmetries
    const size_t rows = sizeof(uint16_t) * 16 * size_t(sec.rmin_size[d]);
    const size_t bytes = rows + refl.size();
    const size_t need = (bytes + 255) & ~size_t(255);
    if (need > C.left) {
      const size_t chunk = std::max<size_t>(need, size_t(4) << 20);
      void* p = nullptr;
      if (hipMalloc(&p, chunk) != hipSuccess) { *err = "hipMalloc(line move tables) failed"; return PE_ERR_DEVICE; }
      C.chunk = static_cast<uint8_t*>(p);
      C.left = chunk;
    }
    if (!C.upload && hipStreamCreateWithFlags(&C.upload, hipStreamNonBlocking) != hipSuccess) {
      C.upload = nullptr;
      *err = "hipStreamCreate(line tables) failed";
      return PE_ERR_DEVICE;
    }
    std::vector<uint8_t> blob(bytes);
    std::memcpy(blob.data(), mt.data() + rows_off, rows);
    if (!refl.empty()) std::memcpy(blob.data() + rows, refl.data(), refl.size());
    if (hipMemcpyAsync(C.chunk, blob.data(), bytes, hipMemcpyHostToDevice, C.upload) != hipSuccess ||
        hipStreamSynchronize(C.upload) != hipSuccess) {
      *err = "hipMemcpy(line move tables) failed";
      return PE_ERR_DEVICE;
    }
    t.moves = reinterpret_cast<const uint16_t*>(C.chunk);
    t.refl = C.chunk + rows;
    sec.move_off[d] = 0;
    C.chunk += need;
    C.left -= need;
  }
  C.tabs.emplace(key, t);
  *out = t;
  return PE_OK;
}

}  // namespace

pe_status fused_line_launch(const LineHost& h, const void* input, void* output, int accumulate, hipStream_t stream,
                            std::string* err) {
  DevTables T;
  if (device_tables(&T, err) != PE_OK) return PE_ERR_DEVICE;
  LineTabs t;
  const pe_status st = line_tabs(h, &t, err);
  if (st != PE_OK) return st;
  LineLaunch a;
  std::memset(&a, 0, sizeof a);
  a.line = h.dev;
  a.line.input = static_cast<const uint64_t*>(input);
  a.line.output = static_cast<uint64_t*>(output);
  a.line.accumulate = accumulate ? 1 : 0;
  a.line.in_base = 0;
  a.line.in_kstride = int64_t(h.dev.block_stride) + 1;   // pentago/end/compute.cc:134-139
  a.line.in_off = nullptr;
  a.line.out_off = nullptr;
  a.line.refl_base = 0;
  // the symmetries of cross axis x are those of child axis c = perm^-1(x) from the line's
  // child block on, as all_reflection_symmetries holds them
  for (int x = 0; x < 4; x++) {
    a.line.ax_rs[x] = 0;
    if (x == h.dev.dim || !t.refl) continue;
    int c = 0;
    while (h.dev.perm[c] != x) c++;
    a.line.ax_rs[x] = t.refl_off[c] + 8 * h.dev.first_child_block[c];
  }
  a.sec = t.sec;
  const unsigned grid = unsigned(h.dev.length) * kParts;
  hipLaunchKernelGGL(h.dev.slice35 ? fused_line_kernel<true> : fused_line_kernel<false>, dim3(grid), dim3(kThreads), 0,
                     stream, a, T.fields, t.moves, t.refl);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { *err = std::string("fused_line_kernel launch: ") + hipGetErrorString(e); return PE_ERR_DEVICE; }
  return PE_OK;
}

}  // namespace pe
