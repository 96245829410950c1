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
//     Per-section tables hold F0, F1 and Fm per (quadrant, rotation-minimal index).
//   * rmax(A) | rmax(B) = rmax(A | B) and T_m = R_g o shift_{y_m}: the children are shifted
//     and ORed per dimension, rearranged once per dimension, and rmax runs once per node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "device.h"
#include "end.h"
#include "super.h"
#include "tables.h"

namespace pe {

pe_status plan_create(const std::vector<LineHost>& lines, EndPlan* plan);

namespace {

__device__ __forceinline__ int local_add(int a, int b) {
  return ((((a & 0x33) + (b & 0x33)) & 0x33) + (((a & 0xcc) + (b & 0xcc)) & 0xcc)) & 0xff;
}
__device__ __forceinline__ int dot4(const int a[4], const int b[4]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}
__device__ __forceinline__ void strides4(const int s[4], int st[4]) {
  st[3] = 1; st[2] = s[3]; st[1] = s[2] * s[3]; st[0] = s[1] * s[2] * s[3];
}
__device__ __forceinline__ S4 ld(const uint64_t* p) {
  const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
  const ulonglong2 a = q[0], b = q[1];
  S4 r; r.w[0] = a.x; r.w[1] = a.y; r.w[2] = b.x; r.w[3] = b.y;
  return r;
}
__device__ __forceinline__ void st(uint64_t* p, S4 v) {
  ulonglong2* q = reinterpret_cast<ulonglong2*>(p);
  q[0] = make_ulonglong2(v.w[0], v.w[1]);
  q[1] = make_ulonglong2(v.w[2], v.w[3]);
}

enum { FH = 0, FV = 1, FLO = 2, FHI = 3, FAS = 4 };

// Field f of table t (0: F0, 1: F1, 2: Fm) for quadrant q of this node.
struct NodeFields {
  const uint64_t* base[4];
  __device__ __forceinline__ S4 get(int q, int t, int f) const { return ld(base[q] + 4 * (5 * t + f)); }
};

// Three group evaluations at once (X0: wins0, X1: wins1, XM: wins of all moves) for a
// two-quadrant group {a, b}.
__device__ __forceinline__ void group2(const NodeFields& F, int a, int fa, int b, int fb, S4& x0, S4& x1, S4& xm) {
  const S4 a0 = F.get(a, 0, fa), b0 = F.get(b, 0, fb);
  x0 = s4_and(a0, b0);
  x1 = s4_and(F.get(a, 1, fa), F.get(b, 1, fb));
  xm = s4_or(s4_and(F.get(a, 2, fa), b0), s4_and(a0, F.get(b, 2, fb)));
}
__device__ __forceinline__ void group3(const NodeFields& F, int a, int fa, int b, int fb, int c, int fc,
                                       S4& x0, S4& x1, S4& xm) {
  const S4 a0 = F.get(a, 0, fa), b0 = F.get(b, 0, fb), c0 = F.get(c, 0, fc);
  const S4 bc = s4_and(b0, c0);
  x0 = s4_and(a0, bc);
  x1 = s4_and(s4_and(F.get(a, 1, fa), F.get(b, 1, fb)), F.get(c, 1, fc));
  xm = s4_or(s4_and(F.get(a, 2, fa), bc),
             s4_and(a0, s4_or(s4_and(F.get(b, 2, fb), c0), s4_and(b0, F.get(c, 2, fc)))));
}

constexpr int kThreads = 256;
constexpr int kParts = 4096 / kThreads;

__global__ __launch_bounds__(kThreads)
void fused_blocks_kernel(const BlockDev* __restrict__ blocks, const LineDev* __restrict__ lines,
                         const SecDev* __restrict__ secs, const uint64_t* __restrict__ fields,
                         const uint8_t* __restrict__ refl, uint64_t* __restrict__ output, DevTables T) {
  const BlockDev& B = blocks[blockIdx.x / kParts];
  const SecDev& S = secs[B.sec];
  const int p = (blockIdx.x % kParts) * kThreads + threadIdx.x;
  int bs[4];
  for (int a = 0; a < 4; a++) bs[a] = min(8, S.rmin_size[a] - 8 * B.b[a]);
  if (p >= bs[0] * bs[1] * bs[2] * bs[3]) return;
  int loc[4];
  {
    int rem = p;
    loc[3] = rem % bs[3]; rem /= bs[3];
    loc[2] = rem % bs[2]; rem /= bs[2];
    loc[1] = rem % bs[1];
    loc[0] = rem / bs[1];
  }
  const int turn = S.turn;
  int gi[4];
  NodeFields F;
  uint32_t qs[4];
  for (int a = 0; a < 4; a++) {
    gi[a] = 8 * B.b[a] + loc[a];
    qs[a] = T.rmin_sides[S.rmin_off[a] + gi[a]];
    F.base[a] = fields + 4 * (int64_t(S.field_off[a]) + 15 * int64_t(gi[a]));
  }

  // ---- wins0, wins1 and the wins of every move, grouped by outer reduction ----
  S4 w0, w1, wm;
  {
    S4 x0, x1, xm, y0, y1, ym, z0, z1, zm;
    group2(F, 0, FH, 2, FH, x0, x1, xm);                // red3 then red1
    group2(F, 2, FV, 3, FV, y0, y1, ym);                // red0 then red1
    group3(F, 0, FLO, 2, FAS, 3, FLO, z0, z1, zm);      // red1
    w0 = reduce1(s4_or(s4_or(reduce3(x0), reduce0(y0)), z0));
    w1 = reduce1(s4_or(s4_or(reduce3(x1), reduce0(y1)), z1));
    wm = reduce1(s4_or(s4_or(reduce3(xm), reduce0(ym)), zm));
    group2(F, 0, FV, 1, FV, x0, x1, xm);                // red3 then red2
    group2(F, 1, FH, 3, FH, y0, y1, ym);                // red0 then red2
    group3(F, 0, FHI, 1, FAS, 3, FHI, z0, z1, zm);      // red2
    w0 = s4_or(w0, reduce2(s4_or(s4_or(reduce3(x0), reduce0(y0)), z0)));
    w1 = s4_or(w1, reduce2(s4_or(s4_or(reduce3(x1), reduce0(y1)), z1)));
    wm = s4_or(wm, reduce2(s4_or(s4_or(reduce3(xm), reduce0(ym)), zm)));
    group3(F, 1, FLO, 0, FAS, 2, FLO, x0, x1, xm);      // red3
    group3(F, 1, FHI, 3, FAS, 2, FHI, y0, y1, ym);      // red0
    w0 = s4_or(w0, s4_or(reduce3(x0), reduce0(y0)));
    w1 = s4_or(w1, s4_or(reduce3(x1), reduce0(y1)));
    wm = s4_or(wm, s4_or(reduce3(xm), reduce0(ym)));
  }

  // ---- children, one line per open dimension ----
  S4 t0 = s4_zero(), t1 = s4_zero();
  for (int d = 0; d < 4; d++) {
    if (B.line[d] < 0) continue;
    const LineDev& L = lines[B.line[d]];
    const int cd = L.child_dim;
    int full_base[4];
    for (int a = 0; a < 4; a++) full_base[a] = a == d ? 0 : loc[a];
    int obs[4];
    for (int a = 0; a < 4; a++) obs[a] = a == d ? 8 : bs[a];
    int ibs[4], fcb[4], rcb[4];
    for (int a = 0; a < 4; a++) { ibs[a] = obs[L.perm[a]]; fcb[a] = full_base[L.perm[a]]; }
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
    const uint8_t* syms = refl + L.refl_base;
    int local_sym = 0;
    if (L.rs_off[4] > 0)
      for (int a = 0; a < 4; a++)
        if (a != cd) local_sym |= syms[L.rs_off[a] + fcb[a]];
    const int g = L.inv_global;
    const int base_local = local_add(L.inv_local, local_sym);
    const uint32_t sq0 = turn ? qs[d] >> 16 : qs[d] & 0x1ff, sq1 = turn ? qs[d] & 0x1ff : qs[d] >> 16;
    const int quad = T.rmin_flat[S.rmin_off[d] + gi[d]];
    S4 acc0 = s4_zero(), acc1 = s4_zero();
    for (uint32_t m = 511 & ~(sq0 | sq1); m; m &= m - 1) {
      const uint32_t mv = m & (0u - m);
      const int ir = T.rmin_inverse[quad + (1 + turn) * T.pack[mv]];
      const int j = ir >> 2;
      const int y = local_add(local_add(T.commute[g * 256 + ((ir & 3) << (2 * d))], base_local),
                              syms[L.rs_off[cd] + j]);
      const int64_t cidx = L.in_off[j >> 3] + int64_t((j >> 3) == L.child_length - 1 ? last_input_base : input_base) +
                           int64_t(input_stride) * ((j ^ int(j < line_moves)) & 7);
      const uint64_t* c = L.input + 8 * cidx;
      acc0 = s4_or(acc0, shift_local(s4_not(ld(c)), y));
      acc1 = s4_or(acc1, shift_local(s4_not(ld(c + 4)), y));
    }
    t0 = s4_or(t0, rearrange(acc0, g));
    t1 = s4_or(t1, rearrange(acc1, g));
  }
  const S4 wins = s4_or(wm, rmax(t1));
  const S4 nl = rmax(t0);
  const S4 imm = s4_or(w0, w1);
  uint64_t* o = output + 8 * (B.out + p);
  st(o, s4_or(s4_andnot(wins, imm), s4_andnot(w0, w1)));
  st(o + 4, s4_or(s4_andnot(nl, imm), w0));
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
                            std::map<std::tuple<int, int, int, int>, int>* index, std::string* err) {
  Section s;
  std::memcpy(s.c, out.section.counts, 8);
  int shape[4];
  if (!s.valid_counts() || !s.shape(shape)) { *err = "invalid output section"; return PE_ERR_INVALID; }
  if (s.sum() > 34) { *err = "slab plans need sections with at most 34 stones"; return PE_ERR_INVALID; }
  for (int a = 0; a < 4; a++)
    if (out.lo[a] < 0 || out.hi[a] <= out.lo[a] || 8 * (out.hi[a] - 1) >= shape[a]) { *err = "output box outside the section"; return PE_ERR_INVALID; }
  for (int d = 0; d < 4; d++) {
    if (s.c[d][0] + s.c[d][1] >= 9) continue;
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
          const int cd = h.info.child_dimension;
          const pe_slab* found = nullptr;
          for (int c = 0; c < nch && !found; c++) {
            if (!same_section(ch[c].section, h.info.standard_child)) continue;
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

pe_status fused_plan_create(const pe_slab& out, int nch, const pe_slab* ch, FusedPlan* plan, std::string* err) {
  DevTables T;
  if (device_tables(&T, err) != PE_OK) return PE_ERR_DEVICE;
  std::vector<LineHost> lines;
  std::map<std::tuple<int, int, int, int>, int> index;
  pe_status st = slab_lines(out, nch, ch, &lines, &index, err);
  if (st != PE_OK) return st;
  Section s;
  std::memcpy(s.c, out.section.counts, 8);
  SecDev sec;
  std::memset(&sec, 0, sizeof sec);
  sec.turn = s.sum() & 1;
  std::vector<uint64_t> fields;
  section_field_tables(s, &fields, sec.field_off);
  const auto& Tb = tables();
  for (int a = 0; a < 4; a++) {
    const int b = rmin_bucket(s.c[a][0], s.c[a][1]);
    sec.rmin_off[a] = Tb.rmin_offsets[b];
    sec.rmin_size[a] = Tb.rmin_offsets[b + 1] - Tb.rmin_offsets[b];
  }
  std::vector<BlockDev> blocks;
  int64_t nodes = 0;
  int b[4];
  for (b[0] = out.lo[0]; b[0] < out.hi[0]; b[0]++)
    for (b[1] = out.lo[1]; b[1] < out.hi[1]; b[1]++)
      for (b[2] = out.lo[2]; b[2] < out.hi[2]; b[2]++)
        for (b[3] = out.lo[3]; b[3] < out.hi[3]; b[3]++) {
          BlockDev B;
          B.sec = 0;
          int n = 1;
          for (int a = 0; a < 4; a++) { B.b[a] = b[a]; n *= std::min(8, sec.rmin_size[a] - 8 * b[a]); }
          for (int d = 0; d < 4; d++) {
            B.line[d] = -1;
            if (s.c[d][0] + s.c[d][1] >= 9) continue;
            int o[3], k = 0;
            for (int a = 0; a < 4; a++) if (a != d) o[k++] = b[a];
            B.line[d] = index.at(std::make_tuple(d, o[0], o[1], o[2]));
          }
          B.out = 4096 * slab_index(out, b);
          blocks.push_back(B);
          nodes += n;
        }
  // blob: blocks | lines | section | fields | offsets | symmetries
  std::vector<LineDev> devs(lines.size());
  std::vector<int64_t> offs;
  std::vector<uint8_t> refl;
  std::vector<int64_t> at(lines.size());
  for (size_t k = 0; k < lines.size(); k++) {
    devs[k] = lines[k].dev;
    devs[k].refl_base = int32_t(refl.size());
    refl.insert(refl.end(), lines[k].refl_syms.begin(), lines[k].refl_syms.end());
    at[k] = int64_t(offs.size());
    offs.insert(offs.end(), lines[k].in_off.begin(), lines[k].in_off.end());
  }
  refl.push_back(0);
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t o_lines = up(sizeof(BlockDev) * blocks.size());
  const size_t o_secs = o_lines + up(sizeof(LineDev) * devs.size());
  const size_t o_fields = o_secs + up(sizeof(SecDev));
  const size_t o_offs = o_fields + up(sizeof(uint64_t) * fields.size());
  const size_t o_refl = o_offs + up(sizeof(int64_t) * offs.size());
  const size_t bytes = o_refl + refl.size();
  void* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) { *err = "hipMalloc(fused plan) failed"; return PE_ERR_DEVICE; }
  auto* base = static_cast<uint8_t*>(d);
  for (size_t k = 0; k < devs.size(); k++) devs[k].in_off = reinterpret_cast<const int64_t*>(base + o_offs) + at[k];
  std::vector<uint8_t> blob(bytes, 0);
  std::memcpy(blob.data(), blocks.data(), sizeof(BlockDev) * blocks.size());
  std::memcpy(blob.data() + o_lines, devs.data(), sizeof(LineDev) * devs.size());
  std::memcpy(blob.data() + o_secs, &sec, sizeof sec);
  std::memcpy(blob.data() + o_fields, fields.data(), sizeof(uint64_t) * fields.size());
  std::memcpy(blob.data() + o_offs, offs.data(), sizeof(int64_t) * offs.size());
  std::memcpy(blob.data() + o_refl, refl.data(), refl.size());
  if (hipMemcpy(d, blob.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    *err = "hipMemcpy(fused plan) failed";
    return PE_ERR_DEVICE;
  }
  plan->blob = d;
  plan->nblocks = int(blocks.size());
  plan->off_lines = o_lines;
  plan->off_secs = o_secs;
  plan->off_fields = o_fields;
  plan->off_refl = o_refl;
  plan->output = static_cast<uint64_t*>(out.data);
  plan->output_nodes = nodes;
  return PE_OK;
}

pe_status fused_plan_run(const FusedPlan& plan, hipStream_t stream) {
  if (!plan.nblocks) return PE_OK;
  DevTables T;
  std::string err;
  if (device_tables(&T, &err) != PE_OK) { set_error(err); return PE_ERR_DEVICE; }
  auto* base = static_cast<uint8_t*>(plan.blob);
  const int64_t grid = int64_t(plan.nblocks) * kParts;
  if (grid > 0x7fffffff) { set_error("fused plan too large"); return PE_ERR_CAPACITY; }
  hipLaunchKernelGGL(fused_blocks_kernel, dim3(unsigned(grid)), dim3(kThreads), 0, stream,
                     reinterpret_cast<const BlockDev*>(base), reinterpret_cast<const LineDev*>(base + plan.off_lines),
                     reinterpret_cast<const SecDev*>(base + plan.off_secs),
                     reinterpret_cast<const uint64_t*>(base + plan.off_fields), base + plan.off_refl, plan.output, T);
  PE_HIP(hipGetLastError(), "fused_blocks_kernel launch");
  return PE_OK;
}

void fused_plan_destroy(FusedPlan* plan) {
  if (plan->blob) (void)hipFree(plan->blob);
  plan->blob = nullptr;
  plan->nblocks = 0;
}

// Mode 1: four line passes over the slab, later passes OR-accumulating.
pe_status slab_line_passes(const pe_slab& out, int nch, const pe_slab* ch, std::vector<EndPlan>* passes,
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
  int pass = 0;
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
    EndPlan p;
    st = plan_create(group, &p);
    if (st != PE_OK) return st;
    passes->push_back(p);
    pass++;
  }
  return PE_OK;
}

}  // namespace pe
