// Host-side block line index state (pentago/end/compute.cc:24-123), own implementation.
#include "end.h"

#include <cstring>

#include "tables.h"

namespace pe {

bool Section::valid_counts() const {
  for (int q = 0; q < 4; q++)
    if (c[q][0] + c[q][1] > 9) return false;
  return true;
}

bool Section::shape(int out[4]) const {
  for (int q = 0; q < 4; q++) {
    const int n = rmin_count(c[q][0], c[q][1]);
    if (n < 0) return false;
    out[q] = n;
  }
  return true;
}

// Quadrant q of g(s) is quadrant source[g&3][q] of s; reflection then swaps 0 and 3
// (pentago/base/section.cc:79-87).
Section Section::transform(int g) const {
  static const int source[4][4] = {{0, 1, 2, 3}, {1, 3, 0, 2}, {3, 2, 1, 0}, {2, 0, 3, 1}};
  Section t;
  for (int q = 0; q < 4; q++) std::memcpy(t.c[q], c[source[g & 3][q]], 2);
  if (g & 4)
    for (int k = 0; k < 2; k++) std::swap(t.c[0][k], t.c[3][k]);
  return t;
}

int Section::standardize(Section* out) const {
  Section best = *this;
  int best_g = 0;
  for (int g = 1; g < 8; g++) {
    const Section t = transform(g);
    if (t.sig() < best.sig()) { best = t; best_g = g; }
  }
  *out = best;
  return best_g;
}

namespace {

inline int local_add(int a, int b) {   // Z_4^4 addition of packed local rotations
  return ((((a & 0x33) + (b & 0x33)) & 0x33) + (((a & 0xcc) + (b & 0xcc)) & 0xcc)) & 0xff;
}
// Group product (pentago/base/symmetry.h:83-91).
void sym_mul(int ag, int al, int bg, int bl, int* g, int* l) {
  const int l2 = tables().commute[bg][al];
  *g = ((ag ^ bg) & 4) | (((ag ^ ((ag & (bg >> 2)) << 1)) + bg) & 3);
  *l = local_add(l2, bl);
}
inline int invert_global(int g) { return g ^ (((~g >> 2) & (g & 1)) << 1); }

}  // namespace

int64_t field_pool_words() { return int64_t(16) * tables().rmin_flat_size * 60; }

int32_t field_pool_off(int q, int turn, int moves, int black, int white) {
  const auto& T = tables();
  const int b = rmin_bucket(black, white);
  return int32_t(((q * 2 + turn) * 2 + (moves ? 1 : 0)) * T.rmin_flat_size * 15 + T.rmin_offsets[b] * 15);
}

void build_field_pool(std::vector<uint64_t>* out) {
  const auto& T = tables();
  out->assign(size_t(field_pool_words()), 0);
  for (int q = 0; q < 4; q++)
    for (int turn = 0; turn < 2; turn++)
      for (int moves = 0; moves < 2; moves++) {
        uint64_t* o = out->data() + size_t(((q * 2 + turn) * 2 + moves) * T.rmin_flat_size) * 60;
        for (int i = 0; i < T.rmin_flat_size; i++, o += 60) {
          const uint16_t quad = T.rmin_flat[i];
          const uint16_t mover = T.unpack[quad][turn], other = T.unpack[quad][1 - turn];
          uint64_t fm[5][4] = {};
          for (int e = 0; e < 9 && moves; e++) {
            if ((mover | other) >> e & 1) continue;
            const uint64_t* f = T.superwin + size_t((q * 512 + (mover | 1 << e)) * 5) * 4;
            for (int k = 0; k < 20; k++) fm[k / 4][k % 4] |= f[k];
          }
          const uint64_t* f0 = T.superwin + size_t((q * 512 + mover) * 5) * 4;
          const uint64_t* f1 = T.superwin + size_t((q * 512 + other) * 5) * 4;
          // word-major: for word k, table t (F0, F1, Fm), field f
          for (int k = 0; k < 4; k++)
            for (int f = 0; f < 5; f++) {
              o[15 * k + f] = f0[4 * f + k];
              o[15 * k + 5 + f] = f1[4 * f + k];
              o[15 * k + 10 + f] = fm[f][k];
            }
        }
      }
}

pe_status section_move_tables(const Section& s, std::vector<uint16_t>* out, int32_t move_off[4], int32_t nmoves[4],
                              std::string* err, int only) {
  const auto& T = tables();
  const int turn = s.sum() & 1;
  int shape[4];
  if (!s.shape(shape)) { *err = "invalid section"; return PE_ERR_INVALID; }
  out->clear();
  for (int d = 0; d < 4; d++) {
    move_off[d] = -1;
    nmoves[d] = 9 - s.c[d][0] - s.c[d][1];
    if (nmoves[d] <= 0) { nmoves[d] = 0; continue; }
    if (s.sum() == 35) continue;   // the move fills the board: no child data to address
    if (only >= 0 && d != only) continue;
    // the child side of the moves depends only on (section, d): use the line at block 0
    pe_line L;
    std::memset(&L, 0, sizeof L);
    std::memcpy(L.section.counts, s.c, sizeof s.c);
    L.dimension = uint8_t(d);
    L.length = uint8_t((shape[d] + 7) / 8);
    LineHost h;
    const pe_status st = make_line(L, &h, err);
    if (st != PE_OK) return st;
    const LineDev& D = h.dev;
    const int cd = D.child_dim, g = D.inv_global, lm = D.reflection_moves[cd];
    const uint8_t* cd_syms = h.refl_syms.data() + D.rs_off[cd];
    move_off[d] = int32_t(out->size());
    const int b = rmin_bucket(s.c[d][0], s.c[d][1]);
    for (int i = T.rmin_offsets[b]; i < T.rmin_offsets[b + 1]; i++) {
      const uint16_t quad = T.rmin_flat[i];
      const int filled = T.unpack[quad][0] | T.unpack[quad][1];
      uint16_t row[16] = {};
      int m = 0;
      for (int e = 0; e < 9; e++) {
        if (filled >> e & 1) continue;
        const int ir = T.rmin_inverse[quad + (1 + turn) * T.pack[1 << e]];
        const int j = ir >> 2;
        const int k = ((T.commute[g][(ir & 3) << (2 * d)] + (D.reflected ? cd_syms[j] : 0)) >> (2 * cd)) & 3;
        const int kk = j >> 3, within = (j ^ int(j < lm)) & 7;
        if (kk >= 64) { *err = "internal: child line longer than 64 blocks"; return PE_ERR_INVALID; }
        row[m++] = uint16_t(kk | within << 6 | k << 9);
      }
      if (m != nmoves[d]) { *err = "internal: move count mismatch"; return PE_ERR_INVALID; }
      out->insert(out->end(), row, row + 16);
    }
  }
  return PE_OK;
}

pe_status make_line(const pe_line& line, LineHost* out, std::string* err) {
  const auto& T = tables();
  std::memset(&out->dev, 0, sizeof out->dev);
  std::memset(&out->info, 0, sizeof out->info);
  out->refl_syms.clear();
  out->section = line.section;
  LineDev& D = out->dev;
  pe_line_info& I = out->info;
  Section s;
  std::memcpy(s.c, line.section.counts, sizeof s.c);
  const int dim = line.dimension;
  int shape[4];
  if (dim > 3 || !s.valid_counts() || !s.shape(shape)) { *err = "invalid section"; return PE_ERR_INVALID; }
  if (s.c[dim][0] + s.c[dim][1] >= 9) { *err = "line dimension has no empty cell"; return PE_ERR_INVALID; }
  if (s.sum() > 35) { *err = "section has more than 35 stones"; return PE_ERR_INVALID; }
  const int blocks_d = (shape[dim] + 7) / 8;
  if (line.length != blocks_d) { *err = "line length does not match the section"; return PE_ERR_INVALID; }
  int first_block[4];
  for (int a = 0, k = 0; a < 4; a++) first_block[a] = a == dim ? 0 : line.block_base[k++];
  for (int a = 0; a < 4; a++)
    if (a != dim && 8 * first_block[a] >= shape[a]) { *err = "block base outside the section"; return PE_ERR_INVALID; }

  const Section child = s.child(dim);
  int cshape[4];
  if (!child.shape(cshape)) { *err = "invalid child section"; return PE_ERR_INVALID; }
  const bool slice35 = s.sum() == 35;
  for (int a = 0; a < 4; a++) {
    I.output_shape[a] = a == dim ? shape[a] : std::min(shape[a], 8 * (first_block[a] + 1)) - 8 * first_block[a];
    I.input_shape[a] = a == dim ? cshape[a] : I.output_shape[a];
  }
  if (slice35) std::memset(I.input_shape, 0, sizeof I.input_shape);
  I.dimension = dim;
  I.length = line.length;
  I.block_stride = 8;
  for (int a = 0; a < 4; a++) if (a != dim) I.block_stride *= I.output_shape[a];

  Section stdc;
  const int st = child.standardize(&stdc);
  std::memcpy(I.standard_child.counts, stdc.c, sizeof stdc.c);
  I.section_transform = st;
  // permutation[i] = where quadrant i of the standard child sits in the line's section
  // (section_t::quadrant_permutation(invert_global(st)), pentago/base/section.cc:108-116)
  {
    Section marker;
    for (int q = 0; q < 4; q++) { marker.c[q][0] = uint8_t(q); marker.c[q][1] = 0; }
    const Section t = marker.transform(invert_global(st));
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++)
        if (t.c[j][0] == i) I.permutation[i] = j;
  }
  for (int a = 0; a < 4; a++) if (I.permutation[a] == dim) I.child_dimension = a;
  I.input_blocks = (I.input_shape[dim] + 7) / 8;
  for (int a = 0; a < 4; a++) I.first_child_block[a] = first_block[I.permutation[a]];
  int64_t in_nodes = 1, out_nodes = 1;
  for (int a = 0; a < 4; a++) { in_nodes *= I.input_shape[a]; out_nodes *= I.output_shape[a]; }
  I.input_nodes = slice35 ? 0 : in_nodes + I.input_blocks;
  I.output_nodes = out_nodes;

  // inverse_transform = (-r, r*(1,1,1,1)) * (reflect, 0), r = st & 3 (compute.cc:69)
  int ig, il;
  sym_mul((4 - (st & 3)) & 3, 0x55 * (st & 3), st & 4, 0, &ig, &il);

  // Reflection bookkeeping (compute.cc:96-122)
  int child_moved[4] = {0, 0, 0, 0};
  const uint16_t* child_rmin[4];
  for (int a = 0; a < 4; a++) {
    const int b = rmin_bucket(stdc.c[a][0], stdc.c[a][1]);
    child_rmin[a] = T.rmin_flat + T.rmin_offsets[b];
    if (st & 4) child_moved[a] = T.rmin_moved[b];
  }
  int sis[4];
  for (int a = 0; a < 4; a++) sis[a] = I.input_shape[I.permutation[a]];
  out->refl_syms.assign(sis[0] + sis[1] + sis[2] + sis[3], 0);
  int off = 0;
  D.rs_off[0] = 0;
  for (int a = 0; a < 4; a++) {
    if (st & 4)
      for (int j = 0; j < sis[a]; j++) {
        const uint16_t q = child_rmin[a][8 * I.first_child_block[a] + j];
        const uint16_t r = uint16_t(T.pack[T.refl[T.unpack[q][0]]] + 2 * T.pack[T.refl[T.unpack[q][1]]]);
        out->refl_syms[off + j] = uint8_t((T.rmin_inverse[r] & 3) << (2 * a));
      }
    off += sis[a];
    D.rs_off[a + 1] = off;
  }

  D.dim = dim;
  D.child_dim = I.child_dimension;
  D.turn = s.sum() & 1;
  D.slice35 = slice35;
  D.length = I.length;
  D.child_length = I.input_blocks;
  D.block_stride = I.block_stride;
  D.in_dim_size = I.input_shape[dim];
  for (int a = 0; a < 4; a++) {
    D.out_shape[a] = I.output_shape[a];
    D.perm[a] = I.permutation[a];
    D.first_block[a] = first_block[a];
    D.first_child_block[a] = I.first_child_block[a];
    D.reflection_moves[a] = child_moved[a];
    const int b = rmin_bucket(s.c[a][0], s.c[a][1]);
    D.rmin_off[a] = T.rmin_offsets[b];
    D.rmin_size[a] = T.rmin_offsets[b + 1] - T.rmin_offsets[b];
  }
  for (int a = 0, k = 0; a < 4; a++) if (a != dim) D.cross_shape[k++] = I.output_shape[a];
  D.cross_count = D.cross_shape[0] * D.cross_shape[1] * D.cross_shape[2];
  D.inv_global = ig;
  D.inv_local = il;
  // The local rotation of the moved quadrant lands on one quadrant after conjugation by ig.
  {
    const int c = T.commute[ig][1 << (2 * dim)];
    D.child_coord = -1;
    for (int q = 0; q < 4; q++) if (c >> (2 * q) & 3) D.child_coord = q;
  }
  D.i_begin = 0;
  D.i_end = D.rmin_size[dim];
  D.accumulate = 0;
  D.reflected = (st & 4) ? 1 : 0;
  // unreflected lines: (ig, il) = (-r, r (1,1,1,1)), a coordinate permutation plus a shift of
  // every coordinate by r, which the fused kernel applies as generated straight-line code
  D.fast_t = (st & 4) || il != 0x55 * (st & 3) || ig != ((4 - (st & 3)) & 3) ? -1 : ig;
  if (D.child_coord != I.child_dimension) { *err = "internal: moved quadrant does not map to the child dimension"; return PE_ERR_INVALID; }
  // child block strides per output axis (compute.cc:254-271 with the block shape fixed per line)
  {
    const int cd = I.child_dimension;
    int ibs[4];
    for (int c = 0; c < 4; c++) ibs[c] = c == cd ? 8 : I.output_shape[I.permutation[c]];
    auto strides = [](const int s[4], int out[4]) {
      out[3] = 1; out[2] = s[3]; out[1] = s[2] * s[3]; out[0] = s[1] * s[2] * s[3];
    };
    int ist[4], ist2[4];
    strides(ibs, ist);
    ibs[cd] = (I.input_shape[dim] & 7) ? (I.input_shape[dim] & 7) : 8;
    strides(ibs, ist2);
    D.input_stride = ist[cd];
    for (int a = 0; a < 4; a++) {
      D.ax_stride[a] = D.ax_stride_last[a] = D.ax_fcn[a] = D.ax_moves[a] = D.ax_rs[a] = 0;
      if (a == dim) continue;
      int c = 0;
      while (I.permutation[c] != a) c++;
      D.ax_stride[a] = ist[c];
      D.ax_stride_last[a] = ist2[c];
      D.ax_fcn[a] = 8 * I.first_child_block[c];
      D.ax_moves[a] = child_moved[c];
      D.ax_rs[a] = D.rs_off[c];
    }
  }
  D.in_off = nullptr;
  D.out_off = nullptr;
  return PE_OK;
}

}  // namespace pe
