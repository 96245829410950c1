/* Oracle: symmetry-aware win counts.  TEST INFRASTRUCTURE ONLY (see po.h).
 *
 * popcounts_over_stabilizers: pentago/base/count.cc:235-262.
 * Section counts: count_block_wins (pentago/end/block_store.cc:330-349) summed over every
 * block of a section, with node values from the recursive evaluator (po_eval_node). */
#include <string.h>

#include "po.h"

/* log2 of the number of single-quadrant rotations fixing quadrant q (count.cc:235-245) */
static int log_stabilizers_quadrant(uint16_t q) {
  const uint16_t s0 = po_unpack(q, 0), s1 = po_unpack(q, 1);
  const uint16_t s0r = po_rotate_left(s0), s1r = po_rotate_left(s1);
  if (s0 == s0r && s1 == s1r) return 2;
  return s0 == po_rotate_left(s0r) && s1 == po_rotate_left(s1r) ? 1 : 0;
}

static int popcount256(po_super s) {
  int n = 0;
  for (int i = 0; i < 4; i++) n += __builtin_popcountll(s.w[i]);
  return n;
}

int po_popcounts_over_stabilizers(uint64_t board, const po_super wins[2], uint16_t out[3]) {
  int shift = 0;
  for (int q = 0; q < 4; q++) shift += log_stabilizers_quadrant((uint16_t)(board >> (16 * q) & 0xffff));
  const int mask = (1 << shift) - 1, w0 = popcount256(wins[0]), w1 = popcount256(wins[1]);
  out[0] = (uint16_t)(w0 >> shift);
  out[1] = (uint16_t)(w1 >> shift);
  out[2] = (uint16_t)(256 >> shift);
  return (w0 & mask) || (w1 & mask) ? -1 : 0;
}

int po_section_counts(po_section s, uint64_t out[3]) {
  int shape[4];
  if (po_section_shape(s, shape)) return -1;
  const uint16_t* rm[4];
  for (int q = 0; q < 4; q++) {
    int moved;
    if (po_rmin(s.c[q][0], s.c[q][1], &rm[q], &moved) != shape[q]) return -1;
  }
  memset(out, 0, 3 * sizeof(uint64_t));
  int idx[4];
  for (idx[0] = 0; idx[0] < shape[0]; idx[0]++)
    for (idx[1] = 0; idx[1] < shape[1]; idx[1]++)
      for (idx[2] = 0; idx[2] < shape[2]; idx[2]++)
        for (idx[3] = 0; idx[3] < shape[3]; idx[3]++) {
          po_super v[2];
          if (po_eval_node(s, idx, v)) return -1;
          uint64_t board = 0;
          for (int q = 0; q < 4; q++) board |= (uint64_t)rm[q][idx[q]] << (16 * q);
          uint16_t c[3];
          if (po_popcounts_over_stabilizers(board, v, c)) return -2;
          for (int k = 0; k < 3; k++) out[k] += c[k];
        }
  return 0;
}
