/* Brute-force minimax over real pentago moves (place a stone, then rotate one quadrant by
 * +-90 degrees), used to pin the oracle on positions with few empty cells.  It follows the
 * game rules the engine encodes: a five made by the placement wins at once
 * (pentago/end/compute.cc:325-327); after the rotation both fives decide (win, loss or tie);
 * a full board without fives is a tie (pentago/high/board.cc:141-149).
 * TEST INFRASTRUCTURE ONLY (see po.h). */
#include "po.h"

static const uint64_t SIDE_MASK = 0x01ff01ff01ff01ffull;

static uint64_t rotq(uint64_t side, int q, int left) {
  const uint16_t old = (uint16_t)(side >> 16*q & 0x1ff);
  const uint16_t r = left ? po_rotate_left(old) : po_rotate_right(old);
  return (side & ~((uint64_t)0x1ff << 16*q)) | (uint64_t)r << 16*q;
}

int po_brute_value(uint64_t me, uint64_t them) {
  const uint64_t empty = SIDE_MASK & ~(me | them);
  if (!empty) return 0;
  int best = -1;
  for (uint64_t e = empty; e; e &= e - 1) {
    const uint64_t me2 = me | (e & -e);
    if (po_won(me2)) return 1;
    for (int q = 0; q < 4; q++)
      for (int left = 0; left < 2; left++) {
        const uint64_t m3 = rotq(me2, q, left), t3 = rotq(them, q, left);
        const int mw = po_won(m3), tw = po_won(t3);
        int v;
        if (mw && tw) v = 0;
        else if (mw) v = 1;
        else if (tw) v = -1;
        else if (!(SIDE_MASK & ~(m3 | t3))) v = 0;
        else v = -po_brute_value(t3, m3);
        if (v > best) best = v;
        if (best == 1) return 1;
      }
  }
  return best;
}

/* bit r of the result = value of the position with quadrant q rotated left (r >> 2q) & 3 times,
 * as transform_board(symmetry_t(0, r), .) does (pentago/base/symmetry.cc:21-44). */
void po_brute_super(uint64_t me, uint64_t them, po_super out[2]) {
  for (int i = 0; i < 4; i++) { out[0].w[i] = 0; out[1].w[i] = 0; }
  for (int r = 0; r < 256; r++) {
    const po_sym s = {0, (uint8_t)r};
    const uint64_t m = po_transform_side(s, me), t = po_transform_side(s, them);
    const int mw = po_won(m), tw = po_won(t);
    int win, wot;
    if (mw || tw) { win = mw && !tw; wot = mw || !tw; }
    else { const int v = po_brute_value(m, t); win = v == 1; wot = v >= 0; }
    if (win) out[0].w[r >> 6] |= (uint64_t)1 << (r & 63);
    if (wot) out[1].w[r >> 6] |= (uint64_t)1 << (r & 63);
  }
}
