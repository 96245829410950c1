// Midsolve host side: per-board constants and result traversal (own implementation of
// pentago/mid/internal.h:37-69 and pentago/mid/midengine.cc:127-158).
#include "mid.h"

#include <algorithm>
#include <cstring>

#include "tables.h"

namespace pe {

int choose_host(int n, int k) {
  if (n < 0 || k < 0 || k > n) return 0;
  long long r = 1;
  k = std::min(k, n - k);
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return int(r);
}

// Largest pair of adjacent level sizes (internal.h:37-47)
int bottleneck(int spots) {
  int worst = 0, prev = 1;
  for (int n = 0; n <= spots; n++) {
    int next = choose_host(spots, n + 1);
    if (next) next *= choose_host(n + 1, (n + 1) / 2);
    worst = std::max(worst, prev + next);
    prev = next;
  }
  return worst;
}

int high_count(const pe_high_board& b) { return int((b.ply + 1) >> 1); }

bool make_mid_info(const pe_high_board& b, MidInfo* I) {
  std::memset(I, 0, sizeof *I);
  const uint64_t mask = 0x01ff01ff01ff01ffull;
  if ((b.side[0] & b.side[1]) || ((b.side[0] | b.side[1]) & ~mask)) return false;
  const int count = __builtin_popcountll(b.side[0] | b.side[1]);
  if (high_count(b) != count || b.ply > 72) return false;
  I->side[0] = b.side[0];
  I->side[1] = b.side[1];
  I->ply = b.ply;
  I->spots = 36 - count;
  if (I->spots > 18) return false;
  const uint64_t empty = mask & ~(b.side[0] | b.side[1]);
  for (int i = 0, k = 0; i < 64; i++) if (empty >> i & 1) I->empty[k++] = uint8_t(i);
  const int ws = bottleneck(I->spots);
  I->workspace = ws;
  for (int n = 0; n <= I->spots; n++) {
    const int k0 = n >> 1, k1 = n - k0;
    const int sets0 = choose_host(I->spots, k0), sets1 = choose_host(I->spots, k1);
    const int sets0p = choose_host(I->spots - k1, k0), sets1p = choose_host(I->spots - k0, k1);
    I->sets0_off[n + 1] = I->sets0_off[n] + sets0;
    I->sets1p_off[n + 1] = I->sets1p_off[n] + sets1p;
    I->cs1ps_off[n + 1] = I->cs1ps_off[n] + sets1p * (I->spots - k0);
    I->wins1_off[n + 1] = I->wins1_off[n] + sets1;
    I->spaces[n].ny = sets0p;
    I->spaces[n].size = sets1 * sets0p;
    I->spaces[n].lo = (n & 1) ? ws - sets1 * sets0p : 0;
    I->positions[n] = int64_t(sets0) * sets1p;
  }
  I->spaces[I->spots + 1] = MidGrab{0, 0, 0};
  return true;
}

MidSizes mid_sizes(const MidInfo& I) {
  MidSizes s;
  s.set0 = I.sets0_off[I.spots + 1];
  s.wins1 = I.wins1_off[I.spots + 1];
  s.cs1ps = I.cs1ps_off[I.spots + 1];
  s.set1p = I.sets1p_off[I.spots + 1];
  s.work = I.workspace;
  return s;
}

// ---- high board (pentago/high/board.cc) ----
static bool won_side(uint64_t side) {
  const auto& T = tables();
  const uint64_t c = T.win_contrib[0][side & 0x1ff] + T.win_contrib[1][side >> 16 & 0x1ff] +
                     T.win_contrib[2][side >> 32 & 0x1ff] + T.win_contrib[3][side >> 48 & 0x1ff];
  return (c & (c >> 1) & 0x55) || (c & (0xaaaaaaaaaaaaaaaaull << 8));
}
bool high_done_and_value(const pe_high_board& b, int* value) {
  const bool bw = won_side(b.side[0]), ww = won_side(b.side[1]);
  const int turn = (b.ply >> 1) & 1;
  if (bw || ww) { *value = bw && ww ? 0 : (bw == !turn ? 1 : -1); return true; }
  if (b.ply == 72) { *value = 0; return true; }
  *value = 0;
  return false;
}
pe_high_board high_place(const pe_high_board& b, int bit) {
  pe_high_board r = b;
  r.side[(b.ply >> 1) & 1] |= uint64_t(1) << bit;
  r.ply = b.ply + 1;
  return r;
}
pe_high_board high_rotate(const pe_high_board& b, int q, int d) {
  const auto& T = tables();
  pe_high_board r = b;
  for (int s = 0; s < 2; s++) {
    const uint16_t old = uint16_t(r.side[s] >> (16 * q) & 0x1ff);
    const uint16_t nq = d > 0 ? T.rot[old][0] : T.rot[old][1];
    r.side[s] ^= uint64_t(old ^ nq) << (16 * q);
  }
  r.ply = b.ply + 1;
  return r;
}
uint64_t high_raw(const pe_high_board& b) {
  uint64_t board = 0;
  for (int q = 0; q < 4; q++) {
    uint64_t digits = 0;
    for (int i = 8; i >= 0; i--) {
      const int j = 16 * q + i;
      digits = 3 * digits + (b.side[0] >> j & 1) + 2 * (b.side[1] >> j & 1);
    }
    board |= digits << (16 * q);
  }
  return board | uint64_t(b.ply & 1) << 63;
}

static int half_bit(const uint64_t h[2], int r) { r &= 127; return int(h[r >> 6] >> (r & 63) & 1); }

static int traverse(const pe_high_board& b, const pe_halfsupers* supers, pe_mid_value* out, int* count) {
  int value;
  if (high_done_and_value(b, &value)) {
  } else if (!(b.ply & 1)) {
    value = -1;
    const uint64_t empty = 0x01ff01ff01ff01ffull & ~(b.side[0] | b.side[1]);
    int s = 0;
    for (int bit = 0; bit < 64; bit++)
      if (empty >> bit & 1) value = std::max(value, traverse(high_place(b, bit), supers + ++s, out, count));
  } else {
    const pe_halfsupers& r = supers[0];
    value = -1;
    for (int s = 0; s < 8; s++) {
      const int q = s >> 1, d = s & 1 ? -1 : 1;
      const int idx = ((d & 3) << (2 * q)) >> 1;
      const int v = half_bit(r.win, idx) + half_bit(r.notlose, idx) - 1;
      out[*count].raw = high_raw(high_rotate(b, q, d));
      out[*count].value = v;
      out[*count].pad = 0;
      ++*count;
      value = std::max(value, -v);
    }
  }
  out[*count].raw = high_raw(b);
  out[*count].value = value;
  out[*count].pad = 0;
  ++*count;
  return value;
}

int mid_traverse(const pe_high_board& b, const pe_halfsupers* supers, pe_mid_value* out) {
  int count = 0;
  traverse(b, supers, out, &count);
  return count;
}

}  // namespace pe
