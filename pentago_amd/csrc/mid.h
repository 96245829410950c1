// Midsolve: retrograde sweep over the positions below one board with >= 18 stones.
//
// Layout follows pentago/mid/internal.h: level n (stones placed on the empty cells) holds
// the positions (s1, s0p) of a two-level combinatorial index, stored in one of two
// ping-pong halves of a workspace of bottleneck(spots) entries.  Entries here are 32-byte
// (win, notlose) halfsuper pairs, so both of the reference's aggressive passes
// (midengine.cc:155-160) run as one: the reference's pass a at level n is the win component
// when a ^ (n & 1) == 1 and the notlose component otherwise, and the two passes exchange
// roles through ~ exactly as ~(win, notlose) = (~notlose, ~win) does.
#pragma once
#include <cstdint>

#include "../../include/pentago_amd.h"

namespace pe {

struct MidGrab { int32_t ny, lo, size; };

// Per-board constants (internal.h:138-150 info_t, plus position prefix sums for flat grids).
struct MidInfo {
  uint64_t side[2];
  uint32_t ply;
  int32_t spots;
  int32_t workspace;                 // bottleneck(spots)
  uint8_t empty[20];
  MidGrab spaces[20];                // output = spaces[n], input = spaces[n+1]
  int32_t sets0_off[20], sets1p_off[20], cs1ps_off[20], wins1_off[20];
  int64_t positions[20];             // sets0(n).size * sets1p(n).size
};

// 256-byte per-s0 table (internal.h:182-188 set0_info_t)
struct Set0Info {
  uint64_t wins0[2];
  uint16_t child_s0s[18];
  uint16_t offset0[90];
  uint8_t empty1[18];
  uint8_t pad[2];
};
static_assert(sizeof(Set0Info) == 256, "set0 info is 256 bytes");

struct Wins1 { uint64_t after[2], before[2]; };

int choose_host(int n, int k);
int bottleneck(int spots);
bool make_mid_info(const pe_high_board& b, MidInfo* I);

// Device-side scratch sizes for one board (in elements of each table)
struct MidSizes {
  int64_t set0;      // Set0Info entries over all levels
  int64_t wins1;     // Wins1 entries
  int64_t cs1ps;     // uint16 entries
  int64_t set1p;     // uint32 entries: the set s1p of each level as a bit mask over empty1 slots
  int64_t work;      // 32-byte workspace entries
};
MidSizes mid_sizes(const MidInfo& I);

// high_board_t helpers (pentago/high/board.{h,cc})
int high_count(const pe_high_board& b);
bool high_done_and_value(const pe_high_board& b, int* value);
pe_high_board high_place(const pe_high_board& b, int bit);
pe_high_board high_rotate(const pe_high_board& b, int q, int d);
uint64_t high_raw(const pe_high_board& b);
int mid_traverse(const pe_high_board& b, const pe_halfsupers* supers, pe_mid_value* out);

}  // namespace pe
