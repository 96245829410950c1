// 256-bit rotation sets ("supers") and 128-bit parity halves, as register values.
//
// A super holds one bit per local rotation r = (r0,r1,r2,r3) in Z_4^4 at bit index
// r0 + 4 r1 + 16 r2 + 64 r3 (pentago/base/superscore.h:29-39).  Here it is four u64 words:
// word r3, bit r0 + 4 r1 + 16 r2.  Every coordinate operation below is an in-register
// shift, delta swap or word rename, so one thread owns a whole super and no lane traffic
// is needed.
//
// Coordinate operations, all of the form X'(x) = X(alpha(x)):
//   shift_q(X, k):  alpha(x) = x + k e_q          (local rotation of quadrant q)
//   swap_ab(X):     alpha swaps coordinates a, b
//   negate(X):      alpha(x) = -x                 (quadrant-local reflection)
// transform_super(s, C) for s = (g, y) is C(commute[g][x] + y)
// (pentago/base/symmetry.h:84-91); commute[g] is a signed coordinate permutation, which
// rearrange<g> applies as a fixed sequence of swaps and a negation.
#pragma once
#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define PE_HD __host__ __device__ __forceinline__
#else
#define PE_HD inline
#endif

namespace pe {

struct S4 { uint64_t w[4]; };
struct H2 { uint64_t w[2]; };

PE_HD S4 s4_zero() { S4 r; r.w[0] = r.w[1] = r.w[2] = r.w[3] = 0; return r; }
PE_HD S4 s4_not(S4 a) { for (int i = 0; i < 4; i++) a.w[i] = ~a.w[i]; return a; }
PE_HD S4 s4_or(S4 a, S4 b) { for (int i = 0; i < 4; i++) a.w[i] |= b.w[i]; return a; }
PE_HD S4 s4_and(S4 a, S4 b) { for (int i = 0; i < 4; i++) a.w[i] &= b.w[i]; return a; }
PE_HD S4 s4_andnot(S4 a, S4 b) { for (int i = 0; i < 4; i++) a.w[i] &= ~b.w[i]; return a; }

// ---------------- shifts along one coordinate ----------------
PE_HD uint64_t shift0_word(uint64_t x, int k) {   // rotate each nibble right by k
  if (!k) return x;
  const uint64_t ones = 0x1111111111111111ull;
  return ((x >> k) & (ones * (0xfull >> k))) | ((x << (4 - k)) & (ones * ((0xfull << (4 - k)) & 0xf)));
}
PE_HD uint64_t shift1_word(uint64_t x, int k) {   // rotate nibbles within each 16-bit chunk
  if (!k) return x;
  const uint64_t ones = 0x0001000100010001ull;
  const int s = 4 * k;
  return ((x >> s) & (ones * (0xffffull >> s))) | ((x << (16 - s)) & (ones * ((0xffffull << (16 - s)) & 0xffff)));
}
PE_HD uint64_t shift2_word(uint64_t x, int k) {   // rotate 16-bit chunks within the word
  if (!k) return x;
  const int s = 16 * k;
  return (x >> s) | (x << (64 - s));
}
PE_HD S4 shift0(S4 a, int k) { for (int i = 0; i < 4; i++) a.w[i] = shift0_word(a.w[i], k & 3); return a; }
PE_HD S4 shift1(S4 a, int k) { for (int i = 0; i < 4; i++) a.w[i] = shift1_word(a.w[i], k & 3); return a; }
PE_HD S4 shift2(S4 a, int k) { for (int i = 0; i < 4; i++) a.w[i] = shift2_word(a.w[i], k & 3); return a; }
// m ? b : a bitwise, m all ones or all zeros.  On the device one three-input logic
// instruction per 32-bit half (truth-table index src0 * 4 + src1 * 2 + src2, 0xD8 = "src2 ?
// src1 : src0"); written as ternaries or masked and/or the compiler emits three instructions
// per half, or turns a chain of selects over the words into a private-memory array.
PE_HD uint64_t select64(uint64_t a, uint64_t b, uint32_t m) {
#ifdef __HIP_DEVICE_COMPILE__
  const uint32_t lo = __builtin_amdgcn_bitop3_b32(uint32_t(a), uint32_t(b), m, 0xD8);
  const uint32_t hi = __builtin_amdgcn_bitop3_b32(uint32_t(a >> 32), uint32_t(b >> 32), m, 0xD8);
  return uint64_t(lo) | uint64_t(hi) << 32;
#else
  const uint64_t m64 = uint64_t(0) - uint64_t(m & 1);
  return (a & ~m64) | (b & m64);
#endif
}
// Any bitwise function of three words, given by its truth table TT (bit 4a + 2b + c of TT is
// the result for input bits a, b, c): one v_bitop3_b32 per 32-bit half on the device.  On
// gfx950 it issues at the rate of a two-input and/or/xor, twice the rate of v_or3_b32,
// v_and_or_b32 and v_bfi_b32, which the compiler would otherwise pick for these forms.
template <int TT>
PE_HD uint32_t lop3_32(uint32_t a, uint32_t b, uint32_t c) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_bitop3_b32(a, b, c, TT);
#else
  uint32_t r = 0;
  for (int i = 0; i < 8; i++)
    if (TT >> i & 1) r |= ((i & 4) ? a : ~a) & ((i & 2) ? b : ~b) & ((i & 1) ? c : ~c);
  return r;
#endif
}
template <int TT>
PE_HD uint64_t lop3(uint64_t a, uint64_t b, uint64_t c) {
  return uint64_t(lop3_32<TT>(uint32_t(a), uint32_t(b), uint32_t(c))) |
         uint64_t(lop3_32<TT>(uint32_t(a >> 32), uint32_t(b >> 32), uint32_t(c >> 32))) << 32;
}
enum : int {
  kOr3 = 0xFE,      // a | b | c
  kSel = 0xE4,      // c ? a : b
  kOrAnd = 0xF8,    // a | (b & c)
  kAndXor = 0x28,   // (a ^ b) & c
  kXor3 = 0x96,     // a ^ b ^ c
  kAndAnd = 0x80,   // a & b & c
};

// Byte gather from the 8-byte value (hi:lo): byte k of the result is byte sel_k of it
// (0-3 from lo, 4-7 from hi) or zero (sel_k = 12); v_perm_b32 on the device.
PE_HD uint32_t pe_perm(uint32_t hi, uint32_t lo, uint32_t sel) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_perm(hi, lo, sel);
#else
  const uint64_t v = uint64_t(hi) << 32 | lo;
  uint32_t r = 0;
  for (int k = 0; k < 4; k++) {
    const uint32_t s = sel >> (8 * k) & 0xff;
    r |= (s < 8 ? uint32_t(v >> (8 * s) & 0xff) : 0u) << (8 * k);
  }
  return r;
#endif
}
PE_HD S4 shift3(S4 a, int k) {
  // word rotation by k as two stages of selects (a per-thread index into the words would be
  // lowered to a private-memory array)
  const uint32_t m1 = 0u - uint32_t(k & 1), m2 = 0u - uint32_t(k >> 1 & 1);
  S4 r, t;
  for (int i = 0; i < 4; i++) r.w[i] = select64(a.w[i], a.w[(i + 1) & 3], m1);
  for (int i = 0; i < 4; i++) t.w[i] = select64(r.w[i], r.w[(i + 2) & 3], m2);
  return t;
}
PE_HD S4 shift_coord_var(S4 a, int q, int k);
// X'(z) = X(z + y) for a packed local rotation y (2 bits per quadrant), branch-free.
PE_HD S4 shift_local(S4 a, int y) {
  a = shift_coord_var(a, 0, y & 3);
  a = shift_coord_var(a, 1, y >> 2 & 3);
  a = shift_coord_var(a, 2, y >> 4 & 3);
  return shift3(a, y >> 6 & 3);
}
// X'(z) = X(z + k e_q) for a per-thread amount k in 0..3 and a coordinate q that is the
// same for every thread of a wave (branch-free in k).
PE_HD S4 shift_coord_var(S4 a, int q, int k) {
  k &= 3;
  if (q == 0) {
    const uint64_t m = k == 0 ? ~uint64_t(0) : k == 1 ? 0x7777777777777777ull : k == 2 ? 0x3333333333333333ull : 0x1111111111111111ull;
    for (int i = 0; i < 4; i++) a.w[i] = lop3<kSel>(a.w[i] >> k, a.w[i] << (4 - k), m);
  } else if (q == 1) {
    const uint64_t m = k == 0 ? ~uint64_t(0) : k == 1 ? 0x0FFF0FFF0FFF0FFFull : k == 2 ? 0x00FF00FF00FF00FFull : 0x000F000F000F000Full;
    const int s = 4 * k;
    for (int i = 0; i < 4; i++) a.w[i] = lop3<kSel>(a.w[i] >> s, a.w[i] << (16 - s), m);
  } else if (q == 2) {
    const int s = 16 * k;
    for (int i = 0; i < 4; i++) a.w[i] = (a.w[i] >> s) | (a.w[i] << ((64 - s) & 63));
  } else {
    return shift3(a, k);
  }
  return a;
}
PE_HD S4 shift_coord(S4 a, int q, int k) {
  switch (q) {
    case 0: return shift0(a, k);
    case 1: return shift1(a, k);
    case 2: return shift2(a, k);
    default: return shift3(a, k & 3);
  }
}

// ---------------- coordinate swaps and negation ----------------
PE_HD uint64_t delta_swap(uint64_t x, uint64_t mask, int s) {
  const uint64_t t = (x ^ (x >> s)) & mask;
  return x ^ t ^ (t << s);
}
PE_HD S4 swap01(S4 a) {
  for (int i = 0; i < 4; i++) {
    a.w[i] = delta_swap(a.w[i], 0x0A0A0A0A0A0A0A0Aull, 3);
    a.w[i] = delta_swap(a.w[i], 0x00CC00CC00CC00CCull, 6);
  }
  return a;
}
PE_HD S4 swap02(S4 a) {
  for (int i = 0; i < 4; i++) {
    a.w[i] = delta_swap(a.w[i], 0x0000AAAA0000AAAAull, 15);
    a.w[i] = delta_swap(a.w[i], 0x00000000CCCCCCCCull, 30);
  }
  return a;
}
PE_HD S4 swap12(S4 a) {
  for (int i = 0; i < 4; i++) {
    a.w[i] = delta_swap(a.w[i], 0x0000F0F00000F0F0ull, 12);
    a.w[i] = delta_swap(a.w[i], 0x00000000FF00FF00ull, 24);
  }
  return a;
}
PE_HD S4 swap23(S4 a) {
  S4 r;
  for (int j = 0; j < 4; j++) {
    uint64_t v = 0;
    for (int i = 0; i < 4; i++) v |= (a.w[i] >> (16 * j) & 0xffff) << (16 * i);
    r.w[j] = v;
  }
  return r;
}
// Exchange bits between a word pair: hi bit p <- lo bit p+s and lo bit p+s <- hi bit p, p in mask.
PE_HD void pair_swap(uint64_t& lo, uint64_t& hi, uint64_t mask, int s) {
  const uint64_t t = ((lo >> s) ^ hi) & mask;
  hi ^= t;
  lo ^= t << s;
}
PE_HD S4 swap03(S4 a) {
  pair_swap(a.w[0], a.w[1], 0x5555555555555555ull, 1);
  pair_swap(a.w[2], a.w[3], 0x5555555555555555ull, 1);
  pair_swap(a.w[0], a.w[2], 0x3333333333333333ull, 2);
  pair_swap(a.w[1], a.w[3], 0x3333333333333333ull, 2);
  return a;
}
PE_HD S4 swap13(S4 a) {
  pair_swap(a.w[0], a.w[1], 0x0F0F0F0F0F0F0F0Full, 4);
  pair_swap(a.w[2], a.w[3], 0x0F0F0F0F0F0F0F0Full, 4);
  pair_swap(a.w[0], a.w[2], 0x00FF00FF00FF00FFull, 8);
  pair_swap(a.w[1], a.w[3], 0x00FF00FF00FF00FFull, 8);
  return a;
}
PE_HD S4 negate(S4 a) {
  for (int i = 0; i < 4; i++) {
    uint64_t x = a.w[i];
    x = delta_swap(x, 0x2222222222222222ull, 2);    // r0: 1 <-> 3
    x = delta_swap(x, 0x00F000F000F000F0ull, 8);    // r1: 1 <-> 3
    x = delta_swap(x, 0x00000000FFFF0000ull, 32);   // r2: 1 <-> 3
    a.w[i] = x;
  }
  const uint64_t t = a.w[1]; a.w[1] = a.w[3]; a.w[3] = t;   // r3: 1 <-> 3
  return a;
}

// X'(x) = X(commute[g][x]).  Sequences found by exhaustive search over the primitives above
// and checked bit-exactly against the commute table; applied in order.
PE_HD S4 rearrange(S4 a, int g) {
  switch (g & 7) {
    case 0: return a;
    case 1: return swap13(swap12(swap01(a)));
    case 2: return swap12(swap03(a));
    case 3: return swap23(swap02(swap01(a)));
    case 4: return negate(swap03(a));
    case 5: return negate(swap13(swap02(a)));
    case 6: return negate(swap12(a));
    default: return negate(swap23(swap01(a)));
  }
}
PE_HD S4 transform_super(int g, int y, S4 c) { return rearrange(shift_local(c, y), g); }

// rmax(f)(a) = OR over the 8 single-quadrant rotations b of f(a+b)
// (pentago/base/superscore.h:245-306).
// Per word: the +-1 rotations of r0 (inside nibbles) and r1 (nibbles inside 16-bit chunks) as
// masked selects; those of r2 (16-bit chunks c0..c3) OR to (c1|c3, c0|c2) in both halves,
// a 16-bit rotation of the OR of the halves; those of r3 are the neighbouring words.
PE_HD S4 rmax(S4 a) {
  S4 r;
  for (int i = 0; i < 4; i++) {
    const uint64_t x = a.w[i];
    const uint64_t s0 = lop3<kSel>(x >> 1, x << 3, 0x7777777777777777ull) |
                        lop3<kSel>(x >> 3, x << 1, 0x1111111111111111ull);
    const uint64_t s1 = lop3<kSel>(x >> 4, x << 12, 0x0FFF0FFF0FFF0FFFull) |
                        lop3<kSel>(x >> 12, x << 4, 0x000F000F000F000Full);
    const uint32_t h = uint32_t(x) | uint32_t(x >> 32);
    const uint32_t h2 = h >> 16 | h << 16;
    const uint64_t s2 = uint64_t(h2) | uint64_t(h2) << 32;
    r.w[i] = lop3<kOr3>(lop3<kOr3>(s0, s1, s2), a.w[(i + 1) & 3], a.w[(i + 3) & 3]);
  }
  return r;
}

// ---------------- super_wins (pentago/base/superscore.cc:38-96) ----------------
// Table: per (quadrant, state) five supers (horizontal, vertical, diagonal lo, hi, assist).
// OR-reductions over one coordinate axis:
// (each stage ORs in the partner bit at distance s: mask ? x >> s : x << s)
PE_HD uint64_t red0(uint64_t x) {
  x |= lop3<kSel>(x >> 1, x << 1, 0x5555555555555555ull);
  x |= lop3<kSel>(x >> 2, x << 2, 0x3333333333333333ull);
  return x;
}
PE_HD uint64_t red1(uint64_t x) {
  x |= lop3<kSel>(x >> 4, x << 4, 0x0F0F0F0F0F0F0F0Full);
  x |= lop3<kSel>(x >> 8, x << 8, 0x00FF00FF00FF00FFull);
  return x;
}
PE_HD uint64_t red2(uint64_t x) {
  x = lop3<kOr3>(x, x >> 16, x << 48);
  const uint64_t h = (x | (x >> 32)) & 0xFFFFFFFFull;
  return h | h << 32;
}
PE_HD S4 reduce0(S4 a) { for (int i = 0; i < 4; i++) a.w[i] = red0(a.w[i]); return a; }
PE_HD S4 reduce1(S4 a) { for (int i = 0; i < 4; i++) a.w[i] = red1(a.w[i]); return a; }
PE_HD S4 reduce2(S4 a) { for (int i = 0; i < 4; i++) a.w[i] = red2(a.w[i]); return a; }
PE_HD S4 reduce3(S4 a) {
  const uint64_t x = a.w[0] | a.w[1] | a.w[2] | a.w[3];
  a.w[0] = a.w[1] = a.w[2] = a.w[3] = x;
  return a;
}

PE_HD S4 load_s4(const uint64_t* p) { S4 r; for (int i = 0; i < 4; i++) r.w[i] = p[i]; return r; }

// Compiler-only memory fence: keeps the table loads of later pattern groups from being
// hoisted above earlier groups, which would keep dozens of loaded words live at once.
#define PE_GROUP_FENCE() \
  do {                   \
    if (kLean) asm volatile("" ::: "memory"); \
  } while (0)

// table: 4*512*5 supers as u64 words, entry ((q*512+state)*5+field)*4.  kLean evaluates the
// pattern groups one after another (fewer live registers, less overlap of table loads).
template <bool kLean = false>
PE_HD S4 super_wins(uint64_t side, const uint64_t* table) {
  const uint64_t* t0 = table + size_t((0 * 512 + (side & 0x1ff)) * 5) * 4;
  const uint64_t* t1 = table + size_t((1 * 512 + (side >> 16 & 0x1ff)) * 5) * 4;
  const uint64_t* t2 = table + size_t((2 * 512 + (side >> 32 & 0x1ff)) * 5) * 4;
  const uint64_t* t3 = table + size_t((3 * 512 + (side >> 48 & 0x1ff)) * 5) * 4;
  enum { H = 0, V = 4, LO = 8, HI = 12, AS = 16 };
  S4 wins = s4_zero();
  wins = s4_or(wins, reduce2(reduce3(s4_and(load_s4(t0 + V), load_s4(t1 + V)))));
  PE_GROUP_FENCE();
  wins = s4_or(wins, reduce1(reduce3(s4_and(load_s4(t0 + H), load_s4(t2 + H)))));
  PE_GROUP_FENCE();
  wins = s4_or(wins, reduce2(reduce0(s4_and(load_s4(t1 + H), load_s4(t3 + H)))));
  PE_GROUP_FENCE();
  wins = s4_or(wins, reduce1(reduce0(s4_and(load_s4(t2 + V), load_s4(t3 + V)))));
  PE_GROUP_FENCE();
  wins = s4_or(wins, reduce1(s4_and(s4_and(load_s4(t0 + LO), load_s4(t2 + AS)), load_s4(t3 + LO))));
  PE_GROUP_FENCE();
  wins = s4_or(wins, reduce2(s4_and(s4_and(load_s4(t0 + HI), load_s4(t1 + AS)), load_s4(t3 + HI))));
  PE_GROUP_FENCE();
  wins = s4_or(wins, reduce3(s4_and(s4_and(load_s4(t1 + LO), load_s4(t0 + AS)), load_s4(t2 + LO))));
  PE_GROUP_FENCE();
  wins = s4_or(wins, reduce0(s4_and(s4_and(load_s4(t1 + HI), load_s4(t3 + AS)), load_s4(t2 + HI))));
  return wins;
}

// ---------------- halfsupers (pentago/mid/halfsuper.h) ----------------
// h(p)(a,b,c,d) = s(2a+(b+c+d+p)%2, b, c, d) at bit a + 2b + 8c + 32d.
PE_HD H2 h2_not(H2 a) { a.w[0] = ~a.w[0]; a.w[1] = ~a.w[1]; return a; }
PE_HD H2 h2_or(H2 a, H2 b) { a.w[0] |= b.w[0]; a.w[1] |= b.w[1]; return a; }
PE_HD H2 h2_and(H2 a, H2 b) { a.w[0] &= b.w[0]; a.w[1] &= b.w[1]; return a; }

PE_HD H2 split_half(S4 s, int parity) {
  H2 h; h.w[0] = h.w[1] = 0;
  const uint64_t odd_nibbles = 0x0F0FF0F00F0FF0F0ull;   // nibbles r1 + 4 r2 with r1 + r2 odd
  for (int d = 0; d < 4; d++) {
    const uint64_t x = s.w[d];
    const uint64_t take_odd = ((d + parity) & 1) ? ~odd_nibbles : odd_nibbles;
    uint64_t y = (x & 0x5555555555555555ull & ~take_odd) | ((x >> 1) & 0x5555555555555555ull & take_odd);
    y = (y | (y >> 1)) & 0x3333333333333333ull;
    y = (y | (y >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    y = (y | (y >> 4)) & 0x00FF00FF00FF00FFull;
    y = (y | (y >> 8)) & 0x0000FFFF0000FFFFull;
    y = (y | (y >> 16)) & 0x00000000FFFFFFFFull;
    h.w[d >> 1] |= y << (32 * (d & 1));
  }
  return h;
}
PE_HD H2 halfsuper_wins(uint64_t side, int parity, const uint64_t* table) {
  return split_half(super_wins<true>(side, table), parity);
}
// rmax on halves; flips parity (pentago/mid/halfsuper.h:18-42, 155-203).
PE_HD uint64_t hrmax_lo(uint64_t x) {
  const uint64_t t = (x | (x >> 1)) & 0x5555555555555555ull;
  const uint64_t y0 = t | (t << 1);                               // parity quadrant
  const uint64_t y1 = ((x << 2) & 0xFCFCFCFCFCFCFCFCull) | ((x << 6) & 0xC0C0C0C0C0C0C0C0ull) |
                      ((x >> 2) & 0x3F3F3F3F3F3F3F3Full) | ((x >> 6) & 0x0303030303030303ull);
  const uint64_t y2 = ((x << 8) & 0xFFFFFF00FFFFFF00ull) | ((x << 24) & 0xFF000000FF000000ull) |
                      ((x >> 8) & 0x00FFFFFF00FFFFFFull) | ((x >> 24) & 0x000000FF000000FFull);
  return y0 | y1 | y2;
}
PE_HD H2 hrmax(H2 h) {
  const uint64_t a = h.w[0], b = h.w[1];
  const uint64_t y3 = (a << 32) | (a >> 32) | (b << 32) | (b >> 32);
  H2 r; r.w[0] = hrmax_lo(a) | y3; r.w[1] = hrmax_lo(b) | y3;
  return r;
}

}  // namespace pe
