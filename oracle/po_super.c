/* Oracle super_t operations: restates pentago/base/superscore.{h,cc} and
 * pentago/base/symmetry.{h,cc} (64-bit forms).  TEST INFRASTRUCTURE ONLY. */
#include "po.h"
#include <string.h>

const uint64_t* po_win_contrib_row(int q);
uint8_t po_sstd(int i);

static uint16_t quad(uint64_t s, int q) { return (uint16_t)(s >> 16*q); }
static uint64_t quads4(uint64_t q0, uint64_t q1, uint64_t q2, uint64_t q3) {
  return q0 | q1 << 16 | q2 << 32 | q3 << 48;
}

/* rmax(f)(a) = OR_{b in single rotations} f(a+b).  pentago/base/superscore.h:283-301 */
static uint64_t first_three(uint64_t x) {
  const uint64_t e0 = 0x1111111111111111ull, e1 = 0x000f000f000f000full, e2 = 0x000000000000ffffull;
  return ((x << 1) & ~e0) | ((x >> 3) & e0)                    /* quadrant 0 left */
       | ((x >> 1) & (~e0 >> 1)) | ((x << 3) & (e0 << 3))      /* quadrant 0 right */
       | ((x << 4) & ~e1) | ((x >> 12) & e1)                   /* quadrant 1 left */
       | ((x >> 4) & (~e1 >> 4)) | ((x << 12) & (e1 << 12))    /* quadrant 1 right */
       | ((x << 16) & ~e2) | ((x >> 48) & e2)                  /* quadrant 2 left */
       | ((x >> 16) & (~e2 >> 16)) | ((x << 48) & (e2 << 48)); /* quadrant 2 right */
}
po_super po_rmax(po_super f) {
  po_super r;
  r.w[0] = first_three(f.w[0]) | f.w[3] | f.w[1];
  r.w[1] = first_three(f.w[1]) | f.w[0] | f.w[2];
  r.w[2] = first_three(f.w[2]) | f.w[1] | f.w[3];
  r.w[3] = first_three(f.w[3]) | f.w[2] | f.w[0];
  return r;
}

/* pentago/base/superscore.cc:38-96 (non-SSE reductions) */
static void or3(po_super* w) { uint64_t x = w->w[0] | w->w[1] | w->w[2] | w->w[3]; w->w[0] = w->w[1] = w->w[2] = w->w[3] = x; }
static uint64_t or2p(uint64_t x) {
  x = (x | x >> 32) & 0x00000000ffffffffull; x |= x << 32;
  x = (x | x >> 16) & 0x0000ffff0000ffffull; x |= x << 16; return x;
}
static uint64_t or1p(uint64_t x) {
  x = (x | x >> 8) & 0x00ff00ff00ff00ffull; x |= x << 8;
  x = (x | x >> 4) & 0x0f0f0f0f0f0f0f0full; x |= x << 4; return x;
}
static uint64_t or0p(uint64_t x) {
  x = (x | x >> 2) & 0x3333333333333333ull; x |= x << 2;
  x = (x | x >> 1) & 0x5555555555555555ull; x |= x << 1; return x;
}
static void or2(po_super* w) { for (int i = 0; i < 4; i++) w->w[i] = or2p(w->w[i]); }
static void or1(po_super* w) { for (int i = 0; i < 4; i++) w->w[i] = or1p(w->w[i]); }
static void or0(po_super* w) { for (int i = 0; i < 4; i++) w->w[i] = or0p(w->w[i]); }
static po_super sand(po_super a, po_super b) { for (int i = 0; i < 4; i++) a.w[i] &= b.w[i]; return a; }
static void sor(po_super* a, po_super b) { for (int i = 0; i < 4; i++) a->w[i] |= b.w[i]; }

po_super po_super_wins(uint64_t side) {
  const po_super* t = (const po_super*)po_superwin_info();
  /* fields: 0 horizontal, 1 vertical, 2 diagonal lo, 3 diagonal hi, 4 diagonal assist */
  const po_super* i0 = t + 5*(0*512 + quad(side, 0));
  const po_super* i1 = t + 5*(1*512 + quad(side, 1));
  const po_super* i2 = t + 5*(2*512 + quad(side, 2));
  const po_super* i3 = t + 5*(3*512 + quad(side, 3));
  po_super wins; memset(&wins, 0, sizeof wins);
  po_super w;
  w = sand(i0[1], i1[1]); or3(&w); or2(&w); sor(&wins, w);            /* vertical q0-q1 */
  w = sand(i0[0], i2[0]); or3(&w); or1(&w); sor(&wins, w);            /* horizontal q0-q2 */
  w = sand(i1[0], i3[0]); or0(&w); or2(&w); sor(&wins, w);            /* horizontal q1-q3 */
  w = sand(i2[1], i3[1]); or0(&w); or1(&w); sor(&wins, w);            /* vertical q2-q3 */
  w = sand(sand(i0[2], i2[4]), i3[2]); or1(&w); sor(&wins, w);        /* diagonal q0-q3 lo */
  w = sand(sand(i0[3], i1[4]), i3[3]); or2(&w); sor(&wins, w);        /* diagonal q0-q3 hi */
  w = sand(sand(i1[2], i0[4]), i2[2]); or3(&w); sor(&wins, w);        /* diagonal q1-q2 lo */
  w = sand(sand(i1[3], i3[4]), i2[3]); or0(&w); sor(&wins, w);        /* diagonal q1-q2 hi */
  return wins;
}

/* pentago/base/score.h:39-51 */
int po_won(uint64_t side) {
  const uint64_t c = po_win_contrib_row(0)[quad(side, 0)] + po_win_contrib_row(1)[quad(side, 1)]
                   + po_win_contrib_row(2)[quad(side, 2)] + po_win_contrib_row(3)[quad(side, 3)];
  return (c & (c >> 1) & 0x55) || (c & (0xaaaaaaaaaaaaaaaaull << 8));
}

/* ---- symmetries: pentago/base/symmetry.h:18-100 ---- */
static uint8_t local_mul(uint8_t a, uint8_t b) {
  return (uint8_t)((((a & 0x33) + (b & 0x33)) & 0x33) + (((a & 0xcc) + (b & 0xcc)) & 0xcc));
}
static uint8_t local_inv(uint8_t a) { return (uint8_t)(a ^ (a & 0x55) << 1); }
static uint8_t invert_global(uint8_t g) { return (uint8_t)(g ^ ((~g >> 2 & (g & 1)) << 1)); }

po_sym po_sym_mul(po_sym a, po_sym b) {
  const uint8_t l2 = po_commute(b.global, a.local);
  po_sym r;
  r.global = (uint8_t)((((a.global ^ b.global) & 4) |
                        (((a.global ^ (a.global & b.global >> 2) << 1) + b.global) & 3)) & 7);
  r.local = local_mul(l2, b.local);
  return r;
}
po_sym po_sym_inverse(po_sym s) {
  const uint8_t li = local_inv(s.local), gi = invert_global(s.global);
  po_sym r = {gi, po_commute(gi, li)};
  return r;
}

/* pentago/base/symmetry.cc:21-44 */
static const uint8_t rotate_quadrants[4][4] = {{0,1,2,3},{1,3,0,2},{3,2,1,0},{2,0,3,1}};
uint64_t po_transform_side(po_sym s, uint64_t side) {
  uint16_t q[4];
  for (int i = 0; i < 4; i++) {
    q[i] = quad(side, i);
    switch ((s.global + (s.local >> 2*i)) & 3) {
      case 1: q[i] = po_rotate_left(q[i]); break;
      case 2: q[i] = po_rotate_left(po_rotate_left(q[i])); break;
      case 3: q[i] = po_rotate_right(q[i]); break;
    }
  }
  const int g = s.global & 3;
  uint16_t qr[4];
  for (int i = 0; i < 4; i++) qr[i] = q[rotate_quadrants[g][i]];
  if (s.global & 4) {
    uint16_t t = qr[0]; qr[0] = qr[3]; qr[3] = t;
    for (int i = 0; i < 4; i++) qr[i] = po_reflect(qr[i]);
  }
  return quads4(qr[0], qr[1], qr[2], qr[3]);
}

uint64_t po_pack_board(uint64_t s0, uint64_t s1) {
  uint64_t b = 0;
  for (int q = 0; q < 4; q++)
    b |= (uint64_t)(uint16_t)(po_pack(quad(s0, q)) + 2*po_pack(quad(s1, q))) << 16*q;
  return b;
}
uint64_t po_unpack_board(uint64_t board, int s) {
  uint64_t r = 0;
  for (int q = 0; q < 4; q++) r |= (uint64_t)po_unpack(quad(board, q), s) << 16*q;
  return r;
}
uint64_t po_transform_board(po_sym s, uint64_t board) {
  return po_pack_board(po_transform_side(s, po_unpack_board(board, 0)),
                       po_transform_side(s, po_unpack_board(board, 1)));
}

/* transform_super: pentago/base/symmetry.cc:146-307, 64-bit form.
 * s(C) = {g z y' g' | z in C} for s = g y. */
#define BIT(i) ((uint64_t)1 << (i))
static uint64_t rr4(uint64_t x, int k) {
  return ((x >> k) & (0x1111111111111111ull * (0xfull >> k))) |
         ((x << ((4 - k) & 3)) & (0x1111111111111111ull * ((0xfull << (4 - k)) & 0xf)));
}
static uint64_t rr16(uint64_t x, int k) {
  return ((x >> k) & (0x0001000100010001ull * (0xffffull >> k))) |
         ((x << ((16 - k) & 15)) & (0x0001000100010001ull * ((0xffffull << (16 - k)) & 0xffff)));
}
static uint64_t rr64(uint64_t x, int k) {
  if (!k) return x;
  return ((x >> k) & ((~(uint64_t)0) >> k)) | ((x << ((64 - k) & 63)) & ((~(uint64_t)0) << (64 - k)));
}
static uint64_t low_quarter_transpose(uint64_t x, int i, int j) {
  const int ii = 1 << 2*i, jj = 1 << 2*j, kk = 1 << 2*(3 - i - j), sh = jj - ii;
  const uint64_t other = 1 | BIT(kk) | BIT(2*kk) | BIT(3*kk);
  uint64_t t = (x ^ x >> sh) & (other * (BIT(ii) | BIT(3*ii) | BIT(ii + 2*jj) | BIT(3*ii + 2*jj)));
  x ^= t ^ t << sh;
  t = (x ^ x >> 2*sh) & (other * (BIT(2*ii) | BIT(3*ii) | BIT(2*ii + jj) | BIT(3*ii + jj)));
  x ^= t ^ t << 2*sh;
  return x;
}
static void transpose23(po_super* c) {
  uint64_t r[4];
  for (int i = 0; i < 4; i++)
    r[i] = (c->w[0] >> 16*i & 0xffff) | (c->w[1] >> 16*i & 0xffff) << 16 |
           (c->w[2] >> 16*i & 0xffff) << 32 | (c->w[3] >> 16*i & 0xffff) << 48;
  memcpy(c->w, r, sizeof r);
}
static uint64_t quarter_negate(uint64_t x) {
  uint64_t t = (x ^ x >> 2) & 0x2222222222222222ull;
  x ^= t ^ t << 2;
  t = (x ^ x >> 2*4) & 0x00f000f000f000f0ull;
  x ^= t ^ t << 2*4;
  t = (x ^ x >> 2*16) & 0x00000000ffff0000ull;
  x ^= t ^ t << 2*16;
  return x;
}

po_super po_transform_super(po_sym s, po_super C) {
  /* local part first: C = C local' */
  const int k0 = s.local & 3, k1 = 4*(s.local >> 2 & 3), k2 = 16*(s.local >> 4 & 3);
  for (int i = 0; i < 4; i++) C.w[i] = rr4(C.w[i], k0);
  for (int i = 0; i < 4; i++) C.w[i] = rr16(C.w[i], k1);
  for (int i = 0; i < 4; i++) C.w[i] = rr64(C.w[i], k2);
  if (s.local & 1 << 6) { po_super t = C; C.w[0] = t.w[1]; C.w[1] = t.w[2]; C.w[2] = t.w[3]; C.w[3] = t.w[0]; }
  if (s.local & 1 << 7) { po_super t = C; C.w[0] = t.w[2]; C.w[1] = t.w[3]; C.w[2] = t.w[0]; C.w[3] = t.w[1]; }
  /* conjugate by the global part via quadrant transpositions (symmetry.cc:231-261) */
#define TT(i,j) ((j)==3 ? 3 : ((j)*((j)-1)/2+(i)))
  static const uint8_t offsets[] = {0,0,3,7,10,13,17,18,20};
  static const uint8_t transp[] = {
    TT(0,1),TT(0,2),TT(2,3),
    TT(0,2),TT(0,1),TT(2,3),TT(0,2),
    TT(2,3),TT(0,2),TT(0,1),
    TT(0,2),TT(2,3),TT(0,2),
    TT(0,2),TT(1,2),TT(2,3),TT(1,2),
    TT(1,2),
    TT(0,1),TT(2,3)};
  for (int i = offsets[s.global + 1] - 1; i >= (int)offsets[s.global]; i--) {
    switch (transp[i]) {
      case TT(0,1): for (int k = 0; k < 4; k++) C.w[k] = low_quarter_transpose(C.w[k], 0, 1); break;
      case TT(0,2): for (int k = 0; k < 4; k++) C.w[k] = low_quarter_transpose(C.w[k], 0, 2); break;
      case TT(1,2): for (int k = 0; k < 4; k++) C.w[k] = low_quarter_transpose(C.w[k], 1, 2); break;
      case TT(2,3): transpose23(&C); break;
    }
  }
#undef TT
  if (s.global & 4) {
    for (int k = 0; k < 4; k++) C.w[k] = quarter_negate(C.w[k]);
    uint64_t t = C.w[1]; C.w[1] = C.w[3]; C.w[3] = t;
  }
  return C;
}

/* pentago/base/symmetry.cc:311-335 */
static int meaningless_hash(uint64_t key) {
  key = (~key) + (key << 18);
  key = key ^ (key >> 31);
  key = key * 21;
  key = key ^ (key >> 11);
  key = key + (key << 6);
  key = key ^ (key >> 22);
  return (int)key;
}
int po_meaningless(uint64_t board, uint64_t salt) {
  int r = 0;
  for (int g = 0; g < 8; g++) {
    po_sym s = {(uint8_t)g, 0};
    r ^= meaningless_hash(po_transform_board(s, board) ^ salt) & 1;
  }
  return r;
}
po_super po_super_meaningless(uint64_t board, uint64_t salt) {
  po_super r; memset(&r, 0, sizeof r);
  for (int l = 0; l < 256; l++) {
    po_sym s = {0, (uint8_t)l};
    if (po_meaningless(po_transform_board(s, board), salt)) r.w[l >> 6] |= (uint64_t)1 << (l & 63);
  }
  return r;
}
