/* Oracle lookup tables: restates the mathematics of pentago/base/precompute.cc.
 * TEST INFRASTRUCTURE ONLY (see po.h).  Each table is pinned by the SHA-1 digest
 * the reference records for it; po_table_hash reproduces the reference's hashing
 * (portable_hash over little-endian flat bytes, pentago/utility/portable_hash.h:10-17,
 * optionally after a cast to u64, precompute.cc:148-155). */
#include "po.h"
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* ---------------- SHA-1 (FIPS 180-4) ---------------- */
static uint32_t rol32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

void po_sha1_hex(const void* data, size_t n, char out[41]) {
  uint32_t h[5] = {0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u, 0xC3D2E1F0u};
  const uint8_t* p = (const uint8_t*)data;
  const uint64_t bits = (uint64_t)n * 8;
  size_t total = ((n + 8) / 64 + 1) * 64;   /* padded length */
  for (size_t off = 0; off < total; off += 64) {
    uint8_t block[64];
    for (int i = 0; i < 64; i++) {
      size_t k = off + i;
      uint8_t v;
      if (k < n) v = p[k];
      else if (k == n) v = 0x80;
      else if (k >= total - 8) v = (uint8_t)(bits >> (8 * (total - 1 - k)));
      else v = 0;
      block[i] = v;
    }
    uint32_t w[80];
    for (int t = 0; t < 16; t++)
      w[t] = (uint32_t)block[4*t] << 24 | (uint32_t)block[4*t+1] << 16 | (uint32_t)block[4*t+2] << 8 | block[4*t+3];
    for (int t = 16; t < 80; t++) w[t] = rol32(w[t-3] ^ w[t-8] ^ w[t-14] ^ w[t-16], 1);
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4];
    for (int t = 0; t < 80; t++) {
      uint32_t f, k;
      if (t < 20) { f = (b & c) | (~b & d); k = 0x5A827999u; }
      else if (t < 40) { f = b ^ c ^ d; k = 0x6ED9EBA1u; }
      else if (t < 60) { f = (b & c) | (b & d) | (c & d); k = 0x8F1BBCDCu; }
      else { f = b ^ c ^ d; k = 0xCA62C1D6u; }
      uint32_t tmp = rol32(a, 5) + f + e + k + w[t];
      e = d; d = c; c = rol32(b, 30); b = a; a = tmp;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e;
  }
  for (int i = 0; i < 5; i++) sprintf(out + 8*i, "%08x", h[i]);
  out[40] = 0;
}

/* ---------------- tables ---------------- */
static int inited = 0;
static uint16_t T_pack[512];                 /* precompute.cc:160-172 */
static uint16_t T_unpack[19683][2];          /* precompute.cc:174-197 */
static uint64_t T_win_patterns[32];          /* precompute.cc:199-254 */
static uint64_t T_win_contrib[4][512];       /* precompute.cc:256-277 */
static uint16_t T_rot[512][2];               /* precompute.cc:279-296 (left, right) */
static uint16_t T_allrot[512][4];            /* precompute.cc:299-307 */
static uint16_t T_refl[512];                 /* precompute.cc:606-620 */
static uint64_t T_superwin[4*512*5*4];       /* precompute.cc:762-820 */
static uint8_t  T_commute[8][256];           /* precompute.cc:883-931 */
static uint8_t  T_sstd[256];                 /* precompute.cc:933-965 */
static uint16_t T_rmin_offsets[56];          /* precompute.cc:967-1043 */
static uint16_t T_rmin_flat[19683];
static int      T_rmin_flat_size;
static int      T_rmin_inverse[19683];
static uint16_t T_rmin_inverse16[19683];
static int      T_rmin_moved[55];

static int popc(uint64_t x) { return __builtin_popcountll(x); }
static uint16_t qbit(int x, int y) { return (uint16_t)(1 << (3*x + y)); }

static uint64_t win_bit(int x, int y, int flip, int rotate) {
  if (flip) { int t = x; x = y; y = t; }
  if (rotate) { int t = 5 - y; y = x; x = t; }
  return (uint64_t)1 << (16*(2*(x/3) + y/3) + 3*(x%3) + y%3);
}

static void make_win_patterns(void) {
  int n = 0;
  for (int x = 0; x < 2; x++)          /* off-centre diagonals (three quadrants) */
    for (int y = 0; y < 2; y++) {
      uint64_t p = 0;
      for (int i = -2; i <= 2; i++) p |= win_bit(2+x+i, 2+y+i*(x==y ? -1 : 1), 0, 0);
      T_win_patterns[n++] = p;
    }
  for (int flip = 0; flip < 2; flip++)  /* axis aligned */
    for (int x = 0; x < 2; x++)
      for (int y = 0; y < 6; y++) {
        uint64_t p = 0;
        for (int i = 0; i < 5; i++) p |= win_bit(x+i, y, flip, 0);
        T_win_patterns[n++] = p;
      }
  for (int rotate = 0; rotate < 2; rotate++)  /* centred diagonals */
    for (int x = 0; x < 2; x++) {
      uint64_t p = 0;
      for (int i = 0; i < 5; i++) p |= win_bit(i+x, i+x, 0, rotate);
      T_win_patterns[n++] = p;
    }
}

/* View of the 7-D boolean superwin array: (quadrant, state, field, r0, r1, r2, r3). */
static uint8_t* SW;   /* 4*512*5*256 bytes */
static size_t SWI(int q, int v, int f, int r0, int r1, int r2, int r3) {
  return ((((((size_t)q*512 + v)*5 + f)*4 + r0)*4 + r1)*4 + r2)*4 + r3;
}

static void make_superwin(void) {
  /* Per pattern group (quadrant roles "h v l u a" or unused '-'), precompute.cc:769-805.
   * For group G, used quadrant q with field f, and the i-th way w of G, the entry at
   * (q, v, f, r) is set when rotating state v by r_q covers w's cells in quadrant q and the
   * unused-quadrant rotation coordinates encode i. */
  static const char* groups[8] = {"vv--", "h-h-", "-h-h", "--vv", "l-al", "ua-u", "all-", "-uua"};
  static const char* types = "hvlua";
  SW = (uint8_t*)calloc((size_t)4*512*5*256, 1);
  for (int gi = 0; gi < 8; gi++) {
    const char* g = groups[gi];
    int used[4], nu = 0, unused[4], nn = 0;
    for (int i = 0; i < 4; i++) { if (g[i] != '-') used[nu++] = i; else unused[nn++] = i; }
    uint16_t ws[32][4]; int nw = 0;
    for (int k = 0; k < 32; k++) {
      uint16_t w[4];
      for (int j = 0; j < 4; j++) w[j] = (uint16_t)(T_win_patterns[k] >> 16*j & 0x1ff);
      int ok = 1;
      for (int i = 0; i < 4; i++)
        if (!((w[i] != 0) == (g[i] != '-') || g[i] == 'a')) ok = 0;
      if (ok) { memcpy(ws[nw], w, sizeof w); nw++; }
    }
    if (!(nw == 6 || nw == 3)) abort();
    for (int a = 0; a < nu; a++) {
      const int q = used[a];
      const int f = (int)(strchr(types, g[q]) - types);
      for (int i = 0; i < nw; i++) {
        const uint16_t w = ws[i][q];
        for (int r = 0; r < 256; r++) {
          int rr[4] = {r & 3, r >> 2 & 3, r >> 4 & 3, r >> 6 & 3};
          /* unused coordinates must encode way i */
          if (nn == 1) { if (rr[unused[0]] != i) continue; }
          else { if (rr[unused[0]] != i/4 || rr[unused[1]] != i%4) continue; }
          for (int v = 0; v < 512; v++)
            if ((T_allrot[v][rr[q]] & w) == w)
              SW[SWI(q, v, f, rr[0], rr[1], rr[2], rr[3])] = 1;
        }
      }
    }
  }
  /* pack bit (q,v,f,r) at flat index ((q*512+v)*5+f)*256 + r0+4r1+16r2+64r3, precompute.cc:807-812 */
  memset(T_superwin, 0, sizeof T_superwin);
  for (int q = 0; q < 4; q++)
    for (int v = 0; v < 512; v++)
      for (int f = 0; f < 5; f++)
        for (int r = 0; r < 256; r++) {
          int r0 = r & 3, r1 = r >> 2 & 3, r2 = r >> 4 & 3, r3 = r >> 6 & 3;
          if (SW[SWI(q, v, f, r0, r1, r2, r3)]) {
            size_t bit = (((size_t)q*512 + v)*5 + f)*256 + r;
            T_superwin[bit/64] |= (uint64_t)1 << (bit % 64);
          }
        }
  free(SW);
}

/* Permutations of the 36 cells, cell index 6x+y (precompute.cc:822-881). */
typedef struct { uint8_t p[36]; } sym36;
static sym36 s_id(void) { sym36 g; for (int i = 0; i < 36; i++) g.p[i] = (uint8_t)i; return g; }
static sym36 s_mul(sym36 g, sym36 h) { sym36 r; for (int i = 0; i < 36; i++) r.p[i] = g.p[h.p[i]]; return r; }
static sym36 s_inv(sym36 g) { sym36 r; for (int i = 0; i < 36; i++) r.p[g.p[i]] = (uint8_t)i; return r; }
static int s_eq(sym36 a, sym36 b) { return !memcmp(a.p, b.p, 36); }

static void make_commute(void) {
  sym36 reflect = s_id(), gr = s_id(), lr[4];
  for (int x = 0; x < 6; x++)
    for (int y = 0; y < 6; y++) {
      reflect.p[6*x+y] = (uint8_t)(6*(5-y) + (5-x));
      gr.p[6*x+y] = (uint8_t)(6*(5-y) + x);
    }
  for (int q = 0; q < 4; q++) {
    lr[q] = s_id();
    const int d = 3*(6*(q/2) + q%2);
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) lr[q].p[d + 6*x + y] = (uint8_t)(d + 6*(2-y) + x);
  }
  sym36 plr[4][4];
  for (int q = 0; q < 4; q++) {
    plr[q][0] = s_id();
    for (int i = 0; i < 3; i++) plr[q][i+1] = s_mul(lr[q], plr[q][i]);
  }
  sym36 local[256];
  for (int i0 = 0; i0 < 4; i0++) for (int i1 = 0; i1 < 4; i1++)
    for (int i2 = 0; i2 < 4; i2++) for (int i3 = 0; i3 < 4; i3++)
      local[((i0*4 + i1)*4 + i2)*4 + i3] =
        s_mul(s_mul(s_mul(plr[0][i0], plr[1][i1]), plr[2][i2]), plr[3][i3]);
  sym36 global[8];
  global[0] = s_id();
  for (int i = 0; i < 3; i++) global[i+1] = s_mul(gr, global[i]);
  for (int i = 0; i < 4; i++) global[4+i] = s_mul(reflect, global[i]);
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 256; j++) {
      sym36 c = s_mul(s_mul(s_inv(global[i]), local[j]), global[i]);
      int found = -1;
      for (int k = 0; k < 256; k++) if (s_eq(local[k], c)) { found = k; break; }
      if (found < 0) abort();
      T_commute[i][j] = (uint8_t)found;
    }
}

static void make_sstd(void) {
  int rot[4][4] = {{0,1,2,3},{2,0,3,1},{0},{0}};
  for (int p = 0; p < 2; p++)
    for (int i = 0; i < 4; i++) rot[p+2][i] = rot[1][rot[p+1][i]];
  for (int idx = 0; idx < 256; idx++) {
    int pp[4] = {idx & 3, idx >> 2 & 3, idx >> 4 & 3, idx >> 6 & 3};
    int best = 1 << 16, best_r = -1;
    for (int r = 0; r < 4; r++) {
      int v = 0;
      for (int i = 0; i < 4; i++) v += pp[i] << 2*rot[r][i];
      if (best > v) { best = v; best_r = r; }
    }
    T_sstd[idx] = (uint8_t)best_r;
  }
}

static uint16_t pack_q(uint16_t s0, uint16_t s1) { return (uint16_t)(T_pack[s0] + 2*T_pack[s1]); }

static uint16_t* g_minq;
static uint16_t min_reflect(uint16_t q) {
  return g_minq[pack_q(T_refl[T_unpack[q][0]], T_refl[T_unpack[q][1]])];
}
static int rmin_sig(uint16_t q) {
  uint16_t r = min_reflect(q);
  return (q < r ? q : r) + 19683*(q == r);
}
static int cmp_sig(const void* a, const void* b) {
  int sa = rmin_sig(*(const uint16_t*)a), sb = rmin_sig(*(const uint16_t*)b);
  if (sa != sb) return sa < sb ? -1 : 1;
  /* stable: original order is increasing quadrant value */
  uint16_t qa = *(const uint16_t*)a, qb = *(const uint16_t*)b;
  return qa < qb ? -1 : qa > qb;
}

static void make_rmin(void) {
  static uint16_t minq[19683];
  g_minq = minq;
  uint16_t all[19683]; int na = 0;
  for (int q = 0; q < 19683; q++) {
    const uint16_t s0 = T_unpack[q][0], s1 = T_unpack[q][1];
    uint16_t m = (uint16_t)q;
    for (int r = 1; r < 4; r++) {
      uint16_t v = pack_q(T_allrot[s0][r], T_allrot[s1][r]);
      if (v < m) m = v;
    }
    minq[q] = m;
    if (m == q) all[na++] = (uint16_t)q;
  }
  qsort(all, na, sizeof(uint16_t), cmp_sig);   /* stable by construction (ties broken by q) */
  /* bucket by (black, white) counts */
  int counts[55] = {0};
  uint16_t buckets[55][420];
  for (int i = 0; i < na; i++) {
    const int b = popc(T_unpack[all[i]][0]), w = popc(T_unpack[all[i]][1]);
    const int k = b*(21-b)/2 + w;
    buckets[k][counts[k]++] = all[i];
  }
  int off = 0;
  for (int k = 0; k < 55; k++) {
    T_rmin_offsets[k] = (uint16_t)off;
    int m = 0;
    for (int j = 0; j < counts[k]; j++) {
      T_rmin_flat[off + j] = buckets[k][j];
      m += min_reflect(buckets[k][j]) != buckets[k][j];
    }
    T_rmin_moved[k] = m;
    off += counts[k];
  }
  T_rmin_offsets[55] = (uint16_t)off;
  T_rmin_flat_size = off;
  for (int q = 0; q < 19683; q++) T_rmin_inverse[q] = 1 << 20;
  for (int k = 0; k < 55; k++)
    for (int j = 0; j < counts[k]; j++) {
      const uint16_t s0 = T_unpack[buckets[k][j]][0], s1 = T_unpack[buckets[k][j]][1];
      for (int r = 0; r < 4; r++) {
        const uint16_t q = pack_q(T_allrot[s0][r], T_allrot[s1][r]);
        if (4*j + r < T_rmin_inverse[q]) T_rmin_inverse[q] = 4*j + r;
      }
    }
  for (int q = 0; q < 19683; q++) T_rmin_inverse16[q] = (uint16_t)T_rmin_inverse[q];
}

void po_init(void) {
  if (inited) return;
  for (int v = 0; v < 512; v++) {
    int pv = 0, p3 = 1;
    for (int i = 0; i < 9; i++, p3 *= 3) if (v >> i & 1) pv += p3;
    T_pack[v] = (uint16_t)pv;
  }
  for (int v = 0; v < 19683; v++) {
    int vv = v; uint16_t p0 = 0, p1 = 0;
    for (int i = 0; i < 9; i++) {
      const int c = vv % 3;
      if (c == 1) p0 |= (uint16_t)(1 << i);
      else if (c == 2) p1 |= (uint16_t)(1 << i);
      vv /= 3;
    }
    T_unpack[v][0] = p0; T_unpack[v][1] = p1;
  }
  make_win_patterns();
  for (int q = 0; q < 4; q++) {
    const uint64_t qb = (uint64_t)0x1ff << 16*q;
    for (int v = 0; v < 512; v++) {
      const uint64_t b = (uint64_t)v << 16*q;
      uint64_t t = 0;
      for (int i = 0; i < 32; i++) {
        const uint64_t w = T_win_patterns[i];
        if ((w & qb) && !(~b & (w & qb))) t |= (uint64_t)1 << 2*i;
      }
      T_win_contrib[q][v] = t;
    }
  }
  for (int v = 0; v < 512; v++) {
    uint16_t left = 0, right = 0, refl = 0;
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++)
        if (v & qbit(x, y)) {
          left |= qbit(2-y, x);
          right |= qbit(y, 2-x);
          refl |= qbit(2-y, 2-x);
        }
    T_rot[v][0] = left; T_rot[v][1] = right; T_refl[v] = refl;
  }
  for (int v = 0; v < 512; v++) {
    T_allrot[v][0] = (uint16_t)v;
    for (int i = 0; i < 3; i++) T_allrot[v][i+1] = T_rot[T_allrot[v][i]][0];
  }
  make_superwin();
  make_commute();
  make_sstd();
  make_rmin();
  inited = 1;
}

static void hash_u64_cast_u16(const uint16_t* a, int n, char out[41]) {
  uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t)*(size_t)n);
  for (int i = 0; i < n; i++) t[i] = a[i];
  po_sha1_hex(t, sizeof(uint64_t)*(size_t)n, out); free(t);
}
static void hash_u64_cast_int(const int* a, int n, char out[41]) {
  uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t)*(size_t)n);
  for (int i = 0; i < n; i++) t[i] = (uint64_t)(int64_t)a[i];
  po_sha1_hex(t, sizeof(uint64_t)*(size_t)n, out); free(t);
}
static void hash_u64_cast_u8(const uint8_t* a, int n, char out[41]) {
  uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t)*(size_t)n);
  for (int i = 0; i < n; i++) t[i] = a[i];
  po_sha1_hex(t, sizeof(uint64_t)*(size_t)n, out); free(t);
}

int po_table_hash(const char* name, char out[41]) {
  po_init();
  if (!strcmp(name, "pack")) po_sha1_hex(T_pack, sizeof T_pack, out);
  else if (!strcmp(name, "unpack")) po_sha1_hex(T_unpack, sizeof T_unpack, out);
  else if (!strcmp(name, "win_contributions")) po_sha1_hex(T_win_contrib, sizeof T_win_contrib, out);
  else if (!strcmp(name, "rotations")) po_sha1_hex(T_rot, sizeof T_rot, out);
  else if (!strcmp(name, "reflections")) po_sha1_hex(T_refl, sizeof T_refl, out);
  else if (!strcmp(name, "superwin_info")) po_sha1_hex(T_superwin, sizeof T_superwin, out);
  else if (!strcmp(name, "commute_global_local_symmetries")) hash_u64_cast_u8(&T_commute[0][0], 8*256, out);
  else if (!strcmp(name, "superstandardize_table")) hash_u64_cast_u8(T_sstd, 256, out);
  else if (!strcmp(name, "rotation_minimal_quadrants_offsets")) hash_u64_cast_u16(T_rmin_offsets, 56, out);
  else if (!strcmp(name, "rotation_minimal_quadrants_flat")) hash_u64_cast_u16(T_rmin_flat, T_rmin_flat_size, out);
  else if (!strcmp(name, "rotation_minimal_quadrants_inverse")) hash_u64_cast_int(T_rmin_inverse, 19683, out);
  else if (!strcmp(name, "rotation_minimal_quadrants_reflect_moved")) hash_u64_cast_int(T_rmin_moved, 55, out);
  else return -1;
  return 0;
}

uint16_t po_pack(uint16_t s) { po_init(); return T_pack[s]; }
uint16_t po_unpack(uint16_t q, int s) { po_init(); return T_unpack[q][s]; }
uint16_t po_rotate_left(uint16_t s) { po_init(); return T_rot[s][0]; }
uint16_t po_rotate_right(uint16_t s) { po_init(); return T_rot[s][1]; }
uint16_t po_reflect(uint16_t s) { po_init(); return T_refl[s]; }
uint8_t po_commute(int g, int l) { po_init(); return T_commute[g][l]; }
uint16_t po_rmin_inverse(uint16_t q) { po_init(); return T_rmin_inverse16[q]; }
const uint64_t* po_superwin_info(void) { po_init(); return T_superwin; }
const uint16_t* po_rmin_flat(void) { po_init(); return T_rmin_flat; }
const uint16_t* po_rmin_offsets(void) { po_init(); return T_rmin_offsets; }
const uint16_t* po_rmin_inverse_table(void) { po_init(); return T_rmin_inverse16; }
const uint64_t* po_win_contrib_row(int q) { po_init(); return T_win_contrib[q]; }
uint8_t po_sstd(int i) { po_init(); return T_sstd[i]; }

/* pentago/base/section.cc:20-28 */
int po_rmin(int black, int white, const uint16_t** quads, int* moved) {
  po_init();
  if (black < 0 || white < 0 || black + white > 9) return -1;
  const int i = (black*(21-black) >> 1) + white;
  const int lo = T_rmin_offsets[i], hi = T_rmin_offsets[i+1];
  if (quads) *quads = T_rmin_flat + lo;
  if (moved) *moved = T_rmin_moved[i];
  return hi - lo;
}
