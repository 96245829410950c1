/* Oracle mid engine: restates pentago/mid/{halfsuper.{h,cc},subsets.h,internal.h,midengine.cc}
 * and pentago/high/board.{h,cc}.  TEST INFRASTRUCTURE ONLY (see po.h). */
#include "po.h"
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

static const uint64_t SIDE_MASK = 0x01ff01ff01ff01ffull;   /* pentago/base/board_c.h:22 */
static uint16_t quad(uint64_t s, int q) { return (uint16_t)(s >> 16*q); }

/* ---- halfsuper: pentago/mid/halfsuper.h ---- */
static int hget(po_half h, int r) { r &= 127; return (int)(h.w[r >> 6] >> (r & 63) & 1); }
static void hset(po_half* h, int r) { h->w[r >> 6] |= (uint64_t)1 << (r & 63); }
static int sget(po_super s, int r) { return (int)(s.w[r >> 6] >> (r & 63) & 1); }

static uint64_t rep8(uint8_t x) { uint64_t y = x | (uint64_t)x << 8; y |= y << 16; return y | y << 32; }
/* halfsuper.h:161-176 */
static uint64_t rmax_lo(uint64_t x) {
  const uint64_t t = (x | x >> 1) & rep8(0x55);
  const uint64_t y0 = t | t << 1;
  const uint64_t y1 = (x << 2 & rep8(0xfc)) | (x << 6 & rep8(0xc0)) | (x >> 2 & rep8(0x3f)) | (x >> 6 & rep8(0x03));
  const uint64_t y2 = (x << 8 & 0xffffff00ffffff00ull) | (x << 24 & 0xff000000ff000000ull)
                    | (x >> 8 & 0x00ffffff00ffffffull) | (x >> 24 & 0x000000ff000000ffull);
  return y0 | y1 | y2;
}
/* halfsuper.h:180-203 */
po_half po_half_rmax(po_half h) {
  const uint64_t a = h.w[0], b = h.w[1];
  const uint64_t y3 = a << 32 | a >> 32 | b << 32 | b >> 32;
  po_half r = {{rmax_lo(a) | y3, rmax_lo(b) | y3}};
  return r;
}
/* h(p)(a,b,c,d) = s(2a+(b+c+d+p)%2,b,c,d)  (halfsuper.h:9-16; slow_split in mid_test.cc:136-145) */
void po_split(po_super s, po_half out[2]) {
  memset(out, 0, 2*sizeof(po_half));
  for (int p = 0; p < 2; p++)
    for (int a = 0; a < 2; a++) for (int b = 0; b < 4; b++) for (int c = 0; c < 4; c++) for (int d = 0; d < 4; d++)
      if (sget(s, (2*a + ((b + c + d + p) & 1)) + 4*b + 16*c + 64*d)) hset(&out[p], a + 2*b + 8*c + 32*d);
}
po_super po_merge(po_half even, po_half odd) {
  po_super s; memset(&s, 0, sizeof s);
  for (int r = 0; r < 256; r++) {
    const int r0 = r & 3, b = r >> 2 & 3, c = r >> 4 & 3, d = r >> 6 & 3;
    const int p = (r0 + b + c + d) & 1;
    if (hget(p ? odd : even, (r0 >> 1) + 2*b + 8*c + 32*d)) s.w[r >> 6] |= (uint64_t)1 << (r & 63);
  }
  return s;
}
/* halfsuper_wins(side, p) == split(super_wins(side))[p]: the relation pentago/mid/mid_test.cc:203-249
 * asserts for the generated routine (precompute.cc:330-504). */
po_half po_halfsuper_wins(uint64_t side, int parity) {
  po_half h[2]; po_split(po_super_wins(side), h);
  return h[parity & 1];
}
static po_half hnot(po_half a) { a.w[0] = ~a.w[0]; a.w[1] = ~a.w[1]; return a; }
static po_half hand(po_half a, po_half b) { a.w[0] &= b.w[0]; a.w[1] &= b.w[1]; return a; }
static po_half hor(po_half a, po_half b) { a.w[0] |= b.w[0]; a.w[1] |= b.w[1]; return a; }

/* ---- subsets: pentago/mid/subsets.h ---- */
static int choose_table[64][64];
static int choose_ready = 0;
static void choose_init(void) {
  for (int n = 0; n < 64; n++)
    for (int k = 0; k < 64; k++) {
      if (k > n) { choose_table[n][k] = 0; continue; }
      long long r = 1;
      const int kk = k > n - k ? n - k : k;
      for (int i = 1; i <= kk; i++) r = r * (n - kk + i) / i;
      choose_table[n][k] = r > 0x7fffffff ? 0x7fffffff : (int)r;
    }
  choose_ready = 1;
}
int po_choose(int n, int k) {
  if (n < 0 || k < 0 || k > n) return 0;
  if (!choose_ready) choose_init();
  if (n >= 64) return 0;
  return choose_table[n][k];
}
uint64_t po_subset(int n, int k, int index) {
  uint64_t set = 0;
  for (int i = k - 1; i >= 0; i--) {
    int j = i, base = 0;
    for (; j < n; j++) {
      const int next = base + po_choose(j, i);
      if (index < next) break;
      base = next;
    }
    set |= (uint64_t)j << 5*i;
    index -= base;
  }
  return set;
}
uint32_t po_subset_mask(int n, int k, int index) {
  uint32_t mask = 0;
  for (int i = k - 1; i >= 0; i--) {
    int j = i, base = 0;
    for (; j < n; j++) {
      const int next = base + po_choose(j, i);
      if (index < next) break;
      base = next;
    }
    mask |= 1u << j;
    index -= base;
  }
  return mask;
}
int po_subset_index(int n, int k, uint64_t set) {
  (void)n;
  int index = 0;
  for (int i = 0; i < k; i++) index += po_choose((int)(set >> 5*i & 0x1f), i + 1);
  return (uint16_t)index;
}

/* ---- high board: pentago/high/board.{h,cc} ---- */
int po_high_count(po_high b) { return (int)((b.ply + 1) >> 1); }
static int high_turn(po_high b) { return (int)((b.ply >> 1) & 1); }
static int high_middle(po_high b) { return (int)(b.ply & 1); }
static uint64_t high_empty(po_high b) { return SIDE_MASK ^ b.side[0] ^ b.side[1]; }

int po_high_done_and_value(po_high b, int* value) {
  const int bw = po_won(b.side[0]), ww = po_won(b.side[1]);
  if (bw || ww) { *value = bw && ww ? 0 : (bw == !high_turn(b) ? 1 : -1); return 1; }
  if (b.ply == 2*36) { *value = 0; return 1; }
  *value = 0;
  return 0;
}
po_high po_high_place(po_high b, int bit) {
  po_high r = b;
  r.side[high_turn(b)] |= (uint64_t)1 << bit;
  r.ply = b.ply + 1;
  return r;
}
po_high po_high_rotate(po_high b, int q, int d) {
  po_high r = b;
  for (int s = 0; s < 2; s++) {
    const uint16_t old = quad(r.side[s], q);
    uint16_t qq = 0;
    for (int i = 0; i < 9; i++)
      if (old >> i & 1) { const int x = i / 3, y = i % 3; qq |= (uint16_t)(1 << (4 - d*(3*y - x - 2))); }
    r.side[s] ^= (uint64_t)(old ^ qq) << 16*q;
  }
  r.ply = b.ply + 1;
  return r;
}
uint64_t po_high_board(po_high b) {
  uint64_t board = 0;
  for (int q = 0; q < 4; q++) {
    uint32_t qq = 0;
    for (int i = 8; i >= 0; i--) {
      const int j = 16*q + i;
      qq = 3*qq + (uint32_t)(b.side[0] >> j & 1) + 2*(uint32_t)(b.side[1] >> j & 1);
    }
    board |= (uint64_t)qq << 16*q;
  }
  return board;
}
uint64_t po_high_raw(po_high b) { return po_high_board(b) | (uint64_t)high_middle(b) << 63; }
po_high po_high_from_board(uint64_t board, int middle) {
  uint64_t s0 = 0, s1 = 0;
  for (int q = 0; q < 4; q++) {
    uint32_t qq = (uint32_t)(board >> 16*q & 0xffff);
    for (int i = 0; i < 9; i++) {
      const uint32_t s = qq % 3; qq /= 3;
      const int j = 16*q + i;
      s0 |= (uint64_t)(s == 1) << j;
      s1 |= (uint64_t)(s == 2) << j;
    }
  }
  po_high r = {{s0, s1}, (uint32_t)(2*__builtin_popcountll(s0 | s1) - (middle ? 1 : 0))};
  return r;
}

/* ---- mid engine internals: pentago/mid/internal.h ---- */
typedef struct { int ny, lo, size; } grab_t;
typedef struct {
  po_high root;
  int spots;
  uint8_t empty[18];
  grab_t spaces[18+2];
  int sets0_offsets[18+2], sets1p_offsets[18+2], cs1ps_offsets[18+2], wins1_offsets[18+2];
} info_t;
typedef struct { int n, k, size; } sets_t;
static sets_t make_sets(int n, int k) { sets_t s = {n, k, po_choose(n, k)}; return s; }

typedef struct { const info_t* I; int n; } helper_t;
static int H_k0(helper_t H) { return H.n >> 1; }
static int H_k1(helper_t H) { return H.n - H_k0(H); }
static int H_parity(helper_t H) { return (int)((H.n + H.I->root.ply) & 1); }
static uint64_t H_root0(helper_t H) { return H.I->root.side[(H.I->spots + H.n) & 1]; }
static uint64_t H_root1(helper_t H) { return H.I->root.side[(H.I->spots + H.n + 1) & 1]; }
static sets_t H_sets0(helper_t H) { return make_sets(H.I->spots, H_k0(H)); }
static sets_t H_sets1(helper_t H) { return make_sets(H.I->spots, H_k1(H)); }
static sets_t H_sets0p(helper_t H) { return make_sets(H.I->spots - H_k1(H), H_k0(H)); }
static sets_t H_sets1p(helper_t H) { return make_sets(H.I->spots - H_k0(H), H_k1(H)); }

static grab_t make_grab(int end, int nx, int ny, int ws) {
  grab_t g = {ny, end ? ws - nx*ny : 0, nx*ny};
  return g;
}
/* internal.h:37-47 */
int po_bottleneck(int spots) {
  int worst = 0, prev = 1;
  for (int n = 0; n <= spots; n++) {
    int next = po_choose(spots, n + 1);
    if (next) next *= po_choose(n + 1, (n + 1)/2);
    if (prev + next > worst) worst = prev + next;
    prev = next;
  }
  return worst;
}
int po_midsolve_workspace_size(int min_slice) { return po_bottleneck(36 - min_slice); }

static uint64_t side_of(const uint8_t* empty, sets_t sets, uint64_t set) {
  uint64_t s = 0;
  for (int i = 0; i < sets.k; i++) s |= (uint64_t)1 << empty[set >> 5*i & 0x1f];
  return s;
}

/* internal.h:49-69 */
static info_t make_info(po_high root) {
  info_t I; memset(&I, 0, sizeof I);
  I.root = root;
  I.spots = 36 - po_high_count(root);
  const uint64_t free_ = high_empty(root);
  int next = 0;
  for (int i = 0; i < 64; i++) if (free_ >> i & 1) I.empty[next++] = (uint8_t)i;
  const int ws = po_bottleneck(I.spots);
  for (int n = 0; n <= I.spots; n++) {
    helper_t H = {&I, n};
    I.sets0_offsets[n+1] = I.sets0_offsets[n] + H_sets0(H).size;
    I.sets1p_offsets[n+1] = I.sets1p_offsets[n] + H_sets1p(H).size;
    I.cs1ps_offsets[n+1] = I.cs1ps_offsets[n] + H_sets1p(H).size*(I.spots - H_k0(H));
    I.wins1_offsets[n+1] = I.wins1_offsets[n] + H_sets1(H).size;
    I.spaces[n] = make_grab(n & 1, H_sets1(H).size, H_sets0p(H).size, ws);
  }
  I.spaces[I.spots+1] = make_grab(0, 0, 0, 0);
  return I;
}

typedef struct { po_half after, before; } wins1_t;
typedef struct { po_half wins0; uint16_t child_s0s[18]; uint16_t offset0[90]; uint8_t empty1[18]; } set0_info_t;

/* internal.h:123-131 */
static wins1_t mid_wins1(const info_t* I, int n, int s1) {
  helper_t H = {I, n};
  const uint64_t side1 = H_root1(H) | side_of(I->empty, H_sets1(H), po_subset(I->spots, H_k1(H), s1));
  wins1_t w;
  w.after = po_halfsuper_wins(side1, H_parity(H));
  w.before = po_halfsuper_wins(side1, !H_parity(H));
  return w;
}
/* internal.h:133-148 */
static uint16_t make_cs1ps(const info_t* I, const uint64_t* sets1p, int n, int index) {
  helper_t H = {I, n};
  const int k0 = H_k0(H), k1 = H_k1(H);
  const int s1p = index / (I->spots - k0);
  const int j = index - s1p*(I->spots - k0);
  uint16_t c = (uint16_t)s1p;
  for (int a = 0; a < k1; a++) {
    const int s1p_a = (int)(sets1p[s1p] >> 5*a & 0x1f);
    if (j < s1p_a) c = (uint16_t)(c - po_choose(s1p_a - 1, a));
  }
  return c;
}
/* internal.h:174-259 */
static void make_set0_info(const info_t* I, int n, int s0, set0_info_t* I0) {
  memset(I0, 0, sizeof *I0);
  helper_t H = {I, n};
  const sets_t sets0 = H_sets0(H);
  const int k0 = H_k0(H), k1 = H_k1(H);
  const uint64_t set0 = po_subset(sets0.n, sets0.k, s0);
  const uint64_t side0 = H_root0(H) | side_of(I->empty, sets0, set0);
  I0->wins0 = po_halfsuper_wins(side0, H_parity(H));
  {
    const uint64_t free_ = SIDE_MASK & ~side0;
    int next = 0;
    for (int i = 0; i < I->spots; i++) if (free_ >> I->empty[i] & 1) I0->empty1[next++] = (uint8_t)i;
  }
  for (int i = 0; i < I->spots - k0; i++) {
    const int j = I0->empty1[i] - i;
    uint16_t c = (uint16_t)po_choose(I0->empty1[i], j + 1);
    for (int a = 0; a < k0; a++) c = (uint16_t)(c + po_choose((int)(set0 >> 5*a & 0x1f), a + (a >= j) + 1));
    I0->child_s0s[i] = c;
  }
  for (int a = 0; a < k1; a++)
    for (int q = 0; q < I->spots - k0; q++) {
      uint16_t offset = (uint16_t)(a ? 0 : s0);
      for (int i = I0->empty1[q] - q; i < k0; i++) {
        const int v = (int)(set0 >> 5*i & 0x1f);
        if (v > a) offset = (uint16_t)(offset - po_choose(v - a - 1, i));
      }
      I0->offset0[a*(I->spots - k0) + q] = offset;
    }
}

/* internal.h:358-408 */
static void inner(const info_t* I, int n, const uint16_t* cs1ps, const uint64_t* sets1p, const wins1_t* all_wins1,
                  po_half* results, po_half* workspace, const set0_info_t* I0, int s1p, int aggressive) {
  const uint64_t set1p = sets1p[s1p];
  const grab_t gin = I->spaces[n+1], gout = I->spaces[n];
  const int spots = I->spots;
  const int done = spots == n;
  const int k0 = n >> 1, k1 = n - k0;
  const int turn_aggressive = aggressive ^ (n & 1);
  uint32_t filled1p = 0;
  uint16_t s1 = 0, s0p = 0;
  for (int i = 0; i < k1; i++) {
    const int q = (int)(set1p >> 5*i & 0x1f);
    filled1p |= 1u << q;
    s1 = (uint16_t)(s1 + po_choose(I0->empty1[q], i + 1));
    s0p = (uint16_t)(s0p + I0->offset0[i*(spots - k0) + q]);
  }
  po_half us = {{0, 0}};
  if (done && !turn_aggressive) us = hnot(us);
  uint32_t unmoved = ~filled1p;
  for (int m = 0; m < spots - n; m++) {
    const uint32_t bit = unmoved & (0u - unmoved);
    const int i = __builtin_ctz(bit);
    unmoved &= ~bit;
    const po_half* in = workspace + gin.lo + (int64_t)I0->child_s0s[i]*gin.ny + cs1ps[s1p*(spots - k0) + i];
    us = hor(us, *in);
  }
  const wins1_t w1 = all_wins1[s1];
  const po_half inplay = hnot(hor(I0->wins0, w1.after));
  const po_half mask = turn_aggressive ? hnot(w1.after) : hnot((po_half){{0, 0}});
  us = hor(hand(inplay, us), hand(I0->wins0, mask));
  if (n <= 1) results[n + s1p] = us;
  workspace[gout.lo + (int64_t)s1*gout.ny + s0p] = hor(po_half_rmax(hnot(us)), w1.before);
}

/* midengine.cc:93-125 */
int po_midsolve_internal(po_high root, po_half* workspace, po_halves results[19]) {
  const info_t I = make_info(root);
  if (I.spots > 18 || I.spots < 0) return -1;
  uint64_t* sets1p = (uint64_t*)malloc(sizeof(uint64_t)*(size_t)(I.sets1p_offsets[I.spots+1] + 1));
  wins1_t* all_wins1 = (wins1_t*)malloc(sizeof(wins1_t)*(size_t)(I.wins1_offsets[I.spots+1] + 1));
  uint16_t* cs1ps = (uint16_t*)malloc(sizeof(uint16_t)*(size_t)(I.cs1ps_offsets[I.spots+1] + 1));
  for (int n = I.spots; n >= 0; n--) {
    helper_t H = {&I, n};
    const sets_t s1p = H_sets1p(H);
    uint64_t* sp = sets1p + I.sets1p_offsets[n];
    for (int i = 0; i < s1p.size; i++) sp[i] = po_subset(s1p.n, s1p.k, i);
    for (int s = 0; s < H_sets1(H).size; s++) all_wins1[I.wins1_offsets[n] + s] = mid_wins1(&I, n, s);
    const int ncs = s1p.size*(I.spots - H_k0(H));
    for (int i = 0; i < ncs; i++) cs1ps[I.cs1ps_offsets[n] + i] = make_cs1ps(&I, sp, n, i);
  }
  po_half raw[2][19];
  memset(raw, 0, sizeof raw);
  for (int aggressive = 0; aggressive < 2; aggressive++)
    for (int n = I.spots; n >= 0; n--) {
      helper_t H = {&I, n};
      const int ns0 = H_sets0(H).size, ns1p = H_sets1p(H).size;
      for (int s0 = 0; s0 < ns0; s0++) {
        set0_info_t I0; make_set0_info(&I, n, s0, &I0);
        for (int s1p = 0; s1p < ns1p; s1p++)
          inner(&I, n, cs1ps + I.cs1ps_offsets[n], sets1p + I.sets1p_offsets[n], all_wins1 + I.wins1_offsets[n],
                raw[aggressive], workspace, &I0, s1p, aggressive);
      }
    }
  free(sets1p); free(all_wins1); free(cs1ps);
  for (int i = 0; i < 19; i++) {
    results[i].win = raw[i == 0][i];       /* get(results[i], j) = raw[j ^ (i == 0)][i] */
    results[i].notlose = raw[i != 0][i];
  }
  return 0;
}

/* midengine.h:28-30 */
static int mid_value(const po_halves* r, int s) { return hget(r->win, s >> 1) + hget(r->notlose, s >> 1) - 1; }

/* midengine.cc:127-158 */
static int traverse(po_high board, const po_halves* supers, uint64_t* raws, int* values, int* count) {
  int value;
  int imm;
  if (po_high_done_and_value(board, &imm)) {
    value = imm;
  } else if (!high_middle(board)) {
    value = -1;
    const uint64_t empty = high_empty(board);
    int s = 0;
    for (int bit = 0; bit < 64; bit++)
      if (empty >> bit & 1) {
        const int v = traverse(po_high_place(board, bit), supers + ++s, raws, values, count);
        if (v > value) value = v;
      }
  } else {
    const po_halves* r = supers;
    value = -1;
    for (int s = 0; s < 8; s++) {
      const int q = s >> 1, d = s & 1 ? -1 : 1;
      const int v = mid_value(r, (d & 3) << 2*q);
      raws[*count] = po_high_raw(po_high_rotate(board, q, d)); values[*count] = v; (*count)++;
      if (-v > value) value = -v;
    }
  }
  raws[*count] = po_high_raw(board); values[*count] = value; (*count)++;
  return value;
}
int po_midsolve_traverse(po_high root, const po_halves* supers, uint64_t* raws, int* values) {
  int count = 0;
  traverse(root, supers, raws, values, &count);
  return count;
}
int po_midsolve(po_high root, po_half* workspace, uint64_t* raws, int* values) {
  po_halves supers[19];
  if (po_midsolve_internal(root, workspace, supers)) return -1;
  return po_midsolve_traverse(root, supers, raws, values);
}

typedef struct {
  int n; const po_high* roots; po_halves (*results)[19];
  volatile int next; pthread_mutex_t lock; int ws;
} mbatch_t;
static void* mbatch_worker(void* arg) {
  mbatch_t* B = (mbatch_t*)arg;
  po_half* ws = (po_half*)malloc(sizeof(po_half)*(size_t)B->ws);
  for (;;) {
    pthread_mutex_lock(&B->lock);
    const int i = B->next++;
    pthread_mutex_unlock(&B->lock);
    if (i >= B->n) break;
    po_midsolve_internal(B->roots[i], ws, B->results[i]);
  }
  free(ws);
  return 0;
}
int po_midsolve_batch(int n, const po_high* roots, po_halves (*results)[19], int threads) {
  mbatch_t B; memset(&B, 0, sizeof B);
  B.n = n; B.roots = roots; B.results = results;
  int ws = 1;
  for (int i = 0; i < n; i++) {
    const int spots = 36 - po_high_count(roots[i]);
    if (spots < 0 || spots > 18) return -1;
    const int b = po_bottleneck(spots);
    if (b > ws) ws = b;
  }
  B.ws = ws;
  pthread_mutex_init(&B.lock, 0);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; t++) pthread_create(&th[t], 0, mbatch_worker, &B);
  for (int t = 0; t < threads; t++) pthread_join(th[t], 0);
  free(th);
  pthread_mutex_destroy(&B.lock);
  return 0;
}
