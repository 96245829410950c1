/* Oracle end engine: sections (pentago/base/section.cc), line shapes and index state
 * (pentago/end/compute.cc:24-123) and the microline kernel (pentago/end/compute.cc:219-371).
 * TEST INFRASTRUCTURE ONLY (see po.h). */
#include "po.h"
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

static uint16_t quad(uint64_t s, int q) { return (uint16_t)(s >> 16*q); }

/* ---- sections: pentago/base/section.{h,cc} ---- */
static uint64_t section_sig(po_section s) {
  uint64_t r = 0;
  for (int q = 0; q < 4; q++) r |= (uint64_t)s.c[q][0] << 16*q | (uint64_t)s.c[q][1] << (16*q + 8);
  return r;
}
static int section_sum(po_section s) {
  int t = 0;
  for (int q = 0; q < 4; q++) t += s.c[q][0] + s.c[q][1];
  return t;
}
int po_section_shape(po_section s, int shape[4]) {
  for (int q = 0; q < 4; q++) {
    const int n = po_rmin(s.c[q][0], s.c[q][1], 0, 0);
    if (n < 0) return -1;
    shape[q] = n;
  }
  return 0;
}
/* section.cc:79-87 */
void po_section_transform(po_section s, int g, po_section* out) {
  static const uint8_t source[4][4] = {{0,1,2,3},{1,3,0,2},{3,2,1,0},{2,0,3,1}};
  const int r = g & 3;
  po_section t;
  for (int i = 0; i < 4; i++) memcpy(t.c[i], s.c[source[r][i]], 2);
  if (g & 4) { uint8_t tmp[2]; memcpy(tmp, t.c[0], 2); memcpy(t.c[0], t.c[3], 2); memcpy(t.c[3], tmp, 2); }
  *out = t;
}
/* section.cc:89-103 (standardize<8>) */
int po_section_standardize(po_section s, po_section* out) {
  po_section best = s; int best_g = 0;
  for (int g = 1; g < 8; g++) {
    po_section t; po_section_transform(s, g, &t);
    if (section_sig(best) > section_sig(t)) { best = t; best_g = g; }
  }
  *out = best;
  return best_g;
}
static po_section section_child(po_section s, int q) {
  po_section c = s;
  c.c[q][section_sum(s) & 1]++;
  return c;
}
/* section.cc:108-116 */
static void quadrant_permutation(int symmetry, int p[4]) {
  po_section s; memset(&s, 0, sizeof s);
  for (int i = 0; i < 4; i++) s.c[i][0] = (uint8_t)i;
  po_section t; po_section_transform(s, symmetry, &t);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) if (t.c[j][0] == i && t.c[j][1] == 0) { p[i] = j; break; }
}
static uint8_t invert_global(uint8_t g) { return (uint8_t)(g ^ ((~g >> 2 & (g & 1)) << 1)); }

/* ---- line state: pentago/end/compute.cc:24-123 ---- */
typedef struct {
  po_line line;
  po_line_info info;
  const uint16_t* rmin[4];
  int rmin_size[4];
  po_sym inverse_transform;
  int reflection_moves[4];
  uint8_t* refl_syms;        /* all_reflection_symmetries */
  int rs_offsets[5];
  int slice35;
} line_state;

static void ceil_shape(const int shape[4], int blocks[4]) {
  for (int i = 0; i < 4; i++) blocks[i] = (shape[i] + 7) / 8;
}
static void block_shape4(const int shape[4], const int block[4], int out[4]) {
  for (int i = 0; i < 4; i++) {
    int hi = 8*(block[i] + 1); if (hi > shape[i]) hi = shape[i];
    out[i] = hi - 8*block[i];
  }
}
static void line_block(const po_line* l, int i, int out[4]) {
  int k = 0;
  for (int a = 0; a < 4; a++) out[a] = a == l->dimension ? i : l->block_base[k++];
}

static int line_state_init(const po_line* l, line_state* S) {
  memset(S, 0, sizeof *S);
  S->line = *l;
  po_line_info* I = &S->info;
  const int dim = l->dimension;
  int shape[4];
  if (po_section_shape(l->section, shape)) return -1;
  if (l->section.c[dim][0] + l->section.c[dim][1] >= 9) return -2;
  const po_section child = section_child(l->section, dim);
  int cshape[4];
  if (po_section_shape(child, cshape)) return -3;
  int b0[4]; line_block(l, 0, b0);
  int blocks[4]; ceil_shape(shape, blocks);
  for (int a = 0; a < 4; a++) if (a != dim && b0[a] >= blocks[a]) return -4;
  if (l->length != blocks[dim]) return -5;
  block_shape4(shape, b0, I->output_shape);
  I->output_shape[dim] = shape[dim];
  memcpy(I->input_shape, I->output_shape, sizeof I->input_shape);
  I->input_shape[dim] = cshape[dim];
  S->slice35 = section_sum(l->section) == 35;
  if (S->slice35) memset(I->input_shape, 0, sizeof I->input_shape);
  I->dim = dim;
  I->length = l->length;
  /* block_stride = 8 * product of the cross-section block shape */
  int bs = 8;
  for (int a = 0; a < 4; a++) if (a != dim) bs *= I->output_shape[a];
  I->block_stride = bs;
  I->section_transform = po_section_standardize(child, &I->standard_child);
  quadrant_permutation(invert_global((uint8_t)I->section_transform), I->permutation);
  for (int a = 0; a < 4; a++) if (I->permutation[a] == dim) I->child_dim = a;
  I->input_blocks = (I->input_shape[dim] + 7) / 8;
  for (int a = 0; a < 4; a++) I->first_child_block[a] = b0[I->permutation[a]];
  {
    const int st = I->section_transform;
    po_sym a = {(uint8_t)((4 - (st & 3)) & 3), (uint8_t)((1+4+16+64)*(st & 3))};
    po_sym b = {(uint8_t)(st & 4), 0};
    S->inverse_transform = po_sym_mul(a, b);
  }
  for (int a = 0; a < 4; a++)
    S->rmin_size[a] = po_rmin(l->section.c[a][0], l->section.c[a][1], &S->rmin[a], 0);
  const uint16_t* child_rmin[4]; int child_moved[4];
  for (int a = 0; a < 4; a++)
    po_rmin(I->standard_child.c[a][0], I->standard_child.c[a][1], &child_rmin[a], &child_moved[a]);
  if (I->section_transform & 4)
    for (int a = 0; a < 4; a++) S->reflection_moves[a] = child_moved[a];
  int sis[4];
  for (int a = 0; a < 4; a++) sis[a] = I->input_shape[I->permutation[a]];
  const int total = sis[0] + sis[1] + sis[2] + sis[3];
  if (total) {
    S->refl_syms = (uint8_t*)calloc((size_t)total, 1);
    int off = 0;
    S->rs_offsets[0] = 0;
    for (int a = 0; a < 4; a++) {
      if (I->section_transform & 4)
        for (int j = 0; j < sis[a]; j++) {
          const uint16_t q = child_rmin[a][8*I->first_child_block[a] + j];
          const uint16_t rq = (uint16_t)(po_pack(po_reflect(po_unpack(q, 0))) + 2*po_pack(po_reflect(po_unpack(q, 1))));
          S->refl_syms[off + j] = (uint8_t)((po_rmin_inverse(rq) & 3) << 2*a);
        }
      off += sis[a];
      S->rs_offsets[a+1] = off;
    }
  }
  int64_t in_nodes = 1, out_nodes = 1;
  for (int a = 0; a < 4; a++) { in_nodes *= I->input_shape[a]; out_nodes *= I->output_shape[a]; }
  I->input_nodes = S->slice35 ? 0 : in_nodes + I->input_blocks;   /* +1 per input block */
  I->output_nodes = out_nodes;
  return 0;
}
static void line_state_free(line_state* S) { free(S->refl_syms); S->refl_syms = 0; }

int po_line_describe(const po_line* line, po_line_info* info) {
  line_state S;
  const int r = line_state_init(line, &S);
  if (!r) *info = S.info;
  line_state_free(&S);
  return r;
}

static void strides4(const int shape[4], int st[4]) {
  st[3] = 1; st[2] = shape[3]; st[1] = shape[2]*shape[3]; st[0] = shape[1]*shape[2]*shape[3];
}
static int dot4(const int a[4], const int b[4]) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2] + a[3]*b[3]; }
static po_super snot(po_super a) { for (int i = 0; i < 4; i++) a.w[i] = ~a.w[i]; return a; }
static po_super sand(po_super a, po_super b) { for (int i = 0; i < 4; i++) a.w[i] &= b.w[i]; return a; }
static po_super sorr(po_super a, po_super b) { for (int i = 0; i < 4; i++) a.w[i] |= b.w[i]; return a; }
static uint8_t local_mul(uint8_t a, uint8_t b) {
  return (uint8_t)((((a & 0x33) + (b & 0x33)) & 0x33) + (((a & 0xcc) + (b & 0xcc)) & 0xcc));
}
/* local * general (symmetry.h:96-99) */
static po_sym local_times(uint8_t a, po_sym b) {
  po_sym r = {b.global, local_mul(po_commute(b.global, a), b.local)};
  return r;
}

/* pentago/end/compute.cc:219-356; base = the microline's 3-D cross-section coordinates */
static void compute_microline_range(const line_state* S, const int base3[3], const po_super* input, po_super* output,
                                   int i_lo, int i_hi) {
  const po_line_info* I = &S->info;
  const int dim = I->dim, length = I->length, child_dim = I->child_dim;
  int full_base[4];
  { int k = 0; for (int a = 0; a < 4; a++) full_base[a] = a == dim ? 0 : base3[k++]; }
  const int child_length = I->input_blocks;
  int first_block[4]; line_block(&S->line, 0, first_block);
  int first_node[4];
  for (int a = 0; a < 4; a++) first_node[a] = 8*first_block[a] + full_base[a];
  uint64_t base_board = 0;
  for (int a = 0; a < 4; a++) if (a != dim) base_board |= (uint64_t)S->rmin[a][first_node[a]] << 16*a;
  const int turn = section_sum(S->line.section) & 1;
  const uint64_t base_side0 = po_unpack_board(base_board, turn), base_side1 = po_unpack_board(base_board, 1 - turn);
  const uint16_t* rmin = S->rmin[dim];
  const int rmin_size = S->rmin_size[dim];
  const int dim_shift = 16*dim;
  const int block_stride = I->block_stride;

  int obs[4]; memcpy(obs, I->output_shape, sizeof obs); obs[dim] = 8;
  int ost[4]; strides4(obs, ost);
  const int output_stride = ost[dim];
  const int output_base = dot4(ost, full_base);

  int ibs[4]; for (int a = 0; a < 4; a++) ibs[a] = obs[I->permutation[a]];
  int ist[4]; strides4(ibs, ist);
  const int input_stride = ist[child_dim];
  int fcb[4]; for (int a = 0; a < 4; a++) fcb[a] = full_base[I->permutation[a]];
  int fcn[4]; for (int a = 0; a < 4; a++) fcn[a] = 8*I->first_child_block[a] + fcb[a];
  int rcb[4];
  for (int a = 0; a < 4; a++) rcb[a] = fcb[a] ^ (child_dim != a && fcn[a] < S->reflection_moves[a]);
  const int input_base = dot4(ist, rcb);
  const int line_moves = S->reflection_moves[child_dim];

  obs[dim] = (rmin_size & 7) ? (rmin_size & 7) : 8;
  int ost2[4]; strides4(obs, ost2);
  const int last_output_base = dot4(ost2, full_base);
  ibs[child_dim] = (I->input_shape[dim] & 7) ? (I->input_shape[dim] & 7) : 8;
  int ist2[4]; strides4(ibs, ist2);
  const int last_input_base = dot4(ist2, rcb);

  const uint8_t* symmetries = S->refl_syms ? S->refl_syms + S->rs_offsets[child_dim] : 0;
  uint8_t local_symmetry = 0;
  if (S->refl_syms)
    for (int a = 0; a < 4; a++)
      if (a != child_dim) local_symmetry = (uint8_t)(local_symmetry + S->refl_syms[S->rs_offsets[a] + fcb[a]]);
  po_sym base_transform = {S->inverse_transform.global, local_mul(S->inverse_transform.local, local_symmetry)};

  if (i_hi > rmin_size) i_hi = rmin_size;
  for (int i = i_lo; i < i_hi; i++) {
    const uint16_t q = rmin[i];
    const uint16_t sq0 = po_unpack(q, turn), sq1 = po_unpack(q, 1 - turn);
    const uint64_t side0 = base_side0 | (uint64_t)sq0 << dim_shift, side1 = base_side1 | (uint64_t)sq1 << dim_shift;
    const po_super wins0 = po_super_wins(side0), wins1 = po_super_wins(side1);
    const po_super immediate = sorr(wins0, wins1);
    po_super wins = {{0,0,0,0}}, not_loss = {{0,0,0,0}};
    uint16_t moves = (uint16_t)(511 & ~(sq0 | sq1));
    while (moves) {
      const uint16_t move = (uint16_t)(moves & -moves);
      moves ^= move;
      const uint64_t new_side0 = side0 | (uint64_t)move << dim_shift;
      const po_super new_wins = po_super_wins(new_side0);
      wins = sorr(wins, new_wins);
      if (S->slice35) {
        wins = sorr(wins, po_rmax(sand(new_wins, snot(wins1))));
        not_loss = sorr(not_loss, po_rmax(sorr(new_wins, snot(wins1))));
      } else {
        const uint16_t new_q = (uint16_t)(q + (1 + turn)*po_pack(move));
        const uint16_t ir = po_rmin_inverse(new_q);
        const int j = ir / 4;
        const po_sym t = local_times((uint8_t)((ir & 3) << 2*dim), base_transform);
        const po_sym sym = {t.global, local_mul(t.local, symmetries[j])};
        const int64_t idx = (int64_t)((j >> 3) == child_length - 1 ? last_input_base : input_base)
                          + (int64_t)(block_stride + 1)*(j >> 3) + (int64_t)input_stride*((j ^ (j < line_moves)) & 7);
        const po_super* child = input + 2*idx;
        wins = sorr(wins, po_rmax(po_transform_super(sym, snot(child[1]))));
        not_loss = sorr(not_loss, po_rmax(po_transform_super(sym, snot(child[0]))));
      }
    }
    const int64_t oidx = (int64_t)((i >> 3) == length - 1 ? last_output_base : output_base)
                       + (int64_t)block_stride*(i >> 3) + (int64_t)output_stride*(i & 7);
    po_super* result = output + 2*oidx;
    result[0] = sorr(sand(wins, snot(immediate)), sand(wins0, snot(wins1)));
    result[1] = sorr(sand(not_loss, snot(immediate)), wins0);
  }
}

static void compute_microline(const line_state* S, const int base3[3], const po_super* input, po_super* output) {
  compute_microline_range(S, base3, input, output, 0, 1 << 30);
}

static int microline_count(const line_state* S) {
  int n = 1;
  for (int a = 0; a < 4; a++) if (a != S->info.dim) n *= S->info.output_shape[a];
  return n;
}
static void microline_base(const line_state* S, int m, int base3[3]) {
  int cs[3], k = 0;
  for (int a = 0; a < 4; a++) if (a != S->info.dim) cs[k++] = S->info.output_shape[a];
  base3[2] = m % cs[2]; m /= cs[2];
  base3[1] = m % cs[1]; m /= cs[1];
  base3[0] = m;
}

int po_compute_microlines(const po_line* line, const po_super* input, po_super* output, int lo, int hi) {
  line_state S;
  const int r = line_state_init(line, &S);
  if (r) return r;
  const int n = microline_count(&S);
  if (hi > n) hi = n;
  for (int m = lo; m < hi; m++) {
    int b[3]; microline_base(&S, m, b);
    compute_microline(&S, b, input, output);
  }
  line_state_free(&S);
  return 0;
}
int po_compute_line(const po_line* line, const po_super* input, po_super* output) {
  return po_compute_microlines(line, input, output, 0, 1 << 30);
}

/* Thread pool over microlines of many lines (pentago/end/compute.cc:394-404: one job per microline). */
typedef struct {
  int nlines;
  line_state* states;
  const po_super* const* inputs;
  po_super* const* outputs;
  int64_t* prefix;     /* microline prefix sums */
  int64_t total;
  volatile int64_t next;
  pthread_mutex_t lock;
} pool_t;

static void* pool_worker(void* arg) {
  pool_t* P = (pool_t*)arg;
  for (;;) {
    pthread_mutex_lock(&P->lock);
    const int64_t lo = P->next;
    P->next = lo + 1;
    pthread_mutex_unlock(&P->lock);
    if (lo >= P->total) break;
    int l = 0;
    while (P->prefix[l + 1] <= lo) l++;
    int b[3]; microline_base(&P->states[l], (int)(lo - P->prefix[l]), b);
    compute_microline(&P->states[l], b, P->inputs[l], P->outputs[l]);
  }
  return 0;
}

int po_compute_lines_threaded(int n, const po_line* lines, const po_super* const* inputs,
                              po_super* const* outputs, int threads) {
  pool_t P; memset(&P, 0, sizeof P);
  P.nlines = n;
  P.states = (line_state*)calloc((size_t)n, sizeof(line_state));
  P.prefix = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  for (int i = 0; i < n; i++) {
    const int r = line_state_init(&lines[i], &P.states[i]);
    if (r) { for (int k = 0; k < i; k++) line_state_free(&P.states[k]); free(P.states); free(P.prefix); return r; }
    P.prefix[i + 1] = P.prefix[i] + microline_count(&P.states[i]);
  }
  P.total = P.prefix[n];
  P.inputs = inputs; P.outputs = outputs;
  pthread_mutex_init(&P.lock, 0);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; t++) pthread_create(&th[t], 0, pool_worker, &P);
  for (int t = 0; t < threads; t++) pthread_join(th[t], 0);
  free(th);
  pthread_mutex_destroy(&P.lock);
  for (int i = 0; i < n; i++) line_state_free(&P.states[i]);
  free(P.states); free(P.prefix);
  return 0;
}
int po_compute_line_threaded(const po_line* line, const po_super* input, po_super* output, int threads) {
  return po_compute_lines_threaded(1, line, &input, &output, threads);
}

/* ---------------- recursive node evaluator (tests only) ----------------
 * Real game values for any node of a standard section.  For every open dimension d the
 * reference's microline kernel runs on the line through the node (restricted to the node's
 * position), with the child block line filled the way a caller fills line.input
 * (pentago/end/compute.h:81, blocks of block_stride+1 nodes, C-order within each child block)
 * but only at the cross positions the microline can read: the node's own cross position and
 * its reflection partner (index^1) in each non-line dimension.  Child values are evaluated
 * recursively and memoised; dimension contributions are OR-merged as in
 * pentago/end/block_store.cc:173-229. */
typedef struct node_entry { uint64_t key; po_super v[2]; struct node_entry* next; } node_entry;
#define EVAL_BUCKETS (1 << 20)
static node_entry** eval_table;
static pthread_mutex_t eval_lock = PTHREAD_MUTEX_INITIALIZER;

void po_eval_reset(void) {
  if (!eval_table) return;
  for (int i = 0; i < EVAL_BUCKETS; i++) {
    node_entry* e = eval_table[i];
    while (e) { node_entry* n = e->next; free(e); e = n; }
    eval_table[i] = 0;
  }
}
static uint64_t node_key(po_section s, const int idx[4]) {
  uint64_t k = 0;
  for (int q = 0; q < 4; q++) k = k*100 + s.c[q][0]*10 + s.c[q][1];
  for (int a = 0; a < 4; a++) k = k*512 + (uint64_t)idx[a];
  return k;
}

static void eval_node(po_section s, const int idx[4], po_super out[2]);

static void eval_node_dim(po_section s, const int idx[4], int d, po_super out[2]) {
  int shape[4]; po_section_shape(s, shape);
  po_line L; L.section = s; L.dimension = (uint8_t)d; L.length = (uint8_t)((shape[d] + 7) / 8);
  { int k = 0; for (int a = 0; a < 4; a++) if (a != d) L.block_base[k++] = (uint8_t)(idx[a] / 8); }
  line_state S; line_state_init(&L, &S);
  const po_line_info* I = &S.info;
  po_super* input = 0;
  if (!S.slice35) {
    input = (po_super*)calloc((size_t)2*I->input_nodes, sizeof(po_super));
    int cshape[4]; po_section_shape(I->standard_child, cshape);
    int full_base[4];
    for (int a = 0; a < 4; a++) full_base[a] = a == d ? 0 : idx[a] % 8;
    int fcb[4]; for (int a = 0; a < 4; a++) fcb[a] = full_base[I->permutation[a]];
    const int cd = I->child_dim;
    for (int combo = 0; combo < 16; combo++) {
      if (combo >> cd & 1) continue;
      int loc[4], ok = 1;
      for (int a = 0; a < 4; a++) loc[a] = fcb[a] ^ (combo >> a & 1);
      int cblock[4]; memcpy(cblock, I->first_child_block, sizeof cblock);
      for (int jj = 0; jj < cshape[cd] && ok; jj++) {
        cblock[cd] = jj / 8;
        int bsh[4]; block_shape4(cshape, cblock, bsh);
        int local[4]; memcpy(local, loc, sizeof local); local[cd] = jj % 8;
        int cidx[4];
        for (int a = 0; a < 4; a++) {
          if (local[a] >= bsh[a]) { ok = 0; break; }
          cidx[a] = 8*cblock[a] + local[a];
        }
        if (!ok) break;
        po_super v[2]; eval_node(I->standard_child, cidx, v);
        int st[4]; strides4(bsh, st);
        const int64_t pos = (int64_t)(I->block_stride + 1)*cblock[cd] + dot4(st, local);
        input[2*pos] = v[0]; input[2*pos + 1] = v[1];
      }
    }
  }
  po_super* lout = (po_super*)calloc((size_t)2*I->output_nodes, sizeof(po_super));
  int b3[3]; { int k = 0; for (int a = 0; a < 4; a++) if (a != d) b3[k++] = idx[a] % 8; }
  compute_microline_range(&S, b3, input, lout, idx[d], idx[d] + 1);
  /* locate the node in the line output: block idx[d]/8, C-order within the block */
  int bsh[4]; int blk[4];
  for (int a = 0; a < 4; a++) blk[a] = idx[a] / 8;
  block_shape4(shape, blk, bsh);
  int st[4]; strides4(bsh, st);
  int local[4]; for (int a = 0; a < 4; a++) local[a] = idx[a] % 8;
  const int64_t pos = (int64_t)I->block_stride*blk[d] + dot4(st, local);
  out[0] = lout[2*pos]; out[1] = lout[2*pos + 1];
  free(lout); free(input); line_state_free(&S);
}

static void eval_node(po_section s, const int idx[4], po_super out[2]) {
  if (!eval_table) eval_table = (node_entry**)calloc(EVAL_BUCKETS, sizeof(node_entry*));
  const uint64_t key = node_key(s, idx);
  const int h = (int)((key * 0x9E3779B97F4A7C15ull) >> 44);
  for (node_entry* e = eval_table[h]; e; e = e->next)
    if (e->key == key) { out[0] = e->v[0]; out[1] = e->v[1]; return; }
  po_super acc[2]; memset(acc, 0, sizeof acc);
  for (int d = 0; d < 4; d++) {
    if (s.c[d][0] + s.c[d][1] >= 9) continue;
    po_super v[2]; eval_node_dim(s, idx, d, v);
    acc[0] = sorr(acc[0], v[0]); acc[1] = sorr(acc[1], v[1]);
  }
  node_entry* e = (node_entry*)malloc(sizeof(node_entry));
  e->key = key; e->v[0] = acc[0]; e->v[1] = acc[1]; e->next = eval_table[h]; eval_table[h] = e;
  out[0] = acc[0]; out[1] = acc[1];
}

int po_eval_node(po_section standard, const int index[4], po_super out[2]) {
  int shape[4];
  if (po_section_shape(standard, shape)) return -1;
  for (int a = 0; a < 4; a++) if (index[a] < 0 || index[a] >= shape[a]) return -1;
  pthread_mutex_lock(&eval_lock);
  eval_node(standard, index, out);
  pthread_mutex_unlock(&eval_lock);
  return 0;
}
