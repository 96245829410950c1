/* Oracle: a plain-C restatement of the reference's retrograde superscore path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (pentago_amd/) links,
 * imports or executes this code.  It is used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker, never as the thing measured or shipped.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference repository root.
 *
 * Parity pins (what checks each part of this restatement):
 *   - lookup tables: the reference's SHA-1 digests (pentago/base/precompute.cc:169-1038)
 *   - super_wins:    brute force won() over all 256 rotations (pentago/base/super_test.cc:20-69)
 *   - rmax:          OR over single rotations (pentago/base/super_test.cc:71-87)
 *   - transform:     group-theoretic definition + meaningless equivariance
 *                    (pentago/base/super_test.cc:187-212, golden 223-240)
 *   - halfsuper:     split/merge/rmax/halfsuper_wins relations (pentago/mid/mid_test.cc:174-272)
 *   - subsets:       enumerative definitions (pentago/mid/mid_test.cc:274-347)
 *   - midsolve:      163-value known answer (pentago/mid/benchmark_test.cc:39-58),
 *                    bottleneck sizes (pentago/mid/mid_test.cc:129-134), brute force
 *   - end line:      production sparse samples at slices 33-35
 *                    (pentago/high/high_test.cc:41-58) through a recursive block evaluator,
 *                    brute-force minimax
 */
#ifndef PO_H
#define PO_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 256-bit set of local rotations; bit r0+4*(r1+4*(r2+4*r3)), little-endian words.
 * pentago/base/superscore.h:29-39 */
typedef struct { uint64_t w[4]; } po_super;
/* 128-bit parity half.  pentago/mid/halfsuper_c.h:9-14 */
typedef struct { uint64_t w[2]; } po_half;
/* (win, notlose) pair.  pentago/mid/halfsuper_c.h:16-18 */
typedef struct { po_half win, notlose; } po_halves;
/* Global (3 bits) + local (8 bits) symmetry.  pentago/base/symmetry.h:36-47 */
typedef struct { uint8_t global, local; } po_sym;
/* section counts[q][color].  pentago/base/section.h:32-33 */
typedef struct { uint8_t c[4][2]; } po_section;
/* Block line.  pentago/end/line.h:15-35 */
typedef struct { po_section section; uint8_t dimension, length; uint8_t block_base[3]; } po_line;
/* high_board_s.  pentago/high/board_c.h:9-12 */
typedef struct { uint64_t side[2]; uint32_t ply; } po_high;

/* ---- tables (pentago/base/precompute.cc) ---- */
void po_init(void);
/* SHA-1 hex digest of a named table as the reference hashes it (check / check<uint64_t>). */
int po_table_hash(const char* name, char out[41]);
void po_sha1_hex(const void* data, size_t n, char out[41]);
uint16_t po_pack(uint16_t side_quadrant);
uint16_t po_unpack(uint16_t q, int s);
uint16_t po_rotate_left(uint16_t side_quadrant);
uint16_t po_rotate_right(uint16_t side_quadrant);
uint16_t po_reflect(uint16_t side_quadrant);
uint8_t po_commute(int g, int l);
int po_rmin(int black, int white, const uint16_t** quads, int* moved);
uint16_t po_rmin_inverse(uint16_t q);
const uint64_t* po_superwin_info(void);    /* 4*512*5 supers as u64 */
const uint16_t* po_rmin_flat(void);
const uint16_t* po_rmin_offsets(void);
const uint16_t* po_rmin_inverse_table(void);

/* ---- supers (pentago/base/superscore.{h,cc}, symmetry.{h,cc}) ---- */
po_super po_rmax(po_super f);
po_super po_super_wins(uint64_t side);
int po_won(uint64_t side);
po_sym po_sym_mul(po_sym a, po_sym b);
po_sym po_sym_inverse(po_sym s);
po_super po_transform_super(po_sym s, po_super c);
uint64_t po_transform_side(po_sym s, uint64_t side);
uint64_t po_transform_board(po_sym s, uint64_t board);
int po_meaningless(uint64_t board, uint64_t salt);
po_super po_super_meaningless(uint64_t board, uint64_t salt);
uint64_t po_pack_board(uint64_t side0, uint64_t side1);
uint64_t po_unpack_board(uint64_t board, int s);

/* ---- sections and lines (pentago/base/section.cc, pentago/end/compute.cc) ---- */
int po_section_shape(po_section s, int shape[4]);
int po_section_standardize(po_section s, po_section* out);   /* returns g */
void po_section_transform(po_section s, int g, po_section* out);
/* Shapes of a line's input (unstandardized) and output, node counts, block stride. */
typedef struct {
  int input_shape[4], output_shape[4];
  int block_stride, input_blocks, length, dim, child_dim;
  int64_t input_nodes, output_nodes;   /* including the +1 per input block pad */
  po_section standard_child;
  int section_transform;
  int permutation[4];
  int first_child_block[4];
} po_line_info;
int po_line_describe(const po_line* line, po_line_info* info);
/* Run every microline of a line (single thread).  input: 2 supers per node, reference layout. */
int po_compute_line(const po_line* line, const po_super* input, po_super* output);
/* Threaded: one job per microline over `threads` pthreads; returns 0 on success. */
int po_compute_line_threaded(const po_line* line, const po_super* input, po_super* output, int threads);
/* Threaded over many lines (one job per microline across all lines). */
int po_compute_lines_threaded(int n, const po_line* lines, const po_super* const* inputs,
                              po_super* const* outputs, int threads);
/* Microlines [lo,hi) of one line, for bounded samples. */
int po_compute_microlines(const po_line* line, const po_super* input, po_super* output, int lo, int hi);

/* Recursive, memoised block evaluator over real game values (tests only).
 * Value of one node (we-win, we-win-or-tie) of a standard section. */
void po_eval_reset(void);
int po_eval_node(po_section standard, const int index[4], po_super out[2]);

/* ---- counts (pentago/base/count.cc, pentago/end/block_store.cc) ---- */
/* (wins, wins-or-ties, total) of one node over its distinct rotated positions:
 * popcounts_over_stabilizers (pentago/base/count.cc:235-262).  Returns -1 where the reference
 * asserts (popcounts not multiples of the stabilizer size). */
int po_popcounts_over_stabilizers(uint64_t board, const po_super wins[2], uint16_t out[3]);
/* Sum over every node of a standard section, node values from po_eval_node: count_block_wins
 * (pentago/end/block_store.cc:330-349) summed over the section's blocks. */
int po_section_counts(po_section s, uint64_t out[3]);

/* ---- supertensor filter (pentago/data/filter.h) ---- */
/* n nodes of (black, white) supers, in place. */
void po_interleave(int64_t n, po_super* nodes);
void po_uninterleave(int64_t n, po_super* nodes);

/* ---- brute force (tests only) ---- */
/* value in {-1,0,1} of side-to-move `me` vs `them`, next action a stone placement. */
int po_brute_value(uint64_t me, uint64_t them);
/* brute-force super: (we-win, we-win-or-tie) for all 256 local rotations of the position */
void po_brute_super(uint64_t me, uint64_t them, po_super out[2]);

/* ---- mid engine (pentago/mid/) ---- */
po_half po_half_rmax(po_half h);
void po_split(po_super s, po_half out[2]);
po_super po_merge(po_half even, po_half odd);
po_half po_halfsuper_wins(uint64_t side, int parity);
int po_choose(int n, int k);
uint64_t po_subset(int n, int k, int index);
uint32_t po_subset_mask(int n, int k, int index);
int po_subset_index(int n, int k, uint64_t set);
int po_bottleneck(int spots);
int po_midsolve_workspace_size(int min_slice);
/* results: 1+18 (win, notlose) halfsuper pairs.  workspace: >= bottleneck entries.
 * pentago/mid/midengine.cc:93-125 */
int po_midsolve_internal(po_high root, po_half* workspace, po_halves results[19]);
/* Full midsolve + traverse: writes up to 163 (raw, value) pairs; returns count.
 * pentago/mid/midengine.cc:161-169 (midsolve) and :127-158 (midsolve_traverse) */
int po_midsolve(po_high root, po_half* workspace, uint64_t* raws, int* values);
int po_midsolve_traverse(po_high root, const po_halves* supers, uint64_t* raws, int* values);
/* Threaded batch: one board per thread.  returns 0 on success. */
int po_midsolve_batch(int n, const po_high* roots, po_halves (*results)[19], int threads);

/* high_board_t (pentago/high/board.{h,cc}) */
po_high po_high_from_board(uint64_t board, int middle);
uint64_t po_high_board(po_high b);
uint64_t po_high_raw(po_high b);
po_high po_high_place(po_high b, int bit);
po_high po_high_rotate(po_high b, int q, int d);
int po_high_done_and_value(po_high b, int* value);
int po_high_count(po_high b);

#ifdef __cplusplus
}
#endif
#endif
