/* pentago_amd: MI355X (gfx950) engine for pentago's retrograde superscore path.
 *
 * Plain C ABI: plain pointers and sizes, no C++ or framework types.  Device pointers are
 * HIP device addresses; `stream` is a hipStream_t passed as void* (NULL = default stream).
 * Every entry point returns a pe_status; pe_last_error() describes the last failure on the
 * calling thread.  The reference interface each entry point replaces is cited (paths are
 * relative to the girving/pentago repository).
 */
#ifndef PENTAGO_AMD_H
#define PENTAGO_AMD_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PE_OK = 0,
  PE_ERR_INVALID = 1,     /* bad argument (section, line, board, size) */
  PE_ERR_DEVICE = 2,      /* HIP runtime error or no device */
  PE_ERR_TABLES = 3,      /* lookup table self-check failed */
  PE_ERR_CAPACITY = 4,    /* request exceeds a fixed capacity or budget */
} pe_status;

/* ---------------- library ---------------- */

/* Build and self-check the lookup tables (SHA-1 against the reference digests of
 * pentago/base/precompute.cc) and upload them to the current device.
 * Replaces the generated tables of pentago/base/gen/tables.h. */
pe_status pe_init(void);
const char* pe_last_error(void);
/* SHA-1 digest (40 hex chars + NUL) of a named table, hashed as the reference hashes it. */
pe_status pe_table_digest(const char* name, char out[41]);

/* ---------------- end engine: block lines ---------------- */

/* section_t: per-quadrant (black, white) stone counts.  pentago/base/section.h:32-33 */
typedef struct { uint8_t counts[4][2]; } pe_section;

/* line_t: one block line.  pentago/end/line.h:15-35 (dimension:2, length:6, block_base[3]) */
typedef struct {
  pe_section section;
  uint8_t dimension;
  uint8_t length;          /* ceil(shape[dimension] / 8) */
  uint8_t block_base[3];   /* block indices of the other three dimensions, in order */
} pe_line;

/* Rotation-minimal quadrant counts along each dimension: section_t::shape(),
 * pentago/base/section.cc:43-48. */
pe_status pe_section_shape(const pe_section* section, int32_t shape[4]);
/* Standard form under the 8 global symmetries and the symmetry used:
 * section_t::standardize<8>(), pentago/base/section.cc:89-103. */
pe_status pe_section_standardize(const pe_section* section, pe_section* standard, int32_t* symmetry);

/* Shapes and sizes derived from a line: pentago/end/compute.cc:24-123
 * (line_data_t and line_details_t). */
typedef struct {
  int32_t input_shape[4];       /* unstandardized child block line; zero at slice 35 */
  int32_t output_shape[4];
  int32_t block_stride;         /* nodes per output block (last block may be shorter) */
  int32_t input_blocks;
  int32_t length;
  int32_t dimension;
  int32_t child_dimension;
  int64_t input_nodes;          /* input buffer nodes: input_shape product + input_blocks */
  int64_t output_nodes;
  pe_section standard_child;
  int32_t section_transform;
  int32_t permutation[4];
  int32_t first_child_block[4];
} pe_line_info;

/* Nodes are two 256-bit supers (we-win, we-win-or-tie), 64 bytes, little-endian words.
 * Input layout: input_blocks child blocks, block k at node offset k*(block_stride+1),
 * C-order within each block (pentago/end/compute.cc:134-139).  Output layout: length
 * blocks at node offset k*block_stride (compute.cc:147-152). */
pe_status pe_line_describe(const pe_line* line, pe_line_info* info);

/* Compute one block line on the device: one kernel launch on `stream`, returning without
 * waiting for it.  Replaces schedule_compute_line(line_details_t&), pentago/end/compute.h:123
 * (compute.cc:394-404).  Thread-safe (the reference calls it from pool threads,
 * pentago/mpi/flow.cc:465-494).  The line's descriptor travels as kernel arguments, so a
 * call takes no device memory; the per-(section, dimension) move and reflection tables are
 * built on a section's first line and kept for the process (a few KB each).  d_input may be
 * NULL only for 35-stone sections (no child data). */
pe_status pe_end_compute_line(const pe_line* line, const void* d_input, void* d_output, void* stream);

/* Same, then call wakeup(ctx) exactly once, on a HIP runtime thread, once the output is
 * complete: the reference's completion callback line_details_t::wakeup
 * (pentago/end/compute.h:84-86, compute.cc:359-370).  wakeup must not call HIP. */
pe_status pe_end_schedule_line(const pe_line* line, const void* d_input, void* d_output, void* stream,
                               void (*wakeup)(void* ctx), void* ctx);

/* Many lines in one launch (lines of different sections and lengths may be mixed), built as
 * a line plan and run; synchronous: returns after the lines are computed.  The batched form
 * of pentago/end/compute.cc:394-404. */
pe_status pe_end_compute_lines(int n, const pe_line* lines, const void* const* d_inputs,
                               void* const* d_outputs, void* stream);

/* Line plans: a device-resident set of line jobs built once and launched any number of
 * times (the batched form of the reference's per-microline thread-pool scheduling,
 * pentago/end/compute.cc:378-404).  A job may read its child blocks and write its output
 * blocks at arbitrary node offsets (for example straight out of section-resident slabs),
 * restrict the computed indices along the line dimension, and OR its result into the output
 * (the block-store merge of pentago/end/block_store.cc:173-229, done in place on the device).
 * Jobs in one plan must not write the same output nodes. */
typedef struct pe_end_plan pe_end_plan;
typedef struct {
  pe_line line;
  const void* input;                    /* device base of the child data (NULL at slice 35) */
  void* output;                         /* device base of the output */
  const int64_t* input_block_offsets;   /* host: node offset of input block k (input_blocks
                                           entries), or NULL for k*(block_stride+1) */
  const int64_t* output_block_offsets;  /* host: node offset of output block k (length
                                           entries), or NULL for k*block_stride */
  int32_t i_begin, i_end;               /* indices along the line dimension; 0,0 = all */
  int32_t accumulate;                   /* nonzero: OR into the output */
  int32_t pad;
} pe_line_job;
pe_status pe_end_plan_create(int n, const pe_line_job* jobs, pe_end_plan** plan);
pe_status pe_end_plan_run(const pe_end_plan* plan, void* stream);
/* Output nodes one run computes. */
int64_t pe_end_plan_nodes(const pe_end_plan* plan);
pe_status pe_end_plan_destroy(pe_end_plan* plan);
/* Launch form of a plan's fused blocks (the blocks of every section of the plan, of any
 * size, in one launch).  0: a static grid, 16 workgroups per block.  n > 0: a persistent
 * grid of n workgroups per CU fed from device work queues, one per XCD over that XCD's run
 * of blocks, workgroups taking from the other XCDs' queues once their own is empty.  The
 * batched form of the reference's per-microline job pool (pentago/end/compute.cc:378-404).
 * Runs of one plan must not overlap (they share the queue heads). */
pe_status pe_end_plan_set_launch(pe_end_plan* plan, int queue_wgs_per_cu);
/* Time the launch forms {grid, 2, 3, 4, 5, 6 workgroups per CU} on this plan (warm-up, then the
 * faster of two runs each, on `stream`; the plan's output is recomputed, identical each
 * time), keep the fastest for the plan and as the current device's default for plans created
 * later.  PE_ERR_INVALID for a finishing pass (pe_end_slabs_plan_create_dims with
 * accumulate 3), whose run is not repeatable.  No counterpart in the reference. */
typedef struct {
  int32_t queue_wgs_per_cu;   /* the choice (0: grid) */
  int32_t candidates;
  int32_t forms[8];
  float ms[8];                /* timed milliseconds per candidate */
} pe_end_tuning;
pe_status pe_end_plan_autotune(pe_end_plan* plan, void* stream, pe_end_tuning* out);
/* The launch form the current device's later plans start with (what pe_end_plan_autotune
 * sets). */
pe_status pe_end_set_default_launch(int queue_wgs_per_cu);

/* Section slabs: the device-resident storage unit.  A dense box of blocks [lo, hi) per axis
 * of one section, block (b - lo) in C-order over (hi - lo), one 4096-node slot per block,
 * nodes in C-order over the block's own shape (edge blocks use a prefix of their slot). */
typedef struct {
  pe_section section;
  int32_t lo[4], hi[4];
  void* data;
} pe_slab;

/* Plan computing every output block of `output` completely (all open dimensions,
 * OR-merged), reading child blocks from `children` (slabs of the standard child sections;
 * every child block any line through the box reads must be covered, else PE_ERR_INVALID).
 * mode 0: one fused kernel per run (each node's dimensions computed together and written
 *         once -- the line kernel of pentago/end/compute.cc:219-371 for all four lines
 *         through a block, merged as pentago/end/block_store.cc:173-229);
 * mode 1: four line passes, the later ones OR-accumulating (for comparison).
 * Output sections with 35 stones (one open dimension, no child data; the move fills the
 * board, compute.cc:329-332) run on the fused kernel's no-children form in mode 0. */
pe_status pe_end_slab_plan_create(const pe_slab* output, int nchildren, const pe_slab* children, int mode,
                                  pe_end_plan** plan);
/* The same for several output slabs (of one or several sections) in one plan: mode 0 runs
 * every block of every output in one fused launch. */
pe_status pe_end_slabs_plan_create(int noutputs, const pe_slab* outputs, int nchildren, const pe_slab* children,
                                   int mode, pe_end_plan** plan);

/* The fused plan of pe_end_slabs_plan_create (mode 0) for a subset of the open dimensions:
 * bit d of `dims` set = dimension d computed (its moves and its child lines), the others
 * skipped, so only the computed dimensions' child slabs are needed.  For sections too large
 * to hold with all their children at once; sections with at most 34 stones.  `accumulate`:
 *   0  store the result of the computed dimensions;
 *   1  OR it into the output (the block-store merge, pentago/end/block_store.cc:173-229), so
 *      runs over disjoint subsets covering every open dimension give the full result;
 *   2  store a partial instead: the computed dimensions' children only, transformed and ORed
 *      (the output slots then hold an intermediate, not a result);
 *   4  OR the computed dimensions' children into the partial in the output slots;
 *   3  finish: the partial in the slots plus the computed dimensions' children, then the
 *      result, stored.
 * Runs 2, 4 ..., 3 over disjoint subsets give the same output as one plan over every
 * dimension and evaluate the win-pattern groups once, in the finishing run.  Runs 0, 1, 2
 * and 4 may be repeated with the same output; a finishing run replaces the partial it reads
 * by the result, so it runs once per partial. */
pe_status pe_end_slabs_plan_create_dims(int noutputs, const pe_slab* outputs, int nchildren, const pe_slab* children,
                                        uint32_t dims, int accumulate, pe_end_plan** plan);

/* Synchronous host-buffer convenience: copies the input to the device, computes, copies the
 * output back.  Same layouts as above. */
pe_status pe_end_compute_line_host(const pe_line* line, const void* h_input, void* h_output);

/* Seeded synthetic child data on the device: for node i (counted from `first_node`) and
 * super word w, known = Bernoulli(0.9) bits (16-term binary expansion, p = 58982/65536) and
 * win = Bernoulli(1/2) & known, stored as node = (win, known).  Deterministic in
 * (seed, node index, word). */
pe_status pe_fill_seeded(void* d_nodes, int64_t nodes, uint64_t seed, int64_t first_node, void* stream);

/* ---------------- end engine: counts, samples, slices ---------------- */

/* Win counts of every block of a slab: count_block_wins (pentago/end/block_store.cc:330-349),
 * i.e. the sum over the block's nodes of popcounts_over_stabilizers
 * (pentago/base/count.cc:254-262).  d_counts (device) receives 4 u64 per block, C-order over
 * the box: wins, wins-or-ties (for the player to move), total distinct positions, and the
 * number of nodes whose popcounts are not multiples of their stabilizer size (the reference
 * asserts zero).  Asynchronous on `stream`. */
pe_status pe_end_count_blocks(const pe_slab* slab, uint64_t* d_counts, void* stream);
/* The same summed per slab, synchronous: counts[k] = totals over slab k's blocks. */
pe_status pe_end_count_slabs(int n, const pe_slab* slabs, uint64_t (*counts)[4]);
/* pe_end_count_blocks over n slabs in one kernel launch: slab k's per-block counts start at
 * d_counts + 4 * (blocks of slabs 0..k-1).  d_workspace is device memory of
 * pe_end_count_workspace_bytes(n) bytes for the launch descriptors; it must stay unused by
 * anything else until the kernel has run.  Returns once the descriptors are uploaded (which
 * first waits for the work already on `stream`); the kernel runs asynchronously. */
pe_status pe_end_count_blocks_many(int n, const pe_slab* slabs, uint64_t* d_counts, void* d_workspace, void* stream);
int64_t pe_end_count_workspace_bytes(int n);

/* Sparse samples: the node of each board (a rotation-standardized board of the slab's
 * section inside the slab's box; quadrant q in bits 16q..16q+15, radix 3) copied to
 * h_nodes (64 bytes each).  The block store's sample capture, pentago/end/block_store.cc:
 * 190-194, with the index map of pentago/end/check.cc:110-121.  Synchronous. */
pe_status pe_end_slab_samples(const pe_slab* slab, int n, const uint64_t* boards, void* h_nodes);

/* Supertensor (.pentago) block data <-> nodes, on the device.  A file block holds per node
 * (black wins, white wins), filter 1 interleaving the two supers bit by bit
 * (pentago/data/filter.h, applied by pentago/data/supertensor.cc:230-243 and 319-328);
 * nodes hold (we-win, we-win-or-tie) for the player to move, turn = stones & 1
 * (orientation as pentago/end/check.cc:149-158).  decode: decompressed file words -> nodes;
 * encode: nodes -> file words ready for compression.  `nodes` 64-byte entries, in place
 * allowed, asynchronous.  Compression and file I/O stay on the host. */
pe_status pe_supertensor_decode(int filter, int turn, const void* d_in, int64_t nodes, void* d_out, void* stream);
pe_status pe_supertensor_encode(int filter, int turn, const void* d_in, int64_t nodes, void* d_out, void* stream);

/* All standardized sections with `slice` stones that `root` depends on, sorted by sig():
 * slice `slice` of descendent_sections (pentago/end/sections.cc:18-49).  Writes at most cap
 * sections and the full count to *count (call with cap 0 to size). */
pe_status pe_descendent_sections(const pe_section* root, int slice, int cap, pe_section* out, int* count);

/* A slice store: whole sections of one slice resident on the current device, one slab per
 * section covering all its blocks -- the one-GPU form of the block store the reference's
 * slice loop fills (pentago/mpi/toplevel.cc:263-421, pentago/end/block_store.h:113-160). */
typedef struct pe_slice pe_slice;
/* Allocate device slabs for `n` sections with the same stone count (<= 35). */
pe_status pe_slice_create(int n, const pe_section* sections, pe_slice** out);
/* The same with the slabs carved, in order, from caller device memory `arena` of
 * `arena_bytes` (nothing allocated; the arena must outlive the store). */
pe_status pe_slice_create_in(int n, const pe_section* sections, void* arena, int64_t arena_bytes, pe_slice** out);
/* Device bytes the slabs of these sections take (no allocation). */
int64_t pe_slice_bytes(int n, const pe_section* sections);
/* Compute every section of `out` from `children` (the store of the next slice holding every
 * child section they read; NULL when `out` is slice 35): one fused plan over all sections,
 * run on `stream` and waited for. */
pe_status pe_slice_compute(pe_slice* out, const pe_slice* children, void* stream);
/* Section k's slab (k < sections), valid until pe_slice_destroy. */
pe_status pe_slice_slab(const pe_slice* s, int k, pe_slab* slab);
int pe_slice_sections(const pe_slice* s);
/* pe_end_count_slabs over every section of the store. */
pe_status pe_slice_counts(const pe_slice* s, uint64_t (*counts)[4]);
pe_status pe_slice_destroy(pe_slice* s);
/* The slice loop for a slice too large to keep whole (pentago/mpi/toplevel.cc:263-421 on one
 * device): sections[0..n) of one slice are computed in the groups of pe_slice_plan_groups,
 * each group's slabs
 * (and its regenerated children) within `budget` bytes.  Children come from `children`, the
 * whole store of the next slice, or -- when it is NULL -- are regenerated per group from
 * `grandchildren`, the whole store of the slice after it.  counts[k] = pe_slice_counts of
 * sections[k]; visit(ctx, group), if not NULL, sees every computed group before it is freed
 * (for samples).  *groups (may be NULL) receives the number of groups.  Every group lives in
 * one device arena of pe_slice_arena_bytes, allocated and freed by the call. */
pe_status pe_slice_grouped_counts(int n, const pe_section* sections, int64_t budget, const pe_slice* children,
                                  const pe_slice* grandchildren, uint64_t (*counts)[4],
                                  void (*visit)(void* ctx, const pe_slice* group), void* ctx, int* groups,
                                  void* stream);
/* The same inside caller device memory (nothing allocated for slabs, so the arena can be
 * reused from slice to slice); arena_bytes must be at least pe_slice_arena_bytes. */
pe_status pe_slice_grouped_counts_in(int n, const pe_section* sections, int64_t budget, const pe_slice* children,
                                     const pe_slice* grandchildren, uint64_t (*counts)[4],
                                     void (*visit)(void* ctx, const pe_slice* group), void* ctx, int* groups,
                                     void* arena, int64_t arena_bytes, void* stream);
/* The groups pe_slice_grouped_counts uses (host only): the sections are taken in the order
 * sections[order[0]], sections[order[1]], ... and group g is positions [ends[g-1], ends[g])
 * of that order, ends[-1] = 0; order and ends need room for n entries.  Without
 * regeneration the order is the given one and each group as long as it fits; with it,
 * sections that share child sections are grouped together (each child is regenerated once
 * per group that needs it). */
pe_status pe_slice_plan_groups(int n, const pe_section* sections, int64_t budget, int regenerate, int* order,
                               int* ends, int* ngroups);
/* Device bytes of the largest of those groups (outputs plus regenerated children), -1 on
 * invalid arguments (host only). */
int64_t pe_slice_arena_bytes(int n, const pe_section* sections, int64_t budget, int regenerate);

/* ---------------- midsolve ---------------- */

/* high_board_s: pentago/high/board_c.h:9-12 */
typedef struct { uint64_t side[2]; uint32_t ply; } pe_high_board;
/* halfsupers_t: (win, notlose) halves.  pentago/mid/halfsuper_c.h:16-18 */
typedef struct { uint64_t win[2]; uint64_t notlose[2]; } pe_halfsupers;
/* (raw board, value) pair of mid_values_t.  pentago/mid/midengine.h:36 */
typedef struct { uint64_t raw; int32_t value; int32_t pad; } pe_mid_value;

/* Whether the game is over and its value for the player to move (1 win, 0 tie, -1 loss).
 * high_board_t::done_and_value, pentago/high/board.h:56-60. */
pe_status pe_high_done(const pe_high_board* board, int* done, int* value);

/* Workspace entries needed for boards with at least min_slice stones.
 * pentago/mid/midengine.cc:47-50 (entries are 32-byte (win, notlose) pairs here). */
int64_t pe_midsolve_workspace_size(int min_slice);

/* Solve one board (>= 18 stones) on the device: results[0..spots] as in
 * midsolve_internal, pentago/mid/midengine.h:40.  Synchronous. */
pe_status pe_midsolve_internal(const pe_high_board* board, pe_halfsupers results[19]);

/* pe_midsolve_internal for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Allocates nothing: the board's tables, workspace, results and launch
 * descriptors live in a fixed device array of pe_midsolve_noalloc_scratch_bytes() (64 MiB)
 * that every GPU the library runs on holds until the process ends, so boards with at most
 * 15 empty cells (at least 21 stones) fit; larger boards return PE_ERR_CAPACITY.  pe_init
 * must have run on the device (PE_ERR_INVALID otherwise).  Runs on `stream` and returns once
 * results[0..18] are written (entries past spots zero), equal to pe_midsolve_internal's.
 * Every call on a GPU shares that GPU's array: this library serializes the calls of one
 * process, and calls from several processes that overlap on one GPU race. */
pe_status pe_midsolve_internal_noalloc(const pe_high_board* board, pe_halfsupers results[19], void* stream);
int64_t pe_midsolve_noalloc_scratch_bytes(void);

/* Solve and traverse: values of the board, its children and (if not middle) grandchildren.
 * Returns the count through *count (<= 163).  pentago/mid/midengine.cc:171-179 (wasm
 * midsolve) and ios/shared/c_board.h:30-34 (board_midsolve). */
pe_status pe_midsolve(const pe_high_board* board, pe_mid_value* values, int* count);

/* Batched: n boards of uneven stone counts in one device pass.  results[i][0..18].
 * mode 0 = level-synchronous launches, 1 = persistent work-queue kernel. */
pe_status pe_midsolve_batch(int n, const pe_high_board* boards, pe_halfsupers (*results)[19], int mode,
                            void* stream);

/* Batched plans.  A batch of boards of any mix of sizes runs as: small boards (at most
 * small_spots empty cells) in one persistent kernel whose workgroups pop boards from a device
 * queue and solve each whole board (every level, workgroup barriers between levels) in a
 * workspace slot of their own; large boards level by level, one full-device launch per
 * level, on `streams` streams of their own that overlap with the small-board kernel.
 * Memory is held within workspace_budget (fewer workgroups / streams when it is tight).
 * No counterpart in the reference, which solves one board per call
 * (pentago/mid/midengine.cc:57-125); each board's results equal pe_midsolve_internal's.
 * Zero fields take the device defaults (set by pe_mid_autotune, else built in). */
typedef struct {
  int64_t workspace_budget;   /* bytes held by one run (0: 60% of free HBM, <= 200 GB) */
  int32_t small_spots;        /* boards with at most this many empty cells: one workgroup each (0: 14) */
  int32_t workgroups;         /* persistent workgroups for small boards (0: the tuned count when
                                 threads is the tuned size, else a full device of `threads`) */
  int32_t threads;            /* threads per small-board workgroup: 256, 512, 1024 (0: 512) */
  int32_t streams;            /* streams for large boards (0: 2; negative: all on the run stream) */
  int32_t use_graph;          /* > 0: pe_mid_plan_run replays a captured hipGraph; < 0: eager
                                 launches; 0: the device default (pe_mid_autotune's choice) */
  int32_t pad;
} pe_mid_config;
typedef struct {
  int64_t small_boards, large_boards, launches, workspace_bytes;
  int32_t workgroups, streams;
} pe_mid_plan_stats;
typedef struct pe_mid_plan pe_mid_plan;
pe_status pe_mid_plan_create(int n, const pe_high_board* boards, const pe_mid_config* config, pe_mid_plan** plan);
/* Asynchronous on `stream`; a plan may be run any number of times (not concurrently). */
pe_status pe_mid_plan_run(pe_mid_plan* plan, void* stream);
/* Waits for the last run on `stream` and copies results[i][0..18] (entries past spots zero). */
pe_status pe_mid_plan_results(pe_mid_plan* plan, pe_halfsupers (*results)[19], void* stream);
/* How the batch was split and the configuration in effect. */
pe_status pe_mid_plan_info(const pe_mid_plan* plan, pe_mid_plan_stats* stats, pe_mid_config* config);
pe_status pe_mid_plan_destroy(pe_mid_plan* plan);
/* Host-only: the launches per run and the device memory a plan with this (complete)
 * configuration would use, or PE_ERR_CAPACITY.  No device calls. */
pe_status pe_mid_schedule_info(int n, const pe_high_board* boards, const pe_mid_config* config, int64_t* launches,
                               int64_t* workspace_bytes);
/* Time candidate configurations on these boards (a warm-up, then the faster of two runs of the
 * whole batch each): the split small_spots in {12, 14, 16} at 512 threads, then 256 and 1024
 * threads at the best split, then the winner replayed as a captured hipGraph against eager
 * launches; the fastest becomes the current device's default.  Hand in the batch itself (or one like it): the best split
 * depends on how many boards of each size there are.  The chosen configuration is returned
 * through *best (use_graph 1 or -1). */
pe_status pe_mid_autotune(int n, const pe_high_board* boards, int64_t workspace_budget, pe_mid_config* best);

/* Serving loop: single-position solve requests admitted as they arrive.  A worker thread
 * batches the pending requests whose device memory fits the part of the HBM budget not held
 * by batches in flight and runs each batch on one of several streams as a batched plan, so
 * requests that arrive while others are on the device start as soon as memory and a stream
 * free up; batches overlap on the device.  Results are those of pe_midsolve_internal.  The
 * reference serves one wasm midsolve per request (web/server/index.js:16-58); this has no
 * direct counterpart.  All calls are thread-safe; each ticket is taken by exactly one
 * successful poll (done = 1) or wait. */
typedef struct pe_mid_server pe_mid_server;
typedef struct {
  int64_t workspace_budget;   /* bytes all batches in flight may hold (0: 40% of free HBM, <= 100 GB) */
  int32_t streams;            /* batches in flight at once (0: 4) */
  int32_t max_batch;          /* requests per batch (0: 512) */
  int32_t small_spots;        /* batched-plan small-board limit (0: 14) */
  int32_t workgroups;         /* small-board workgroups per batch (0: a device's worth / streams) */
} pe_mid_server_config;
typedef struct {
  int64_t submitted, completed, batches, peak_workspace, pending, in_flight;
} pe_mid_server_info;
pe_status pe_mid_server_create(const pe_mid_server_config* config, pe_mid_server** server);
/* Returns PE_ERR_INVALID for a board that is not solvable here, PE_ERR_CAPACITY when its
 * workspace exceeds the whole pool. */
pe_status pe_mid_server_submit(pe_mid_server* server, const pe_high_board* board, int64_t* ticket);
/* Nonblocking.  *done = 1 once finished: results[0..18] and the latency (submit to
 * completion, ms) are written and the ticket is released. */
pe_status pe_mid_server_poll(pe_mid_server* server, int64_t ticket, pe_halfsupers results[19], int* done,
                             double* latency_ms);
/* Blocks until the request has finished, then as poll. */
pe_status pe_mid_server_wait(pe_mid_server* server, int64_t ticket, pe_halfsupers results[19], double* latency_ms);
pe_status pe_mid_server_stats(pe_mid_server* server, pe_mid_server_info* stats);
/* Finishes every submitted request, then releases the server. */
pe_status pe_mid_server_destroy(pe_mid_server* server);

/* Traverse precomputed results into (raw, value) pairs.  pentago/mid/midengine.cc:127-158 */
pe_status pe_midsolve_traverse(const pe_high_board* board, const pe_halfsupers* results,
                               pe_mid_value* values, int* count);

#ifdef __cplusplus
}
#endif
#endif
