// End-engine block lines: host-side index state and device descriptors.
//
// A block line (pentago/end/line.h:15-35) is one section, one dimension d and a 3-D block
// base; it computes, for every node of the line, the moves in quadrant d.  Its input is the
// child block line in the standardized child section, stored the way the reference stores
// line.input: input_blocks blocks of block_stride+1 nodes, C-order within each block
// (pentago/end/compute.cc:84-94,134-139).  Its output is `length` blocks of block_stride
// nodes (compute.cc:147-152).  A node is two supers (we-win, we-win-or-tie), 64 bytes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pentago_amd.h"

namespace pe {

struct Section {
  uint8_t c[4][2];
  int sum() const { int t = 0; for (int q = 0; q < 4; q++) t += c[q][0] + c[q][1]; return t; }
  uint64_t sig() const {
    uint64_t r = 0;
    for (int q = 0; q < 4; q++) r |= uint64_t(c[q][0]) << (16 * q) | uint64_t(c[q][1]) << (16 * q + 8);
    return r;
  }
  bool valid_counts() const;
  bool shape(int out[4]) const;
  Section transform(int g) const;
  int standardize(Section* out) const;   // returns g
  Section child(int q) const { Section s = *this; s.c[q][sum() & 1]++; return s; }
};

// Uniform per-line state consumed by the device kernel.
struct LineDev {
  const uint64_t* input;       // child block line (null at slice 35)
  uint64_t* output;            // output block line
  int32_t dim, child_dim, turn, slice35;
  int32_t length, child_length, block_stride, in_dim_size;
  int32_t out_shape[4];
  int32_t perm[4];
  int32_t first_block[4];
  int32_t first_child_block[4];
  int32_t reflection_moves[4];
  int32_t rmin_off[4], rmin_size[4];   // into the device rmin tables
  int32_t rs_off[5];                   // reflection symmetry offsets, relative to refl_base
  int32_t refl_base;                   // offset of this line's symmetries in the launch's byte array
  int32_t cross_shape[3];              // output shape without the line dimension
  int32_t cross_count;
  int32_t inv_global, inv_local;       // inverse_transform (compute.cc:69)
  int32_t child_coord;                 // quadrant that the move's local rotation lands on after conjugation
  int32_t i_begin, i_end;              // range of indices along the line dimension to compute
  int32_t accumulate;                  // OR into the output instead of storing
  int32_t reflected;                   // child standardized through a reflection
  int32_t fast_t;                      // unreflected: inv_global (0-3, y = -g (1,1,1,1)); -1 otherwise
  // Per output axis a (zero for the line dimension): the child axis c it maps to is
  // perm^-1(a); stride of c inside a child block (full / last block), 8*first_child_block[c],
  // reflection_moves[c] and the offset of c's reflection symmetries.
  int32_t ax_stride[4], ax_stride_last[4], ax_fcn[4], ax_moves[4], ax_rs[4];
  int32_t input_stride;                // stride of the child dimension inside a child block
  const int64_t* in_off;               // node offset of input block k (null: k*(block_stride+1))
  int64_t in_base, in_kstride;         // fused plans: in_off[k] = in_base + k * in_kstride
  const int64_t* out_off;              // node offset of output block k (null: k*block_stride)
};

struct LineHost {
  LineDev dev;
  pe_line_info info;
  pe_section section;                  // the line's (standard) section
  std::vector<uint8_t> refl_syms;      // all_reflection_symmetries (compute.cc:105-122)
  std::vector<int64_t> in_off, out_off; // optional block offset tables
};

// A device-resident set of line jobs, built once and launched any number of times.
struct EndPlan {
  void* blob = nullptr;        // descriptors, task prefix, symmetries, offset tables
  size_t bytes = 0;
  int nlines = 0;
  int64_t tasks = 0;
  int device = -1;
  int64_t output_nodes = 0;    // nodes written per run (sum of the i ranges' cross sections)
  size_t off_prefix = 0, off_refl = 0;
};

// Build the index state of a line (pentago/end/compute.cc:24-123).  Returns a status and
// fills `err` on failure.
pe_status make_line(const pe_line& line, LineHost* out, std::string* err);

// ---- fused block plans ----
// Per-section state of the fused kernel.
struct SecDev {
  int32_t turn;
  int32_t rmin_off[4], rmin_size[4];
  int32_t field_off[4];     // start (in supers) of quadrant q's entries in the field pool: 15 supers each
  int32_t move_off[4];      // start (in u16 entries) of dimension d's move table rows
  int32_t nmoves[4];        // moves per node along dimension d (0: closed)
};
// One output block: its section, coordinates, the line through it per dimension, output offset.
struct BlockDev {
  int32_t sec;
  int32_t b[4];
  int32_t line[4];          // -1 if the dimension is closed (or not computed)
  int32_t accumulate;       // OR into the block's existing output instead of storing
  uint64_t* out;            // the block's output: C-order over its shape
};

struct FusedPlan {
  void* blob = nullptr;
  int nblocks = 0;
  bool s35 = false;   // blocks of 35-stone sections (the kernel's no-children form)
  size_t off_lines = 0, off_secs = 0, off_moves = 0, off_refl = 0, off_heads = 0;
  int queue_wgs_per_cu = 0;   // 0: one workgroup per (block, part); else persistent queue form
  int64_t output_nodes = 0;
};

// Fused plan over block lines (one dimension each, the reference's line buffers): every
// job must cover its whole line (all indices along the dimension), have child blocks at
// evenly spaced offsets and at most 34 stones; jobs that do not qualify stay on the line
// kernel (fused_line_eligible).
bool fused_line_eligible(const LineHost& h);
pe_status fused_line_plan_create(const std::vector<LineHost>& jobs, FusedPlan* plan, std::string* err);

// One whole block line on the fused kernel, asynchronous on `stream`, its descriptor passed
// by value (no device memory per call): input block k at node k * (block_stride + 1),
// output block k at node k * block_stride (pentago/end/compute.cc:134-152).  Thread-safe.
pe_status fused_line_launch(const LineHost& h, const void* input, void* output, int accumulate, hipStream_t stream,
                            std::string* err);

// Field pool (one per device, built once): for quadrant q, player to move `turn` and
// moves flag m, one entry of 60 words per rotation-minimal quadrant (all buckets, in
// rmin_flat order): for each super word k, tables [F0 (mover's side), F1 (other side), Fm
// (OR over the mover's moves in q; zero when m = 0)] x 5 fields (word k of each
// superwin_info field, pentago/base/superscore.cc:38-96).  A section's field_off[q] is
// field_pool_off(q, turn, m, counts[q]) in supers (4 words).
int64_t field_pool_words();
int32_t field_pool_off(int q, int turn, int moves, int black, int white);
void build_field_pool(std::vector<uint64_t>* out);

// Move tables of one section (fused kernel): for each open dimension d (move_off[d] >= 0,
// nmoves[d] = empty cells of quadrant d) and rotation-minimal index i of quadrant d, a row
// of 16 u16 entries, one per move in ascending cell order: child block along the child
// dimension (bits 0-5), node index inside that block along it after the reflection swap
// (bits 6-8) and the move's local rotation restricted to the child dimension (bits 9-10).
// Everything here depends only on (section, d, i, move), not on the line's other blocks.
// only >= 0: rows of that dimension only (the others keep move_off -1).
pe_status section_move_tables(const Section& s, std::vector<uint16_t>* out, int32_t move_off[4], int32_t nmoves[4],
                              std::string* err, int only = -1);

}  // namespace pe
