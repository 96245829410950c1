// Batched midsolve plans (host side); the kernels live in mid_queue.hip.
//
// A batch is split by size.  Small boards (at most small_spots empty cells) are solved whole
// by one workgroup each: a persistent kernel whose workgroups pop board indices from a device
// queue (one atomic counter) and sweep every level of the board with workgroup barriers
// between levels, in a workspace slot owned by the workgroup.  No workgroup ever waits for
// another.  Large boards run level by level on one or more streams of their own (table pass,
// then one launch per level over the whole device), boards of one stream reusing that
// stream's slot in stream order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device.h"
#include "mid.h"

namespace pe {

struct MidBoardDev;

struct MidConfig {
  int64_t budget;        // bytes one run may hold (tables, workspaces, descriptors)
  int32_t small_spots;   // boards with at most this many empty cells: one workgroup each
  int32_t workgroups;    // persistent workgroups of the small-board kernel (upper bound)
  int32_t threads;       // threads per small-board workgroup: 256, 512 or 1024
  int32_t streams;       // streams carrying large boards (0: all work on the run stream)
  bool graph;            // replay the whole launch sequence through a captured hipGraph
};

// Memory of one run, as offsets into a single device allocation.
struct MidLayout {
  std::vector<int> small, big;         // board indices (big: largest first)
  std::vector<int> big_stream;         // stream of each big board (parallel to big)
  int32_t workgroups = 0;              // small-board kernel grid, within the budget
  int32_t streams = 0;                 // large-board streams (0: the run stream)
  std::vector<size_t> table_off;       // per small board
  size_t small_work = 0, small_stride = 0;   // workspace slot of each workgroup
  size_t big_slots = 0, big_stride = 0;      // per stream: tables, then workspace
  size_t bytes = 0;                    // total
  // descriptor blob at offset 0: small boards | big boards | prep prefixes | results | head
  size_t o_big = 0, o_prefix = 0, o_res = 0, o_head = 0, blob_bytes = 0;
};
// Returns null on success, else an error message.
const char* mid_layout(const std::vector<MidInfo>& infos, const MidConfig& cfg, MidLayout* L);

struct MidPlan {
  int device = -1;
  int nb = 0;
  MidConfig cfg{};
  MidLayout layout;
  std::vector<MidInfo> infos;
  hipStream_t alloc_stream = nullptr;  // stream-ordered allocation (serving) or null
  uint8_t* mem = nullptr;              // layout.bytes
  bool owns_mem = true;                // false: caller memory, not freed by the plan
  const MidBoardDev* small_dev = nullptr;
  const MidBoardDev* big_dev = nullptr;
  const int64_t* prefix = nullptr;     // small boards' prep prefix, then [0, total] per big board
  pe_halfsupers* results = nullptr;    // 19 per board, in board order
  unsigned int* head = nullptr;
  int64_t small_prep_total = 0;
  std::vector<int64_t> big_prep_total;
  DevTables tables{};
  std::vector<hipStream_t> streams;    // owned, layout.streams of them
  std::vector<hipEvent_t> events;      // fork + one join per stream
  hipGraphExec_t graph_exec = nullptr;
  int64_t launches = 0;                // kernel launches per run
};

// alloc_stream null: synchronous hipMalloc/hipMemcpy; else stream-ordered on it.  With
// pinned host memory `upload` of at least the layout's o_res bytes (which must stay untouched
// until alloc_stream has passed the plan's first run), the descriptor upload is queued
// without waiting; otherwise the call waits for the upload.
// With `memory` of at least the layout's bytes the plan lives there (nothing is allocated,
// the caller keeps the memory until the plan is destroyed and its runs have finished).
pe_status mid_plan_create(const std::vector<MidInfo>& infos, const MidConfig& cfg, MidPlan* plan,
                          hipStream_t alloc_stream = nullptr, void* upload = nullptr, size_t upload_bytes = 0,
                          void* memory = nullptr, size_t memory_bytes = 0);
pe_status mid_plan_run(MidPlan& plan, hipStream_t stream);
// Queue the copy of all results to host memory on `stream`; after the stream has been
// synchronized, mid_plan_zero_tail clears the entries past each board's spots.
pe_status mid_plan_copy_results(const MidPlan& plan, pe_halfsupers (*out)[19], hipStream_t stream);
void mid_plan_zero_tail(const MidPlan& plan, pe_halfsupers (*out)[19]);
// Frees the plan's memory (stream-ordered on alloc_stream when it was allocated that way).
void mid_plan_destroy(MidPlan* plan);
int mid_board_blocks_per_cu(int threads);

// Serving loop (mid_server.cc)
struct MidServer;
pe_status mid_server_create(const pe_mid_server_config* c, MidServer** out);
pe_status mid_server_submit(MidServer* S, const pe_high_board& board, int64_t* ticket);
pe_status mid_server_poll(MidServer* S, int64_t ticket, pe_halfsupers* results, int* done, double* latency_ms);
pe_status mid_server_wait(MidServer* S, int64_t ticket, pe_halfsupers* results, double* latency_ms);
void mid_server_stats(MidServer* S, pe_mid_server_info* out);
void mid_server_destroy(MidServer* S);

// Byte sizes of one board's tables, workspace, and both plus its 19 results.
size_t mid_table_bytes(const MidInfo& I);
size_t mid_work_bytes(const MidInfo& I);
size_t mid_board_bytes(const MidInfo& I);
// Lay out one board's tables (mid_table_bytes) at `base`; or tables, workspace and results
// (mid_board_bytes) at `base`.
void mid_carve_tables(MidBoardDev* B, uint8_t* base);
void mid_carve(MidBoardDev* B, uint8_t* base);
// One launch building the tables of every board: prefix[b] = prep entries before board b.
pe_status mid_launch_prep(const MidBoardDev* boards, const int64_t* prefix, int nb, int64_t total,
                          const DevTables& T, hipStream_t stream);
// Launch shape of one level of one board: `band` > 0 spreads the level over the XCDs by bands
// of s1p columns (blocks counts the padded grid), 0 keeps the linear order.  Measured on
// 2048 boards of 18..30 stones, batched: linear 1.14 s; bands of 4 / 8 / 12 / 16 / 24 / 32
// columns 1.22 / 0.99 / 0.96 / 0.95 / 0.98 / 0.99 s, one band per XCD 1.01 s; bands 16 wide
// dealt to the XCDs in contiguous runs instead of in turn 1.05 s.
constexpr uint32_t kLevelBand = 16;
struct LevelGrid {
  unsigned blocks;
  uint32_t band, nsets1p, nsets0;
};
LevelGrid level_grid(const MidInfo& I, int n);
// One launch sweeping level n of one board (all positions).
pe_status mid_launch_level(const MidBoardDev* board, const MidInfo& info, int n, int64_t positions,
                           const DevTables& T, hipStream_t stream);

}  // namespace pe
