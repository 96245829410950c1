// Work-queue midsolve plans (host side); the kernels live in mid_queue.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device.h"
#include "mid.h"

namespace pe {

struct MidBoardDev;

// Step codes of a task: the slot gate, the table-building pass, or a level step t >= 0
// (level n = spots - t).
constexpr int32_t kStepGate = -2;
constexpr int32_t kStepPrep = -1;

struct MidTask {
  int32_t board, step;
  int64_t begin;             // first entry (table entry or position index within the level)
  int32_t count;             // entries in this task
  int32_t dep_off, dep_cnt;  // dependency list: (counter index, required count) pairs
  int32_t done_slot;         // counter incremented when the task has finished
};
static_assert(sizeof(MidTask) == 32, "task records are 32 bytes");

struct MidConfig {
  int64_t budget;            // bytes of workspace the batch may hold at once
  int32_t tile;              // minimum entries per task
  int32_t workgroups;        // persistent workgroups (256 threads each)
  int64_t wait_limit;        // wall-clock ticks a dependency wait may last before failing
  bool graph;                // replay the launch pair through a captured hipGraph
};

struct MidPlan {
  int device = -1;
  int nb = 0;
  int64_t ntasks = 0;
  int32_t nslots = 0;
  MidConfig cfg{};
  std::vector<int> spots;
  void* blob = nullptr;      // boards | tasks | deps | results | counters | head | error
  void* arena = nullptr;     // workspace slots
  int64_t arena_bytes = 0;
  const MidBoardDev* boards = nullptr;
  const MidTask* tasks = nullptr;
  const int2* deps = nullptr;
  pe_halfsupers* results = nullptr;
  int32_t* counters = nullptr;
  unsigned long long* head = nullptr;
  int32_t* error = nullptr;
  DevTables tables{};
  hipGraphExec_t graph_exec = nullptr;
};

// Order the tasks of a batch (simulated list schedule) and assign workspace slots within
// cfg.budget, or use the given per-board arena offsets (every board resident at once).
// Returns null on success, else an error message.
const char* mid_plan_build_schedule(const std::vector<MidInfo>& infos, const MidConfig& cfg,
                                    const std::vector<int64_t>* fixed_offsets, std::vector<MidTask>* order,
                                    std::vector<int2>* deps, std::vector<int64_t>* offsets, int64_t* arena_bytes,
                                    int32_t* nslots);

// A schedule plus the layout of the plan's device blob
// (boards | tasks | deps | results | counters | head, error).
struct MidPrepared {
  std::vector<MidTask> order;
  std::vector<int2> deps;
  std::vector<int64_t> offsets;
  int64_t arena_bytes = 0;
  int32_t nslots = 0;
  size_t o_tasks = 0, o_deps = 0, o_res = 0, o_state = 0, o_head = 0, blob_bytes = 0;
};
pe_status mid_plan_prepare(const std::vector<MidInfo>& infos, const MidConfig& cfg,
                           const std::vector<int64_t>* fixed_offsets, MidPrepared* p);
// Host image of the blob (the first o_state bytes) for a blob at dev_blob and workspace at
// arena; sets the plan's device pointers (blob and arena ownership stay with the caller).
void mid_plan_fill(const std::vector<MidInfo>& infos, const MidConfig& cfg, const MidPrepared& p, void* dev_blob,
                   uint8_t* arena, const DevTables& T, std::vector<uint8_t>* host_blob, MidPlan* plan);
pe_status mid_plan_create(const std::vector<MidInfo>& infos, const MidConfig& cfg, MidPlan* plan);
pe_status mid_plan_launch(const MidPlan& plan, hipStream_t stream);
pe_status mid_plan_run(MidPlan& plan, hipStream_t stream);
pe_status mid_plan_results(const MidPlan& plan, pe_halfsupers (*out)[19], hipStream_t stream);
void mid_plan_destroy(MidPlan* plan);

// Serving loop (mid_server.cc)
struct MidServer;
pe_status mid_server_create(const pe_mid_server_config* c, MidServer** out);
pe_status mid_server_submit(MidServer* S, const pe_high_board& board, int64_t* ticket);
pe_status mid_server_poll(MidServer* S, int64_t ticket, pe_halfsupers* results, int* done, double* latency_ms);
pe_status mid_server_wait(MidServer* S, int64_t ticket, pe_halfsupers* results, double* latency_ms);
void mid_server_stats(MidServer* S, pe_mid_server_info* out);
void mid_server_destroy(MidServer* S);

// Lay out one board's tables and workspace at `base` (mid_board_bytes(I) bytes).
void mid_carve(MidBoardDev* B, uint8_t* base);

}  // namespace pe
