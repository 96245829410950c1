// Work-queue midsolve: a whole batch of boards of uneven size in one persistent launch.
//
// The work of a board is a chain of steps: a gate (waits until the board's workspace slot
// has been released by the slot's previous board), the table-building pass
// (pentago/mid/internal.h:122-259) and one sweep per level n = spots..0
// (internal.h:358-408).  Every step is cut into tasks of at least `tile` entries (and at
// most 2048 tasks per step); a task may start once every task of the board's previous step
// has finished.
//
// The host orders all tasks of the batch into one array by simulating a list schedule on
// `workgroups` identical workers (earliest-admitted board first), assigning each board a
// workspace slot inside an HBM budget as it is admitted.  Workgroups of the persistent
// kernel pop tasks in array order from one atomic head counter.  A task's dependencies are
// always earlier in the array, so the earliest unfinished task is always runnable and the
// queue drains; each wait is additionally bounded by a wall-clock limit that reports an
// error instead of spinning forever.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <queue>
#include <string>
#include <vector>

#include "mid_device.h"
#include "mid_queue.h"

namespace pe {

size_t mid_board_bytes(const MidInfo& I);

namespace {

__global__ void mid_queue_reset_kernel(int32_t* __restrict__ counters, int64_t n,
                                       unsigned long long* __restrict__ head, int32_t* __restrict__ error) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) counters[i] = 0;
  if (i == 0) { *head = 0; *error = 0; }
}

__global__ __launch_bounds__(256) void mid_queue_kernel(const MidBoardDev* __restrict__ boards,
                                                        const MidTask* __restrict__ tasks, int64_t ntasks,
                                                        const int2* __restrict__ deps, int32_t* counters,
                                                        unsigned long long* head, int32_t* error,
                                                        int64_t wait_limit, DevTables T) {
  __shared__ int64_t s_id;
  __shared__ int s_ok;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_id = int64_t(atomicAdd(head, 1ull));
    __syncthreads();
    const int64_t id = s_id;
    if (id >= ntasks) return;
    const MidTask tk = tasks[id];
    if (threadIdx.x == 0) {
      int ok = 1;
      for (int d = 0; d < tk.dep_cnt && ok; d++) {
        const int2 dp = deps[tk.dep_off + d];
        const int64_t t0 = wall_clock64();
        while (__hip_atomic_load(counters + dp.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < dp.y) {
          if (wall_clock64() - t0 > wait_limit) {
            __hip_atomic_fetch_or(error, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(8);
        }
      }
      __threadfence();   // acquire: the finished tasks' writes are visible to this workgroup
      // after any failed wait every later task skips its work (its inputs may be incomplete),
      // but is still counted so that the queue drains
      if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
      s_ok = ok;
    }
    __syncthreads();
    if (s_ok) {
      const MidBoardDev& B = boards[tk.board];
      if (tk.step == kStepPrep) {
        for (int i = threadIdx.x; i < tk.count; i += blockDim.x) mid_prep_entry(B, tk.begin + i, T);
      } else if (tk.step >= 0) {
        const int n = B.I.spots - tk.step;
        for (int i = threadIdx.x; i < tk.count; i += blockDim.x) mid_position(B, n, tk.begin + i, T);
      }
    }
    __threadfence();     // release this task's writes before it is counted
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(counters + tk.done_slot, 1);
  }
}

// ---- host-side schedule ----

struct Step {
  int64_t entries = 0;     // table entries (prep) or positions (level)
  int64_t chunk = 1;       // entries per task
  int32_t tasks = 0;
  double weight = 1;       // relative cost per entry
};

struct BoardSched {
  std::vector<Step> steps;   // index 0 gate, 1 prep, 2 + t level spots - t
  int32_t slot_base = 0;     // counter index of step 0
  int64_t bytes = 0;
  int64_t offset = -1;       // arena offset
  int32_t prev_board = -1;   // previous occupant of the slot
  int admit_seq = -1;
};

struct Slot { int64_t offset, bytes; int32_t last_board; };

struct Ready {
  int admit_seq, step, chunk, board;
  bool operator<(const Ready& o) const {   // max-heap: the smallest key on top
    if (admit_seq != o.admit_seq) return admit_seq > o.admit_seq;
    if (step != o.step) return step > o.step;
    return chunk > o.chunk;
  }
};

struct Event {
  double time;
  int kind;     // 0: step of `board` became ready, 1: board finished
  int board, step;
  bool operator<(const Event& o) const { return time > o.time; }
};

}  // namespace

const char* mid_plan_build_schedule(const std::vector<MidInfo>& infos, const MidConfig& cfg,
                                    const std::vector<int64_t>* fixed_offsets, std::vector<MidTask>* order,
                                    std::vector<int2>* deps, std::vector<int64_t>* offsets, int64_t* arena_bytes,
                                    int32_t* nslots) {
  const int nb = int(infos.size());
  std::vector<BoardSched> B(nb);
  int32_t slot_count = 0;
  for (int b = 0; b < nb; b++) {
    const MidInfo& I = infos[b];
    const MidSizes s = mid_sizes(I);
    B[b].bytes = int64_t(mid_board_bytes(I));
    if (!fixed_offsets && B[b].bytes > cfg.budget) return "a board needs more workspace than the budget";
    B[b].steps.resize(I.spots + 3);
    B[b].steps[0].entries = 0;
    B[b].steps[1].entries = s.set0 + s.wins1 + s.cs1ps;
    B[b].steps[1].weight = 8;
    for (int t = 0; t <= I.spots; t++) {
      const int n = I.spots - t;
      B[b].steps[2 + t].entries = I.positions[n];
      B[b].steps[2 + t].weight = double(I.spots - n + 1);
    }
    for (auto& st : B[b].steps) {
      // at most 2048 tasks per step; at least `tile` entries per task
      st.chunk = std::max<int64_t>(cfg.tile, (st.entries + 2047) / 2048);
      st.tasks = st.entries ? int32_t((st.entries + st.chunk - 1) / st.chunk) : 1;
    }
    B[b].slot_base = slot_count;
    slot_count += int32_t(B[b].steps.size());
  }
  // admission: largest workspace first, smaller boards backfill; a board takes the
  // smallest free slot that holds it, else carves a new slot while the budget allows
  std::map<int64_t, std::deque<int>, std::greater<int64_t>> pending;   // bytes -> boards
  for (int b = 0; b < nb; b++) pending[B[b].bytes].push_back(b);
  std::vector<Slot> slots;
  std::multimap<int64_t, int> free_slots;   // bytes -> slot
  std::vector<int> board_slot(nb, -1);
  int64_t carved = 0;
  int admit_seq = 0;

  std::priority_queue<Ready> ready;
  std::priority_queue<Event> events;
  std::vector<std::vector<double>> finish(nb);   // per step: latest finish of emitted tasks
  std::vector<std::vector<int32_t>> emitted(nb);
  for (int b = 0; b < nb; b++) { finish[b].assign(B[b].steps.size(), 0.0); emitted[b].assign(B[b].steps.size(), 0); }

  auto make_ready = [&](int b, int step) {
    for (int c = 0; c < B[b].steps[step].tasks; c++) ready.push(Ready{B[b].admit_seq, step, c, b});
  };
  auto try_admit = [&]() {
    if (fixed_offsets) {   // every board has its own region already: all admitted at once
      for (auto& kv : pending)
        for (int b : kv.second) {
          slots.push_back(Slot{(*fixed_offsets)[b], B[b].bytes, b});
          carved = std::max(carved, (*fixed_offsets)[b] + B[b].bytes);
          B[b].offset = (*fixed_offsets)[b];
          board_slot[b] = int(slots.size()) - 1;
          B[b].admit_seq = admit_seq++;
          make_ready(b, 0);
        }
      pending.clear();
      return;
    }
    for (auto it = pending.begin(); it != pending.end();) {
      const int64_t bytes = it->first;
      auto& q = it->second;
      while (!q.empty()) {
        int slot = -1;
        auto f = free_slots.lower_bound(bytes);
        if (f != free_slots.end()) {
          slot = f->second;
          free_slots.erase(f);
        } else if (carved + bytes <= cfg.budget) {
          slots.push_back(Slot{carved, bytes, -1});
          carved += bytes;
          slot = int(slots.size()) - 1;
        } else {
          break;
        }
        const int b = q.front();
        q.pop_front();
        B[b].offset = slots[slot].offset;
        B[b].prev_board = slots[slot].last_board;
        slots[slot].last_board = b;
        board_slot[b] = slot;
        B[b].admit_seq = admit_seq++;
        make_ready(b, 0);
      }
      it = q.empty() ? pending.erase(it) : std::next(it);
    }
  };

  const int P = std::max(1, cfg.workgroups);
  std::priority_queue<double, std::vector<double>, std::greater<double>> workers;
  for (int p = 0; p < P; p++) workers.push(0.0);
  try_admit();
  int64_t total_tasks = 0;
  for (int b = 0; b < nb; b++) for (auto& st : B[b].steps) total_tasks += st.tasks;
  order->clear();
  order->reserve(size_t(total_tasks));
  deps->clear();
  // dependency entries: one per (board, step >= 1) on the previous step, one per gate on the
  // previous slot occupant's last step
  std::vector<int32_t> dep_index(size_t(slot_count), -1);
  auto step_dep = [&](int b, int step) -> int32_t {   // index into deps of "step finished"
    int32_t& d = dep_index[size_t(B[b].slot_base + step)];
    if (d < 0) {
      d = int32_t(deps->size());
      deps->push_back(make_int2(B[b].slot_base + step, B[b].steps[step].tasks));
    }
    return d;
  };

  const double rate = 1.0;
  while (int64_t(order->size()) < total_tasks) {
    double now = workers.top();
    workers.pop();
    // apply every event up to `now`
    while (!events.empty() && events.top().time <= now) {
      const Event e = events.top();
      events.pop();
      if (e.kind == 0) {
        make_ready(e.board, e.step);
      } else {
        const int s = board_slot[e.board];
        free_slots.emplace(slots[s].bytes, s);
        try_admit();
      }
    }
    if (ready.empty()) {
      if (events.empty()) return "internal: schedule stalled";
      workers.push(events.top().time);   // idle until the next event
      continue;
    }
    const Ready r = ready.top();
    ready.pop();
    BoardSched& bs = B[r.board];
    const Step& st = bs.steps[r.step];
    MidTask t;
    t.board = r.board;
    t.step = r.step == 0 ? kStepGate : r.step == 1 ? kStepPrep : r.step - 2;
    t.begin = int64_t(r.chunk) * st.chunk;
    t.count = int32_t(std::min<int64_t>(st.chunk, st.entries - t.begin));
    if (st.entries == 0) { t.begin = 0; t.count = 0; }
    t.done_slot = bs.slot_base + r.step;
    if (r.step == 0) {
      if (bs.prev_board >= 0) {
        const int pb = bs.prev_board;
        t.dep_off = step_dep(pb, int(B[pb].steps.size()) - 1);
        t.dep_cnt = 1;
      } else {
        t.dep_off = 0;
        t.dep_cnt = 0;
      }
    } else {
      t.dep_off = step_dep(r.board, r.step - 1);
      t.dep_cnt = 1;
    }
    order->push_back(t);
    const double end = now + (double(t.count) * st.weight + 64.0) / rate;
    finish[r.board][r.step] = std::max(finish[r.board][r.step], end);
    if (++emitted[r.board][r.step] == st.tasks) {
      const double done = finish[r.board][r.step];
      if (r.step + 1 < int(bs.steps.size())) events.push(Event{done, 0, r.board, r.step + 1});
      else events.push(Event{done, 1, r.board, 0});
    }
    workers.push(end);
  }
  offsets->resize(nb);
  for (int b = 0; b < nb; b++) (*offsets)[b] = B[b].offset;
  *arena_bytes = carved;
  *nslots = slot_count;
  // integrity: every dependency is met by tasks earlier in the order, every counter reaches
  // its task count, and boards sharing a slot are chained through their gates
  {
    std::vector<int32_t> seen(size_t(slot_count), 0);
    for (const MidTask& t : *order) {
      for (int d = 0; d < t.dep_cnt; d++) {
        const int2 dp = (*deps)[size_t(t.dep_off + d)];
        if (seen[size_t(dp.x)] < dp.y) return "internal: task order violates a dependency";
      }
      seen[size_t(t.done_slot)]++;
    }
    for (int b = 0; b < nb; b++) {
      for (size_t k = 0; k < B[b].steps.size(); k++)
        if (seen[size_t(B[b].slot_base) + k] != B[b].steps[k].tasks) return "internal: a step's tasks are missing";
      if (B[b].offset < 0 || B[b].offset + B[b].bytes > carved) return "internal: board outside the arena";
    }
    for (size_t k = 0; k < slots.size(); k++) {
      int chain = 0;
      for (int b = slots[k].last_board; b >= 0; b = B[b].prev_board) {
        if (B[b].offset != slots[k].offset || ++chain > nb) return "internal: slot chain broken";
      }
    }
    if (!fixed_offsets && carved > cfg.budget) return "internal: workspace over budget";
  }
  return nullptr;
}

pe_status mid_plan_prepare(const std::vector<MidInfo>& infos, const MidConfig& cfg,
                           const std::vector<int64_t>* fixed_offsets, MidPrepared* p) {
  if (const char* e = mid_plan_build_schedule(infos, cfg, fixed_offsets, &p->order, &p->deps, &p->offsets,
                                              &p->arena_bytes, &p->nslots)) {
    set_error(e);
    return std::strncmp(e, "internal", 8) ? PE_ERR_CAPACITY : PE_ERR_INVALID;
  }
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t nb = infos.size();
  p->o_tasks = up(sizeof(MidBoardDev) * nb);
  p->o_deps = p->o_tasks + up(sizeof(MidTask) * p->order.size());
  p->o_res = p->o_deps + up(sizeof(int2) * std::max<size_t>(p->deps.size(), 1));
  p->o_state = p->o_res + up(sizeof(pe_halfsupers) * 19 * std::max<size_t>(nb, 1));
  p->o_head = p->o_state + up(sizeof(int32_t) * size_t(p->nslots));
  p->blob_bytes = p->o_head + 256;
  return PE_OK;
}

void mid_plan_fill(const std::vector<MidInfo>& infos, const MidConfig& cfg, const MidPrepared& p, void* dev_blob,
                   uint8_t* arena, const DevTables& T, std::vector<uint8_t>* host_blob, MidPlan* plan) {
  const int nb = int(infos.size());
  auto* base = static_cast<uint8_t*>(dev_blob);
  std::vector<MidBoardDev> boards(nb);
  for (int b = 0; b < nb; b++) {
    boards[b].I = infos[b];
    mid_carve(&boards[b], arena + p.offsets[b]);
    boards[b].results = reinterpret_cast<pe_halfsupers*>(base + p.o_res) + 19 * size_t(b);
  }
  host_blob->assign(p.o_state, 0);
  std::memcpy(host_blob->data(), boards.data(), sizeof(MidBoardDev) * nb);
  std::memcpy(host_blob->data() + p.o_tasks, p.order.data(), sizeof(MidTask) * p.order.size());
  if (!p.deps.empty()) std::memcpy(host_blob->data() + p.o_deps, p.deps.data(), sizeof(int2) * p.deps.size());
  plan->nb = nb;
  plan->ntasks = int64_t(p.order.size());
  plan->nslots = p.nslots;
  plan->cfg = cfg;
  plan->spots.resize(nb);
  for (int b = 0; b < nb; b++) plan->spots[b] = infos[b].spots;
  plan->arena_bytes = p.arena_bytes;
  plan->boards = reinterpret_cast<const MidBoardDev*>(base);
  plan->tasks = reinterpret_cast<const MidTask*>(base + p.o_tasks);
  plan->deps = reinterpret_cast<const int2*>(base + p.o_deps);
  plan->results = reinterpret_cast<pe_halfsupers*>(base + p.o_res);
  plan->counters = reinterpret_cast<int32_t*>(base + p.o_state);
  plan->head = reinterpret_cast<unsigned long long*>(base + p.o_head);
  plan->error = reinterpret_cast<int32_t*>(base + p.o_head + 64);
  plan->tables = T;
}

pe_status mid_plan_create(const std::vector<MidInfo>& infos, const MidConfig& cfg, MidPlan* plan) {
  DevTables T;
  std::string err;
  if (device_tables(&T, &err) != PE_OK) { set_error(err); return PE_ERR_DEVICE; }
  MidPrepared p;
  pe_status st = mid_plan_prepare(infos, cfg, nullptr, &p);
  if (st != PE_OK) return st;
  plan->device = 0;
  PE_HIP(hipGetDevice(&plan->device), "hipGetDevice");
  void* d = nullptr;
  PE_HIP(hipMalloc(&d, p.blob_bytes), "hipMalloc(mid plan)");
  void* arena = nullptr;
  if (hipMalloc(&arena, size_t(std::max<int64_t>(p.arena_bytes, 256))) != hipSuccess) {
    (void)hipFree(d);
    set_error("hipMalloc(mid workspace arena) failed");
    return PE_ERR_DEVICE;
  }
  std::vector<uint8_t> blob;
  mid_plan_fill(infos, cfg, p, d, static_cast<uint8_t*>(arena), T, &blob, plan);
  if (hipMemcpy(d, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    (void)hipFree(arena);
    set_error("hipMemcpy(mid plan) failed");
    return PE_ERR_DEVICE;
  }
  plan->blob = d;
  plan->arena = arena;
  return PE_OK;
}

int mid_queue_blocks_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mid_queue_kernel, 256, 0) != hipSuccess) return 1;
  return n;
}

pe_status mid_plan_launch(const MidPlan& plan, hipStream_t stream) {
  const int64_t n = plan.nslots;
  hipLaunchKernelGGL(mid_queue_reset_kernel, dim3(unsigned((std::max<int64_t>(n, 1) + 255) / 256)), dim3(256), 0,
                     stream, plan.counters, n, plan.head, plan.error);
  PE_HIP(hipGetLastError(), "mid_queue_reset_kernel launch");
  hipLaunchKernelGGL(mid_queue_kernel, dim3(unsigned(std::max(1, plan.cfg.workgroups))), dim3(256), 0, stream,
                     plan.boards, plan.tasks, plan.ntasks, plan.deps, plan.counters, plan.head, plan.error,
                     plan.cfg.wait_limit, plan.tables);
  PE_HIP(hipGetLastError(), "mid_queue_kernel launch");
  return PE_OK;
}

pe_status mid_plan_run(MidPlan& plan, hipStream_t stream) {
  if (!plan.cfg.graph) return mid_plan_launch(plan, stream);
  if (!plan.graph_exec) {
    hipStream_t cap = nullptr;
    PE_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate(capture)");
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const pe_status st = mid_plan_launch(plan, cap);
      e = hipStreamEndCapture(cap, &g);
      if (st != PE_OK) e = hipErrorUnknown;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&plan.graph_exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cap);
    if (e != hipSuccess) {
      plan.graph_exec = nullptr;
      set_error(std::string("mid plan graph capture failed: ") + hipGetErrorString(e));
      return PE_ERR_DEVICE;
    }
  }
  PE_HIP(hipGraphLaunch(plan.graph_exec, stream), "hipGraphLaunch(mid plan)");
  return PE_OK;
}

pe_status mid_plan_results(const MidPlan& plan, pe_halfsupers (*out)[19], hipStream_t stream) {
  int32_t error = 0;
  PE_HIP(hipMemcpyAsync(&error, plan.error, sizeof error, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(error)");
  if (plan.nb)
    PE_HIP(hipMemcpyAsync(out, plan.results, sizeof(pe_halfsupers) * 19 * size_t(plan.nb), hipMemcpyDeviceToHost,
                          stream), "hipMemcpyAsync(results)");
  PE_HIP(hipStreamSynchronize(stream), "mid plan results");
  if (error) { set_error("midsolve work queue: a dependency wait exceeded the wait limit"); return PE_ERR_DEVICE; }
  for (int b = 0; b < plan.nb; b++)
    for (int i = plan.spots[b] + 1; i < 19; i++) std::memset(&out[b][i], 0, sizeof(pe_halfsupers));
  return PE_OK;
}

void mid_plan_destroy(MidPlan* plan) {
  if (plan->graph_exec) (void)hipGraphExecDestroy(plan->graph_exec);
  if (plan->blob) (void)hipFree(plan->blob);
  if (plan->arena) (void)hipFree(plan->arena);
  plan->graph_exec = nullptr;
  plan->blob = plan->arena = nullptr;
}

}  // namespace pe
