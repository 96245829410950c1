// Batched midsolve: small boards through a persistent one-workgroup-per-board kernel fed
// by a device queue, large boards level by level on their own streams (see mid_queue.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "mid_device.h"
#include "mid_queue.h"

namespace pe {

namespace {

__global__ void mid_head_reset_kernel(unsigned int* head) { *head = 0; }

// Workgroups pop small boards from `head` and sweep all their levels; workgroup w owns the
// workspace slot work_base + w * work_stride.  Levels are separated by workgroup barriers,
// which order the global-memory writes of one level before the reads of the next.
template <int kThreads>
__global__ __launch_bounds__(kThreads) void mid_boards_kernel(const MidBoardDev* __restrict__ boards, int nboards,
                                                              uint8_t* __restrict__ work_base, int64_t work_stride,
                                                              unsigned int* __restrict__ head, DevTables T) {
  __shared__ int s_b;
  uint64_t* work = reinterpret_cast<uint64_t*>(work_base + int64_t(blockIdx.x) * work_stride);
  for (;;) {
    if (threadIdx.x == 0) s_b = int(atomicAdd(head, 1u));
    __syncthreads();
    const int b = s_b;
    __syncthreads();
    if (b >= nboards) return;
    const MidBoardDev& B = boards[b];
    const int spots = B.I.spots;
    for (int n = spots; n >= 0; n--) {
      const int positions = int(B.I.positions[n]);
      for (int p = threadIdx.x; p < positions; p += kThreads) mid_position(B, n, p, work, T);
      __syncthreads();
    }
  }
}

size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

int mid_board_blocks_per_cu(int threads) {
  int n = 0;
  hipError_t e = hipErrorInvalidValue;
  if (threads == 256) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mid_boards_kernel<256>, 256, 0);
  if (threads == 512) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mid_boards_kernel<512>, 512, 0);
  if (threads == 1024) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mid_boards_kernel<1024>, 1024, 0);
  return e == hipSuccess ? std::max(n, 1) : 1;
}

const char* mid_layout(const std::vector<MidInfo>& infos, const MidConfig& cfg, MidLayout* L) {
  *L = MidLayout();
  const int nb = int(infos.size());
  if (cfg.threads != 256 && cfg.threads != 512 && cfg.threads != 1024) return "threads must be 256, 512 or 1024";
  for (int b = 0; b < nb; b++) (infos[b].spots <= cfg.small_spots ? L->small : L->big).push_back(b);
  std::stable_sort(L->big.begin(), L->big.end(), [&](int a, int b) { return infos[a].spots > infos[b].spots; });
  // descriptor blob
  L->o_big = up256(sizeof(MidBoardDev) * L->small.size());
  L->o_prefix = L->o_big + up256(sizeof(MidBoardDev) * L->big.size());
  L->o_res = L->o_prefix + up256(sizeof(int64_t) * (L->small.size() + 1 + 2 * L->big.size()));
  L->o_head = L->o_res + up256(sizeof(pe_halfsupers) * 19 * size_t(std::max(nb, 1)));
  L->blob_bytes = L->o_head + 256;
  size_t off = L->blob_bytes;
  // small boards: resident tables, one workspace slot per workgroup
  L->table_off.resize(L->small.size());
  size_t small_stride = 0;
  for (size_t k = 0; k < L->small.size(); k++) {
    L->table_off[k] = off;
    off += mid_table_bytes(infos[size_t(L->small[k])]);
    small_stride = std::max(small_stride, mid_work_bytes(infos[size_t(L->small[k])]));
  }
  // large boards: one slot (tables + workspace) per stream
  size_t big_stride = 0;
  for (int b : L->big) big_stride = std::max(big_stride, mid_table_bytes(infos[size_t(b)]) + mid_work_bytes(infos[size_t(b)]));
  const size_t fixed = off;
  if (fixed > size_t(cfg.budget)) return "the batch's small-board tables exceed the workspace budget";
  const int big_min = L->big.empty() ? 0 : 1;
  int streams = L->big.empty() ? 0 : std::max(cfg.streams, 0);
  int wgs = L->small.empty() ? 0 : std::max(1, std::min<int>(cfg.workgroups, int(L->small.size())));
  auto total = [&](int w, int s) { return fixed + size_t(w) * small_stride + size_t(std::max(s, big_min)) * big_stride; };
  while (total(wgs, streams) > size_t(cfg.budget) && streams > 1) streams--;
  while (total(wgs, streams) > size_t(cfg.budget) && wgs > 1) wgs = std::max(1, wgs / 2);
  if (total(wgs, streams) > size_t(cfg.budget)) return "a board needs more workspace than the budget";
  L->workgroups = wgs;
  L->streams = streams;
  L->small_work = off;
  L->small_stride = small_stride;
  off += size_t(wgs) * small_stride;
  L->big_slots = off;
  L->big_stride = big_stride;
  off += size_t(std::max(streams, big_min)) * big_stride;
  L->bytes = off;
  // large boards to streams: each to the stream with the least work so far
  L->big_stream.assign(L->big.size(), 0);
  if (streams > 1) {
    std::vector<double> load(size_t(streams), 0.0);
    for (size_t k = 0; k < L->big.size(); k++) {
      const int s = int(std::min_element(load.begin(), load.end()) - load.begin());
      L->big_stream[k] = s;
      load[size_t(s)] += double(mid_work_bytes(infos[size_t(L->big[k])]));
    }
  }
  return nullptr;
}

pe_status mid_plan_create(const std::vector<MidInfo>& infos, const MidConfig& cfg, MidPlan* plan,
                          hipStream_t alloc_stream, void* upload, size_t upload_bytes, void* memory,
                          size_t memory_bytes) {
  DevTables T;
  std::string err;
  if (device_tables(&T, &err) != PE_OK) { set_error(err); return PE_ERR_DEVICE; }
  MidLayout& L = plan->layout;
  if (const char* e = mid_layout(infos, cfg, &L)) {
    set_error(e);
    return PE_ERR_CAPACITY;
  }
  plan->nb = int(infos.size());
  plan->cfg = cfg;
  plan->infos = infos;
  plan->tables = T;
  plan->alloc_stream = alloc_stream;
  PE_HIP(hipGetDevice(&plan->device), "hipGetDevice");
  void* mem = nullptr;
  plan->owns_mem = !(memory && L.bytes <= memory_bytes);
  if (!plan->owns_mem) mem = memory;
  else if (alloc_stream) PE_HIP(hipMallocAsync(&mem, L.bytes, alloc_stream), "hipMallocAsync(mid batch)");
  else PE_HIP(hipMalloc(&mem, L.bytes), "hipMalloc(mid batch)");
  plan->mem = static_cast<uint8_t*>(mem);
  uint8_t* base = plan->mem;
  // descriptors
  std::vector<uint8_t> blob(L.o_res, 0);
  auto* sdev = reinterpret_cast<MidBoardDev*>(blob.data());
  auto* bdev = reinterpret_cast<MidBoardDev*>(blob.data() + L.o_big);
  auto* prefix = reinterpret_cast<int64_t*>(blob.data() + L.o_prefix);
  auto* res = reinterpret_cast<pe_halfsupers*>(base + L.o_res);
  prefix[0] = 0;
  for (size_t k = 0; k < L.small.size(); k++) {
    MidBoardDev B;
    std::memset(&B, 0, sizeof B);
    B.I = infos[size_t(L.small[k])];
    mid_carve_tables(&B, base + L.table_off[k]);
    B.work = nullptr;   // the workgroup's slot
    B.results = res + 19 * size_t(L.small[k]);
    std::memcpy(&sdev[k], &B, sizeof B);
    const MidSizes s = mid_sizes(B.I);
    prefix[k + 1] = prefix[k] + s.set0 + s.wins1 + s.set1p;
  }
  plan->small_prep_total = prefix[L.small.size()];
  plan->big_prep_total.resize(L.big.size());
  for (size_t k = 0; k < L.big.size(); k++) {
    MidBoardDev B;
    std::memset(&B, 0, sizeof B);
    B.I = infos[size_t(L.big[k])];
    uint8_t* slot = base + L.big_slots + size_t(L.big_stream[k]) * L.big_stride;
    mid_carve_tables(&B, slot);
    B.work = reinterpret_cast<uint64_t*>(slot + mid_table_bytes(B.I));
    B.results = res + 19 * size_t(L.big[k]);
    std::memcpy(&bdev[k], &B, sizeof B);
    const MidSizes s = mid_sizes(B.I);
    int64_t* p = prefix + L.small.size() + 1 + 2 * k;
    p[0] = 0;
    p[1] = s.set0 + s.wins1 + s.set1p;
    plan->big_prep_total[k] = p[1];
  }
  if (alloc_stream && upload && blob.size() <= upload_bytes) {
    // pinned staging owned by the caller: the copy is queued, nothing waits
    std::memcpy(upload, blob.data(), blob.size());
    PE_HIP(hipMemcpyAsync(base, upload, blob.size(), hipMemcpyHostToDevice, alloc_stream),
           "hipMemcpyAsync(mid batch)");
  } else if (alloc_stream) {
    PE_HIP(hipMemcpyAsync(base, blob.data(), blob.size(), hipMemcpyHostToDevice, alloc_stream),
           "hipMemcpyAsync(mid batch)");
    PE_HIP(hipStreamSynchronize(alloc_stream), "mid batch upload");   // blob is a local
  } else {
    PE_HIP(hipMemcpy(base, blob.data(), blob.size(), hipMemcpyHostToDevice), "hipMemcpy(mid batch)");
  }
  plan->small_dev = reinterpret_cast<const MidBoardDev*>(base);
  plan->big_dev = reinterpret_cast<const MidBoardDev*>(base + L.o_big);
  plan->prefix = reinterpret_cast<const int64_t*>(base + L.o_prefix);
  plan->results = res;
  plan->head = reinterpret_cast<unsigned int*>(base + L.o_head);
  for (int s = 0; s < L.streams; s++) {
    hipStream_t st = nullptr;
    PE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate(mid batch)");
    plan->streams.push_back(st);
  }
  for (int e = 0; e < L.streams + 1; e++) {
    hipEvent_t ev = nullptr;
    PE_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate(mid batch)");
    plan->events.push_back(ev);
  }
  int64_t launches = L.small.empty() ? 0 : 3;
  for (int b : L.big) launches += 1 + infos[size_t(b)].spots + 1;
  plan->launches = launches;
  return PE_OK;
}

namespace {

pe_status launch_small(const MidPlan& plan, hipStream_t s) {
  const MidLayout& L = plan.layout;
  if (L.small.empty()) return PE_OK;
  hipLaunchKernelGGL(mid_head_reset_kernel, dim3(1), dim3(1), 0, s, plan.head);
  PE_HIP(hipGetLastError(), "mid_head_reset_kernel launch");
  pe_status st = mid_launch_prep(plan.small_dev, plan.prefix, int(L.small.size()), plan.small_prep_total,
                                 plan.tables, s);
  if (st != PE_OK) return st;
  uint8_t* work = plan.mem + L.small_work;
  const int n = int(L.small.size());
  const int64_t stride = int64_t(L.small_stride);
  const dim3 grid(unsigned(L.workgroups));
  switch (plan.cfg.threads) {
    case 256:
      hipLaunchKernelGGL(mid_boards_kernel<256>, grid, dim3(256), 0, s, plan.small_dev, n, work, stride, plan.head,
                         plan.tables);
      break;
    case 512:
      hipLaunchKernelGGL(mid_boards_kernel<512>, grid, dim3(512), 0, s, plan.small_dev, n, work, stride, plan.head,
                         plan.tables);
      break;
    default:
      hipLaunchKernelGGL(mid_boards_kernel<1024>, grid, dim3(1024), 0, s, plan.small_dev, n, work, stride, plan.head,
                         plan.tables);
  }
  PE_HIP(hipGetLastError(), "mid_boards_kernel launch");
  return PE_OK;
}

pe_status launch_big(const MidPlan& plan, size_t k, hipStream_t s) {
  const MidLayout& L = plan.layout;
  const MidInfo& I = plan.infos[size_t(L.big[k])];
  const MidBoardDev* B = plan.big_dev + k;
  pe_status st = mid_launch_prep(B, plan.prefix + L.small.size() + 1 + 2 * k, 1, plan.big_prep_total[k], plan.tables, s);
  if (st != PE_OK) return st;
  for (int n = I.spots; n >= 0; n--) {
    st = mid_launch_level(B, I, n, I.positions[n], plan.tables, s);
    if (st != PE_OK) return st;
  }
  return PE_OK;
}

// The whole run on `stream`, forking the large boards onto the plan's streams.
pe_status launch_all(const MidPlan& plan, hipStream_t stream) {
  const MidLayout& L = plan.layout;
  if (L.streams == 0) {
    pe_status st = launch_small(plan, stream);
    for (size_t k = 0; k < L.big.size() && st == PE_OK; k++) st = launch_big(plan, k, stream);
    return st;
  }
  PE_HIP(hipEventRecord(plan.events[0], stream), "hipEventRecord(fork)");
  for (int s = 0; s < L.streams; s++)
    PE_HIP(hipStreamWaitEvent(plan.streams[size_t(s)], plan.events[0], 0), "hipStreamWaitEvent(fork)");
  for (size_t k = 0; k < L.big.size(); k++) {
    const pe_status st = launch_big(plan, k, plan.streams[size_t(L.big_stream[k])]);
    if (st != PE_OK) return st;
  }
  const pe_status st = launch_small(plan, stream);
  if (st != PE_OK) return st;
  for (int s = 0; s < L.streams; s++) {
    PE_HIP(hipEventRecord(plan.events[size_t(s) + 1], plan.streams[size_t(s)]), "hipEventRecord(join)");
    PE_HIP(hipStreamWaitEvent(stream, plan.events[size_t(s) + 1], 0), "hipStreamWaitEvent(join)");
  }
  return PE_OK;
}

}  // namespace

pe_status mid_plan_run(MidPlan& plan, hipStream_t stream) {
  if (!plan.cfg.graph) return launch_all(plan, stream);
  if (!plan.graph_exec) {
    hipStream_t cap = nullptr;
    PE_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate(capture)");
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const pe_status st = launch_all(plan, cap);
      e = hipStreamEndCapture(cap, &g);
      if (st != PE_OK && e == hipSuccess) e = hipErrorUnknown;
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

pe_status mid_plan_copy_results(const MidPlan& plan, pe_halfsupers (*out)[19], hipStream_t stream) {
  if (plan.nb)
    PE_HIP(hipMemcpyAsync(out, plan.results, sizeof(pe_halfsupers) * 19 * size_t(plan.nb), hipMemcpyDeviceToHost,
                          stream), "hipMemcpyAsync(results)");
  return PE_OK;
}

void mid_plan_zero_tail(const MidPlan& plan, pe_halfsupers (*out)[19]) {
  for (int b = 0; b < plan.nb; b++)
    for (int i = plan.infos[size_t(b)].spots + 1; i < 19; i++) std::memset(&out[b][i], 0, sizeof(pe_halfsupers));
}

void mid_plan_destroy(MidPlan* plan) {
  if (plan->graph_exec) (void)hipGraphExecDestroy(plan->graph_exec);
  for (hipStream_t s : plan->streams) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  for (hipEvent_t e : plan->events) (void)hipEventDestroy(e);
  if (plan->mem && plan->owns_mem) {
    if (plan->alloc_stream) (void)hipFreeAsync(plan->mem, plan->alloc_stream);
    else (void)hipFree(plan->mem);
  }
  plan->graph_exec = nullptr;
  plan->streams.clear();
  plan->events.clear();
  plan->mem = nullptr;
}

}  // namespace pe
