// Serving loop for position solves: requests are admitted as they arrive, while earlier
// ones are still on the device (continuous batching).
//
// One worker thread owns the device work.  Each pass it (1) retires finished batches:
// copies their results out, releases their memory and marks the requests done, and (2)
// while a stream is free, forms a batch from the pending requests in arrival order whose
// device memory (tables, workspace, descriptors) fits in the part of the HBM budget not held
// by batches in flight (a request that does not fit waits; later, smaller requests may pass
// it until it has waited kHeadOfLine, after which admission stops until it fits), and runs
// the batch on that stream as a batched plan.  A board with more empty cells than the
// small-board limit is a batch of its own and such batches hold at most all streams but
// one, so the small boards' batches (one persistent kernel each, about a millisecond) are
// never queued behind a large board's level launches.  Batches on different streams
// overlap on the device.  Each request's state stays in device memory from admission to
// retirement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "device.h"
#include "mid.h"
#include "mid_queue.h"

namespace pe {

namespace {

using Clock = std::chrono::steady_clock;
constexpr auto kHeadOfLine = std::chrono::milliseconds(20);

struct Request {
  MidInfo info;
  Clock::time_point submitted, done_at;
  bool done = false;
  pe_status status = PE_OK;
  std::string error;
  pe_halfsupers results[19];
};

struct Batch {
  std::vector<int64_t> tickets;
  MidPlan plan;
  int64_t bytes = 0;
  bool large = false;   // one board with more than small_spots empty cells
};

}  // namespace

struct MidServer {
  int device = 0;
  int32_t max_batch = 512;
  int64_t budget = 0, in_use = 0;
  int64_t footprint_peak = 0;            // largest device bytes held by the arenas together
  MidConfig cfg{};
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<pe_halfsupers*> staging;   // pinned, max_batch * 19 per stream
  std::vector<void*> upload;             // pinned descriptor staging per stream
  std::vector<size_t> upload_cap;
  std::vector<void*> arena;              // device memory per stream, grown to the largest batch
  std::vector<size_t> arena_cap;
  std::vector<char> stream_busy;
  std::vector<Batch> batch_of_stream;

  std::mutex mu;
  std::condition_variable cv_done, cv_work;
  std::unordered_map<int64_t, Request> requests;
  std::deque<int64_t> pending;
  int64_t next_ticket = 1, submitted = 0, completed = 0, batches = 0;
  bool stop = false;
  bool broken = false;   // the worker stopped on an error: no request is admitted any more
  std::thread worker;

  // Mark requests that never reached a stream as failed.  Caller holds no lock.
  void fail_tickets(const std::vector<int64_t>& tickets, const std::string& e) {
    std::lock_guard<std::mutex> lock(mu);
    const auto now = Clock::now();
    for (int64_t t : tickets) {
      const auto it = requests.find(t);
      if (it == requests.end()) continue;
      it->second.status = PE_ERR_CAPACITY;
      it->second.error = e;
      it->second.done = true;
      it->second.done_at = now;
      completed++;
    }
    cv_done.notify_all();
  }

  // Every request not finished yet fails (the worker could not continue).
  void fail_all(const std::string& e) {
    std::lock_guard<std::mutex> lock(mu);
    const auto now = Clock::now();
    for (auto& kv : requests)
      if (!kv.second.done) {
        kv.second.status = PE_ERR_CAPACITY;
        kv.second.error = e;
        kv.second.done = true;
        kv.second.done_at = now;
        completed++;
      }
    pending.clear();
    stop = broken = true;
    cv_done.notify_all();
  }

  // Finish the requests of a batch (results from staging when ok).  Caller holds no lock.
  void finish(int k, pe_status st, const std::string& e) {
    Batch& b = batch_of_stream[size_t(k)];
    const pe_halfsupers* res = staging[size_t(k)];
    {
      std::lock_guard<std::mutex> lock(mu);
      const auto now = Clock::now();
      for (size_t i = 0; i < b.tickets.size(); i++) {
        Request& r = requests[b.tickets[i]];
        if (st == PE_OK) {
          std::memcpy(r.results, res + 19 * i, sizeof r.results);
          for (int j = r.info.spots + 1; j < 19; j++) std::memset(&r.results[j], 0, sizeof(pe_halfsupers));
        }
        r.status = st;
        r.error = e;
        r.done = true;
        r.done_at = now;
        completed++;
      }
      in_use -= b.bytes;
    }
    cv_done.notify_all();
    mid_plan_destroy(&b.plan);   // stream-ordered free on the batch's stream
    b = Batch();
    stream_busy[size_t(k)] = 0;
  }

  // Device bytes held by the arenas of the busy streams other than k, and by every arena
  // other than k (k = -1: all).  Only the worker thread changes arenas and busy flags.
  int64_t held_busy_except(int k) const {
    int64_t h = 0;
    for (size_t j = 0; j < arena.size(); j++)
      if (int(j) != k && stream_busy[j]) h += int64_t(arena_cap[j]);
    return h;
  }
  size_t held_all_except(int k) const {
    size_t h = 0;
    for (size_t j = 0; j < arena.size(); j++)
      if (int(j) != k) h += arena_cap[j];
    return h;
  }

  // Launch the batch on stream k: plan (in the stream's arena), run, results to staging.
  pe_status launch(int k) {
    Batch& b = batch_of_stream[size_t(k)];
    const hipStream_t s = streams[size_t(k)];
    std::vector<MidInfo> infos;
    {
      std::lock_guard<std::mutex> lock(mu);
      for (int64_t t : b.tickets) infos.push_back(requests[t].info);
    }
    // descriptors go up through this stream's pinned staging, so creating the plan never
    // waits for the device
    MidLayout L;
    if (const char* e = mid_layout(infos, cfg, &L)) {
      set_error(e);
      return PE_ERR_CAPACITY;
    }
    if (L.o_res > upload_cap[size_t(k)]) {
      if (upload[size_t(k)]) (void)hipHostFree(upload[size_t(k)]);
      upload[size_t(k)] = nullptr;
      upload_cap[size_t(k)] = 0;
      const size_t cap = std::max<size_t>(L.o_res, size_t(1) << 20);
      PE_HIP(hipHostMalloc(&upload[size_t(k)], cap, hipHostMallocDefault), "hipHostMalloc(server upload)");
      upload_cap[size_t(k)] = cap;
    }
    // each stream keeps one device arena (batches on a stream run one after another), so a
    // batch allocates nothing once the arena has grown to the largest batch it runs;
    // per-batch stream-ordered allocations made concurrent streams wait on each other.
    // Admission keeps every arena of a busy stream plus this one within the budget, so the
    // arena can always grow here once idle streams give theirs up.
    if (L.bytes > arena_cap[size_t(k)]) {
      const size_t room = size_t(budget) - size_t(held_busy_except(k));
      if (L.bytes > room) {
        set_error("batch exceeds the workspace budget left by the batches in flight");
        return PE_ERR_CAPACITY;
      }
      const size_t cap = std::min(room, std::max<size_t>(L.bytes + L.bytes / 4, size_t(64) << 20));
      for (size_t j = 0; j < arena.size(); j++)
        if (j != size_t(k) && !stream_busy[j] && arena[j] && held_all_except(k) + cap > size_t(budget)) {
          (void)hipFree(arena[j]);
          arena[j] = nullptr;
          arena_cap[j] = 0;
        }
      if (arena[size_t(k)]) (void)hipFree(arena[size_t(k)]);
      arena[size_t(k)] = nullptr;
      arena_cap[size_t(k)] = 0;
      PE_HIP(hipMalloc(&arena[size_t(k)], cap), "hipMalloc(server arena)");
      arena_cap[size_t(k)] = cap;
      footprint_peak = std::max(footprint_peak, int64_t(held_all_except(-1)));
    }
    pe_status st = mid_plan_create(infos, cfg, &b.plan, s, upload[size_t(k)], upload_cap[size_t(k)], arena[size_t(k)],
                                   arena_cap[size_t(k)]);
    if (st != PE_OK) return st;
    st = mid_plan_run(b.plan, s);
    if (st != PE_OK) return st;
    st = mid_plan_copy_results(b.plan, reinterpret_cast<pe_halfsupers (*)[19]>(staging[size_t(k)]), s);
    if (st != PE_OK) return st;
    PE_HIP(hipEventRecord(events[size_t(k)], s), "hipEventRecord(batch)");
    return PE_OK;
  }

  // Bytes a batch of these requests would hold.
  int64_t batch_bytes(const std::vector<MidInfo>& infos) const {
    MidConfig c = cfg;
    c.budget = int64_t(1) << 62;
    MidLayout L;
    if (mid_layout(infos, c, &L)) return int64_t(1) << 62;
    return int64_t(L.bytes);
  }

  void run() {
    (void)hipSetDevice(device);
    for (;;) {
      bool progressed = false;
      for (size_t k = 0; k < streams.size(); k++) {
        if (!stream_busy[k]) continue;
        const hipError_t q = hipEventQuery(events[k]);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) finish(int(k), PE_ERR_DEVICE, std::string("batch failed: ") + hipGetErrorString(q));
        else finish(int(k), PE_OK, "");
        progressed = true;
      }
      for (size_t k = 0; k < streams.size(); k++) {
        if (stream_busy[k]) continue;
        Batch b;
        std::vector<MidInfo> infos;
        try {
          std::lock_guard<std::mutex> lock(mu);
          const auto now = Clock::now();
          // A large board (more empty cells than small_spots: full-device level launches,
          // milliseconds) runs alone, on at most all streams but one; small boards batch
          // together on the others, so they never wait behind a large board's levels.
          int large_busy = 0;
          for (size_t j = 0; j < streams.size(); j++) large_busy += stream_busy[j] && batch_of_stream[j].large;
          const bool large_ok = streams.size() == 1 || large_busy + 1 < int(streams.size());
          // bound >= the batch's exact bytes (every small board counted with a full workspace
          // slot, descriptors with their alignment), updated in O(1) per request; the exact
          // layout is computed only when the bound does not fit
          int64_t tables = 0, max_work = 0, max_slot = 0, nsmall = 0;
          for (auto it = pending.begin(); it != pending.end() && int32_t(b.tickets.size()) < max_batch;) {
            Request& r = requests[*it];
            const bool large = r.info.spots > cfg.small_spots;
            if (large && (!b.tickets.empty() || !large_ok)) {
              ++it;
              continue;
            }
            infos.push_back(r.info);
            int64_t t2 = tables, w2 = max_work, s2 = max_slot, n2 = nsmall;
            if (large) {
              s2 = std::max<int64_t>(s2, int64_t(mid_table_bytes(r.info) + mid_work_bytes(r.info)));
            } else {
              t2 += int64_t(mid_table_bytes(r.info));
              w2 = std::max<int64_t>(w2, int64_t(mid_work_bytes(r.info)));
              n2++;
            }
            // per board: a descriptor (MidInfo + 6 pointers, padded), 3 prefix entries and
            // 19 results; 256-byte alignment of the 4 blob parts and the queue head
            const int64_t per_board = int64_t(sizeof(MidInfo) + 8 * sizeof(void*) + 3 * sizeof(int64_t) +
                                              19 * sizeof(pe_halfsupers));
            const int64_t bound = t2 + n2 * w2 + s2 + int64_t(infos.size()) * per_board + 5 * 256 + 4096;
            // the device footprint once this batch runs: the arenas of the busy streams and
            // this stream's arena grown to the batch (idle streams give theirs up)
            const int64_t others = held_busy_except(int(k)), own = int64_t(arena_cap[k]);
            if (others + std::max(own, bound) > budget && others + std::max(own, batch_bytes(infos)) > budget) {
              infos.pop_back();
              // the oldest waiting request blocks further admission once it has waited long
              if (it == pending.begin() && now - r.submitted > kHeadOfLine) break;
              ++it;
              continue;
            }
            tables = t2, max_work = w2, max_slot = s2, nsmall = n2;
            b.tickets.push_back(*it);
            it = pending.erase(it);
            if (large) {
              b.large = true;
              break;
            }
          }
          if (!b.tickets.empty()) b.bytes = batch_bytes(infos);
          if (!b.tickets.empty()) {
            in_use += b.bytes;
            batches++;
          }
        } catch (const std::exception& e) {   // host memory exhausted while forming the batch
          fail_tickets(b.tickets, std::string("serving loop: ") + e.what());
          if (b.bytes) {
            std::lock_guard<std::mutex> lock(mu);
            in_use -= b.bytes;
          }
          progressed = true;
          continue;
        }
        if (b.tickets.empty()) break;
        batch_of_stream[k] = std::move(b);
        stream_busy[k] = 1;
        pe_status st = PE_OK;
        std::string why = "batch launch failed";
        try {
          st = launch(int(k));
        } catch (const std::bad_alloc&) {
          st = PE_ERR_CAPACITY;
          why = "host memory exhausted while launching a batch";
        } catch (const std::exception& e) {
          st = PE_ERR_INVALID;
          why = std::string("batch launch failed: ") + e.what();
        }
        if (st != PE_OK) {
          (void)hipStreamSynchronize(streams[k]);
          finish(int(k), st, why);
        }
        progressed = true;
      }
      std::unique_lock<std::mutex> lock(mu);
      const bool idle = std::none_of(stream_busy.begin(), stream_busy.end(), [](char c) { return c != 0; });
      if (stop && pending.empty() && idle) break;
      if (!progressed) {
        if (idle && pending.empty()) cv_work.wait(lock, [&] { return stop || !pending.empty(); });
        else cv_work.wait_for(lock, std::chrono::microseconds(40));
      }
    }
  }
};

pe_status mid_server_create(const pe_mid_server_config* c, MidServer** out) {
  auto* S = new MidServer;
  auto fail = [&](pe_status st, const std::string& e) { set_error(e); delete S; return st; };
  DevTables T;
  std::string err;
  if (device_tables(&T, &err) != PE_OK) return fail(PE_ERR_DEVICE, err);
  if (hipGetDevice(&S->device) != hipSuccess) return fail(PE_ERR_DEVICE, "hipGetDevice failed");
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, S->device);
  int64_t budget = c ? c->workspace_budget : 0;
  if (budget <= 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(PE_ERR_DEVICE, "hipMemGetInfo failed");
    budget = std::min<int64_t>(int64_t(100e9), int64_t(0.4 * double(free_b)));
  }
  const int nstreams = c && c->streams > 0 ? c->streams : 4;
  S->max_batch = c && c->max_batch > 0 ? c->max_batch : 512;
  S->budget = budget;
  S->cfg.budget = budget;
  S->cfg.small_spots = c && c->small_spots > 0 ? c->small_spots : 14;
  S->cfg.threads = 512;
  S->cfg.workgroups = c && c->workgroups > 0 ? c->workgroups : cus * mid_board_blocks_per_cu(512) / nstreams;
  S->cfg.workgroups = std::max(S->cfg.workgroups, 1);
  S->cfg.streams = 0;   // a batch runs entirely on its server stream
  S->cfg.graph = false;

  for (int k = 0; k < nstreams; k++) {
    hipStream_t s = nullptr;
    hipEvent_t e = nullptr;
    void* h = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(&h, sizeof(pe_halfsupers) * 19 * size_t(S->max_batch), hipHostMallocDefault) != hipSuccess)
      return fail(PE_ERR_DEVICE, "server stream setup failed");
    S->streams.push_back(s);
    S->events.push_back(e);
    S->staging.push_back(static_cast<pe_halfsupers*>(h));
    S->upload.push_back(nullptr);
    S->upload_cap.push_back(0);
    S->arena.push_back(nullptr);
    S->arena_cap.push_back(0);
  }
  S->stream_busy.assign(size_t(nstreams), 0);
  S->batch_of_stream.resize(size_t(nstreams));
  S->worker = std::thread([S] {
    try {
      S->run();
    } catch (const std::exception& e) {   // no exception leaves the worker: fail what is left
      S->fail_all(std::string("serving loop stopped: ") + e.what());
    }
  });
  *out = S;
  return PE_OK;
}

pe_status mid_server_submit(MidServer* S, const pe_high_board& board, int64_t* ticket) {
  Request r;
  if (!make_mid_info(board, &r.info)) { set_error("not a valid board with >= 18 stones"); return PE_ERR_INVALID; }
  if (S->batch_bytes({r.info}) > S->budget) { set_error("board needs more workspace than the server holds"); return PE_ERR_CAPACITY; }
  {
    std::lock_guard<std::mutex> lock(S->mu);
    if (S->broken) { set_error("the serving loop stopped after an error"); return PE_ERR_DEVICE; }
    r.submitted = Clock::now();
    *ticket = S->next_ticket++;
    S->requests.emplace(*ticket, std::move(r));
    S->pending.push_back(*ticket);
    S->submitted++;
  }
  S->cv_work.notify_one();
  return PE_OK;
}

// Copy out a finished request and forget it.  Caller holds the lock.
static pe_status take(MidServer* S, int64_t ticket, pe_halfsupers* results, double* latency_ms) {
  auto it = S->requests.find(ticket);
  Request& r = it->second;
  const pe_status st = r.status;
  if (st == PE_OK && results) std::memcpy(results, r.results, sizeof r.results);
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.done_at - r.submitted).count();
  if (st != PE_OK) set_error(r.error);
  S->requests.erase(it);
  return st;
}

pe_status mid_server_poll(MidServer* S, int64_t ticket, pe_halfsupers* results, int* done, double* latency_ms) {
  std::lock_guard<std::mutex> lock(S->mu);
  auto it = S->requests.find(ticket);
  if (it == S->requests.end()) { set_error("unknown ticket"); return PE_ERR_INVALID; }
  *done = it->second.done ? 1 : 0;
  if (!*done) return PE_OK;
  return take(S, ticket, results, latency_ms);
}

pe_status mid_server_wait(MidServer* S, int64_t ticket, pe_halfsupers* results, double* latency_ms) {
  std::unique_lock<std::mutex> lock(S->mu);
  auto it = S->requests.find(ticket);
  if (it == S->requests.end()) { set_error("unknown ticket"); return PE_ERR_INVALID; }
  S->cv_done.wait(lock, [&] {
    auto i = S->requests.find(ticket);
    return i == S->requests.end() || i->second.done;
  });
  if (S->requests.find(ticket) == S->requests.end()) { set_error("ticket taken by another caller"); return PE_ERR_INVALID; }
  return take(S, ticket, results, latency_ms);
}

void mid_server_stats(MidServer* S, pe_mid_server_info* out) {
  std::lock_guard<std::mutex> lock(S->mu);
  out->submitted = S->submitted;
  out->completed = S->completed;
  out->batches = S->batches;
  out->peak_workspace = S->footprint_peak;
  out->pending = int64_t(S->pending.size());
  out->in_flight = S->submitted - S->completed - int64_t(S->pending.size());
}

void mid_server_destroy(MidServer* S) {
  {
    std::lock_guard<std::mutex> lock(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (size_t k = 0; k < S->streams.size(); k++) {
    (void)hipStreamSynchronize(S->streams[k]);
    (void)hipStreamDestroy(S->streams[k]);
    (void)hipEventDestroy(S->events[k]);
    (void)hipHostFree(S->staging[k]);
    if (S->upload[k]) (void)hipHostFree(S->upload[k]);
    if (S->arena[k]) (void)hipFree(S->arena[k]);
  }
  delete S;
}

}  // namespace pe
