// Serving loop for position solves: requests are admitted as they arrive, while earlier
// ones are still on the device (continuous batching).
//
// One worker thread owns the device work.  Each pass it (1) retires finished batches:
// copies their results out, releases their workspace and marks the requests done, and
// (2) while a stream is free, forms a batch from the pending requests in arrival order whose
// workspace fits in the free part of the HBM pool (a request that does not fit waits; later,
// smaller requests may pass it until it has waited kHeadOfLine, after which admission stops
// until it fits), and launches the batch on that stream through the work-queue kernel.
// Batches on different streams overlap on the device.  Each request's state (tables,
// workspace, results) stays in device memory from admission to retirement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "device.h"
#include "mid.h"
#include "mid_queue.h"

namespace pe {

size_t mid_board_bytes(const MidInfo& I);
int mid_queue_blocks_per_cu();

namespace {

using Clock = std::chrono::steady_clock;
constexpr auto kHeadOfLine = std::chrono::milliseconds(20);

// First-fit allocator over [0, size) with coalescing, 256-byte granules.
class Pool {
 public:
  explicit Pool(int64_t size) : size_(size) { free_[0] = size; }
  int64_t alloc(int64_t bytes) {
    bytes = (bytes + 255) & ~int64_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const int64_t off = it->first, rem = it->second - bytes;
      free_.erase(it);
      if (rem) free_[off + bytes] = rem;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return off;
    }
    return -1;
  }
  void release(int64_t off, int64_t bytes) {
    bytes = (bytes + 255) & ~int64_t(255);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t size() const { return size_; }
  int64_t peak() const { return peak_; }

 private:
  int64_t size_, used_ = 0, peak_ = 0;
  std::map<int64_t, int64_t> free_;
};

struct Request {
  pe_high_board board;
  MidInfo info;
  int64_t bytes = 0, offset = -1;
  Clock::time_point submitted, done_at;
  bool done = false;
  pe_status status = PE_OK;
  std::string error;
  pe_halfsupers results[19];
};

struct Batch {
  std::vector<int64_t> tickets;
  MidPlan plan;
  void* blob = nullptr;
  std::vector<uint8_t> host_blob;   // source of the blob upload, kept until the batch retires
  int stream = -1;
};

}  // namespace

struct MidServer {
  int device = 0;
  int32_t max_batch = 512;
  MidConfig cfg{};
  uint8_t* arena = nullptr;
  Pool pool{0};
  DevTables tables{};
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<pe_halfsupers*> staging;   // pinned, max_batch * 19 per stream
  std::vector<char> stream_busy;
  std::vector<Batch> batch_of_stream;

  std::mutex mu;
  std::condition_variable cv_done, cv_work;
  std::unordered_map<int64_t, Request> requests;
  std::deque<int64_t> pending;
  int64_t next_ticket = 1, submitted = 0, completed = 0, batches = 0;
  bool stop = false;
  std::thread worker;

  void fail_batch(Batch& b, pe_status st, const std::string& e) {
    std::lock_guard<std::mutex> lock(mu);
    for (int64_t t : b.tickets) {
      Request& r = requests[t];
      r.status = st;
      r.error = e;
      r.done = true;
      r.done_at = Clock::now();
      pool.release(r.offset, r.bytes);
      completed++;
    }
    cv_done.notify_all();
  }

  // Launch the batch on its stream: upload, reset, kernel, results to pinned staging, event.
  pe_status launch(Batch& b) {
    std::vector<MidInfo> infos;
    std::vector<int64_t> offsets;
    {
      std::lock_guard<std::mutex> lock(mu);
      for (int64_t t : b.tickets) {
        infos.push_back(requests[t].info);
        offsets.push_back(requests[t].offset);
      }
    }
    MidPrepared p;
    pe_status st = mid_plan_prepare(infos, cfg, &offsets, &p);
    if (st != PE_OK) return st;
    const hipStream_t s = streams[size_t(b.stream)];
    PE_HIP(hipMallocAsync(&b.blob, p.blob_bytes, s), "hipMallocAsync(batch)");
    mid_plan_fill(infos, cfg, p, b.blob, arena, tables, &b.host_blob, &b.plan);
    PE_HIP(hipMemcpyAsync(b.blob, b.host_blob.data(), b.host_blob.size(), hipMemcpyHostToDevice, s),
           "hipMemcpyAsync(batch)");
    st = mid_plan_launch(b.plan, s);
    if (st != PE_OK) return st;
    PE_HIP(hipMemcpyAsync(staging[size_t(b.stream)], b.plan.results, sizeof(pe_halfsupers) * 19 * infos.size(),
                          hipMemcpyDeviceToHost, s), "hipMemcpyAsync(batch results)");
    PE_HIP(hipMemcpyAsync(staging[size_t(b.stream)] + 19 * size_t(max_batch), b.plan.error, sizeof(int32_t),
                          hipMemcpyDeviceToHost, s), "hipMemcpyAsync(batch error)");
    PE_HIP(hipFreeAsync(b.blob, s), "hipFreeAsync(batch)");
    PE_HIP(hipEventRecord(events[size_t(b.stream)], s), "hipEventRecord(batch)");
    return PE_OK;
  }

  void retire(int k) {
    Batch& b = batch_of_stream[size_t(k)];
    const pe_halfsupers* res = staging[size_t(k)];
    int32_t err = 0;
    std::memcpy(&err, res + 19 * size_t(max_batch), sizeof err);
    {
      std::lock_guard<std::mutex> lock(mu);
      const auto now = Clock::now();
      for (size_t i = 0; i < b.tickets.size(); i++) {
        Request& r = requests[b.tickets[i]];
        std::memcpy(r.results, res + 19 * i, sizeof r.results);
        for (int j = r.info.spots + 1; j < 19; j++) std::memset(&r.results[j], 0, sizeof(pe_halfsupers));
        if (err) { r.status = PE_ERR_DEVICE; r.error = "dependency wait exceeded the wait limit"; }
        r.done = true;
        r.done_at = now;
        pool.release(r.offset, r.bytes);
        completed++;
      }
    }
    cv_done.notify_all();
    b = Batch();
    stream_busy[size_t(k)] = 0;
  }

  void run() {
    (void)hipSetDevice(device);
    for (;;) {
      bool progressed = false;
      // retire finished batches
      for (size_t k = 0; k < streams.size(); k++) {
        if (!stream_busy[k]) continue;
        const hipError_t q = hipEventQuery(events[k]);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) {
          fail_batch(batch_of_stream[k], PE_ERR_DEVICE, std::string("batch failed: ") + hipGetErrorString(q));
          batch_of_stream[k] = Batch();
          stream_busy[k] = 0;
        } else {
          retire(int(k));
        }
        progressed = true;
      }
      // admit pending requests onto free streams
      for (size_t k = 0; k < streams.size(); k++) {
        if (stream_busy[k]) continue;
        Batch b;
        b.stream = int(k);
        {
          std::lock_guard<std::mutex> lock(mu);
          const auto now = Clock::now();
          bool head_blocked = false;
          for (auto it = pending.begin(); it != pending.end() && int32_t(b.tickets.size()) < max_batch;) {
            Request& r = requests[*it];
            const int64_t off = head_blocked ? -1 : pool.alloc(r.bytes);
            if (off < 0) {
              // the oldest waiting request blocks further admission once it has waited long
              if (it == pending.begin() && now - r.submitted > kHeadOfLine) head_blocked = true;
              ++it;
              continue;
            }
            r.offset = off;
            b.tickets.push_back(*it);
            it = pending.erase(it);
          }
        }
        if (b.tickets.empty()) break;
        batch_of_stream[k] = std::move(b);
        stream_busy[k] = 1;
        const pe_status st = launch(batch_of_stream[k]);
        if (st != PE_OK) {
          // the stream may hold part of the batch: wait for it before reusing memory
          (void)hipStreamSynchronize(streams[k]);
          fail_batch(batch_of_stream[k], st, "batch launch failed");
          batch_of_stream[k] = Batch();
          stream_busy[k] = 0;
        } else {
          std::lock_guard<std::mutex> lock(mu);
          batches++;
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
  std::string err;
  auto* S = new MidServer;
  auto fail = [&](pe_status st, const std::string& e) { set_error(e); delete S; return st; };
  if (device_tables(&S->tables, &err) != PE_OK) return fail(PE_ERR_DEVICE, err);
  if (hipGetDevice(&S->device) != hipSuccess) return fail(PE_ERR_DEVICE, "hipGetDevice failed");
  int cus = 0, rate_khz = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, S->device);
  (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, S->device);
  int64_t budget = c ? c->workspace_budget : 0;
  if (budget <= 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(PE_ERR_DEVICE, "hipMemGetInfo failed");
    budget = std::min<int64_t>(int64_t(100e9), int64_t(0.4 * double(free_b)));
  }
  const int nstreams = c && c->streams > 0 ? c->streams : 4;
  S->max_batch = c && c->max_batch > 0 ? c->max_batch : 512;
  S->cfg.budget = budget;
  S->cfg.tile = c && c->tile > 0 ? c->tile : 1024;
  S->cfg.workgroups = c && c->workgroups > 0 ? c->workgroups : cus * std::max(1, mid_queue_blocks_per_cu());
  S->cfg.wait_limit = int64_t(10) * 1000 * int64_t(std::max(rate_khz, 1));   // 10 s
  S->cfg.graph = false;
  void* arena = nullptr;
  if (hipMalloc(&arena, size_t(budget)) != hipSuccess) return fail(PE_ERR_DEVICE, "hipMalloc(server workspace) failed");
  S->arena = static_cast<uint8_t*>(arena);
  S->pool = Pool(budget);
  for (int k = 0; k < nstreams; k++) {
    hipStream_t s = nullptr;
    hipEvent_t e = nullptr;
    void* h = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(&h, sizeof(pe_halfsupers) * 19 * size_t(S->max_batch) + 64, hipHostMallocDefault) != hipSuccess)
      return fail(PE_ERR_DEVICE, "server stream setup failed");
    S->streams.push_back(s);
    S->events.push_back(e);
    S->staging.push_back(static_cast<pe_halfsupers*>(h));
  }
  S->stream_busy.assign(size_t(nstreams), 0);
  S->batch_of_stream.resize(size_t(nstreams));
  S->worker = std::thread([S] { S->run(); });
  *out = S;
  return PE_OK;
}

pe_status mid_server_submit(MidServer* S, const pe_high_board& board, int64_t* ticket) {
  Request r;
  r.board = board;
  if (!make_mid_info(board, &r.info)) { set_error("not a valid board with >= 18 stones"); return PE_ERR_INVALID; }
  r.bytes = int64_t(mid_board_bytes(r.info));
  if (r.bytes > S->pool.size()) { set_error("board needs more workspace than the server holds"); return PE_ERR_CAPACITY; }
  {
    std::lock_guard<std::mutex> lock(S->mu);
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
  out->peak_workspace = S->pool.peak();
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
  }
  if (S->arena) (void)hipFree(S->arena);
  delete S;
}

}  // namespace pe
