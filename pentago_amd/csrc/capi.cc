// C ABI entry points (include/pentago_amd.h) and device table management.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "device.h"
#include "end.h"
#include "mid.h"
#include "mid_queue.h"
#include "tables.h"

namespace pe {

pe_status plan_create(const std::vector<LineHost>& lines, EndPlan* plan);
pe_status plan_run(const EndPlan& plan, hipStream_t stream);
void plan_destroy(EndPlan* plan);
pe_status fused_plan_create(int nout, const pe_slab* outs, int nch, const pe_slab* ch, FusedPlan* plan,
                            std::string* err, uint32_t dims = 0xf, int accumulate = 0);
pe_status fused_plan_run(const FusedPlan& plan, hipStream_t stream);
void fused_plan_destroy(FusedPlan* plan);
pe_status slab_line_passes(const pe_slab& out, int nch, const pe_slab* ch, std::vector<std::vector<LineHost>>* passes,
                           int64_t* nodes_out, std::string* err);
pe_status launch_fill_seeded(void* d_nodes, int64_t nodes, uint64_t seed, int64_t first, hipStream_t stream);

pe_status mid_solve_levels(const std::vector<MidInfo>& infos, uint8_t* arena, std::vector<size_t>* result_offsets,
                           hipStream_t stream, DeviceBuffer* desc_cache);
int mid_queue_blocks_per_cu();
pe_status mid_solve_noalloc(const pe_high_board& board, pe_halfsupers* results, hipStream_t stream);
int64_t mid_noalloc_scratch_bytes();

namespace {
thread_local std::string g_error;
std::mutex g_dev_mutex;
constexpr int kMaxDevices = 64;
DevTables g_dev[kMaxDevices];
bool g_dev_ready[kMaxDevices];
// launch form of new end-engine plans per device (pe_end_plan_autotune's choice; 0: grid)
std::atomic<int> g_end_form[kMaxDevices];
}  // namespace

void set_error(const std::string& e) { g_error = e; }

pe_status device_tables(DevTables* out, std::string* err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) { *err = "no HIP device"; return PE_ERR_DEVICE; }
  std::lock_guard<std::mutex> lock(g_dev_mutex);
  if (g_dev_ready[dev]) { *out = g_dev[dev]; return PE_OK; }
  const Tables* T;
  try {
    T = &tables();
  } catch (const std::exception& e) {
    *err = e.what();
    return PE_ERR_TABLES;
  }
  // unpacked rotation-minimal quadrants and choose(n, k)
  std::vector<uint32_t> sides(T->rmin_flat_size);
  for (int i = 0; i < T->rmin_flat_size; i++)
    sides[i] = uint32_t(T->unpack[T->rmin_flat[i]][0]) | uint32_t(T->unpack[T->rmin_flat[i]][1]) << 16;
  std::vector<int32_t> choose(32 * 32);
  for (int n = 0; n < 32; n++)
    for (int k = 0; k < 32; k++) choose[n * 32 + k] = choose_host(n, k);
  // log2 of the number of local rotations fixing a quadrant: 2 if a quarter turn fixes both
  // sides, 1 if only a half turn does, else 0 (pentago/base/count.cc:235-245)
  std::vector<uint8_t> stab(T->rmin_flat_size);
  for (int i = 0; i < T->rmin_flat_size; i++) {
    const int s0 = T->unpack[T->rmin_flat[i]][0], s1 = T->unpack[T->rmin_flat[i]][1];
    const int r0 = T->rot[s0][0], r1 = T->rot[s1][0];
    stab[i] = (r0 == s0 && r1 == s1) ? 2 : (T->rot[r0][0] == s0 && T->rot[r1][0] == s1) ? 1 : 0;
  }
  // field pool of the fused kernels (see end.h)
  std::vector<uint64_t> pool;
  build_field_pool(&pool);
  const size_t sizes[] = {sizeof T->superwin, sizeof(uint16_t) * size_t(T->rmin_flat_size),
                          sizeof(uint32_t) * sides.size(), sizeof T->rmin_inverse, sizeof T->pack,
                          sizeof T->commute, sizeof(int32_t) * choose.size(), stab.size(),
                          sizeof(uint64_t) * pool.size()};
  const void* srcs[] = {T->superwin, T->rmin_flat, sides.data(), T->rmin_inverse, T->pack, T->commute, choose.data(),
                        stab.data(), pool.data()};
  constexpr int kParts = 9;
  size_t total = 0, offs[kParts];
  for (int i = 0; i < kParts; i++) { offs[i] = total; total += (sizes[i] + 255) & ~size_t(255); }
  void* d = nullptr;
  if (hipMalloc(&d, total) != hipSuccess) { *err = "hipMalloc(tables) failed"; return PE_ERR_DEVICE; }
  for (int i = 0; i < kParts; i++)
    if (hipMemcpy(static_cast<uint8_t*>(d) + offs[i], srcs[i], sizes[i], hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(d);
      *err = "hipMemcpy(tables) failed";
      return PE_ERR_DEVICE;
    }
  auto* b = static_cast<uint8_t*>(d);
  DevTables D;
  D.superwin = reinterpret_cast<const uint64_t*>(b + offs[0]);
  D.rmin_flat = reinterpret_cast<const uint16_t*>(b + offs[1]);
  D.rmin_sides = reinterpret_cast<const uint32_t*>(b + offs[2]);
  D.rmin_inverse = reinterpret_cast<const uint16_t*>(b + offs[3]);
  D.pack = reinterpret_cast<const uint16_t*>(b + offs[4]);
  D.commute = reinterpret_cast<const uint8_t*>(b + offs[5]);
  D.choose = reinterpret_cast<const int32_t*>(b + offs[6]);
  D.rmin_stab = b + offs[7];
  D.fields = reinterpret_cast<const uint64_t*>(b + offs[8]);
  g_dev[dev] = D;
  g_dev_ready[dev] = true;
  *out = D;
  return PE_OK;
}

bool device_tables_ready(DevTables* out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return false;
  std::lock_guard<std::mutex> lock(g_dev_mutex);
  if (!g_dev_ready[dev]) return false;
  *out = g_dev[dev];
  return true;
}

namespace {

pe_status fail(pe_status s, const std::string& e) { g_error = e; return s; }

struct WakeupCtx { void (*fn)(void*); void* ctx; };
void wakeup_trampoline(void* p) {
  auto* w = static_cast<WakeupCtx*>(p);
  w->fn(w->ctx);
  delete w;
}

}  // namespace
}  // namespace pe

using namespace pe;

extern "C" {

pe_status pe_init(void) {
  PE_GUARD_BEGIN
  try {
    tables();
  } catch (const std::exception& e) {
    return fail(PE_ERR_TABLES, e.what());
  }
  DevTables D;
  std::string err;
  const pe_status s = device_tables(&D, &err);
  if (s != PE_OK) return fail(s, err);
  return PE_OK;
  PE_GUARD_END
}

const char* pe_last_error(void) { return g_error.c_str(); }

pe_status pe_table_digest(const char* name, char out[41]) {
  PE_GUARD_BEGIN
  try {
    const std::string d = table_digest(name);
    if (d.empty()) return fail(PE_ERR_INVALID, std::string("unknown table ") + name);
    std::memcpy(out, d.c_str(), 41);
    return PE_OK;
  } catch (const std::exception& e) {
    return fail(PE_ERR_TABLES, e.what());
  }
  PE_GUARD_END
}

pe_status pe_section_shape(const pe_section* section, int32_t shape[4]) {
  PE_GUARD_BEGIN
  if (!section || !shape) return fail(PE_ERR_INVALID, "null argument");
  Section s;
  std::memcpy(s.c, section->counts, sizeof s.c);
  int sh[4];
  if (!s.valid_counts() || !s.shape(sh)) return fail(PE_ERR_INVALID, "invalid section");
  for (int q = 0; q < 4; q++) shape[q] = sh[q];
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_section_standardize(const pe_section* section, pe_section* standard, int32_t* symmetry) {
  PE_GUARD_BEGIN
  if (!section || !standard) return fail(PE_ERR_INVALID, "null argument");
  Section s, t;
  std::memcpy(s.c, section->counts, sizeof s.c);
  if (!s.valid_counts()) return fail(PE_ERR_INVALID, "invalid section");
  const int g = s.standardize(&t);
  std::memcpy(standard->counts, t.c, sizeof t.c);
  if (symmetry) *symmetry = g;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_line_describe(const pe_line* line, pe_line_info* info) {
  PE_GUARD_BEGIN
  if (!line || !info) return fail(PE_ERR_INVALID, "null argument");
  LineHost h;
  std::string err;
  const pe_status s = make_line(*line, &h, &err);
  if (s != PE_OK) return fail(s, err);
  *info = h.info;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_end_compute_lines(int n, const pe_line* lines, const void* const* d_inputs, void* const* d_outputs,
                               void* stream) {
  PE_GUARD_BEGIN
  if (n < 0 || (n && (!lines || !d_inputs || !d_outputs))) return fail(PE_ERR_INVALID, "bad line batch");
  std::vector<pe_line_job> jobs(n);
  for (int k = 0; k < n; k++) {
    std::memset(&jobs[k], 0, sizeof jobs[k]);
    jobs[k].line = lines[k];
    jobs[k].input = d_inputs[k];
    jobs[k].output = d_outputs[k];
  }
  pe_end_plan* plan = nullptr;
  pe_status s = pe_end_plan_create(n, jobs.data(), &plan);
  if (s != PE_OK) return s;
  s = pe_end_plan_run(plan, stream);
  // the plan's device memory is freed once this launch has finished
  if (s == PE_OK && hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
    s = fail(PE_ERR_DEVICE, "line compute failed");
  pe_end_plan_destroy(plan);
  return s;
  PE_GUARD_END
}

struct pe_end_plan {
  FusedPlan fused;               // fused block plan (run first when present)
  FusedPlan fused35;             // fused plan of 35-stone output sections (no child data)
  std::vector<EndPlan> passes;   // then line plans, in order
  int64_t nodes = 0;
  bool finishing = false;        // reads a partial and replaces it by the result: not repeatable
};

namespace {
void apply_default_form(pe_end_plan* plan) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return;
  plan->fused.queue_wgs_per_cu = plan->fused35.queue_wgs_per_cu = g_end_form[dev].load();
}
}  // namespace

pe_status pe_end_plan_create(int n, const pe_line_job* jobs, pe_end_plan** out) {
  PE_GUARD_BEGIN
  if (n < 0 || !out || (n && !jobs)) return fail(PE_ERR_INVALID, "bad plan arguments");
  std::vector<LineHost> hosts(n);
  for (int k = 0; k < n; k++) {
    const pe_line_job& J = jobs[k];
    std::string err;
    const pe_status s = make_line(J.line, &hosts[k], &err);
    if (s != PE_OK) return fail(s, "job " + std::to_string(k) + ": " + err);
    LineHost& h = hosts[k];
    if (!J.output) return fail(PE_ERR_INVALID, "null output for job " + std::to_string(k));
    if (!h.dev.slice35 && !J.input) return fail(PE_ERR_INVALID, "null input for job " + std::to_string(k));
    h.dev.input = static_cast<const uint64_t*>(J.input);
    h.dev.output = static_cast<uint64_t*>(J.output);
    const int size = h.dev.rmin_size[h.dev.dim];
    if (J.i_begin || J.i_end) {
      if (J.i_begin < 0 || J.i_end > size || J.i_begin >= J.i_end)
        return fail(PE_ERR_INVALID, "bad index range for job " + std::to_string(k));
      h.dev.i_begin = J.i_begin;
      h.dev.i_end = J.i_end;
    }
    h.dev.accumulate = J.accumulate ? 1 : 0;
    if (J.input_block_offsets && !h.dev.slice35)
      h.in_off.assign(J.input_block_offsets, J.input_block_offsets + h.info.input_blocks);
    if (J.output_block_offsets) h.out_off.assign(J.output_block_offsets, J.output_block_offsets + h.info.length);
  }
  // jobs covering whole lines run on the fused kernel (one dimension open per block);
  // partial index ranges, slice 35 and irregular child offsets on the line kernel
  // jobs covering whole lines run on the fused kernel (one dimension open per block; 35-stone
  // lines on its no-children form); partial index ranges and irregular child offsets on the
  // line kernel
  std::vector<LineHost> fused, fused35, lined;
  for (auto& h : hosts) (!fused_line_eligible(h) ? lined : h.dev.slice35 ? fused35 : fused).push_back(std::move(h));
  auto* plan = new pe_end_plan();
  std::string err;
  pe_status s = PE_OK;
  if (!fused.empty()) {
    s = fused_line_plan_create(fused, &plan->fused, &err);
    plan->nodes += plan->fused.output_nodes;
  }
  if (s == PE_OK && !fused35.empty()) {
    s = fused_line_plan_create(fused35, &plan->fused35, &err);
    plan->nodes += plan->fused35.output_nodes;
  }
  if (s == PE_OK && !lined.empty()) {
    plan->passes.resize(1);
    s = plan_create(lined, &plan->passes[0]);
    if (s != PE_OK) err = pe_last_error();
    else plan->nodes += plan->passes[0].output_nodes;
  }
  if (s != PE_OK) {
    for (auto& p : plan->passes) plan_destroy(&p);
    fused_plan_destroy(&plan->fused);
    fused_plan_destroy(&plan->fused35);
    delete plan;
    return fail(s, err);
  }
  apply_default_form(plan);
  *out = plan;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_end_slab_plan_create(const pe_slab* output, int nchildren, const pe_slab* children, int mode,
                                  pe_end_plan** out) {
  PE_GUARD_BEGIN
  return pe_end_slabs_plan_create(1, output, nchildren, children, mode, out);
  PE_GUARD_END
}

pe_status pe_end_slabs_plan_create(int noutputs, const pe_slab* outputs, int nchildren, const pe_slab* children,
                                   int mode, pe_end_plan** out) {
  PE_GUARD_BEGIN
  if (noutputs <= 0 || !outputs || !out || nchildren < 0 || (nchildren && !children))
    return fail(PE_ERR_INVALID, "bad slab plan arguments");
  for (int k = 0; k < noutputs; k++)
    if (!outputs[k].data) return fail(PE_ERR_INVALID, "output slab without memory");
  if (mode != 0 && mode != 1) return fail(PE_ERR_INVALID, "unknown slab plan mode");
  auto* plan = new pe_end_plan();
  std::string err;
  pe_status s = PE_OK;
  // mode 0: every section goes to the fused kernel, 35-stone sections (one open dimension,
  // no child data) in a plan of their own; mode 1: line passes
  std::vector<pe_slab> fused, fused35, lined;
  for (int k = 0; k < noutputs; k++) {
    int sum = 0;
    for (int q = 0; q < 4; q++) sum += outputs[k].section.counts[q][0] + outputs[k].section.counts[q][1];
    (mode == 1 ? lined : sum == 35 ? fused35 : fused).push_back(outputs[k]);
  }
  if (!fused.empty()) {
    s = fused_plan_create(int(fused.size()), fused.data(), nchildren, children, &plan->fused, &err);
    plan->nodes = plan->fused.output_nodes;
  }
  if (!fused35.empty() && s == PE_OK) {
    s = fused_plan_create(int(fused35.size()), fused35.data(), nchildren, children, &plan->fused35, &err);
    plan->nodes += plan->fused35.output_nodes;
  }
  // line passes: pass p of every such slab in one launch
  std::vector<std::vector<LineHost>> groups;
  for (size_t k = 0; k < lined.size() && s == PE_OK; k++) {
    int64_t nodes = 0;
    s = slab_line_passes(lined[k], nchildren, children, &groups, &nodes, &err);
    plan->nodes += nodes;
  }
  for (size_t p = 0; p < groups.size() && s == PE_OK; p++) {
    plan->passes.emplace_back();
    s = plan_create(groups[p], &plan->passes.back());
    if (s != PE_OK) err = pe_last_error();
  }
  if (s != PE_OK) {
    for (auto& p : plan->passes) plan_destroy(&p);
    fused_plan_destroy(&plan->fused);
    fused_plan_destroy(&plan->fused35);
    delete plan;
    return fail(s, err.empty() ? pe_last_error() : err);
  }
  apply_default_form(plan);
  *out = plan;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_end_slabs_plan_create_dims(int noutputs, const pe_slab* outputs, int nchildren, const pe_slab* children,
                                        uint32_t dims, int accumulate, pe_end_plan** out) {
  PE_GUARD_BEGIN
  if (noutputs <= 0 || !outputs || !out || nchildren < 0 || (nchildren && !children) || !(dims & 0xf) || dims > 0xf)
    return fail(PE_ERR_INVALID, "bad slab plan arguments");
  if (accumulate < 0 || accumulate > 4) return fail(PE_ERR_INVALID, "accumulate must be 0..4");
  for (int k = 0; k < noutputs; k++) {
    if (!outputs[k].data) return fail(PE_ERR_INVALID, "output slab without memory");
    int sum = 0;
    for (int q = 0; q < 4; q++) sum += outputs[k].section.counts[q][0] + outputs[k].section.counts[q][1];
    if (sum >= 35) return fail(PE_ERR_INVALID, "dimension subsets need sections with at most 34 stones");
  }
  auto* plan = new pe_end_plan();
  std::string err;
  const pe_status s = fused_plan_create(noutputs, outputs, nchildren, children, &plan->fused, &err, dims, accumulate);
  if (s != PE_OK) {
    fused_plan_destroy(&plan->fused);
    delete plan;
    return fail(s, err);
  }
  plan->nodes = plan->fused.output_nodes;
  plan->finishing = accumulate == 3;
  apply_default_form(plan);
  *out = plan;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_end_plan_run(const pe_end_plan* plan, void* stream) {
  PE_GUARD_BEGIN
  if (!plan) return fail(PE_ERR_INVALID, "null plan");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  for (const FusedPlan* f : {&plan->fused, &plan->fused35})
    if (f->nblocks) {
      const pe_status s = fused_plan_run(*f, st);
      if (s != PE_OK) return s;
    }
  for (const auto& p : plan->passes) {
    const pe_status s = plan_run(p, st);
    if (s != PE_OK) return s;
  }
  return PE_OK;
  PE_GUARD_END
}

int64_t pe_end_plan_nodes(const pe_end_plan* plan) { return plan ? plan->nodes : -1; }

pe_status pe_end_plan_set_launch(pe_end_plan* plan, int queue_wgs_per_cu) {
  PE_GUARD_BEGIN
  if (!plan || queue_wgs_per_cu < 0 || queue_wgs_per_cu > 16) return fail(PE_ERR_INVALID, "bad launch form");
  plan->fused.queue_wgs_per_cu = plan->fused35.queue_wgs_per_cu = queue_wgs_per_cu;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_end_set_default_launch(int queue_wgs_per_cu) {
  PE_GUARD_BEGIN
  int dev = 0;
  if (queue_wgs_per_cu < 0 || queue_wgs_per_cu > 16) return fail(PE_ERR_INVALID, "bad launch form");
  PE_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(PE_ERR_DEVICE, "device index out of range");
  g_end_form[dev] = queue_wgs_per_cu;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_end_plan_autotune(pe_end_plan* plan, void* stream, pe_end_tuning* out) {
  PE_GUARD_BEGIN
  if (!plan) return fail(PE_ERR_INVALID, "null plan");
  if (plan->finishing) return fail(PE_ERR_INVALID, "a finishing pass consumes its partial: it cannot be run repeatedly to time it");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int32_t forms[] = {0, 2, 3, 4, 5, 6};
  constexpr int kForms = 6;
  pe_end_tuning t;
  std::memset(&t, 0, sizeof t);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PE_HIP(hipEventCreate(&e0), "hipEventCreate");
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return fail(PE_ERR_DEVICE, "hipEventCreate"); }
  pe_status st = PE_OK;
  float best = -1;
  for (int c = 0; c < kForms && st == PE_OK; c++) {
    plan->fused.queue_wgs_per_cu = plan->fused35.queue_wgs_per_cu = forms[c];
    st = pe_end_plan_run(plan, stream);   // warm-up
    float ms = -1;
    for (int rep = 0; rep < 2 && st == PE_OK; rep++) {   // the faster of two timed runs
      float x = 0;
      if (hipEventRecord(e0, s) != hipSuccess) st = fail(PE_ERR_DEVICE, "hipEventRecord");
      if (st == PE_OK) st = pe_end_plan_run(plan, stream);
      if (st == PE_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                          hipEventElapsedTime(&x, e0, e1) != hipSuccess))
        st = fail(PE_ERR_DEVICE, "autotune timing failed");
      if (st == PE_OK && (ms < 0 || x < ms)) ms = x;
    }
    t.forms[c] = forms[c];
    t.ms[c] = ms;
    if (st == PE_OK && (best < 0 || ms < best)) {
      best = ms;
      t.queue_wgs_per_cu = forms[c];
    }
  }
  t.candidates = kForms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  plan->fused.queue_wgs_per_cu = plan->fused35.queue_wgs_per_cu = t.queue_wgs_per_cu;
  if (st != PE_OK) return st;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kMaxDevices) g_end_form[dev] = t.queue_wgs_per_cu;
  if (out) *out = t;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_end_plan_destroy(pe_end_plan* plan) {
  PE_GUARD_BEGIN
  if (!plan) return PE_OK;
  for (auto& p : plan->passes) plan_destroy(&p);
  fused_plan_destroy(&plan->fused);
  fused_plan_destroy(&plan->fused35);
  delete plan;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_end_compute_line(const pe_line* line, const void* d_input, void* d_output, void* stream) {
  PE_GUARD_BEGIN
  return pe_end_schedule_line(line, d_input, d_output, stream, nullptr, nullptr);
  PE_GUARD_END
}

pe_status pe_end_schedule_line(const pe_line* line, const void* d_input, void* d_output, void* stream,
                               void (*wakeup)(void*), void* ctx) {
  PE_GUARD_BEGIN
  if (!line || !d_output) return fail(PE_ERR_INVALID, "null argument");
  LineHost h;
  std::string err;
  pe_status s = make_line(*line, &h, &err);
  if (s != PE_OK) return fail(s, err);
  if (!h.dev.slice35 && !d_input) return fail(PE_ERR_INVALID, "null input for a line with child data");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  s = fused_line_launch(h, d_input, d_output, 0, st, &err);
  if (s != PE_OK) return fail(s, err);
  if (!wakeup) return PE_OK;
  auto* w = new WakeupCtx{wakeup, ctx};
  if (hipLaunchHostFunc(st, wakeup_trampoline, w) != hipSuccess) {
    delete w;
    return fail(PE_ERR_DEVICE, "hipLaunchHostFunc failed");
  }
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_end_compute_line_host(const pe_line* line, const void* h_input, void* h_output) {
  PE_GUARD_BEGIN
  if (!line || !h_output) return fail(PE_ERR_INVALID, "null argument");
  LineHost h;
  std::string err;
  pe_status s = make_line(*line, &h, &err);
  if (s != PE_OK) return fail(s, err);
  const size_t in_bytes = size_t(h.info.input_nodes) * 64, out_bytes = size_t(h.info.output_nodes) * 64;
  if (in_bytes && !h_input) return fail(PE_ERR_INVALID, "null input");
  void *din = nullptr, *dout = nullptr;
  struct Free {   // the device buffers are released on every way out
    void*& a; void*& b;
    ~Free() { if (a) (void)hipFree(a); if (b) (void)hipFree(b); }
  } free_bufs{din, dout};
  PE_HIP(hipMalloc(&dout, out_bytes), "hipMalloc(output)");
  if (in_bytes) {
    PE_HIP(hipMalloc(&din, in_bytes), "hipMalloc(input)");
    PE_HIP(hipMemcpy(din, h_input, in_bytes, hipMemcpyHostToDevice), "hipMemcpy(input)");
  }
  PE_HIP(hipMemset(dout, 0, out_bytes), "hipMemset(output)");
  s = pe_end_compute_line(line, din, dout, nullptr);
  if (s != PE_OK) return s;
  PE_HIP(hipStreamSynchronize(nullptr), "line compute");
  PE_HIP(hipMemcpy(h_output, dout, out_bytes, hipMemcpyDeviceToHost), "hipMemcpy(output)");
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_fill_seeded(void* d_nodes, int64_t nodes, uint64_t seed, int64_t first_node, void* stream) {
  PE_GUARD_BEGIN
  if (nodes < 0 || (nodes && !d_nodes)) return fail(PE_ERR_INVALID, "bad fill");
  return launch_fill_seeded(d_nodes, nodes, seed, first_node, static_cast<hipStream_t>(stream));
  PE_GUARD_END
}

int64_t pe_midsolve_workspace_size(int min_slice) {
  if (min_slice < 18 || min_slice > 36) return -1;
  return bottleneck(36 - min_slice);
}

}  // extern "C"

namespace {

struct TunedMid { bool set = false; int32_t small_spots = 0, threads = 0, workgroups = 0, use_graph = 0; };
TunedMid g_tuned[kMaxDevices];

pe_status mid_infos(int n, const pe_high_board* boards, std::vector<MidInfo>* infos) {
  infos->resize(n);
  for (int b = 0; b < n; b++)
    if (!make_mid_info(boards[b], &(*infos)[b]))
      return fail(PE_ERR_INVALID, "board " + std::to_string(b) + " is not a valid board with >= 18 stones");
  return PE_OK;
}

// Fill the zero fields of a configuration with the current device's defaults.
pe_status mid_resolve_config(const pe_mid_config* in, MidConfig* out) {
  int dev = 0;
  PE_HIP(hipGetDevice(&dev), "hipGetDevice");
  int cus = 0;
  PE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute(CUs)");
  pe_mid_config c{};
  if (in) c = *in;
  if (c.workspace_budget <= 0) {
    size_t free_b = 0, total_b = 0;
    PE_HIP(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    c.workspace_budget = std::min<int64_t>(int64_t(200e9), int64_t(0.6 * double(free_b)));
  }
  const TunedMid& t = g_tuned[dev];
  if (c.small_spots <= 0) c.small_spots = t.set ? t.small_spots : 14;
  if (c.threads <= 0) c.threads = t.set ? t.threads : 512;
  if (c.threads != 256 && c.threads != 512 && c.threads != 1024) return fail(PE_ERR_INVALID, "threads must be 256, 512 or 1024");
  // the tuned workgroup count belongs to the tuned workgroup size; another size given by the
  // caller gets a full device of its own
  if (c.workgroups <= 0)
    c.workgroups = t.set && c.threads == t.threads ? t.workgroups : cus * mid_board_blocks_per_cu(c.threads);
  if (c.streams < 0) c.streams = 0;
  else if (c.streams == 0) c.streams = 2;
  out->budget = c.workspace_budget;
  out->small_spots = c.small_spots;
  out->workgroups = c.workgroups;
  out->threads = c.threads;
  out->streams = c.streams;
  out->graph = c.use_graph > 0 || (c.use_graph == 0 && t.set && t.use_graph);
  return PE_OK;
}

pe_status mid_run_sync(MidPlan& plan, pe_halfsupers (*results)[19], hipStream_t s) {
  pe_status st = mid_plan_run(plan, s);
  if (st == PE_OK) st = mid_plan_copy_results(plan, results, s);
  PE_HIP(hipStreamSynchronize(s), "midsolve plan");
  if (st == PE_OK) mid_plan_zero_tail(plan, results);
  return st;
}

}  // namespace

struct pe_mid_plan {
  MidPlan plan;
};

extern "C" {

pe_status pe_mid_plan_create(int n, const pe_high_board* boards, const pe_mid_config* config, pe_mid_plan** plan) {
  PE_GUARD_BEGIN
  if (n < 0 || (n && !boards) || !plan) return fail(PE_ERR_INVALID, "bad plan arguments");
  *plan = nullptr;
  std::vector<MidInfo> infos;
  pe_status st = mid_infos(n, boards, &infos);
  if (st != PE_OK) return st;
  MidConfig cfg;
  st = mid_resolve_config(config, &cfg);
  if (st != PE_OK) return st;
  auto* p = new pe_mid_plan;
  st = mid_plan_create(infos, cfg, &p->plan);
  if (st != PE_OK) {
    mid_plan_destroy(&p->plan);
    delete p;
    return st;
  }
  *plan = p;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_mid_schedule_info(int n, const pe_high_board* boards, const pe_mid_config* config, int64_t* launches,
                               int64_t* workspace_bytes) {
  PE_GUARD_BEGIN
  if (n < 0 || (n && !boards) || !config || config->workspace_budget <= 0 || config->small_spots <= 0 ||
      config->workgroups <= 0 || config->threads <= 0)
    return fail(PE_ERR_INVALID, "schedule info needs a board list and a complete configuration");
  std::vector<MidInfo> infos;
  pe_status st = mid_infos(n, boards, &infos);
  if (st != PE_OK) return st;
  MidConfig cfg;
  cfg.budget = config->workspace_budget;
  cfg.small_spots = config->small_spots;
  cfg.workgroups = config->workgroups;
  cfg.threads = config->threads;
  cfg.streams = std::max(config->streams, 0);
  cfg.graph = false;
  MidLayout L;
  if (const char* e = mid_layout(infos, cfg, &L)) return fail(PE_ERR_CAPACITY, e);
  int64_t l = L.small.empty() ? 0 : 3;
  for (int b : L.big) l += 2 + infos[size_t(b)].spots;
  if (launches) *launches = l;
  if (workspace_bytes) *workspace_bytes = int64_t(L.bytes);
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_mid_plan_run(pe_mid_plan* plan, void* stream) {
  PE_GUARD_BEGIN
  if (!plan) return fail(PE_ERR_INVALID, "null plan");
  return mid_plan_run(plan->plan, static_cast<hipStream_t>(stream));
  PE_GUARD_END
}

pe_status pe_mid_plan_results(pe_mid_plan* plan, pe_halfsupers (*results)[19], void* stream) {
  PE_GUARD_BEGIN
  if (!plan || (plan->plan.nb && !results)) return fail(PE_ERR_INVALID, "null argument");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  pe_status st = mid_plan_copy_results(plan->plan, results, s);
  PE_HIP(hipStreamSynchronize(s), "mid plan results");
  if (st == PE_OK) mid_plan_zero_tail(plan->plan, results);
  return st;
  PE_GUARD_END
}

pe_status pe_mid_plan_info(const pe_mid_plan* plan, pe_mid_plan_stats* info, pe_mid_config* config) {
  PE_GUARD_BEGIN
  if (!plan) return fail(PE_ERR_INVALID, "null plan");
  const MidPlan& P = plan->plan;
  if (info) {
    std::memset(info, 0, sizeof *info);
    info->small_boards = int64_t(P.layout.small.size());
    info->large_boards = int64_t(P.layout.big.size());
    info->launches = P.launches;
    info->workspace_bytes = int64_t(P.layout.bytes);
    info->workgroups = P.layout.workgroups;
    info->streams = P.layout.streams;
  }
  if (config) {
    std::memset(config, 0, sizeof *config);
    config->workspace_budget = P.cfg.budget;
    config->small_spots = P.cfg.small_spots;
    config->workgroups = P.cfg.workgroups;
    config->threads = P.cfg.threads;
    config->streams = P.cfg.streams;
    config->use_graph = P.cfg.graph;
  }
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_mid_plan_destroy(pe_mid_plan* plan) {
  PE_GUARD_BEGIN
  if (!plan) return PE_OK;
  mid_plan_destroy(&plan->plan);
  delete plan;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_mid_autotune(int n, const pe_high_board* boards, int64_t workspace_budget, pe_mid_config* best) {
  PE_GUARD_BEGIN
  if (n <= 0 || !boards || !best) return fail(PE_ERR_INVALID, "bad autotune arguments");
  std::vector<MidInfo> infos;
  pe_status st = mid_infos(n, boards, &infos);
  if (st != PE_OK) return st;
  int dev = 0, cus = 0;
  PE_HIP(hipGetDevice(&dev), "hipGetDevice");
  PE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute(CUs)");
  hipStream_t s = nullptr;
  PE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // one untimed run and the faster of two timed runs of the whole batch per candidate (a new
  // plan's first run touches freshly allocated workspace, and a graph candidate captures in
  // it; candidates differ by a few percent, about what one run varies by), so the boards
  // handed in can be the real batch: the best split between the persistent small-board
  // kernel and the large boards' level launches depends on how many boards of each kind
  // there are
  auto trial = [&](const pe_mid_config& c, float* ms) -> pe_status {
    MidConfig cfg;
    pe_status r = mid_resolve_config(&c, &cfg);
    if (r != PE_OK) return r;
    MidPlan plan;
    r = mid_plan_create(infos, cfg, &plan);
    if (r == PE_OK) r = mid_plan_run(plan, s);   // untimed
    float best_t = -1;
    for (int rep = 0; rep < 2 && r == PE_OK; rep++) {
      float t = 0;
      if (hipEventRecord(e0, s) != hipSuccess) r = fail(PE_ERR_DEVICE, "autotune timing failed");
      if (r == PE_OK) r = mid_plan_run(plan, s);
      if (r == PE_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        r = fail(PE_ERR_DEVICE, "autotune timing failed");
      if (r == PE_OK && (best_t < 0 || t < best_t)) best_t = t;
    }
    (void)hipStreamSynchronize(s);
    mid_plan_destroy(&plan);
    *ms = best_t;
    return r;
  };
  auto config = [&](int32_t sm, int32_t th, int32_t graph) {
    pe_mid_config c{};
    c.workspace_budget = workspace_budget;
    c.small_spots = sm;
    c.threads = th;
    c.workgroups = cus * mid_board_blocks_per_cu(th);
    c.use_graph = graph;
    return c;
  };
  // stage 1: the small/large split at 512 threads; stage 2: threads per small-board
  // workgroup at that split; stage 3: the winner replayed as a captured graph
  float best_ms = -1;
  pe_mid_config chosen{};
  for (int32_t sm : {12, 14, 16}) {
    float ms = 0;
    const pe_mid_config c = config(sm, 512, -1);
    if ((st = trial(c, &ms)) != PE_OK) break;
    if (best_ms < 0 || ms < best_ms) { best_ms = ms; chosen = c; }
  }
  if (st == PE_OK)
    for (int32_t th : {256, 1024}) {
      float ms = 0;
      const pe_mid_config c = config(chosen.small_spots, th, -1);
      if ((st = trial(c, &ms)) != PE_OK) break;
      if (ms < best_ms) { best_ms = ms; chosen = c; }
    }
  if (st == PE_OK) {
    float ms = 0;
    pe_mid_config c = chosen;
    c.use_graph = 1;
    if ((st = trial(c, &ms)) == PE_OK && ms < best_ms) chosen.use_graph = 1;
  }
  if (st == PE_OK) {
    MidConfig cfg;
    if ((st = mid_resolve_config(&chosen, &cfg)) == PE_OK) chosen.streams = cfg.streams;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  if (st != PE_OK) return st;
  g_tuned[dev].set = true;
  g_tuned[dev].small_spots = chosen.small_spots;
  g_tuned[dev].threads = chosen.threads;
  g_tuned[dev].workgroups = chosen.workgroups;
  g_tuned[dev].use_graph = chosen.use_graph > 0;
  *best = chosen;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_mid_server_create(const pe_mid_server_config* config, pe_mid_server** server) {
  PE_GUARD_BEGIN
  if (!server) return fail(PE_ERR_INVALID, "null argument");
  MidServer* S = nullptr;
  const pe_status st = mid_server_create(config, &S);
  *server = reinterpret_cast<pe_mid_server*>(S);
  return st;
  PE_GUARD_END
}

pe_status pe_mid_server_submit(pe_mid_server* server, const pe_high_board* board, int64_t* ticket) {
  PE_GUARD_BEGIN
  if (!server || !board || !ticket) return fail(PE_ERR_INVALID, "null argument");
  return mid_server_submit(reinterpret_cast<MidServer*>(server), *board, ticket);
  PE_GUARD_END
}

pe_status pe_mid_server_poll(pe_mid_server* server, int64_t ticket, pe_halfsupers results[19], int* done,
                             double* latency_ms) {
  PE_GUARD_BEGIN
  if (!server || !done) return fail(PE_ERR_INVALID, "null argument");
  return mid_server_poll(reinterpret_cast<MidServer*>(server), ticket, results, done, latency_ms);
  PE_GUARD_END
}

pe_status pe_mid_server_wait(pe_mid_server* server, int64_t ticket, pe_halfsupers results[19], double* latency_ms) {
  PE_GUARD_BEGIN
  if (!server) return fail(PE_ERR_INVALID, "null argument");
  return mid_server_wait(reinterpret_cast<MidServer*>(server), ticket, results, latency_ms);
  PE_GUARD_END
}

pe_status pe_mid_server_stats(pe_mid_server* server, pe_mid_server_info* stats) {
  PE_GUARD_BEGIN
  if (!server || !stats) return fail(PE_ERR_INVALID, "null argument");
  mid_server_stats(reinterpret_cast<MidServer*>(server), stats);
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_mid_server_destroy(pe_mid_server* server) {
  PE_GUARD_BEGIN
  if (server) mid_server_destroy(reinterpret_cast<MidServer*>(server));
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_midsolve_batch(int n, const pe_high_board* boards, pe_halfsupers (*results)[19], int mode,
                            void* stream) {
  PE_GUARD_BEGIN
  if (n < 0 || (n && (!boards || !results))) return fail(PE_ERR_INVALID, "bad batch");
  if (mode != 0 && mode != 1) return fail(PE_ERR_INVALID, "unsupported batch mode");
  if (mode == 1) {
    std::vector<MidInfo> infos;
    pe_status st = mid_infos(n, boards, &infos);
    if (st != PE_OK) return st;
    MidConfig cfg;
    st = mid_resolve_config(nullptr, &cfg);
    if (st != PE_OK) return st;
    MidPlan plan;
    st = mid_plan_create(infos, cfg, &plan);
    if (st == PE_OK) st = mid_run_sync(plan, results, static_cast<hipStream_t>(stream));
    mid_plan_destroy(&plan);
    return st;
  }
  std::vector<MidInfo> infos(n);
  size_t bytes = 0;
  for (int b = 0; b < n; b++) {
    if (!make_mid_info(boards[b], &infos[b])) return fail(PE_ERR_INVALID, "board " + std::to_string(b) + " is not a valid board with >= 18 stones");
    bytes += mid_board_bytes(infos[b]);
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  // per-device scratch kept between calls (a caller's workspace, as midsolve's
  // RawArray<halfsuper_s> workspace argument is), so a call costs launches, not allocations
  int dev = 0;
  PE_HIP(hipGetDevice(&dev), "hipGetDevice");
  static std::mutex cache_mu[kMaxDevices];
  static DeviceBuffer arena_cache[kMaxDevices], desc_cache[kMaxDevices];
  std::lock_guard<std::mutex> lock(cache_mu[dev]);
  PE_HIP(hipStreamSynchronize(s), "midsolve batch (stream idle)");
  // large batches get a one-off arena so that the cache stays small (<= 8 GB)
  const bool cached = bytes <= (size_t(8) << 30);
  DeviceBuffer once;
  struct FreeOnce { DeviceBuffer& b; ~FreeOnce() { if (b.ptr) (void)hipFree(b.ptr); } } free_once{once};
  DeviceBuffer& ab = cached ? arena_cache[dev] : once;
  PE_HIP(ab.reserve(std::max<size_t>(bytes, 256)), "hipMalloc(mid arena)");
  void* arena = ab.ptr;
  std::vector<size_t> offs;
  pe_status st = mid_solve_levels(infos, static_cast<uint8_t*>(arena), &offs, s, &desc_cache[dev]);
  if (st == PE_OK) {
    for (int b = 0; b < n; b++) {
      PE_HIP(hipMemcpyAsync(results[b], static_cast<uint8_t*>(arena) + offs[b], sizeof(pe_halfsupers) * 19,
                            hipMemcpyDeviceToHost, s), "hipMemcpyAsync(results)");
    }
    PE_HIP(hipStreamSynchronize(s), "midsolve batch");
    for (int b = 0; b < n; b++)
      for (int i = infos[b].spots + 1; i < 19; i++) std::memset(&results[b][i], 0, sizeof(pe_halfsupers));
  }
  PE_HIP(hipStreamSynchronize(s), "midsolve batch");
  return st;
  PE_GUARD_END
}

pe_status pe_midsolve_internal(const pe_high_board* board, pe_halfsupers results[19]) {
  PE_GUARD_BEGIN
  if (!board || !results) return fail(PE_ERR_INVALID, "null argument");
  pe_halfsupers (*r)[19] = reinterpret_cast<pe_halfsupers (*)[19]>(results);
  return pe_midsolve_batch(1, board, r, 0, nullptr);
  PE_GUARD_END
}

pe_status pe_midsolve_internal_noalloc(const pe_high_board* board, pe_halfsupers results[19], void* stream) {
  PE_GUARD_BEGIN
  if (!board || !results) return fail(PE_ERR_INVALID, "null argument");
  return mid_solve_noalloc(*board, results, static_cast<hipStream_t>(stream));
  PE_GUARD_END
}

int64_t pe_midsolve_noalloc_scratch_bytes(void) { return mid_noalloc_scratch_bytes(); }

pe_status pe_high_done(const pe_high_board* board, int* done, int* value) {
  PE_GUARD_BEGIN
  if (!board || !done || !value) return fail(PE_ERR_INVALID, "null argument");
  if ((board->side[0] & board->side[1]) || board->ply > 72) return fail(PE_ERR_INVALID, "invalid board");
  *done = high_done_and_value(*board, value) ? 1 : 0;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_midsolve_traverse(const pe_high_board* board, const pe_halfsupers* results, pe_mid_value* values,
                               int* count) {
  PE_GUARD_BEGIN
  if (!board || !results || !values || !count) return fail(PE_ERR_INVALID, "null argument");
  MidInfo I;
  if (!make_mid_info(*board, &I)) return fail(PE_ERR_INVALID, "invalid board");
  *count = mid_traverse(*board, results, values);
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_midsolve(const pe_high_board* board, pe_mid_value* values, int* count) {
  PE_GUARD_BEGIN
  if (!board || !values || !count) return fail(PE_ERR_INVALID, "null argument");
  pe_halfsupers results[19];
  const pe_status s = pe_midsolve_internal(board, results);
  if (s != PE_OK) return s;
  return pe_midsolve_traverse(board, results, values, count);
  PE_GUARD_END
}

}  // extern "C"
