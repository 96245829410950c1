// Slice stores, section enumeration, block counts and sparse samples: the one-GPU form of
// the reference's slice loop (pentago/mpi/toplevel.cc:263-421) and block store
// (pentago/end/block_store.h:113-160), without MPI, compression or files.
//
// A store holds whole sections of one slice on the device, one slab per section covering
// every block (a 4096-node slot per block).  Computing a slice is one slab plan over all of
// its sections (the fused kernel) reading the next slice's store.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <future>
#include <set>
#include <unordered_map>
#include <unordered_set>
#include <string>
#include <vector>

#include "device.h"
#include "end.h"
#include "tables.h"

namespace pe {
pe_status count_blocks_launch(int n, const pe_slab* slabs, uint64_t* const* d_counts, void* desc, hipStream_t stream,
                              std::string* err);
size_t count_desc_bytes(int n);
int64_t slab_blocks(const pe_slab& slab);
pe_status launch_supertensor_format(bool decode, int filter, int turn, const void* in, int64_t nodes, void* out,
                                    hipStream_t stream, std::string* err);

namespace {

pe_status failed(pe_status s, const std::string& e) {
  set_error(e);
  return s;
}

bool section_of(const pe_section& p, Section* s, int shape[4]) {
  std::memcpy(s->c, p.counts, 8);
  return s->valid_counts() && s->shape(shape);
}

Section section_from_sig(uint64_t sig) {
  Section c;
  for (int q = 0; q < 4; q++) {
    c.c[q][0] = uint8_t(sig >> (16 * q));
    c.c[q][1] = uint8_t(sig >> (16 * q + 8));
  }
  return c;
}

// Standardized child sections of sections [begin, end), sorted by sig.
std::vector<pe_section> group_kids(const pe_section* sections, int begin, int end) {
  std::set<uint64_t> sigs;
  for (int k = begin; k < end; k++) {
    Section s;
    std::memcpy(s.c, sections[k].counts, 8);
    for (int q = 0; q < 4; q++)
      if (s.c[q][0] + s.c[q][1] < 9) {
        Section t;
        s.child(q).standardize(&t);
        sigs.insert(t.sig());
      }
  }
  std::vector<pe_section> out;
  for (uint64_t sig : sigs) {
    const Section c = section_from_sig(sig);
    pe_section p;
    std::memcpy(p.counts, c.c, 8);
    out.push_back(p);
  }
  return out;
}

pe_slab whole_slab(const pe_section& sec, const int shape[4], void* data) {
  pe_slab slab;
  slab.section = sec;
  for (int a = 0; a < 4; a++) {
    slab.lo[a] = 0;
    slab.hi[a] = (shape[a] + 7) / 8;
  }
  slab.data = data;
  return slab;
}

// Counts of n slabs summed per slab: one launch over every block, one device buffer (per-block
// counts, then the launch descriptors) and one copy back.
pe_status count_slabs(int n, const pe_slab* slabs, uint64_t (*counts)[4]) {
  std::vector<int64_t> first(n + 1, 0);
  for (int k = 0; k < n; k++) first[k + 1] = first[k] + slab_blocks(slabs[k]);
  std::fill(&counts[0][0], &counts[0][0] + 4 * size_t(n), uint64_t(0));
  if (!first[n]) return PE_OK;
  const size_t count_bytes = sizeof(uint64_t) * 4 * size_t(first[n]);
  void* d = nullptr;
  if (hipMalloc(&d, count_bytes + count_desc_bytes(n)) != hipSuccess)
    return failed(PE_ERR_DEVICE, "hipMalloc(counts) failed");
  std::vector<uint64_t*> dst(n);
  for (int k = 0; k < n; k++) dst[k] = static_cast<uint64_t*>(d) + 4 * first[k];
  std::string err;
  pe_status s = count_blocks_launch(n, slabs, dst.data(), static_cast<char*>(d) + count_bytes, nullptr, &err);
  std::vector<uint64_t> h(4 * first[n]);
  if (s == PE_OK && hipMemcpy(h.data(), d, count_bytes, hipMemcpyDeviceToHost) != hipSuccess) {
    s = PE_ERR_DEVICE;
    err = "counting blocks failed";
  }
  (void)hipFree(d);
  if (s != PE_OK) return failed(s, err);
  for (int k = 0; k < n; k++)
    for (int64_t b = first[k]; b < first[k + 1]; b++)
      for (int j = 0; j < 4; j++) counts[k][j] += h[4 * b + j];
  return PE_OK;
}

}  // namespace
}  // namespace pe

using namespace pe;

struct pe_slice {
  std::vector<pe_slab> slabs;
  std::vector<void*> data;
  int slice = 0;
  int device = 0;
};

extern "C" {

pe_status pe_end_count_blocks(const pe_slab* slab, uint64_t* d_counts, void* stream) {
  PE_GUARD_BEGIN
  if (!slab) return failed(PE_ERR_INVALID, "null slab");
  std::string err;
  const pe_status s = count_blocks_launch(1, slab, &d_counts, nullptr, static_cast<hipStream_t>(stream), &err);
  return s == PE_OK ? s : failed(s, err);
  PE_GUARD_END
}

pe_status pe_end_count_blocks_many(int n, const pe_slab* slabs, uint64_t* d_counts, void* d_workspace, void* stream) {
  PE_GUARD_BEGIN
  if (n < 0 || (n && (!slabs || !d_counts || !d_workspace))) return failed(PE_ERR_INVALID, "bad count arguments");
  std::vector<uint64_t*> dst(n);
  int64_t first = 0;
  for (int k = 0; k < n; k++) {
    dst[k] = d_counts + 4 * first;
    first += slab_blocks(slabs[k]);
  }
  std::string err;
  const pe_status s = count_blocks_launch(n, slabs, dst.data(), d_workspace, static_cast<hipStream_t>(stream), &err);
  return s == PE_OK ? s : failed(s, err);
  PE_GUARD_END
}

int64_t pe_end_count_workspace_bytes(int n) { return int64_t(count_desc_bytes(n)); }

pe_status pe_end_count_slabs(int n, const pe_slab* slabs, uint64_t (*counts)[4]) {
  PE_GUARD_BEGIN
  if (n < 0 || (n && (!slabs || !counts))) return failed(PE_ERR_INVALID, "bad count arguments");
  return count_slabs(n, slabs, counts);
  PE_GUARD_END
}

pe_status pe_supertensor_decode(int filter, int turn, const void* d_in, int64_t nodes, void* d_out, void* stream) {
  PE_GUARD_BEGIN
  std::string err;
  const pe_status s = launch_supertensor_format(true, filter, turn, d_in, nodes, d_out,
                                                static_cast<hipStream_t>(stream), &err);
  return s == PE_OK ? s : failed(s, err);
  PE_GUARD_END
}

pe_status pe_supertensor_encode(int filter, int turn, const void* d_in, int64_t nodes, void* d_out, void* stream) {
  PE_GUARD_BEGIN
  std::string err;
  const pe_status s = launch_supertensor_format(false, filter, turn, d_in, nodes, d_out,
                                                static_cast<hipStream_t>(stream), &err);
  return s == PE_OK ? s : failed(s, err);
  PE_GUARD_END
}

pe_status pe_end_slab_samples(const pe_slab* slab, int n, const uint64_t* boards, void* h_nodes) {
  PE_GUARD_BEGIN
  if (!slab || n < 0 || (n && (!boards || !h_nodes))) return failed(PE_ERR_INVALID, "bad sample arguments");
  Section s;
  int shape[4];
  if (!section_of(slab->section, &s, shape)) return failed(PE_ERR_INVALID, "invalid slab section");
  const Tables& T = tables();
  for (int i = 0; i < n; i++) {
    int idx[4];
    for (int q = 0; q < 4; q++) {
      const int quad = int(boards[i] >> (16 * q) & 0xffff);
      if (quad >= 19683) return failed(PE_ERR_INVALID, "sample " + std::to_string(i) + ": invalid quadrant");
      const int black = __builtin_popcount(T.unpack[quad][0]), white = __builtin_popcount(T.unpack[quad][1]);
      if (black != s.c[q][0] || white != s.c[q][1])
        return failed(PE_ERR_INVALID, "sample " + std::to_string(i) + ": board is not in the slab's section");
      const int ir = T.rmin_inverse[quad];
      if (ir & 3) return failed(PE_ERR_INVALID, "sample " + std::to_string(i) + ": board is not rotation standardized");
      idx[q] = ir >> 2;
    }
    int b[4], n_[4];
    int64_t slot = 0, within = 0;
    for (int a = 0; a < 4; a++) {
      b[a] = idx[a] / 8;
      if (b[a] < slab->lo[a] || b[a] >= slab->hi[a])
        return failed(PE_ERR_INVALID, "sample " + std::to_string(i) + ": block outside the slab's box");
      n_[a] = std::min(8, shape[a] - 8 * b[a]);
      slot = slot * (slab->hi[a] - slab->lo[a]) + (b[a] - slab->lo[a]);
      within = within * n_[a] + idx[a] % 8;
    }
    const auto* src = static_cast<const uint8_t*>(slab->data) + 64 * (4096 * slot + within);
    if (hipMemcpy(static_cast<uint8_t*>(h_nodes) + 64 * size_t(i), src, 64, hipMemcpyDeviceToHost) != hipSuccess)
      return failed(PE_ERR_DEVICE, "sample copy failed");
  }
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_descendent_sections(const pe_section* root, int slice, int cap, pe_section* out, int* count) {
  PE_GUARD_BEGIN
  if (!root || !count || cap < 0 || (cap && !out)) return failed(PE_ERR_INVALID, "bad section arguments");
  Section r;
  int shape[4];
  if (!section_of(*root, &r, shape)) return failed(PE_ERR_INVALID, "invalid root section");
  if (slice < r.sum() || slice > 35) return failed(PE_ERR_INVALID, "slice outside [root stones, 35]");
  // depth-first over standardized sections (every child adds one stone of the player to move)
  std::set<uint64_t> seen;
  std::vector<Section> stack{r};
  std::vector<Section> found;
  while (!stack.empty()) {
    Section s;
    stack.back().standardize(&s);
    stack.pop_back();
    if (!seen.insert(s.sig()).second) continue;
    if (s.sum() == slice) {
      found.push_back(s);
      continue;
    }
    for (int q = 0; q < 4; q++)
      if (s.c[q][0] + s.c[q][1] < 9) stack.push_back(s.child(q));
  }
  std::sort(found.begin(), found.end(), [](const Section& a, const Section& b) { return a.sig() < b.sig(); });
  *count = int(found.size());
  for (int k = 0; k < std::min(cap, int(found.size())); k++) std::memcpy(out[k].counts, found[k].c, 8);
  return PE_OK;
  PE_GUARD_END
}

int64_t pe_slice_bytes(int n, const pe_section* sections) {
  if (n < 0 || (n && !sections)) return -1;
  int64_t bytes = 0;
  for (int k = 0; k < n; k++) {
    Section s;
    int shape[4];
    if (!section_of(sections[k], &s, shape)) return -1;
    int64_t blocks = 1;
    for (int a = 0; a < 4; a++) blocks *= (shape[a] + 7) / 8;
    bytes += blocks * 4096 * 64;
  }
  return bytes;
}

pe_status pe_slice_create_in(int n, const pe_section* sections, void* arena, int64_t arena_bytes, pe_slice** out) {
  PE_GUARD_BEGIN
  if (n <= 0 || !sections || !out || (arena && arena_bytes < 0)) return failed(PE_ERR_INVALID, "bad slice arguments");
  auto* st = new pe_slice();
  (void)hipGetDevice(&st->device);
  int64_t used = 0;
  for (int k = 0; k < n; k++) {
    Section s, t;
    int shape[4];
    if (!section_of(sections[k], &s, shape) || s.sum() > 35 || (k && s.sum() != st->slice)) {
      pe_slice_destroy(st);
      return failed(PE_ERR_INVALID, "section " + std::to_string(k) + " is invalid or of another slice");
    }
    s.standardize(&t);
    if (t.sig() != s.sig()) {
      pe_slice_destroy(st);
      return failed(PE_ERR_INVALID, "section " + std::to_string(k) + " is not standardized");
    }
    st->slice = s.sum();
    int64_t blocks = 1;
    for (int a = 0; a < 4; a++) blocks *= (shape[a] + 7) / 8;
    const int64_t bytes = blocks * 4096 * 64;
    void* d = nullptr;
    if (arena) {   // carve from the caller's memory (no allocation)
      if (used + bytes > arena_bytes) {
        pe_slice_destroy(st);
        return failed(PE_ERR_CAPACITY, "sections do not fit the arena");
      }
      d = static_cast<uint8_t*>(arena) + used;
      used += bytes;
    } else {
      if (hipMalloc(&d, size_t(bytes)) != hipSuccess) {
        pe_slice_destroy(st);
        return failed(PE_ERR_CAPACITY, "hipMalloc of section " + std::to_string(k) + " failed");
      }
      st->data.push_back(d);
    }
    st->slabs.push_back(whole_slab(sections[k], shape, d));
  }
  *out = st;
  return PE_OK;
  PE_GUARD_END
}

pe_status pe_slice_create(int n, const pe_section* sections, pe_slice** out) {
  PE_GUARD_BEGIN
  return pe_slice_create_in(n, sections, nullptr, 0, out);
  PE_GUARD_END
}

pe_status pe_slice_compute(pe_slice* out, const pe_slice* children, void* stream) {
  PE_GUARD_BEGIN
  if (!out) return failed(PE_ERR_INVALID, "null slice");
  if (out->slice < 35 && !children) return failed(PE_ERR_INVALID, "slices below 35 need the next slice's store");
  if (children && children->slice != out->slice + 1) return failed(PE_ERR_INVALID, "children are not the next slice");
  pe_end_plan* plan = nullptr;
  const int nch = out->slice < 35 ? int(children->slabs.size()) : 0;
  pe_status s = pe_end_slabs_plan_create(int(out->slabs.size()), out->slabs.data(), nch,
                                         nch ? children->slabs.data() : nullptr, 0, &plan);
  if (s != PE_OK) return s;
  s = pe_end_plan_run(plan, stream);
  if (s == PE_OK && hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
    s = failed(PE_ERR_DEVICE, "slice compute failed");
  pe_end_plan_destroy(plan);
  return s;
  PE_GUARD_END
}

pe_status pe_slice_slab(const pe_slice* s, int k, pe_slab* slab) {
  PE_GUARD_BEGIN
  if (!s || !slab || k < 0 || k >= int(s->slabs.size())) return failed(PE_ERR_INVALID, "bad slab index");
  *slab = s->slabs[k];
  return PE_OK;
  PE_GUARD_END
}

int pe_slice_sections(const pe_slice* s) { return s ? int(s->slabs.size()) : -1; }

pe_status pe_slice_counts(const pe_slice* s, uint64_t (*counts)[4]) {
  PE_GUARD_BEGIN
  if (!s || !counts) return failed(PE_ERR_INVALID, "null argument");
  return count_slabs(int(s->slabs.size()), s->slabs.data(), counts);
  PE_GUARD_END
}

pe_status pe_slice_plan_groups(int n, const pe_section* sections, int64_t budget, int regenerate, int* order,
                               int* ends, int* ngroups) {
  PE_GUARD_BEGIN
  if (n <= 0 || !sections || !order || !ends || !ngroups || budget <= 0)
    return failed(PE_ERR_INVALID, "bad group arguments");
  std::vector<int64_t> own(n);
  std::vector<std::vector<uint64_t>> kids(n);   // standardized child sections (regeneration only)
  std::unordered_map<uint64_t, int64_t> kid_bytes;
  for (int k = 0; k < n; k++) {
    Section s;
    int shape[4];
    if (!section_of(sections[k], &s, shape)) return failed(PE_ERR_INVALID, "invalid section " + std::to_string(k));
    own[k] = pe_slice_bytes(1, &sections[k]);
    if (!regenerate) continue;
    for (int q = 0; q < 4; q++)
      if (s.c[q][0] + s.c[q][1] < 9) {
        Section t;
        s.child(q).standardize(&t);
        const uint64_t sig = t.sig();
        if (std::find(kids[k].begin(), kids[k].end(), sig) != kids[k].end()) continue;
        kids[k].push_back(sig);
        if (!kid_bytes.count(sig)) {
          pe_section p;
          std::memcpy(p.counts, t.c, 8);
          kid_bytes[sig] = pe_slice_bytes(1, &p);
        }
      }
  }
  int count = 0;
  if (!regenerate) {   // sections in the given order, each group as long as it fits
    int64_t used = 0;
    for (int k = 0; k < n; k++) {
      order[k] = k;
      if (k > 0 && used + own[k] > budget) {
        ends[count++] = k;
        used = 0;
      }
      used += own[k];
    }
    ends[count++] = n;
    *ngroups = count;
    return PE_OK;
  }
  // With regeneration a group pays for its sections' children once, so sections that share
  // children go together: each group is filled by repeatedly taking the remaining section
  // that fits and adds the fewest new child bytes net of the child bytes it shares (ties:
  // the larger section, then the earlier one), and is closed when none fits.
  // fresh[k] / shared[k]: child bytes of section k outside / inside the current group, kept
  // up to date through the sections that use each child
  std::unordered_map<uint64_t, std::vector<int>> users;
  std::vector<int64_t> fresh(n, 0), shared(n, 0);
  for (int k = 0; k < n; k++)
    for (uint64_t sig : kids[k]) {
      users[sig].push_back(k);
      fresh[k] += kid_bytes[sig];
    }
  std::vector<char> taken(n, 0);
  std::unordered_set<uint64_t> group;
  int placed = 0, in_group = 0;
  int64_t used = 0;
  while (placed < n) {
    int best = -1;
    for (int k = 0; k < n; k++) {
      if (taken[k] || (in_group && used + own[k] + fresh[k] > budget)) continue;
      const int64_t score = shared[k] - fresh[k];
      const int64_t best_score = best < 0 ? 0 : shared[best] - fresh[best];
      if (best < 0 || score > best_score || (score == best_score && own[k] > own[best])) best = k;
    }
    if (best < 0) {   // nothing fits: close the group
      ends[count++] = placed;
      for (uint64_t sig : group)
        for (int u : users[sig]) {
          fresh[u] += kid_bytes[sig];
          shared[u] -= kid_bytes[sig];
        }
      group.clear();
      used = 0;
      in_group = 0;
      continue;
    }
    taken[best] = 1;
    order[placed++] = best;
    in_group++;
    used += own[best] + fresh[best];
    for (uint64_t sig : kids[best])
      if (group.insert(sig).second)
        for (int u : users[sig]) {
          fresh[u] -= kid_bytes[sig];
          shared[u] += kid_bytes[sig];
        }
  }
  ends[count++] = n;
  *ngroups = count;
  return PE_OK;
  PE_GUARD_END
}

}  // extern "C"

namespace {

// Device bytes of the largest group: its output slabs, and its regenerated children.
int64_t group_arena_bytes(const pe_section* sections, const std::vector<int>& ends, int ng, bool regen) {
  int64_t bytes = 0;
  for (int g = 0, begin = 0; g < ng; begin = ends[g++]) {
    int64_t b = pe_slice_bytes(ends[g] - begin, sections + begin);
    if (regen) {
      const std::vector<pe_section> ks = group_kids(sections, begin, ends[g]);
      b += pe_slice_bytes(int(ks.size()), ks.data());
    }
    bytes = std::max(bytes, b);
  }
  return bytes;
}

}  // namespace

extern "C" {

int64_t pe_slice_arena_bytes(int n, const pe_section* sections, int64_t budget, int regenerate) try {
  std::vector<int> order(std::max(n, 1)), ends(std::max(n, 1));
  int ng = 0;
  if (pe_slice_plan_groups(n, sections, budget, regenerate, order.data(), ends.data(), &ng) != PE_OK) return -1;
  std::vector<pe_section> ps(n);
  for (int k = 0; k < n; k++) ps[k] = sections[order[k]];
  return group_arena_bytes(ps.data(), ends, ng, regenerate != 0);
} catch (const std::exception&) {
  return -1;
}

pe_status pe_slice_grouped_counts_in(int n, const pe_section* sections, int64_t budget, const pe_slice* children,
                                     const pe_slice* grandchildren, uint64_t (*counts)[4],
                                     void (*visit)(void* ctx, const pe_slice* group), void* ctx, int* ngroups,
                                     void* arena, int64_t arena_bytes, void* stream) {
  PE_GUARD_BEGIN
  if (!counts || !arena) return failed(PE_ERR_INVALID, "null counts or arena");
  if (!children && !grandchildren) return failed(PE_ERR_INVALID, "need the next slice's store or the one after it");
  const bool regen = !children;
  std::vector<int> order(std::max(n, 1)), ends(std::max(n, 1));
  int ng = 0;
  pe_status st = pe_slice_plan_groups(n, sections, budget, regen, order.data(), ends.data(), &ng);
  if (st != PE_OK) return st;
  if (ngroups) *ngroups = ng;
  // the sections in plan order; counts go back to the caller's order at the end
  std::vector<pe_section> ps(n);
  for (int k = 0; k < n; k++) ps[k] = sections[order[k]];
  if (group_arena_bytes(ps.data(), ends, ng, regen) > arena_bytes)
    return failed(PE_ERR_CAPACITY, "the arena is smaller than pe_slice_arena_bytes");
  std::vector<uint64_t> pc(4 * size_t(n), 0);
  auto* pcounts = reinterpret_cast<uint64_t(*)[4]>(pc.data());
  sections = ps.data();
  // A group's stores are carved from the one arena (its outputs, then its regenerated
  // children) and its plans are built on a host thread while the previous group's kernels
  // run: building a plan reads addresses only, never the memory behind them.
  struct Group {
    pe_slice* kids = nullptr;
    pe_slice* out = nullptr;
    pe_end_plan* kid_plan = nullptr;
    pe_end_plan* out_plan = nullptr;
    pe_status st = PE_OK;
    std::string err;
    void release() {
      pe_end_plan_destroy(kid_plan);
      pe_end_plan_destroy(out_plan);
      pe_slice_destroy(out);
      pe_slice_destroy(kids);
      kid_plan = out_plan = nullptr;
      out = kids = nullptr;
    }
  };
  int dev = 0;
  (void)hipGetDevice(&dev);
  auto prepare = [&, dev](int g) {
    Group G;
    (void)hipSetDevice(dev);
    const int begin = g ? ends[g - 1] : 0;
    const int64_t out_bytes = pe_slice_bytes(ends[g] - begin, sections + begin);
    if (regen) {
      const std::vector<pe_section> ks = group_kids(sections, begin, ends[g]);
      G.st = pe_slice_create_in(int(ks.size()), ks.data(), static_cast<uint8_t*>(arena) + out_bytes,
                                arena_bytes - out_bytes, &G.kids);
      if (G.st == PE_OK)
        G.st = pe_end_slabs_plan_create(int(G.kids->slabs.size()), G.kids->slabs.data(),
                                        int(grandchildren->slabs.size()), grandchildren->slabs.data(), 0,
                                        &G.kid_plan);
    }
    if (G.st == PE_OK) G.st = pe_slice_create_in(ends[g] - begin, sections + begin, arena, out_bytes, &G.out);
    const pe_slice* ch = regen ? G.kids : children;
    if (G.st == PE_OK)
      G.st = pe_end_slabs_plan_create(int(G.out->slabs.size()), G.out->slabs.data(), int(ch->slabs.size()),
                                      ch->slabs.data(), 0, &G.out_plan);
    if (G.st != PE_OK) G.err = pe_last_error();
    return G;
  };
  std::future<Group> next = std::async(std::launch::async, prepare, 0);
  for (int g = 0, begin = 0; g < ng; begin = ends[g++]) {
    Group G = next.get();
    if (G.st == PE_OK && g + 1 < ng) next = std::async(std::launch::async, prepare, g + 1);
    st = G.st;
    if (st != PE_OK) set_error(G.err);
    if (st == PE_OK && G.kid_plan) st = pe_end_plan_run(G.kid_plan, stream);
    if (st == PE_OK) st = pe_end_plan_run(G.out_plan, stream);
    if (st == PE_OK && hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
      st = failed(PE_ERR_DEVICE, "group compute failed");
    if (st == PE_OK) st = pe_slice_counts(G.out, pcounts + begin);
    if (st == PE_OK && visit) visit(ctx, G.out);
    G.release();
    if (st != PE_OK) break;
  }
  if (next.valid()) {   // a group prepared after a failure: release it
    Group G = next.get();
    G.release();
  }
  if (st == PE_OK)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < 4; j++) counts[order[k]][j] = pcounts[k][j];
  return st;
  PE_GUARD_END
}

pe_status pe_slice_grouped_counts(int n, const pe_section* sections, int64_t budget, const pe_slice* children,
                                  const pe_slice* grandchildren, uint64_t (*counts)[4],
                                  void (*visit)(void* ctx, const pe_slice* group), void* ctx, int* ngroups,
                                  void* stream) {
  PE_GUARD_BEGIN
  if (!counts) return failed(PE_ERR_INVALID, "null counts");
  if (!children && !grandchildren) return failed(PE_ERR_INVALID, "need the next slice's store or the one after it");
  const int64_t bytes = pe_slice_arena_bytes(n, sections, budget, !children);
  if (bytes < 0) return failed(PE_ERR_INVALID, "invalid sections or budget");
  void* arena = nullptr;
  if (hipMalloc(&arena, size_t(std::max<int64_t>(bytes, 1))) != hipSuccess)
    return failed(PE_ERR_CAPACITY, "hipMalloc(group arena) failed");
  const pe_status st = pe_slice_grouped_counts_in(n, sections, budget, children, grandchildren, counts, visit, ctx,
                                                  ngroups, arena, bytes, stream);
  (void)hipFree(arena);
  return st;
  PE_GUARD_END
}

pe_status pe_slice_destroy(pe_slice* s) {
  PE_GUARD_BEGIN
  if (!s) return PE_OK;
  for (void* d : s->data) (void)hipFree(d);
  delete s;
  return PE_OK;
  PE_GUARD_END
}

}  // extern "C"
