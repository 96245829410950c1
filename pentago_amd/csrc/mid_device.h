// Midsolve device code shared by the level-synchronous and the work-queue kernels: the
// per-board descriptor, the table-building entry (pentago/mid/internal.h:122-259) and the
// position sweep (pentago/mid/internal.h:358-408, both aggressive passes at once).
#pragma once
#include <hip/hip_runtime.h>

#include "device.h"
#include "mid.h"
#include "super.h"

namespace pe {

struct MidBoardDev {
  MidInfo I;
  Set0Info* set0;
  Wins1* wins1;
  uint16_t* cs1ps;
  uint32_t* set1p_bits;     // per level and s1p: bit q set for each member q of the subset
  uint64_t* work;           // 4 u64 per entry: win[2], notlose[2]
  pe_halfsupers* results;   // 19 entries
};

// k-subset of [0,n) with the given index, 5 bits per entry, large-entry-major order
// (pentago/mid/subsets.h:33-48)
// Entry i is the largest j with choose(j, i + 1) <= index (the entries below it count
// sum_{j' < j} choose(j', i) = choose(j, i + 1) subsets): counted over j' in (i, n), so the
// table reads do not wait on each other.
__device__ inline uint64_t dsubset(const int32_t* C, int n, int k, int index) {
  uint64_t set = 0;
  for (int i = k - 1; i >= 0; i--) {
    int j = i;
    for (int jj = i + 1; jj < n; jj++) j += C[jj * 32 + i + 1] <= index;
    set |= uint64_t(j) << (5 * i);
    index -= C[j * 32 + i + 1];   // choose(i, i + 1) = 0
  }
  return set;
}

__device__ __forceinline__ uint64_t dside(const MidInfo& I, int k, uint64_t set) {
  uint64_t s = 0;
  for (int i = 0; i < k; i++) s |= uint64_t(1) << I.empty[set >> (5 * i) & 0x1f];
  return s;
}

__device__ __forceinline__ int level_of(const int32_t* off, int spots, int64_t idx) {
  int n = 0;
  while (n < spots && off[n + 1] <= idx) n++;
  return n;
}

// One table entry: index over [set0 | wins1 | cs1ps] of one board.  C: the binomial table
// (T.choose or a copy of it in LDS).
__device__ inline void mid_prep_entry(const MidBoardDev& B, int64_t idx, const DevTables& T, const int32_t* C) {
  const MidInfo& I = B.I;
  const int spots = I.spots;
  const int64_t n_set0 = I.sets0_off[spots + 1], n_wins1 = I.wins1_off[spots + 1];
  if (idx < n_set0) {
    // set0_info (internal.h:174-259)
    const int n = level_of(I.sets0_off, spots, idx);
    const int s0 = int(idx - I.sets0_off[n]);
    const int k0 = n >> 1, k1 = n - k0;
    const int parity = (n + I.ply) & 1;
    const uint64_t set0 = dsubset(C, spots, k0, s0);
    const uint64_t side0 = I.side[(spots + n) & 1] | dside(I, k0, set0);
    // written field by field into the table (a local copy would live in scratch memory);
    // empty1 is also kept packed in registers, 5 bits per entry
    Set0Info& out = B.set0[idx];
    const H2 w0 = halfsuper_wins(side0, parity, T.superwin);
    out.wins0[0] = w0.w[0]; out.wins0[1] = w0.w[1];
    const uint64_t free_ = 0x01ff01ff01ff01ffull & ~side0;
    uint64_t e1[2] = {0, 0};
    for (int i = 0, k = 0; i < spots; i++)
      if (free_ >> I.empty[i] & 1) {
        out.empty1[k] = uint8_t(i);
        e1[k >= 12] |= uint64_t(i) << (5 * (k >= 12 ? k - 12 : k));
        k++;
      }
    auto empty1 = [&](int q) { return int((q < 12 ? e1[0] >> (5 * q) : e1[1] >> (5 * (q - 12))) & 0x1f); };
    auto v = [&](int a) { return int(set0 >> (5 * a) & 0x1f); };
    // child_s0s[i] = choose(empty1(i), j + 1) + sum_{a < j} choose(v_a, a + 1)
    //              + sum_{a >= j} choose(v_a, a + 2),  j = empty1(i) - i (non-decreasing in i),
    // with the two sums carried along as j grows (mod 2^16, as the term-by-term form)
    {
      int P = 0, S = 0, jc = 0;
      for (int a = 0; a < k0; a++) S += C[v(a) * 32 + a + 2];
      for (int i = 0; i < spots - k0; i++) {
        const int j = empty1(i) - i;
        for (; jc < j; jc++) {
          P += C[v(jc) * 32 + jc + 1];
          S -= C[v(jc) * 32 + jc + 2];
        }
        out.child_s0s[i] = uint16_t(C[empty1(i) * 32 + j + 1] + P + S);
      }
    }
    // offset0[a][q] = (a ? 0 : s0) - sum_{i >= empty1(q) - q, v_i > a} choose(v_i - a - 1, i):
    // the suffix sum shrinks as q grows
    for (int a = 0; a < k1; a++) {
      int cur = 0;
      for (int i = 0; i < k0; i++)
        if (v(i) > a) cur += C[(v(i) - a - 1) * 32 + i];
      int i = 0;
      for (int q = 0; q < spots - k0; q++) {
        for (const int t = empty1(q) - q; i < t; i++)
          if (v(i) > a) cur -= C[(v(i) - a - 1) * 32 + i];
        out.offset0[a * (spots - k0) + q] = uint16_t((a ? 0 : s0) - cur);
      }
    }
  } else if (idx < n_set0 + n_wins1) {
    // wins1 (internal.h:122-131)
    const int64_t e = idx - n_set0;
    const int n = level_of(I.wins1_off, spots, e);
    const int s1 = int(e - I.wins1_off[n]);
    const int k0 = n >> 1, k1 = n - k0;
    (void)k0;
    const int parity = (n + I.ply) & 1;
    const uint64_t side1 = I.side[(spots + n + 1) & 1] | dside(I, k1, dsubset(C, spots, k1, s1));
    const S4 w = super_wins<true>(side1, T.superwin);
    const H2 after = split_half(w, parity), before = split_half(w, !parity);
    Wins1 out;
    out.after[0] = after.w[0]; out.after[1] = after.w[1];
    out.before[0] = before.w[0]; out.before[1] = before.w[1];
    B.wins1[e] = out;
  } else {
    // one thread per (level, s1p): the subset s1p as a bit mask, then its cs1ps row
    // (internal.h:133-148: remap s1p when a stone of the mover lands at empty1[j])
    const int64_t e = idx - n_set0 - n_wins1;
    const int n = level_of(I.sets1p_off, spots, e);
    const int s1p = int(e - I.sets1p_off[n]);
    const int k0 = n >> 1, k1 = n - k0;
    const uint64_t set1p = dsubset(C, spots - k0, k1, s1p);
    uint32_t bits = 0;
    for (int a = 0; a < k1; a++) bits |= 1u << (set1p >> (5 * a) & 0x1f);
    B.set1p_bits[e] = bits;
    uint16_t* row = B.cs1ps + I.cs1ps_off[n] + int64_t(s1p) * (spots - k0);
    // row[j] = s1p - sum_{a: v_a > j} choose(v_a - 1, a); the members are increasing, so the
    // sum drops them one by one as j grows
    auto v = [&](int a) { return int(set1p >> (5 * a) & 0x1f); };
    int cur = 0, p = 0;
    for (int a = 0; a < k1; a++)
      if (v(a) > 0) cur += C[(v(a) - 1) * 32 + a];
    while (p < k1 && v(p) <= 0) p++;
    for (int j = 0; j < spots - k0; j++) {
      for (; p < k1 && v(p) <= j; p++) cur -= C[(v(p) - 1) * 32 + p];
      row[j] = uint16_t(s1p - cur);
    }
  }
}

// A level holds fewer than 2^31 positions (at most 17,153,136 for 18 empty cells).
__device__ inline void mid_position(const MidBoardDev& B, int n, int64_t pos64, uint64_t* __restrict__ work,
                                    const DevTables& T) {
  const MidInfo& I = B.I;
  const int spots = I.spots;
  const int k0 = n >> 1;
  const int32_t* C = T.choose;
  const uint32_t pos = uint32_t(pos64);
  const uint32_t nsets1p = uint32_t(I.sets1p_off[n + 1] - I.sets1p_off[n]);
  const uint32_t s0 = pos / nsets1p;
  const int s1p = int(pos - s0 * nsets1p);
  const Set0Info& I0 = B.set0[I.sets0_off[n] + int(s0)];
  const uint32_t filled1p = B.set1p_bits[I.sets1p_off[n] + s1p];
  // members in increasing order (member i is q), four at a time so that their table reads
  // are in flight together
  uint32_t s1w = 0, s0pw = 0;
  {
    uint32_t rest = filled1p;
    for (int i0 = 0; rest; i0 += 4) {
      int qs[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        qs[u] = rest ? __builtin_ctz(rest) : -1;
        rest &= rest - 1;
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (qs[u] >= 0) {
          s1w += uint32_t(C[I0.empty1[qs[u]] * 32 + i0 + u + 1]);
          s0pw += uint32_t(I0.offset0[(i0 + u) * (spots - k0) + qs[u]]);
        }
    }
  }
  const uint16_t s1 = uint16_t(s1w), s0p = uint16_t(s0pw);
  uint64_t win0 = 0, win1 = 0, nl0 = 0, nl1 = 0;
  if (n == spots) nl0 = nl1 = ~uint64_t(0);
  const MidGrab gin = I.spaces[n + 1];
  const uint16_t* cs = B.cs1ps + I.cs1ps_off[n] + s1p * (spots - k0);
  // children: the unfilled cells in increasing order, four gathers issued before any is used
  uint32_t unmoved = ~filled1p;
  const int nm = spots - n;
  for (int m0 = 0; m0 < nm; m0 += 4) {
    const uint64_t* cp[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = __builtin_ctz(unmoved);
      unmoved &= unmoved - 1;
      cp[u] = m0 + u < nm ? work + 4 * (int64_t(gin.lo) + int64_t(I0.child_s0s[i]) * gin.ny + cs[i]) : nullptr;
    }
    ulonglong2 a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (cp[u]) {
        a[u] = reinterpret_cast<const ulonglong2*>(cp[u])[0];
        b[u] = reinterpret_cast<const ulonglong2*>(cp[u])[1];
      } else {
        a[u] = make_ulonglong2(0, 0);
        b[u] = make_ulonglong2(0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      win0 |= a[u].x; win1 |= a[u].y; nl0 |= b[u].x; nl1 |= b[u].y;
    }
  }
  const Wins1 w1 = B.wins1[I.wins1_off[n] + s1];
  const uint64_t ip0 = ~(I0.wins0[0] | w1.after[0]), ip1 = ~(I0.wins0[1] | w1.after[1]);
  win0 = (ip0 & win0) | (I0.wins0[0] & ~w1.after[0]);
  win1 = (ip1 & win1) | (I0.wins0[1] & ~w1.after[1]);
  nl0 = (ip0 & nl0) | I0.wins0[0];
  nl1 = (ip1 & nl1) | I0.wins0[1];
  if (n <= 1) {
    pe_halfsupers& r = B.results[n + s1p];
    r.win[0] = win0; r.win[1] = win1; r.notlose[0] = nl0; r.notlose[1] = nl1;
  }
  H2 nw; nw.w[0] = ~nl0; nw.w[1] = ~nl1;     // ~(win, notlose) = (~notlose, ~win)
  H2 nn; nn.w[0] = ~win0; nn.w[1] = ~win1;
  const H2 sw = hrmax(nw), sn = hrmax(nn);
  const MidGrab gout = I.spaces[n];
  uint64_t* o = work + 4 * (int64_t(gout.lo) + int64_t(s1) * gout.ny + s0p);
  reinterpret_cast<ulonglong2*>(o)[0] = make_ulonglong2(sw.w[0] | w1.before[0], sw.w[1] | w1.before[1]);
  reinterpret_cast<ulonglong2*>(o)[1] = make_ulonglong2(sn.w[0] | w1.before[0], sn.w[1] | w1.before[1]);
}

}  // namespace pe
