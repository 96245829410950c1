// Host build of the device primitives in pentago_amd/csrc/super.h, for CPU tests only:
// the tests compare these functions bit-for-bit with the oracle before any GPU run.
#include <cstring>

#include "../../pentago_amd/csrc/super.h"
#include "../../pentago_amd/csrc/transforms.h"

using namespace pe;

extern "C" {
void hc_rmax(const uint64_t* in, uint64_t* out) { S4 a; std::memcpy(a.w, in, 32); a = rmax(a); std::memcpy(out, a.w, 32); }
void hc_transform(int g, int y, const uint64_t* in, uint64_t* out) {
  S4 a; std::memcpy(a.w, in, 32); a = transform_super(g, y, a); std::memcpy(out, a.w, 32);
}
void hc_super_wins(uint64_t side, const uint64_t* table, uint64_t* out) {
  S4 a = super_wins(side, table); std::memcpy(out, a.w, 32);
}
void hc_split(const uint64_t* in, int parity, uint64_t* out) {
  S4 a; std::memcpy(a.w, in, 32); H2 h = split_half(a, parity); std::memcpy(out, h.w, 16);
}
void hc_shift_var(int q, int k, const uint64_t* in, uint64_t* out) {
  S4 a; std::memcpy(a.w, in, 32); a = shift_coord_var(a, q, k); std::memcpy(out, a.w, 32);
}
void hc_line_transform(int g, const uint64_t* in, uint64_t* out) {
  S4 a; std::memcpy(a.w, in, 32);
  a = g == 1 ? line_transform1(a) : g == 2 ? line_transform2(a) : line_transform3(a);
  std::memcpy(out, a.w, 32);
}
void hc_hrmax(const uint64_t* in, uint64_t* out) { H2 h; std::memcpy(h.w, in, 16); h = hrmax(h); std::memcpy(out, h.w, 16); }
}
