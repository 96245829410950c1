// Lookup table generation (see tables.h).  The mathematics follows
// pentago/base/precompute.cc; the code is this project's own.
#include "tables.h"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <vector>

namespace pe {
namespace {

// Reference digests, pentago/base/precompute.cc:169-1038.
struct Digest { const char* name; const char* sha1; };
const Digest kDigests[] = {
  {"pack", "b86e92ca7f525bd398ba376616219831e3f4f1a5"},
  {"unpack", "99e742106ae60b62c0bb71dee15789ef1eb761a0"},
  {"win_contributions", "4e5cf35e82fceecd464d73c3de35e6af4f75ee34"},
  {"rotations", "195f19d49311f82139a18ae681592de02b9954bc"},
  {"reflections", "2b23dc37f4bc1008eba3df0ee1b7815675b658bf"},
  {"superwin_info", "668eb0a940489f434f804d994698a4fc85f5b576"},
  {"commute_global_local_symmetries", "e051d034c07bfa79fa62273b05839aedf446d499"},
  {"rotation_minimal_quadrants_offsets", "7e450e73e0d54bd3591710e10f4aa76dbcbbd715"},
  {"rotation_minimal_quadrants_flat", "8f48bb94ad675de569b07cca98a2e930b06b45ac"},
  {"rotation_minimal_quadrants_inverse", "339369694f78d4a197db8dc41a1f41300ba4f46c"},
  {"rotation_minimal_quadrants_reflect_moved", "dce212878aaebbcd995a8a0308335972bd1d5ef7"},
};

Tables* g_tables = nullptr;
std::once_flag g_once;

inline int cell_bit(int x, int y) { return 3 * x + y; }   // cell (x,y) of a 3x3 quadrant

// Rotation of a side quadrant: left maps (x,y) -> (2-y,x), right maps (x,y) -> (y,2-x).
uint16_t rotate_side(uint16_t v, bool left) {
  uint16_t r = 0;
  for (int x = 0; x < 3; x++)
    for (int y = 0; y < 3; y++)
      if (v >> cell_bit(x, y) & 1) r |= uint16_t(1) << (left ? cell_bit(2 - y, x) : cell_bit(y, 2 - x));
  return r;
}

// The 32 five-in-a-row patterns in the reference's order (precompute.cc:199-254).
std::vector<uint64_t> win_patterns() {
  auto board_bit = [](int x, int y) {
    return uint64_t(1) << (16 * (2 * (x / 3) + y / 3) + 3 * (x % 3) + y % 3);
  };
  std::vector<uint64_t> pats;
  for (int x = 0; x < 2; x++)
    for (int y = 0; y < 2; y++) {
      uint64_t p = 0;
      const int dy = x == y ? -1 : 1;
      for (int i = -2; i <= 2; i++) p |= board_bit(2 + x + i, 2 + y + i * dy);
      pats.push_back(p);
    }
  for (int flip = 0; flip < 2; flip++)
    for (int x = 0; x < 2; x++)
      for (int y = 0; y < 6; y++) {
        uint64_t p = 0;
        for (int i = 0; i < 5; i++) p |= flip ? board_bit(y, x + i) : board_bit(x + i, y);
        pats.push_back(p);
      }
  for (int rotate = 0; rotate < 2; rotate++)
    for (int x = 0; x < 2; x++) {
      uint64_t p = 0;
      for (int i = 0; i < 5; i++) p |= rotate ? board_bit(5 - (i + x), i + x) : board_bit(i + x, i + x);
      pats.push_back(p);
    }
  return pats;
}

void build_superwin(Tables& T, const std::vector<uint64_t>& pats, const uint16_t allrot[512][4]) {
  // Each pattern group names the role ("hvlua") of every quadrant it uses; '-' is unused.
  // For a group G and its i-th way w, the bit (q, v, field(G,q), r) is set when the unused
  // quadrant coordinates of r spell i (base 4) and state v rotated left r_q times covers w's
  // cells in quadrant q.  AND over G's quadrants then OR over the unused axes gives the wins.
  const char* groups[8] = {"vv--", "h-h-", "-h-h", "--vv", "l-al", "ua-u", "all-", "-uua"};
  const char* roles = "hvlua";
  std::memset(T.superwin, 0, sizeof T.superwin);
  for (const char* g : groups) {
    std::vector<int> unused;
    for (int q = 0; q < 4; q++) if (g[q] == '-') unused.push_back(q);
    std::vector<std::array<uint16_t, 4>> ways;
    for (uint64_t p : pats) {
      std::array<uint16_t, 4> w;
      bool ok = true;
      for (int q = 0; q < 4; q++) {
        w[q] = uint16_t(p >> (16 * q) & 0x1ff);
        if (g[q] != 'a' && (w[q] != 0) != (g[q] != '-')) ok = false;
      }
      if (ok) ways.push_back(w);
    }
    if (ways.size() != 3 && ways.size() != 6) throw std::runtime_error("superwin: bad pattern group");
    for (int q = 0; q < 4; q++) {
      if (g[q] == '-') continue;
      const int field = int(std::strchr(roles, g[q]) - roles);
      for (int r = 0; r < 256; r++) {
        int code = 0;
        for (int u : unused) code = 4 * code + (r >> (2 * u) & 3);
        if (code >= int(ways.size())) continue;
        const uint16_t need = ways[code][q];
        const int rq = r >> (2 * q) & 3;
        for (int v = 0; v < 512; v++)
          if ((allrot[v][rq] & need) == need) {
            const size_t bit = size_t((q * 512 + v) * 5 + field) * 256 + r;
            T.superwin[bit >> 6] |= uint64_t(1) << (bit & 63);
          }
      }
    }
  }
}

// Cells of the 6x6 board as index 6x+y; symmetries as permutations (precompute.cc:883-931).
using Perm = std::array<uint8_t, 36>;
Perm perm_id() { Perm p; for (int i = 0; i < 36; i++) p[i] = uint8_t(i); return p; }
Perm perm_mul(const Perm& a, const Perm& b) { Perm r; for (int i = 0; i < 36; i++) r[i] = a[b[i]]; return r; }
Perm perm_inv(const Perm& a) { Perm r; for (int i = 0; i < 36; i++) r[a[i]] = uint8_t(i); return r; }

void build_commute(Tables& T) {
  Perm reflect = perm_id(), turn = perm_id();
  for (int x = 0; x < 6; x++)
    for (int y = 0; y < 6; y++) {
      reflect[6 * x + y] = uint8_t(6 * (5 - y) + (5 - x));
      turn[6 * x + y] = uint8_t(6 * (5 - y) + x);
    }
  Perm quarter[4];
  for (int q = 0; q < 4; q++) {
    quarter[q] = perm_id();
    const int d = 3 * (6 * (q / 2) + q % 2);
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) quarter[q][d + 6 * x + y] = uint8_t(d + 6 * (2 - y) + x);
  }
  // local[id], id = ((i0*4+i1)*4+i2)*4+i3, is quarter0^i0 quarter1^i1 quarter2^i2 quarter3^i3
  std::vector<Perm> local(256);
  for (int id = 0; id < 256; id++) {
    Perm p = perm_id();
    const int e[4] = {id >> 6 & 3, id >> 4 & 3, id >> 2 & 3, id & 3};
    for (int q = 0; q < 4; q++)
      for (int k = 0; k < e[q]; k++) p = perm_mul(p, quarter[q]);
    local[id] = p;
  }
  Perm global[8];
  global[0] = perm_id();
  for (int i = 0; i < 3; i++) global[i + 1] = perm_mul(turn, global[i]);
  for (int i = 0; i < 4; i++) global[4 + i] = perm_mul(reflect, global[i]);
  for (int g = 0; g < 8; g++)
    for (int l = 0; l < 256; l++) {
      const Perm c = perm_mul(perm_mul(perm_inv(global[g]), local[l]), global[g]);
      const auto it = std::find(local.begin(), local.end(), c);
      if (it == local.end()) throw std::runtime_error("commute: not a local symmetry");
      T.commute[g][l] = uint8_t(it - local.begin());
    }
}

void build_rmin(Tables& T, const uint16_t allrot[512][4]) {
  auto pack2 = [&](uint16_t s0, uint16_t s1) { return uint16_t(T.pack[s0] + 2 * T.pack[s1]); };
  std::vector<uint16_t> minq(19683);
  for (int q = 0; q < 19683; q++) {
    uint16_t m = uint16_t(q);
    for (int r = 1; r < 4; r++) m = std::min(m, pack2(allrot[T.unpack[q][0]][r], allrot[T.unpack[q][1]][r]));
    minq[q] = m;
  }
  auto min_reflect = [&](uint16_t q) { return minq[pack2(T.refl[T.unpack[q][0]], T.refl[T.unpack[q][1]])]; };
  // Order: reflected pairs adjacent and first, then reflection-fixed quadrants; within each
  // class by the smaller member of the pair (precompute.cc:980-997).
  std::vector<uint16_t> mins;
  for (int q = 0; q < 19683; q++) if (minq[q] == q) mins.push_back(uint16_t(q));
  auto key = [&](uint16_t q) { const uint16_t r = min_reflect(q); return std::min(q, r) + 19683 * (q == r); };
  std::stable_sort(mins.begin(), mins.end(), [&](uint16_t a, uint16_t b) { return key(a) < key(b); });
  std::vector<std::vector<uint16_t>> buckets(55);
  for (uint16_t q : mins) {
    const int b = __builtin_popcount(T.unpack[q][0]), w = __builtin_popcount(T.unpack[q][1]);
    buckets[b * (21 - b) / 2 + w].push_back(q);
  }
  int off = 0;
  for (int k = 0; k < 55; k++) {
    T.rmin_offsets[k] = uint16_t(off);
    int moved = 0;
    for (size_t j = 0; j < buckets[k].size(); j++) {
      T.rmin_flat[off + j] = buckets[k][j];
      moved += min_reflect(buckets[k][j]) != buckets[k][j];
    }
    T.rmin_moved[k] = uint16_t(moved);
    off += int(buckets[k].size());
  }
  T.rmin_offsets[55] = uint16_t(off);
  T.rmin_flat_size = off;
  std::vector<int> inv(19683, 1 << 20);
  for (const auto& bucket : buckets)
    for (size_t j = 0; j < bucket.size(); j++)
      for (int r = 0; r < 4; r++) {
        const uint16_t q = pack2(allrot[T.unpack[bucket[j]][0]][r], allrot[T.unpack[bucket[j]][1]][r]);
        inv[q] = std::min(inv[q], int(4 * j + r));
      }
  for (int q = 0; q < 19683; q++) {
    if (inv[q] >= (1 << 16)) throw std::runtime_error("rmin: quadrant without a rotation-minimal form");
    T.rmin_inverse[q] = uint16_t(inv[q]);
  }
}

void build(Tables& T) {
  for (int v = 0; v < 512; v++) {
    int digits = 0;
    for (int i = 8; i >= 0; i--) digits = 3 * digits + (v >> i & 1);
    T.pack[v] = uint16_t(digits);
  }
  for (int v = 0; v < 19683; v++) {
    uint16_t s0 = 0, s1 = 0;
    int x = v;
    for (int i = 0; i < 9; i++, x /= 3) {
      if (x % 3 == 1) s0 |= uint16_t(1) << i;
      if (x % 3 == 2) s1 |= uint16_t(1) << i;
    }
    T.unpack[v][0] = s0;
    T.unpack[v][1] = s1;
  }
  static uint16_t allrot[512][4];
  for (int v = 0; v < 512; v++) {
    T.rot[v][0] = rotate_side(uint16_t(v), true);
    T.rot[v][1] = rotate_side(uint16_t(v), false);
    uint16_t r = 0;
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++)
        if (v >> cell_bit(x, y) & 1) r |= uint16_t(1) << cell_bit(2 - y, 2 - x);
    T.refl[v] = r;
  }
  for (int v = 0; v < 512; v++) {
    allrot[v][0] = uint16_t(v);
    for (int k = 1; k < 4; k++) allrot[v][k] = T.rot[allrot[v][k - 1]][0];
  }
  const auto pats = win_patterns();
  for (int q = 0; q < 4; q++)
    for (int v = 0; v < 512; v++) {
      uint64_t c = 0;
      const uint64_t mine = uint64_t(v) << (16 * q), qmask = uint64_t(0x1ff) << (16 * q);
      for (size_t i = 0; i < pats.size(); i++) {
        const uint64_t part = pats[i] & qmask;
        if (part && (part & mine) == part) c |= uint64_t(1) << (2 * i);
      }
      T.win_contrib[q][v] = c;
    }
  build_superwin(T, pats, allrot);
  build_commute(T);
  build_rmin(T, allrot);
}

}  // namespace

std::string sha1_hex(const void* data, size_t n) {
  // FIPS 180-4
  auto rol = [](uint32_t x, int k) { return (x << k) | (x >> (32 - k)); };
  uint32_t h[5] = {0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u, 0xC3D2E1F0u};
  std::vector<uint8_t> msg(static_cast<const uint8_t*>(data), static_cast<const uint8_t*>(data) + n);
  msg.push_back(0x80);
  while (msg.size() % 64 != 56) msg.push_back(0);
  const uint64_t bits = uint64_t(n) * 8;
  for (int i = 7; i >= 0; i--) msg.push_back(uint8_t(bits >> (8 * i)));
  for (size_t off = 0; off < msg.size(); off += 64) {
    uint32_t w[80];
    for (int t = 0; t < 16; t++)
      w[t] = uint32_t(msg[off + 4 * t]) << 24 | uint32_t(msg[off + 4 * t + 1]) << 16 |
             uint32_t(msg[off + 4 * t + 2]) << 8 | msg[off + 4 * t + 3];
    for (int t = 16; t < 80; t++) w[t] = rol(w[t - 3] ^ w[t - 8] ^ w[t - 14] ^ w[t - 16], 1);
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4];
    for (int t = 0; t < 80; t++) {
      const uint32_t f = t < 20 ? (b & c) | (~b & d) : t < 40 ? b ^ c ^ d : t < 60 ? (b & c) | (b & d) | (c & d) : b ^ c ^ d;
      const uint32_t k = t < 20 ? 0x5A827999u : t < 40 ? 0x6ED9EBA1u : t < 60 ? 0x8F1BBCDCu : 0xCA62C1D6u;
      const uint32_t tmp = rol(a, 5) + f + e + k + w[t];
      e = d; d = c; c = rol(b, 30); b = a; a = tmp;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e;
  }
  char out[41];
  for (int i = 0; i < 5; i++) std::snprintf(out + 8 * i, 9, "%08x", h[i]);
  return std::string(out, 40);
}

namespace {
template <class T> std::string digest_u64(const T* a, int n) {
  std::vector<uint64_t> v(a, a + n);
  return sha1_hex(v.data(), v.size() * 8);
}
std::string digest_of(const Tables& T, const std::string& name) {
  if (name == "pack") return sha1_hex(T.pack, sizeof T.pack);
  if (name == "unpack") return sha1_hex(T.unpack, sizeof T.unpack);
  if (name == "win_contributions") return sha1_hex(T.win_contrib, sizeof T.win_contrib);
  if (name == "rotations") return sha1_hex(T.rot, sizeof T.rot);
  if (name == "reflections") return sha1_hex(T.refl, sizeof T.refl);
  if (name == "superwin_info") return sha1_hex(T.superwin, sizeof T.superwin);
  if (name == "commute_global_local_symmetries") return digest_u64(&T.commute[0][0], 8 * 256);
  if (name == "rotation_minimal_quadrants_offsets") return digest_u64(T.rmin_offsets, 56);
  if (name == "rotation_minimal_quadrants_flat") return digest_u64(T.rmin_flat, T.rmin_flat_size);
  if (name == "rotation_minimal_quadrants_inverse") return digest_u64(T.rmin_inverse, 19683);
  if (name == "rotation_minimal_quadrants_reflect_moved") return digest_u64(T.rmin_moved, 55);
  return "";
}
}  // namespace

const Tables& tables() {
  std::call_once(g_once, [] {
    auto* T = new Tables();
    build(*T);
    for (const auto& d : kDigests)
      if (digest_of(*T, d.name) != d.sha1)
        throw std::runtime_error(std::string("pentago_amd: table digest mismatch for ") + d.name);
    g_tables = T;
  });
  return *g_tables;
}

std::string table_digest(const std::string& name) { return digest_of(tables(), name); }

int rmin_bucket(int black, int white) {
  if (black < 0 || white < 0 || black + white > 9) return -1;
  return black * (21 - black) / 2 + white;
}
int rmin_count(int black, int white) {
  const int b = rmin_bucket(black, white);
  if (b < 0) return -1;
  const auto& T = tables();
  return T.rmin_offsets[b + 1] - T.rmin_offsets[b];
}
const uint16_t* rmin_list(int black, int white) {
  const int b = rmin_bucket(black, white);
  if (b < 0) return nullptr;
  return tables().rmin_flat + tables().rmin_offsets[b];
}

}  // namespace pe
