// Supertensor block data <-> device nodes (pentago/data/supertensor.h:1-50 file format,
// pentago/data/filter.h interleave filter, pentago/end/check.cc:149-158 orientation).
//
// A .pentago file stores per node the pair (black wins, white wins); filter 1 interleaves the
// two supers bit by bit so that in each 64-bit output word bit 2i is black's and bit 2i+1
// white's: output words 2k and 2k+1 hold the low and high halves of word k of both supers.
// Device nodes are (we-win, we-win-or-tie) for the player to move: with turn = stones & 1,
// black = turn ? ~we-win-or-tie : we-win and white = turn ? we-win : ~we-win-or-tie.
//
// One thread per node: 64 bytes read, 64 bytes written, in place allowed.  HBM-bound.
#include <hip/hip_runtime.h>

#include <string>

#include "device.h"

namespace pe {
namespace {

constexpr int kFormatThreads = 256;

__device__ inline uint64_t spread32(uint64_t a) {   // bit i -> bit 2i, i < 32
  a &= 0x00000000ffffffffull;
  a = (a | a << 16) & 0x0000ffff0000ffffull;
  a = (a | a << 8) & 0x00ff00ff00ff00ffull;
  a = (a | a << 4) & 0x0f0f0f0f0f0f0f0full;
  a = (a | a << 2) & 0x3333333333333333ull;
  a = (a | a << 1) & 0x5555555555555555ull;
  return a;
}

__device__ inline uint64_t gather32(uint64_t a) {   // bit 2i -> bit i
  a &= 0x5555555555555555ull;
  a = (a | a >> 1) & 0x3333333333333333ull;
  a = (a | a >> 2) & 0x0f0f0f0f0f0f0f0full;
  a = (a | a >> 4) & 0x00ff00ff00ff00ffull;
  a = (a | a >> 8) & 0x0000ffff0000ffffull;
  a = (a | a >> 16) & 0x00000000ffffffffull;
  return a;
}

// decode: file words (filtered (black, white)) -> (we-win, we-win-or-tie)
__global__ __launch_bounds__(kFormatThreads) void decode_kernel(const uint64_t* in, uint64_t* out,
                                                                int64_t nodes, int filter, int turn) {
  const int64_t i = int64_t(blockIdx.x) * kFormatThreads + threadIdx.x;
  if (i >= nodes) return;
  const ulonglong2* src = reinterpret_cast<const ulonglong2*>(in + 8 * i);
  uint64_t w[8];
  for (int k = 0; k < 4; k++) {
    const ulonglong2 v = src[k];
    w[2 * k] = v.x;
    w[2 * k + 1] = v.y;
  }
  uint64_t black[4], white[4];
  if (filter == 1) {
    for (int k = 0; k < 4; k++) {
      black[k] = gather32(w[2 * k]) | gather32(w[2 * k + 1]) << 32;
      white[k] = gather32(w[2 * k] >> 1) | gather32(w[2 * k + 1] >> 1) << 32;
    }
  } else {
    for (int k = 0; k < 4; k++) {
      black[k] = w[k];
      white[k] = w[4 + k];
    }
  }
  uint64_t o[8];
  for (int k = 0; k < 4; k++) {
    o[k] = turn ? white[k] : black[k];
    o[4 + k] = ~(turn ? black[k] : white[k]);
  }
  ulonglong2* dst = reinterpret_cast<ulonglong2*>(out + 8 * i);
  for (int k = 0; k < 4; k++) dst[k] = make_ulonglong2(o[2 * k], o[2 * k + 1]);
}

// encode: (we-win, we-win-or-tie) -> file words
__global__ __launch_bounds__(kFormatThreads) void encode_kernel(const uint64_t* in, uint64_t* out,
                                                                int64_t nodes, int filter, int turn) {
  const int64_t i = int64_t(blockIdx.x) * kFormatThreads + threadIdx.x;
  if (i >= nodes) return;
  const ulonglong2* src = reinterpret_cast<const ulonglong2*>(in + 8 * i);
  uint64_t w[8];
  for (int k = 0; k < 4; k++) {
    const ulonglong2 v = src[k];
    w[2 * k] = v.x;
    w[2 * k + 1] = v.y;
  }
  uint64_t black[4], white[4];
  for (int k = 0; k < 4; k++) {
    black[k] = turn ? ~w[4 + k] : w[k];
    white[k] = turn ? w[k] : ~w[4 + k];
  }
  uint64_t o[8];
  if (filter == 1) {
    for (int k = 0; k < 4; k++) {
      o[2 * k] = spread32(black[k]) | spread32(white[k]) << 1;
      o[2 * k + 1] = spread32(black[k] >> 32) | spread32(white[k] >> 32) << 1;
    }
  } else {
    for (int k = 0; k < 4; k++) {
      o[k] = black[k];
      o[4 + k] = white[k];
    }
  }
  ulonglong2* dst = reinterpret_cast<ulonglong2*>(out + 8 * i);
  for (int k = 0; k < 4; k++) dst[k] = make_ulonglong2(o[2 * k], o[2 * k + 1]);
}

}  // namespace

pe_status launch_supertensor_format(bool decode, int filter, int turn, const void* in, int64_t nodes, void* out,
                                    hipStream_t stream, std::string* err) {
  if (filter != 0 && filter != 1) { *err = "unknown supertensor filter"; return PE_ERR_INVALID; }
  if (nodes < 0 || (nodes && (!in || !out))) { *err = "bad supertensor buffer arguments"; return PE_ERR_INVALID; }
  if (!nodes) return PE_OK;
  const int64_t blocks = (nodes + kFormatThreads - 1) / kFormatThreads;
  if (blocks > 0x7fffffff) { *err = "supertensor buffer too large"; return PE_ERR_CAPACITY; }
  hipLaunchKernelGGL(decode ? decode_kernel : encode_kernel, dim3(unsigned(blocks)), dim3(kFormatThreads), 0, stream,
                     static_cast<const uint64_t*>(in), static_cast<uint64_t*>(out), nodes, filter, turn & 1);
  if (hipGetLastError() != hipSuccess) { *err = "supertensor format kernel launch failed"; return PE_ERR_DEVICE; }
  return PE_OK;
}

}  // namespace pe
