// Device-resident lookup tables and HIP error plumbing shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/pentago_amd.h"

namespace pe {

// Device copies of the lookup tables (uploaded once per device by pe_init).
struct DevTables {
  const uint64_t* superwin;     // 4*512*5 supers
  const uint16_t* rmin_flat;    // packed radix-3 rotation-minimal quadrants
  const uint32_t* rmin_sides;   // the same quadrants unpacked: side0 | side1 << 16
  const uint16_t* rmin_inverse; // 19683 entries
  const uint16_t* pack;         // 512 entries
  const uint8_t* commute;       // 8*256 entries
  const int32_t* choose;        // choose(n, k) at [n*32 + k], n,k < 32
  const uint8_t* rmin_stab;     // per rotation-minimal quadrant: log2 of its local stabilizer size
  const uint64_t* fields;       // field pool of the fused end kernels (end.h: field_pool_off)
};

// Upload the tables to the current device (idempotent per device).
pe_status device_tables(DevTables* out, std::string* err);
// The current device's tables if they are already uploaded (never allocates).
bool device_tables_ready(DevTables* out);

void set_error(const std::string& e);

// Grow-only device allocation reused across synchronous calls (callers serialize use and
// have no work pending on it when it grows).
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    const hipError_t e = hipMalloc(&ptr, bytes);
    if (e == hipSuccess) cap = bytes;
    return e;
  }
};
// Every C entry point's body runs inside PE_GUARD_BEGIN / PE_GUARD_END: no C++ exception
// (a failed host allocation, above all) crosses the C ABI; it becomes an error status.
#define PE_GUARD_BEGIN try {
#define PE_GUARD_END                                                                \
  }                                                                                 \
  catch (const std::bad_alloc&) {                                                   \
    ::pe::set_error("host memory exhausted");                                       \
    return PE_ERR_CAPACITY;                                                         \
  }                                                                                 \
  catch (const std::exception& e_) {                                                \
    ::pe::set_error(e_.what());                                                     \
    return PE_ERR_INVALID;                                                          \
  }
#define PE_HIP(expr, msg)                                                           \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::pe::set_error(std::string(msg) + ": " + hipGetErrorString(e_));             \
      return PE_ERR_DEVICE;                                                         \
    }                                                                               \
  } while (0)

}  // namespace pe
