// Shared device/host helpers for the neuralangelo_amd HIP library (gfx950).
//
// Everything that crosses the C ABI is a plain pointer, an int64 size or one of
// the POD structs declared in include/neuralangelo_amd.h.  Errors are reported
// as an int status plus a thread-local message (na_last_error()).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/neuralangelo_amd.h"

namespace na {

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

// Launch tunables (na_tune_set / na_tune_get): shapes of launches whose best
// value depends on the device and the workload, chosen by the autotuner and
// read by the host launchers.  Values are atomics: set rarely, read per launch.
enum TuneKey : int {
  TUNE_TAPS_BWD_RUN = 0,   // consecutive samples per thread in the tap scatter
  TUNE_TIB_WG_PER_CU = 1,  // workgroups per CU of the tap-increment backward
  TUNE_PACKED_RAYS = 2,    // rays (threads) per workgroup of the packed compositing
  TUNE_TIB_FUSED = 3,      // tap-increment backward: split (0) or in the product's epilogue (1)
  TUNE_PACKED_WG_PER_CU = 4,  // workgroups per CU of the persistent packed launches
  TUNE_COUNT = 5
};
int64_t tune_value(TuneKey key);

#define NA_CHECK_ARG(cond, msg)                                              \
  do {                                                                       \
    if (!(cond)) {                                                           \
      ::na::set_error(std::string(__func__) + ": " + (msg));                 \
      return NA_ERR_ARGUMENT;                                                \
    }                                                                        \
  } while (0)

#define NA_CHECK_HIP(expr)                                                   \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::na::set_error(std::string(__func__) + ": " #expr " failed: " +       \
                      hipGetErrorString(_e));                                \
      return NA_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)

#define NA_CHECK_LAUNCH() NA_CHECK_HIP(hipGetLastError())

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// bf16 helpers (storage = uint16, round-to-nearest-even, NaN kept quiet)
// ---------------------------------------------------------------------------
typedef uint16_t bf16_t;

__host__ __device__ inline float bf16_to_f32(bf16_t h) {
  union { uint32_t u; float f; } v;
  v.u = ((uint32_t)h) << 16;
  return v.f;
}

__host__ __device__ inline bf16_t f32_to_bf16(float f) {
  union { uint32_t u; float f; } v;
  v.f = f;
  uint32_t u = v.u;
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) {
    return (bf16_t)((u >> 16) | 0x0040u);  // quiet NaN
  }
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (bf16_t)(u >> 16);
}

template <typename T> struct scalar_io;
template <> struct scalar_io<float> {
  __device__ static inline float load(const float* p) { return *p; }
  __device__ static inline void store(float* p, float v) { *p = v; }
};
template <> struct scalar_io<bf16_t> {
  __device__ static inline float load(const bf16_t* p) { return bf16_to_f32(*p); }
  __device__ static inline void store(bf16_t* p, float v) { *p = f32_to_bf16(v); }
};

// Hardware fp32 atomic add (global_atomic_add_f32, no CAS loop).
__device__ inline void atomic_add_f32(float* addr, float v) {
  unsafeAtomicAdd(addr, v);
}

}  // namespace na
