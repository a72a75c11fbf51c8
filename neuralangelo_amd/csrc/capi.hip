// Library-wide C ABI: error reporting, version, and host-side layout logic.
#include <math.h>

#include <string>

#include "common.h"

#include <atomic>
#include <cstring>

namespace na {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

struct Tunable {
  const char* name;
  int64_t lo, hi, step;  // allowed values: lo..hi, multiples of step
  std::atomic<int64_t> value;
};
static Tunable g_tune[TUNE_COUNT] = {
    {"taps_bwd.run", 1, 32, 1, {4}},
    {"taps_increment_bwd.wg_per_cu", 1, 32, 1, {8}},
    {"packed.rays_per_wg", 64, 256, 64, {64}},
    {"taps_increment_bwd.fused", 0, 1, 1, {0}},
    {"packed.wg_per_cu", 1, 16, 1, {4}},
};

int64_t tune_value(TuneKey key) { return g_tune[key].value.load(std::memory_order_relaxed); }

static int tune_index(const char* name) {
  if (!name) return -1;
  for (int k = 0; k < TUNE_COUNT; ++k)
    if (std::strcmp(g_tune[k].name, name) == 0) return k;
  return -1;
}
}  // namespace na

extern "C" int64_t na_tune_set(const char* name, int64_t value) {
  const int k = na::tune_index(name);
  NA_CHECK_ARG(k >= 0, std::string("unknown tunable ") + (name ? name : "(null)"));
  const na::Tunable& t = na::g_tune[k];
  NA_CHECK_ARG(value >= t.lo && value <= t.hi && value % t.step == 0,
               std::string(t.name) + ": value out of range [" + std::to_string(t.lo) + ", " +
                   std::to_string(t.hi) + "] or not a multiple of " + std::to_string(t.step));
  return na::g_tune[k].value.exchange(value);
}

extern "C" int64_t na_tune_get(const char* name) {
  const int k = na::tune_index(name);
  NA_CHECK_ARG(k >= 0, std::string("unknown tunable ") + (name ? name : "(null)"));
  return na::g_tune[k].value.load();
}

extern "C" int na_tune_count(void) { return na::TUNE_COUNT; }

extern "C" const char* na_tune_name(int32_t index) {
  return index >= 0 && index < na::TUNE_COUNT ? na::g_tune[index].name : nullptr;
}

extern "C" const char* na_last_error(void) { return na::g_last_error.c_str(); }

extern "C" int na_version(void) { return 1; }

extern "C" int na_device_synchronize(void) {
  NA_CHECK_HIP(hipDeviceSynchronize());
  return NA_OK;
}

// Hash-grid level table (tiny-cuda-nn GridEncoding constructor semantics; the
// reference passes per_level_scale from projects/neuralangelo/utils/modules.py:41):
//   log2s   = log2f(per_level_scale)                       (float32)
//   scale_l = exp2f(l * log2s) * base_resolution - 1       (float32, each op rounded)
//   res_l   = ceil(scale_l) + 1
//   size_l  = min(next_multiple_of_8(min(res_l^3, 2^31-1)), 2^log2_hashmap_size)
// The float32 steps are written as separate statements and compiled with
// -ffp-contract=off so no FMA fuses them.
#pragma clang fp contract(off)
extern "C" int64_t na_hashgrid_layout(int32_t n_levels, int32_t n_features,
                                      int32_t log2_hashmap_size, int32_t base_resolution,
                                      float per_level_scale, na_grid_desc* desc) {
  NA_CHECK_ARG(desc != nullptr, "desc is null");
  NA_CHECK_ARG(n_levels >= 1 && n_levels <= NA_MAX_LEVELS, "n_levels out of range [1, 32]");
  NA_CHECK_ARG(n_features == 1 || n_features == 2 || n_features == 4 || n_features == 8,
               "n_features must be 1, 2, 4 or 8");
  NA_CHECK_ARG(log2_hashmap_size >= 3 && log2_hashmap_size <= 28,
               "log2_hashmap_size out of range [3, 28]");
  NA_CHECK_ARG(base_resolution >= 1, "base_resolution must be >= 1");
  NA_CHECK_ARG(per_level_scale >= 1.0f, "per_level_scale must be >= 1");
  *desc = na_grid_desc{};
  desc->n_levels = n_levels;
  desc->n_features = n_features;
  desc->log2_hashmap_size = log2_hashmap_size;
  desc->base_resolution = base_resolution;
  desc->per_level_scale = per_level_scale;
  volatile float log2s = log2f(per_level_scale);
  desc->log2_per_level_scale = log2s;
  const uint64_t hash_size = 1ull << log2_hashmap_size;
  uint64_t offset = 0;
  for (int l = 0; l < n_levels; ++l) {
    volatile float a = (float)l * log2s;
    volatile float e = exp2f(a);
    volatile float b = e * (float)base_resolution;
    float scale = b - 1.0f;
    uint32_t res = (uint32_t)ceilf(scale) + 1u;
    uint64_t r = res;
    uint64_t params = r * r * r;
    if (params > 0x7fffffffull) params = 0x7fffffffull;
    params = (params + 7ull) / 8ull * 8ull;
    if (params > hash_size) params = hash_size;
    desc->scale[l] = scale;
    desc->resolution[l] = res;
    desc->offset[l] = (uint32_t)offset;
    offset += params;
    NA_CHECK_ARG(offset < 0xffffffffull, "table too large for 32-bit offsets");
  }
  desc->offset[n_levels] = (uint32_t)offset;
  return (int64_t)offset * n_features;
}
