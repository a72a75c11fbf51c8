// Stable LSD radix sort of (uint32 key, uint32 value) pairs on a stream, for
// the deterministic table-gradient reduction.  A thin wrapper over rocPRIM's
// device radix sort (through its hipCUB interface): the caller brings the
// temporary storage, sized by radix_sort_temp_bytes, so nothing is allocated.
#include <hipcub/hipcub.hpp>

#include "common.h"

namespace na {

int radix_sort_temp_bytes(int64_t count, size_t* bytes) {
  *bytes = 0;
  if (count <= 0) return NA_OK;
  NA_CHECK_ARG(count <= 0x7fffffff, "radix sort: too many items");
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      nullptr, *bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (const uint32_t*)nullptr,
      (uint32_t*)nullptr, (int)count, 0, 32, (hipStream_t)0);
  if (e != hipSuccess) {
    set_error(std::string("radix sort size query: ") + hipGetErrorString(e));
    return NA_ERR_HIP;
  }
  return NA_OK;
}

int radix_sort_pairs_u32(void* temp, size_t temp_bytes, const uint32_t* keys_in,
                         uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                         int64_t count, int end_bit, hipStream_t s) {
  if (count <= 0) return NA_OK;
  NA_CHECK_ARG(count <= 0x7fffffff, "radix sort: too many items");
  NA_CHECK_ARG(end_bit >= 1 && end_bit <= 32, "radix sort: end_bit out of range");
  size_t bytes = temp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys_in, keys_out, vals_in,
                                                    vals_out, (int)count, 0, end_bit, s);
  if (e != hipSuccess) {
    set_error(std::string("radix sort: ") + hipGetErrorString(e));
    return NA_ERR_HIP;
  }
  return NA_OK;
}

}  // namespace na
