// pt_sort.hip -- stable key/value radix sort (rocPRIM) used to order tracer
// indices by cell.  LSD radix sorting is stable, so tracers of one cell keep
// ascending index order: the projection's summation order is fixed.
#include <cstring>
#include <rocprim/rocprim.hpp>
#include "pt_internal.hpp"

namespace pt {

int sort_pairs_temp_bytes(int64_t n, int bits, size_t* bytes) {
  size_t b = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, b, (const int32_t*)nullptr, (int32_t*)nullptr,
                                           (const int32_t*)nullptr, (int32_t*)nullptr, (unsigned)n, 0, bits);
  if (e != hipSuccess) {
    set_error(std::string("rocprim::radix_sort_pairs (size query): ") + hipGetErrorString(e));
    return PT_ERR_HIP;
  }
  *bytes = b;
  return PT_OK;
}

int sort_pairs(void* tmp, size_t bytes, const int32_t* kin, int32_t* kout, const int32_t* vin, int32_t* vout,
               int64_t n, int bits, hipStream_t s) {
  size_t b = bytes;
  hipError_t e = rocprim::radix_sort_pairs(tmp, b, kin, kout, vin, vout, (unsigned)n, 0, bits, s);
  if (e != hipSuccess) {
    set_error(std::string("rocprim::radix_sort_pairs: ") + hipGetErrorString(e));
    return PT_ERR_HIP;
  }
  return PT_OK;
}

}  // namespace pt
