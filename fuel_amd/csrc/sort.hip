// rocPRIM-backed helpers: exclusive scan of per-(query, split) counts and the
// stable segmented sort that orders range results by (distance, index).
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "kernels.h"

namespace fa {

// rocPRIM reports failures through its return value; a successful call must not
// leave a sticky runtime error behind for the next launch check to pick up.
static hipError_t settle(hipError_t e) {
  if (e == hipSuccess) (void)hipGetLastError();
  return e;
}

hipError_t scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t count, void* tmp, size_t& tmp_bytes,
                              hipStream_t st) {
  return settle(rocprim::exclusive_scan(tmp, tmp_bytes, in, out, int64_t(0), (size_t)count,
                                        rocprim::plus<int64_t>(), st));
}

// Input pairs arrive in index order within each segment, so a stable sort by
// the order-preserving key yields fuel's (distance, index) order.
template <typename K>
hipError_t segmented_sort_pairs(const K* keys_in, K* keys_out, const int64_t* vals_in, int64_t* vals_out,
                                int64_t nnz, int64_t nseg, const int64_t* offsets, void* tmp, size_t& tmp_bytes,
                                hipStream_t st) {
  return settle(rocprim::segmented_radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out,
                                                    (unsigned int)nnz, (unsigned int)nseg, offsets, offsets + 1, 0,
                                                    (unsigned int)(8 * sizeof(K)), st));
}

// Stable by construction (LSD radix sort): rows of one list keep their order.
hipError_t sort_pairs_u32(const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                          int64_t n, int bits, void* tmp, size_t& tmp_bytes, hipStream_t st) {
  return settle(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0,
                                          bits, st));
}

template hipError_t segmented_sort_pairs<uint32_t>(const uint32_t*, uint32_t*, const int64_t*, int64_t*, int64_t,
                                                   int64_t, const int64_t*, void*, size_t&, hipStream_t);
template hipError_t segmented_sort_pairs<uint64_t>(const uint64_t*, uint64_t*, const int64_t*, int64_t*, int64_t,
                                                   int64_t, const int64_t*, void*, size_t&, hipStream_t);

}  // namespace fa
