// Stable key/value radix sort of (cell key, particle index) pairs.  The
// deposit and gather kernels are correct in any particle order; sorting the
// bunch by cell gives them cache-resident grid neighbourhoods.  LSD radix
// sort is stable, so equal keys keep their particle order and the
// permutation is unique.
#include <rocprim/device/device_radix_sort.hpp>

#include "kernels.hpp"

namespace opalsc {

size_t sort_temp_bytes(int64_t N) {
    size_t bytes = 0;
    if (N <= 0) return 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                    (const uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)N, 0, 32);
    return bytes;
}

int sort_pairs(hipStream_t s, void *tmp, size_t tmp_bytes, const uint32_t *kin, uint32_t *kout,
               const uint32_t *vin, uint32_t *vout, int64_t N, int end_bit) {
    if (N <= 0) return 0;
    size_t bytes = tmp_bytes;
    hipError_t e = rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, (size_t)N, 0, end_bit, s);
    return e == hipSuccess ? 0 : -1;
}

}  // namespace opalsc
