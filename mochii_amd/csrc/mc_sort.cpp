// Walk order of a batch of launch records: radix sort of the 32-bit
// direction keys written by the launch kernel, producing the permutation the
// walk reads its records through (rocPRIM device radix sort).
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "mc_internal.h"

size_t mc_sort_temp_bytes(int64_t n)
{
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                              rocprim::counting_iterator<int32_t>(0), (int32_t *)nullptr,
                              (unsigned int)n, 0, 32);
    return bytes;
}

int mc_sort_records(mc_ctx *c, int64_t n)
{
    size_t bytes = c->sort_tmp_bytes;
    MC_HIP(rocprim::radix_sort_pairs(c->d_sort_tmp, bytes, c->pk.key, c->d_key_sorted,
                                     rocprim::counting_iterator<int32_t>(0), c->d_perm,
                                     (unsigned int)n, 0, 32, c->stream));
    return 0;
}
