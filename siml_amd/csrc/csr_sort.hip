// csr_sort.hip -- per-row ascending sort of neighbour index lists (CSR).
// Needed when the query set is stored in a permuted order: each row's
// indices come out grouped by storage position, and the reference reports
// them in ascending column order (kernels.py:433-470).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "siml_internal.h"

namespace siml {

int sort_segments_i32(const int32_t* keys_in, int32_t* keys_out, int64_t n, int64_t nseg, const int64_t* d_offsets,
                      void* d_temp, size_t* temp_bytes, hipStream_t stream) {
    if (n > 0x7FFFFFFF || nseg > 0x7FFFFFFF)
        return fail(SIML_ELIMIT, "neighbour sort: more than 2^31-1 entries or rows");
    hipError_t e = hipcub::DeviceSegmentedSort::SortKeys(d_temp, *temp_bytes, keys_in, keys_out, (int)n, (int)nseg,
                                                         d_offsets, d_offsets + 1, stream);
    if (e != hipSuccess) return fail(SIML_EDEVICE, "segmented sort failed: %s", hipGetErrorString(e));
    return SIML_OK;
}

}  // namespace siml
