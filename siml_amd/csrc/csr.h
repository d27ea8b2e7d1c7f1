// csr.h -- device-side assembly of neighbour lists (csr.hip), shared by the
// HIP translation units of libsiml_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace siml {
// Device scratch of csr_scan over nslots slots.
size_t csr_scan_bytes(int64_t nslots);
// Exclusive scan of the counts of nr rows x nw slots (row-major) into slot
// offsets offs (nr * nw) and row offsets rows (nr + 1).  tmp (csr_scan_bytes)
// begins with three int64 totals: {entries, largest slot count, longest
// row}, complete once the stream reaches them.
int csr_scan(const int32_t* counts, int64_t nr, int64_t nw, int64_t* offs, int64_t* rows, void* tmp,
             hipStream_t stream);
// Sorts the indices of each row [rows[r], rows[r+1]) ascending (indices of a
// row distinct), using b (nnz entries) as the second buffer; *sorted is set
// to whichever of a and b holds the result.  longest: the longest row.
int csr_sort_rows(int32_t* a, int32_t* b, int64_t nnz, int64_t nr, const int64_t* rows, int64_t longest,
                  int32_t** sorted, hipStream_t stream);

}  // namespace siml
