// csr.hip -- device-side assembly of neighbour lists in CSR form.
//
//  csr_scan       per-(row, wave) slot counts -> slot offsets, row offsets
//                 and the three totals the host needs to size the index
//                 array, all computed on the device;
//  csr_sort_rows  ascending sort of each row's indices, needed when the
//                 query set is stored in a permuted order: each row's
//                 indices come out grouped by storage position, and the
//                 reference reports them in ascending column order
//                 (kernels.py:433-470).
//
// The sort is a two-level merge sort.  Workgroups sort fixed tiles of
// kTile consecutive entries in LDS (bitonic network over the key
// (row - first row of the tile, index), so every row that lies inside a tile
// is sorted outright and every longer row becomes runs that end at tile
// boundaries).  Merge passes then merge a row's runs pairwise, counted from
// the row's own first run: pass p merges runs [g, g + 2^p) with
// [g + 2^p, g + 2^(p+1)), g a multiple of 2^(p+1).  Every entry finds its
// output position by a binary search in the partner run; indices within a
// row are distinct, so positions never collide.  A row of L entries spans at
// most (L - 1) / kTile + 2 tiles, which fixes the pass count from the
// longest row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <utility>

#include "csr.h"
#include "siml_internal.h"

namespace siml {
namespace {

constexpr int kB = 256;                          // threads per workgroup
constexpr int kWaves = kB / 64;
constexpr int kRounds = 16;                      // scan: kB-entry rounds per workgroup
constexpr int64_t kChunk = (int64_t)kB * kRounds;
constexpr int kTile = 4096;                      // sort: entries per LDS tile

#define CSR_TRY(expr)              \
    do {                           \
        const int rc_ = (expr);    \
        if (rc_ != SIML_OK) return rc_; \
    } while (0)

__device__ __forceinline__ int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }
__device__ __forceinline__ int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }

__device__ __forceinline__ int64_t wave_inclusive(int64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// Exclusive prefix sum of v over the workgroup; *total receives the sum.
// ws holds kWaves partial sums and may be reused after the call.
__device__ int64_t block_exclusive(int64_t v, int64_t* ws, int64_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t inc = wave_inclusive(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    int64_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
        pre += i < w ? ws[i] : 0;
        tot += ws[i];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

__device__ int64_t block_max(int64_t v, int64_t* ws) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = imax(v, __shfl_xor(v, d, 64));
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    int64_t m = ws[0];
#pragma unroll
    for (int i = 1; i < kWaves; ++i) m = imax(m, ws[i]);
    __syncthreads();
    return m;
}

// Pass 1: sum and largest count of each kChunk-entry chunk.
__global__ __launch_bounds__(kB) void k_scan_reduce(const int32_t* __restrict__ counts, int64_t n,
                                                    int64_t* __restrict__ bsum, int64_t* __restrict__ bmax) {
    __shared__ int64_t ws[kWaves];
    const int64_t base = (int64_t)blockIdx.x * kChunk;
    int64_t s = 0, m = 0;
    for (int j = 0; j < kRounds; ++j) {
        const int64_t i = base + (int64_t)j * kB + threadIdx.x;
        const int64_t c = i < n ? counts[i] : 0;
        s += c;
        m = imax(m, c);
    }
    int64_t tot;
    block_exclusive(s, ws, &tot);
    m = block_max(m, ws);
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = tot;
        bmax[blockIdx.x] = m;
    }
}

// Pass 2 (one workgroup): chunk sums -> chunk offsets, and the totals
// stats = {entries, largest slot count, 0 (longest row, filled by pass 4)}.
__global__ __launch_bounds__(kB) void k_scan_top(int64_t* __restrict__ bsum, const int64_t* __restrict__ bmax,
                                                 int64_t nb, int64_t* __restrict__ stats,
                                                 int64_t* __restrict__ rows, int64_t nr) {
    __shared__ int64_t ws[kWaves];
    int64_t carry = 0, m = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kB) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t v = i < nb ? bsum[i] : 0;
        m = imax(m, i < nb ? bmax[i] : 0);
        int64_t tot;
        const int64_t ex = block_exclusive(v, ws, &tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
    m = block_max(m, ws);
    if (threadIdx.x == 0) {
        stats[0] = carry;
        stats[1] = m;
        stats[2] = 0;
        rows[nr] = carry;
    }
}

// Pass 3: slot offsets; the first slot of each row gives the row offset.
__global__ __launch_bounds__(kB) void k_scan_down(const int32_t* __restrict__ counts, int64_t n, int64_t nw,
                                                  const int64_t* __restrict__ boff, int64_t* __restrict__ offs,
                                                  int64_t* __restrict__ rows) {
    __shared__ int64_t ws[kWaves];
    const int64_t base = (int64_t)blockIdx.x * kChunk;
    int64_t carry = boff[blockIdx.x];
    for (int j = 0; j < kRounds; ++j) {
        const int64_t i = base + (int64_t)j * kB + threadIdx.x;
        const int64_t c = i < n ? counts[i] : 0;
        int64_t tot;
        const int64_t o = carry + block_exclusive(c, ws, &tot);
        if (i < n) {
            offs[i] = o;
            if (i % nw == 0) rows[i / nw] = o;
        }
        carry += tot;
    }
}

// Pass 4: longest row into stats[2].
__global__ __launch_bounds__(kB) void k_row_max(const int64_t* __restrict__ rows, int64_t nr,
                                                int64_t* __restrict__ stats) {
    __shared__ int64_t ws[kWaves];
    int64_t m = 0;
    for (int64_t r = (int64_t)blockIdx.x * kB + threadIdx.x; r < nr; r += (int64_t)gridDim.x * kB)
        m = imax(m, rows[r + 1] - rows[r]);
    m = block_max(m, ws);
    if (threadIdx.x == 0 && m > 0)
        atomicMax(reinterpret_cast<unsigned long long*>(stats + 2), (unsigned long long)m);
}

// Largest r in [lo, hi] with rows[r] <= i, given rows[lo] <= i: the
// non-empty row that holds entry i.
__device__ __forceinline__ int64_t row_of(const int64_t* __restrict__ rows, int64_t lo, int64_t hi, int64_t i) {
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (rows[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Sorts each tile of kTile entries by (row, index) in LDS.
__global__ __launch_bounds__(kB) void k_sort_tiles(int32_t* __restrict__ a, int64_t nnz,
                                                   const int64_t* __restrict__ rows, int64_t nr) {
    __shared__ uint64_t key[kTile];
    __shared__ int64_t rr[2];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    const int cnt = (int)imin(kTile, nnz - base);
    if (threadIdx.x == 0) {
        rr[0] = row_of(rows, 0, nr, base);
        rr[1] = row_of(rows, rr[0], nr, base + cnt - 1);
    }
    __syncthreads();
    const int64_t r0 = rr[0], r1 = rr[1];
    int n2 = 64;
    while (n2 < cnt) n2 <<= 1;
    for (int j = threadIdx.x; j < n2; j += kB) {
        uint64_t k = ~0ull;
        if (j < cnt) {
            const int64_t i = base + j;
            const int64_t r = r0 == r1 ? r0 : row_of(rows, r0, r1, i);
            k = ((uint64_t)(r - r0) << 32) | (uint32_t)a[i];
        }
        key[j] = k;
    }
    __syncthreads();
    for (int size = 2; size <= n2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < n2 / 2; t += kB) {
                const int lo = 2 * t - (t & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t x = key[lo], y = key[hi];
                if ((x > y) == up) {
                    key[lo] = y;
                    key[hi] = x;
                }
            }
            __syncthreads();
        }
    for (int j = threadIdx.x; j < cnt; j += kB) a[base + j] = (int32_t)(uint32_t)key[j];
}

// Entries of the sorted range [lo, hi) below x.
__device__ __forceinline__ int64_t count_below(const int32_t* __restrict__ v, int64_t lo, int64_t hi, int32_t x) {
    const int64_t lo0 = lo;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (v[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - lo0;
}

// One merge pass: runs [g, g + 2^p) and [g + 2^p, g + 2^(p+1)) of each row,
// run 0 starting at the row's first entry and run K > 0 at its K-th tile
// boundary.
__global__ __launch_bounds__(kB) void k_merge_runs(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                   int64_t nnz, const int64_t* __restrict__ rows, int64_t nr, int p) {
    __shared__ int64_t rr[2];
    const int64_t base = (int64_t)blockIdx.x * kB;
    if (threadIdx.x == 0) {
        rr[0] = row_of(rows, 0, nr, base);
        rr[1] = row_of(rows, rr[0], nr, imin(base + kB, nnz) - 1);
    }
    __syncthreads();
    const int64_t i = base + threadIdx.x;
    if (i >= nnz) return;
    const int64_t r = row_of(rows, rr[0], rr[1], i);
    const int64_t rs = rows[r], re = rows[r + 1];
    const int64_t t0 = rs / kTile;
    auto run_start = [&](int64_t K) -> int64_t { return K == 0 ? rs : imin(re, (t0 + K) * kTile); };
    const int64_t k = i / kTile - t0;
    const int64_t g = (k >> (p + 1)) << (p + 1);
    const int64_t a0 = run_start(g), a1 = run_start(g + (int64_t(1) << p)), b1 = run_start(g + (int64_t(2) << p));
    const int32_t x = in[i];
    int64_t pos = i;
    if (i < a1)
        pos = i + count_below(in, a1, b1, x);
    else
        pos = a0 + (i - a1) + count_below(in, a0, a1, x);
    out[pos] = x;
}

int launch_check(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SIML_EDEVICE, "%s launch failed: %s", what, hipGetErrorString(e));
    return SIML_OK;
}

}  // namespace

size_t csr_scan_bytes(int64_t nslots) {
    const int64_t nb = (nslots + kChunk - 1) / kChunk;
    return 256 + ((2 * sizeof(int64_t) * (size_t)nb + 255) & ~size_t(255));
}

int csr_scan(const int32_t* counts, int64_t nr, int64_t nw, int64_t* offs, int64_t* rows, void* tmp,
             hipStream_t stream) {
    if (nr < 1 || nw < 1) return fail(SIML_EINVAL, "csr_scan: %lld rows of %lld slots", (long long)nr, (long long)nw);
    const int64_t n = nr * nw;
    const int64_t nb = (n + kChunk - 1) / kChunk;
    if (nb > 0x7FFFFFFF) return fail(SIML_ELIMIT, "csr_scan: %lld slots", (long long)n);
    int64_t* stats = static_cast<int64_t*>(tmp);
    int64_t* bsum = reinterpret_cast<int64_t*>(static_cast<char*>(tmp) + 256);
    int64_t* bmax = bsum + nb;
    k_scan_reduce<<<dim3((unsigned)nb), dim3(kB), 0, stream>>>(counts, n, bsum, bmax);
    CSR_TRY(launch_check("k_scan_reduce"));
    k_scan_top<<<dim3(1), dim3(kB), 0, stream>>>(bsum, bmax, nb, stats, rows, nr);
    CSR_TRY(launch_check("k_scan_top"));
    k_scan_down<<<dim3((unsigned)nb), dim3(kB), 0, stream>>>(counts, n, nw, bsum, offs, rows);
    CSR_TRY(launch_check("k_scan_down"));
    const int64_t rb = std::min<int64_t>((nr + kB - 1) / kB, 2048);
    k_row_max<<<dim3((unsigned)rb), dim3(kB), 0, stream>>>(rows, nr, stats);
    return launch_check("k_row_max");
}

int csr_sort_rows(int32_t* a, int32_t* b, int64_t nnz, int64_t nr, const int64_t* rows, int64_t longest,
                  int32_t** sorted, hipStream_t stream) {
    *sorted = a;
    if (nnz > 0x7FFFFFFF || nr > 0x7FFFFFFF)
        return fail(SIML_ELIMIT, "neighbour sort: more than 2^31-1 entries or rows");
    if (nnz < 2 || longest < 2) return SIML_OK;
    k_sort_tiles<<<dim3((unsigned)((nnz + kTile - 1) / kTile)), dim3(kB), 0, stream>>>(a, nnz, rows, nr);
    CSR_TRY(launch_check("k_sort_tiles"));
    const int64_t runs = (longest - 1) / kTile + 2;
    int32_t *src = a, *dst = b;
    for (int p = 0; (int64_t(1) << p) < runs; ++p) {
        k_merge_runs<<<dim3((unsigned)((nnz + kB - 1) / kB)), dim3(kB), 0, stream>>>(src, dst, nnz, rows, nr, p);
        CSR_TRY(launch_check("k_merge_runs"));
        std::swap(src, dst);
    }
    *sorted = src;
    return SIML_OK;
}

}  // namespace siml
