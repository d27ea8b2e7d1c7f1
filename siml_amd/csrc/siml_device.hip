// siml_device.hip -- MI355X (gfx950) LINGO similarity kernels and the C ABI
// declared in include/siml_amd.h.
//
// Data layout in HBM
//   reference set (rows, staged in LDS one row per workgroup):
//     elems  int2 (id, count), molecule r at [off[r], off[r] + len[r])
//   query set (columns, one molecule per lane):
//     tiles  int2 (id, count) in tiles of 64 molecules; tile t holds
//            tlen[t] = max length in the tile steps, step j of lane l at
//            toff[t] + 64*j + l, so a wave's load of step j is one
//            contiguous 512-byte segment.  Steps past a molecule's length
//            are (0, 0) padding and are never read by the merge.
//   per molecule: len (unique Lingos), mag (magnitude), csum (sum of counts).
//
// Arithmetic (bit-exact with _CPULingo.c:53-71): the intersection is the
// integer sum of min(count) over equal ids; T = isct / ((float)ma + mb - isct)
// with one correctly rounded float32 division (built with
// -fhip-fp32-correctly-rounded-divide-sqrt and -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "groups.h"
#include "siml_internal.h"

namespace siml {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace siml

using siml::fail;

#define HIP_TRY(expr)                                                                                           \
    do {                                                                                                        \
        hipError_t e_ = (expr);                                                                                 \
        if (e_ != hipSuccess)                                                                                   \
            return fail(SIML_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

#define TRY(expr)              \
    do {                       \
        int rc_ = (expr);      \
        if (rc_ != SIML_OK) return rc_; \
    } while (0)

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // 4 waves per workgroup

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

struct RowArgs {
    const int2* elems;
    const int64_t* off;
    const int32_t* len;
    const int32_t* mag;
    const int32_t* csum;
};

struct ColArgs {
    const int2* tiles;
    const int64_t* toff;
    const int32_t* tlen;
    const int32_t* len;
    const int32_t* mag;
    const int32_t* csum;
    int64_t n;
    int64_t ntiles;
};

// multisetTanimoto3_mags' float expression (_CPULingo.c:55,70).
__device__ __forceinline__ float tanimoto_mags(int isct, int ma, int mb) {
    if (ma == 0 || mb == 0) return 0.0f;
    const float u = ((float)ma + (float)mb) - (float)isct;
    return (float)isct / u;
}

// deviceMultisetTanimoto3_bT_tex2's expression (kernels.py:136,157-158):
// zero test on lengths, union formed in int32 (wrapping) then converted.
__device__ __forceinline__ float tanimoto_recomputed(int isct, int La, int Lb, int amag, int bmag) {
    if (La == 0 || Lb == 0) return 0.0f;
    const int u = (int)((unsigned)amag + (unsigned)bmag - (unsigned)isct);
    return (float)isct / (float)u;
}

// Sorted-merge intersection of an LDS-staged row `a` (La entries) with one
// query molecule read with stride 64 from its tile (Lb entries).  The
// j-outer form performs the same state transitions as the three-way merge
// of _CPULingo.c:59-69 on any input, sorted or not.
__device__ __forceinline__ int merge_isct(const int2* __restrict__ a, int La, const int2* __restrict__ p, int Lb) {
    int i = 0, isct = 0;
    for (int j = 0; j < Lb; ++j) {
        const int2 b = p[(size_t)j * kWave];
        while (i < La && a[i].x < b.x) ++i;
        if (i < La && a[i].x == b.x) {
            isct += min(a[i].y, b.y);
            ++i;
        }
    }
    return isct;
}

__device__ __forceinline__ void stage_row(int2* a, const int2* __restrict__ src, int La) {
    for (int k = threadIdx.x; k < La; k += blockDim.x) a[k] = src[k];
}

// Tiles [t0, t1) of the global wave `gw` out of `nw`: contiguous ranges, so
// per-wave outputs concatenate in ascending column order.
__device__ __forceinline__ void wave_tiles(int64_t ntiles, int64_t& t0, int64_t& t1, int64_t& gw, int64_t& nw) {
    const int wpb = blockDim.x / kWave;
    gw = (int64_t)blockIdx.y * wpb + (threadIdx.x / kWave);
    nw = (int64_t)gridDim.y * wpb;
    t0 = (ntiles * gw) / nw;
    t1 = (ntiles * (gw + 1)) / nw;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Number of set bits of `m` below this lane.
__device__ __forceinline__ int rank_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Bin of doHistogram / gpuHistogrammedLingo (kernels.py:657,
// _cpuutil.cpp:65-66); out-of-range products take the documented end bins.
__device__ __forceinline__ int bin101_f32(float t) {
    const float p = t * 100.0f;
    if (!(p >= 0.0f)) return 0;
    if (p >= 100.0f) return 100;
    return (int)(unsigned)p;
}
__device__ __forceinline__ int bin101_f64(double t) {
    const double p = t * 100.0;
    if (!(p >= 0.0)) return 0;
    if (p >= 100.0) return 100;
    return (int)(unsigned)p;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// gpuLingoSim3_multirow (kernels.py:315-352) re-designed for CDNA4: one
// workgroup per reference row (grid.x), query tiles split across waves
// (grid.y * 4 waves), one query molecule per lane, coalesced 512-byte step
// loads from the tile layout.  RECOMPUTE selects the single-row semantics of
// gpuLingoSim3_srgf_T_tex2ns (kernels.py:234-270).
template <bool RECOMPUTE>
__global__ __launch_bounds__(kBlock) void k_rows_f32(RowArgs R, ColArgs C, const int32_t* __restrict__ rowlist,
                                                     int64_t rowbase, float* __restrict__ out, int64_t opitch) {
    extern __shared__ int2 lds_row[];
    const int64_t row = rowlist ? (int64_t)rowlist[blockIdx.x] : rowbase + blockIdx.x;
    const int La = R.len[row];
    const int ma = RECOMPUTE ? R.csum[row] : R.mag[row];
    stage_row(lds_row, R.elems + R.off[row], La);
    __syncthreads();
    int64_t t0, t1, gw, nw;
    wave_tiles(C.ntiles, t0, t1, gw, nw);
    const int lane = lane_id();
    float* orow = out + (row - rowbase) * opitch;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t c = t * kWave + lane;
        if (c >= C.n) continue;
        const int Lb = C.len[c];
        const int isct = merge_isct(lds_row, La, C.tiles + C.toff[t] + lane, Lb);
        orow[c] = RECOMPUTE ? tanimoto_recomputed(isct, La, Lb, ma, C.csum[c]) : tanimoto_mags(isct, ma, C.mag[c]);
    }
}

// ---------------------------------------------------------------------------
// Hashed multi-row kernels (row groups, see groups.h)
// ---------------------------------------------------------------------------

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

struct GroupArgs {
    const siml::GroupHdr* hdr;
    const int2* ekey;   // (id, slot) per entry
    const uint4* ecnt;  // eight u16 counts per entry
};

// acc += min(cv, cb2) on two u16 lanes (v_pk_min_u16 + v_pk_add_u16).
__device__ __forceinline__ uint32_t pk_minadd(uint32_t acc, uint32_t cv, uint32_t cb2) {
    us2 a = __builtin_bit_cast(us2, acc);
    a += __builtin_elementwise_min(__builtin_bit_cast(us2, cv), __builtin_bit_cast(us2, cb2));
    return __builtin_bit_cast(uint32_t, a);
}

struct GroupLds {
    int32_t keys[siml::kSlots];
    uint4 cnt[siml::kSlots];
};

// Fills the group's cuckoo table into LDS: empty slots hold empty_key with
// all-zero counts (slot 0 always stays empty).
__device__ __forceinline__ void stage_group(GroupLds& T, const GroupArgs& G, const siml::GroupHdr& hd) {
    for (int s = threadIdx.x; s < siml::kSlots; s += blockDim.x) {
        T.keys[s] = hd.empty_key;
        T.cnt[s] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < hd.nkeys; k += blockDim.x) {
        const int2 e = G.ekey[hd.entry_off + k];
        T.keys[e.y] = e.x;
        T.cnt[e.y] = G.ecnt[hd.entry_off + k];
    }
    __syncthreads();
}

// Intersections of one query molecule (a lane) with all rows of the group:
// acc[w] holds rows 2w (low u16) and 2w+1 (high u16).  The loop runs over
// the tile's uniform step count; padding entries carry count 0 and add 0.
__device__ __forceinline__ void group_isct(const GroupLds& T, const siml::GroupHdr& hd, const int2* __restrict__ p,
                                           int Lt, uint32_t acc[4]) {
    acc[0] = acc[1] = acc[2] = acc[3] = 0u;
    const uint32_t a1 = hd.a1, a2 = hd.a2;
    for (int j = 0; j < Lt; ++j, p += kWave) {
        const int2 b = *p;
        const uint32_t x = (uint32_t)b.x;
        const uint32_t y = x ^ ((x >> 13) | (x << 19));
        const uint32_t h1 = __umul24(y, a1) >> (32 - siml::kSlotBits);
        const uint32_t h2 = __umul24(y, a2) >> (32 - siml::kSlotBits);
        const uint32_t s = (T.keys[h1] == b.x) ? h1 : ((T.keys[h2] == b.x) ? h2 : 0u);
        const uint4 cv = T.cnt[s];
        uint32_t cb = min((uint32_t)b.y, siml::kCountMax);
        cb |= cb << 16;
        acc[0] = pk_minadd(acc[0], cv.x, cb);
        acc[1] = pk_minadd(acc[1], cv.y, cb);
        acc[2] = pk_minadd(acc[2], cv.z, cb);
        acc[3] = pk_minadd(acc[3], cv.w, cb);
    }
}

__device__ __forceinline__ int acc_row(const uint32_t acc[4], int r) {
    return (int)((acc[r >> 1] >> (16 * (r & 1))) & 0xFFFFu);
}

// Multi-row Tanimoto block: one workgroup per row group (grid.x), query
// tiles split across waves (grid.y * 4).  Each query entry is loaded once
// and probed once for all rows of the group.
__global__ __launch_bounds__(kBlock) void k_hash_rows_f32(GroupArgs G, int64_t g0, const int32_t* __restrict__ rmag,
                                                          ColArgs C, int64_t rb, int64_t rl, float* __restrict__ out,
                                                          int64_t opitch) {
    __shared__ GroupLds T;
    const siml::GroupHdr hd = G.hdr[g0 + blockIdx.x];
    stage_group(T, G, hd);
    int64_t t0, t1, gw, nw;
    wave_tiles(C.ntiles, t0, t1, gw, nw);
    const int lane = lane_id();
    int ma[siml::kGroupRows];
#pragma unroll
    for (int r = 0; r < siml::kGroupRows; ++r) ma[r] = r < hd.nrows ? rmag[hd.row0 + r] : 0;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t c = t * kWave + lane;
        uint32_t acc[4];
        group_isct(T, hd, C.tiles + C.toff[t] + lane, C.tlen[t], acc);
        if (c >= C.n) continue;
        const int mb = C.mag[c];
#pragma unroll
        for (int r = 0; r < siml::kGroupRows; ++r) {
            const int64_t row = (int64_t)hd.row0 + r;
            if (r < hd.nrows && row >= rb && row < rl)
                out[(row - rb) * opitch + c] = tanimoto_mags(acc_row(acc, r), ma[r], mb);
        }
    }
}

// gpuHistogrammedLingo_multirow (kernels.py:606-667): fused Tanimoto and
// per-row 101-bin histogram; per-workgroup LDS bins, flushed with global
// atomics because several workgroups share a row.
__global__ __launch_bounds__(kBlock) void k_hist_rows(RowArgs R, ColArgs C, int64_t rowbase, int32_t* __restrict__ hist) {
    extern __shared__ int2 lds_row[];
    __shared__ int bins[101];
    const int64_t row = rowbase + blockIdx.x;
    const int La = R.len[row];
    const int ma = R.mag[row];
    for (int k = threadIdx.x; k < 101; k += blockDim.x) bins[k] = 0;
    stage_row(lds_row, R.elems + R.off[row], La);
    __syncthreads();
    int64_t t0, t1, gw, nw;
    wave_tiles(C.ntiles, t0, t1, gw, nw);
    const int lane = lane_id();
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t c = t * kWave + lane;
        if (c >= C.n) continue;
        const int isct = merge_isct(lds_row, La, C.tiles + C.toff[t] + lane, C.len[c]);
        atomicAdd(&bins[bin101_f32(tanimoto_mags(isct, ma, C.mag[c]))], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 101; k += blockDim.x)
        if (bins[k]) atomicAdd(&hist[(int64_t)blockIdx.x * 101 + k], bins[k]);
}

// gpuGetNeighbors_multirow (kernels.py:393-475): magnitude-bound prefilter
// (:440-441), exact T only when it passes, lo <= T < hi (:448).  Each wave
// owns a contiguous tile range; neighbours are compacted with a wave ballot
// and written in ascending column order.  Pass 1 (FILL = false) counts per
// (row, wave); pass 2 writes at the scanned offsets.
template <bool FILL>
__global__ __launch_bounds__(kBlock) void k_neighbors(RowArgs R, ColArgs C, int64_t rowbase, float lo, float hi,
                                                      int32_t* __restrict__ counts, const int64_t* __restrict__ offs,
                                                      int32_t* __restrict__ idx) {
    extern __shared__ int2 lds_row[];
    const int64_t row = rowbase + blockIdx.x;
    const int La = R.len[row];
    const int ma = R.mag[row];
    stage_row(lds_row, R.elems + R.off[row], La);
    __syncthreads();
    int64_t t0, t1, gw, nw;
    wave_tiles(C.ntiles, t0, t1, gw, nw);
    const int lane = lane_id();
    const int64_t slot = (int64_t)blockIdx.x * nw + gw;
    int64_t pos = FILL ? offs[slot] : 0;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t c = t * kWave + lane;
        bool nb = false;
        if (c < C.n) {
            const int mb = C.mag[c];
            const float bound = (float)min(ma, mb) / (float)max(ma, mb);
            if (bound >= lo) {
                const int isct = merge_isct(lds_row, La, C.tiles + C.toff[t] + lane, C.len[c]);
                const float tv = tanimoto_mags(isct, ma, mb);
                nb = (tv >= lo) && (tv < hi);
            }
        }
        const uint64_t m = __ballot(nb);
        if (FILL && nb) idx[pos + rank_below(m)] = (int32_t)c;
        pos += __popcll(m);
    }
    if (!FILL && lane == 0) counts[slot] = (int32_t)pos;
}

// rowHistogram101 of a float matrix (_cpuutil.cpp:52-72): one workgroup per
// row, LDS bins.
template <class T>
__global__ __launch_bounds__(kBlock) void k_row_hist101(const T* __restrict__ data, int64_t cols, int64_t pitch,
                                                        uint32_t* __restrict__ hist) {
    __shared__ uint32_t bins[101];
    for (int k = threadIdx.x; k < 101; k += blockDim.x) bins[k] = 0;
    __syncthreads();
    const T* r = data + (int64_t)blockIdx.x * pitch;
    for (int64_t c = threadIdx.x; c < cols; c += blockDim.x) {
        int b;
        if constexpr (sizeof(T) == 4)
            b = bin101_f32((float)r[c]);
        else
            b = bin101_f64((double)r[c]);
        atomicAdd(&bins[b], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 101; k += blockDim.x) hist[(int64_t)blockIdx.x * 101 + k] = bins[k];
}

template <class T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int reserve(size_t bytes) {
        if (bytes <= cap && p) return SIML_OK;
        release();
        if (bytes == 0) bytes = 256;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(SIML_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        }
        cap = bytes;
        return SIML_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Host view of a SMILES set given either as padded matrices or ragged.
struct HostSet {
    const int32_t* ids = nullptr;
    const int32_t* counts = nullptr;
    const int32_t* mags = nullptr;     // may be null
    const int32_t* lengths = nullptr;  // padded form
    const int64_t* offsets = nullptr;  // ragged form (n+1)
    int64_t n = 0, width = 0, pitch = 0;
    int64_t start(int64_t r) const { return offsets ? offsets[r] : r * pitch; }
    int32_t length(int64_t r) const { return offsets ? (int32_t)(offsets[r + 1] - offsets[r]) : lengths[r]; }
};

// Per-molecule facts computed on the host while validating.
struct SetInfo {
    std::vector<int32_t> len, mag, csum;
    int32_t maxlen = 0;
    bool sorted_unique = true;
    bool counts_nonneg = true;
};

template <class F>
void parallel_for(int64_t n, F&& f) {
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<unsigned>(hw ? hw : 1, 32);
    if (n < 65536 || nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto& t : th) t.join();
}

int validate(const HostSet& s, SetInfo& info, const char* what) {
    if (s.n < 0) return fail(SIML_EINVAL, "%s: negative molecule count", what);
    if (s.n > 0 && (!s.ids || !s.counts)) return fail(SIML_EINVAL, "%s: null Lingo or count matrix", what);
    if (s.offsets) {
        if (s.n > 0 && s.offsets[0] < 0) return fail(SIML_EINVAL, "%s: negative offset", what);
        for (int64_t r = 0; r < s.n; ++r)
            if (s.offsets[r + 1] < s.offsets[r]) return fail(SIML_EINVAL, "%s: offsets are not ascending", what);
    } else {
        if (s.n > 0 && !s.lengths) return fail(SIML_EINVAL, "%s: null length vector", what);
        if (s.width < 0 || s.pitch < s.width) return fail(SIML_EINVAL, "%s: pitch smaller than width", what);
        for (int64_t r = 0; r < s.n; ++r)
            if (s.lengths[r] < 0 || s.lengths[r] > s.width)
                return fail(SIML_EINVAL, "%s: length %d of molecule %lld outside [0, %lld]", what, s.lengths[r],
                            (long long)r, (long long)s.width);
    }
    info.len.resize((size_t)s.n);
    info.mag.resize((size_t)s.n);
    info.csum.resize((size_t)s.n);
    std::vector<char> unsorted((size_t)std::max<int64_t>(s.n, 1), 0);
    std::vector<char> negative((size_t)std::max<int64_t>(s.n, 1), 0);
    parallel_for(s.n, [&](int64_t a, int64_t b) {
        for (int64_t r = a; r < b; ++r) {
            const int32_t L = s.length(r);
            const int32_t* id = s.ids + s.start(r);
            const int32_t* ct = s.counts + s.start(r);
            uint32_t sum = 0;
            bool ok = true, neg = false;
            for (int32_t k = 0; k < L; ++k) {
                sum += (uint32_t)ct[k];
                if (k && id[k] <= id[k - 1]) ok = false;
                if (ct[k] < 0) neg = true;
            }
            if (neg) negative[(size_t)r] = 1;
            info.len[(size_t)r] = L;
            info.csum[(size_t)r] = (int32_t)sum;
            info.mag[(size_t)r] = s.mags ? s.mags[r] : (int32_t)sum;
            if (!ok) unsorted[(size_t)r] = 1;
        }
    });
    info.maxlen = 0;
    for (int64_t r = 0; r < s.n; ++r) {
        info.maxlen = std::max(info.maxlen, info.len[(size_t)r]);
        if (unsorted[(size_t)r]) info.sorted_unique = false;
        if (negative[(size_t)r]) info.counts_nonneg = false;
    }
    return SIML_OK;
}

template <class T>
int upload(DevBuf& d, const T* h, size_t count, hipStream_t st) {
    TRY(d.reserve(sizeof(T) * std::max<size_t>(count, 1)));
    if (count) HIP_TRY(hipMemcpyAsync(d.p, h, sizeof(T) * count, hipMemcpyHostToDevice, st));
    return SIML_OK;
}

struct RowSet {
    int64_t n = -1;
    int32_t maxlen = 0;
    bool sorted_unique = true;
    DevBuf elems, off, len, mag, csum;
    std::vector<int32_t> h_len;
    // row groups of the hashed kernels
    DevBuf g_hdr, g_ekey, g_ecnt;
    std::vector<siml::GroupHdr> h_hdr;
    std::vector<int32_t> merge_rows;  // rows served by the merge kernel
    GroupArgs gargs() const { return {g_hdr.as<siml::GroupHdr>(), g_ekey.as<int2>(), g_ecnt.as<uint4>()}; }
    RowArgs args() const {
        return {elems.as<int2>(), off.as<int64_t>(), len.as<int32_t>(), mag.as<int32_t>(), csum.as<int32_t>()};
    }
    void release() {
        elems.release();
        off.release();
        len.release();
        mag.release();
        csum.release();
        g_hdr.release();
        g_ekey.release();
        g_ecnt.release();
        h_hdr.clear();
        merge_rows.clear();
        n = -1;
    }
};

struct ColSet {
    int64_t n = -1, ntiles = 0;
    int32_t maxlen = 0;
    bool sorted_unique = true;
    bool counts_nonneg = true;
    DevBuf tiles, toff, tlen, len, mag, csum;
    ColArgs args() const {
        return {tiles.as<int2>(), toff.as<int64_t>(), tlen.as<int32_t>(), len.as<int32_t>(), mag.as<int32_t>(),
                csum.as<int32_t>(), n, ntiles};
    }
    void release() {
        tiles.release();
        toff.release();
        tlen.release();
        len.release();
        mag.release();
        csum.release();
        n = -1;
    }
};

}  // namespace

struct siml_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    RowSet ref;
    ColSet qry;
    // pending result
    int kind = SIML_RESULT_NONE;
    int64_t rows = 0, cols = 0, nnz = 0;
    DevBuf res, res2, scratch;
    std::vector<int64_t> csr_rows;  // host row offsets of a CSR result
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_ev;
    std::vector<hipEvent_t> ev_pool;
    double kernel_ms = 0.0;
    int64_t launches = 0, pairs = 0, pending_pairs = 0;
    std::string last_kernel = "none";
    int policy = SIML_POLICY_AUTO;
    DevBuf rowlist;
    std::vector<int32_t> h_rowlist;  // host source of the last rowlist upload
};

namespace {

int set_device(siml_ctx* ctx) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(ctx->device));
    return SIML_OK;
}

hipEvent_t take_event(siml_ctx* ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets the compute launches of one operation with events when timing.
struct TimedRegion {
    siml_ctx* ctx;
    hipEvent_t a = nullptr, b = nullptr;
    int64_t pairs;
    TimedRegion(siml_ctx* c, int64_t p) : ctx(c), pairs(p) {
        if (ctx->timing) {
            a = take_event(ctx);
            b = take_event(ctx);
            if (a) (void)hipEventRecord(a, ctx->stream);
        }
    }
    void done() {
        if (ctx->timing && a && b) {
            (void)hipEventRecord(b, ctx->stream);
            ctx->pending_ev.emplace_back(a, b);
            ctx->pending_pairs += pairs;
            a = b = nullptr;
        }
    }
    ~TimedRegion() {
        if (a) ctx->ev_pool.push_back(a);
        if (b) ctx->ev_pool.push_back(b);
    }
};

int collect_timing(siml_ctx* ctx) {
    if (ctx->pending_ev.empty()) return SIML_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (auto& pr : ctx->pending_ev) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second));
        ctx->kernel_ms += ms;
        ctx->launches += 1;
        ctx->ev_pool.push_back(pr.first);
        ctx->ev_pool.push_back(pr.second);
    }
    ctx->pending_ev.clear();
    ctx->pairs += ctx->pending_pairs;
    ctx->pending_pairs = 0;
    return SIML_OK;
}

int upload_rows(siml_ctx* ctx, const HostSet& s) {
    SetInfo info;
    TRY(validate(s, info, "set_refsmiles"));
    if (info.maxlen > SIML_MAX_ROW_LINGOS)
        return fail(SIML_ELIMIT, "set_refsmiles: a reference molecule has %d unique Lingos (limit %d)", info.maxlen,
                    SIML_MAX_ROW_LINGOS);
    TRY(set_device(ctx));
    RowSet& R = ctx->ref;
    std::vector<int64_t> off((size_t)s.n + 1, 0);
    for (int64_t r = 0; r < s.n; ++r) off[(size_t)r + 1] = off[(size_t)r] + info.len[(size_t)r];
    const int64_t total = off[(size_t)s.n];
    std::vector<int2> el((size_t)std::max<int64_t>(total, 1));
    parallel_for(s.n, [&](int64_t a, int64_t b) {
        for (int64_t r = a; r < b; ++r) {
            const int32_t* id = s.ids + s.start(r);
            const int32_t* ct = s.counts + s.start(r);
            int2* d = el.data() + off[(size_t)r];
            for (int32_t k = 0; k < info.len[(size_t)r]; ++k) d[k] = make_int2(id[k], ct[k]);
        }
    });
    TRY(upload(R.elems, el.data(), (size_t)total, ctx->stream));
    TRY(upload(R.off, off.data(), (size_t)s.n, ctx->stream));
    TRY(upload(R.len, info.len.data(), (size_t)s.n, ctx->stream));
    TRY(upload(R.mag, info.mag.data(), (size_t)s.n, ctx->stream));
    TRY(upload(R.csum, info.csum.data(), (size_t)s.n, ctx->stream));
    // row groups and their cuckoo tables for the hashed kernels
    std::vector<int64_t> starts((size_t)std::max<int64_t>(s.n, 1));
    for (int64_t r = 0; r < s.n; ++r) starts[(size_t)r] = s.start(r);
    siml::GroupInput gin{s.ids, s.counts, starts.data(), info.len.data(), info.csum.data(), s.n};
    siml::GroupTables gt;
    siml::build_groups(gin, gt, 0);
    TRY(upload(R.g_hdr, gt.hdr.data(), gt.hdr.size(), ctx->stream));
    TRY(upload(R.g_ekey, reinterpret_cast<const int2*>(gt.ekey.data()), gt.ekey.size(), ctx->stream));
    TRY(upload(R.g_ecnt, reinterpret_cast<const uint4*>(gt.ecnt.data()), gt.ecnt.size() / 4, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    R.n = s.n;
    R.maxlen = info.maxlen;
    R.sorted_unique = info.sorted_unique;
    R.h_len = std::move(info.len);
    R.h_hdr = std::move(gt.hdr);
    R.merge_rows = std::move(gt.merge_rows);
    return SIML_OK;
}

int upload_cols(siml_ctx* ctx, const HostSet& s) {
    SetInfo info;
    TRY(validate(s, info, "set_qsmiles"));
    TRY(set_device(ctx));
    ColSet& C = ctx->qry;
    const int64_t ntiles = (s.n + kWave - 1) / kWave;
    std::vector<int32_t> tlen((size_t)std::max<int64_t>(ntiles, 1), 0);
    std::vector<int64_t> toff((size_t)ntiles + 1, 0);
    for (int64_t t = 0; t < ntiles; ++t) {
        int32_t m = 0;
        for (int64_t c = t * kWave; c < std::min(s.n, (t + 1) * kWave); ++c) m = std::max(m, info.len[(size_t)c]);
        tlen[(size_t)t] = m;
        toff[(size_t)t + 1] = toff[(size_t)t] + (int64_t)m * kWave;
    }
    const int64_t total = toff[(size_t)ntiles];
    std::vector<int2> tiles((size_t)std::max<int64_t>(total, 1), make_int2(0, 0));
    parallel_for(ntiles, [&](int64_t a, int64_t b) {
        for (int64_t t = a; t < b; ++t)
            for (int64_t c = t * kWave; c < std::min(s.n, (t + 1) * kWave); ++c) {
                const int32_t* id = s.ids + s.start(c);
                const int32_t* ct = s.counts + s.start(c);
                int2* d = tiles.data() + toff[(size_t)t] + (c - t * kWave);
                for (int32_t j = 0; j < info.len[(size_t)c]; ++j) d[(size_t)j * kWave] = make_int2(id[j], ct[j]);
            }
    });
    TRY(upload(C.tiles, tiles.data(), (size_t)total, ctx->stream));
    TRY(upload(C.toff, toff.data(), (size_t)ntiles + 1, ctx->stream));
    TRY(upload(C.tlen, tlen.data(), (size_t)ntiles, ctx->stream));
    TRY(upload(C.len, info.len.data(), (size_t)s.n, ctx->stream));
    TRY(upload(C.mag, info.mag.data(), (size_t)s.n, ctx->stream));
    TRY(upload(C.csum, info.csum.data(), (size_t)s.n, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    C.n = s.n;
    C.ntiles = ntiles;
    C.maxlen = info.maxlen;
    C.sorted_unique = info.sorted_unique;
    C.counts_nonneg = info.counts_nonneg;
    return SIML_OK;
}

int check_rows(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, const char* what) {
    if (ctx->ref.n < 0 || ctx->qry.n < 0)
        return fail(SIML_ESTATE, "%s: reference and query sets must both be set", what);
    if (rowbase < 0 || rowlimit > ctx->ref.n || rowlimit < rowbase)
        return fail(SIML_EINVAL, "%s: rows [%lld, %lld) outside [0, %lld)", what, (long long)rowbase,
                    (long long)rowlimit, (long long)ctx->ref.n);
    if (rowlimit - rowbase > 0x7FFFFFFF) return fail(SIML_ELIMIT, "%s: more than 2^31-1 rows in one call", what);
    return SIML_OK;
}

int32_t max_len_in(const RowSet& R, int64_t r0, int64_t r1) {
    int32_t m = 0;
    for (int64_t r = r0; r < r1; ++r) m = std::max(m, R.h_len[(size_t)r]);
    return m;
}

// How rows [rb, rl) are served: hashed groups [g0, g1) plus the rows in
// `merge` on the merge kernel, or every row on the merge kernel.
struct RowPlan {
    bool all_merge = true;
    int64_t g0 = 0, g1 = 0;
    std::vector<int32_t> merge;
};

void plan_rows(const siml_ctx* ctx, int64_t rb, int64_t rl, RowPlan& p) {
    const RowSet& R = ctx->ref;
    const ColSet& C = ctx->qry;
    p = RowPlan();
    if (ctx->policy == SIML_POLICY_MERGE || !C.sorted_unique || !C.counts_nonneg) return;
    p.all_merge = false;
    const auto& H = R.h_hdr;
    p.g0 = std::partition_point(H.begin(), H.end(), [&](const siml::GroupHdr& h) { return h.row0 + h.nrows <= rb; }) -
           H.begin();
    p.g1 = std::partition_point(H.begin(), H.end(), [&](const siml::GroupHdr& h) { return h.row0 < rl; }) - H.begin();
    auto a = std::lower_bound(R.merge_rows.begin(), R.merge_rows.end(), (int32_t)rb);
    auto b = std::lower_bound(R.merge_rows.begin(), R.merge_rows.end(), (int64_t)rl > INT32_MAX ? INT32_MAX : (int32_t)rl);
    p.merge.assign(a, b);
}

int upload_rowlist(siml_ctx* ctx, const std::vector<int32_t>& rows) {
    ctx->h_rowlist = rows;
    return upload(ctx->rowlist, ctx->h_rowlist.data(), ctx->h_rowlist.size(), ctx->stream);
}

// Grid for row-parallel kernels: grid.x = rows, grid.y splits the query
// tiles so the launch has enough workgroups to fill 256 CUs several times.
dim3 row_grid(int64_t nrows, int64_t ntiles, int64_t target_wg = 8192, int64_t max_slots = 0) {
    const int wpb = kBlock / kWave;
    int64_t gy = std::max<int64_t>(1, (target_wg + nrows - 1) / std::max<int64_t>(nrows, 1));
    gy = std::min<int64_t>(gy, std::max<int64_t>(1, (ntiles + wpb - 1) / wpb));
    if (max_slots > 0) gy = std::max<int64_t>(1, std::min<int64_t>(gy, max_slots / std::max<int64_t>(nrows * wpb, 1)));
    gy = std::min<int64_t>(gy, 65535);
    return dim3((unsigned)nrows, (unsigned)gy, 1);
}

size_t row_lds_bytes(int32_t maxlen) { return sizeof(int2) * (size_t)std::max(maxlen, 1); }

int launch_check(const char* name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SIML_EDEVICE, "launch of %s failed: %s", name, hipGetErrorString(e));
    return SIML_OK;
}

void clear_result(siml_ctx* ctx) {
    ctx->kind = SIML_RESULT_NONE;
    ctx->rows = ctx->cols = ctx->nnz = 0;
    ctx->csr_rows.clear();
}

int zero_fill_i32(siml_ctx* ctx, int32_t* p, int64_t n) {
    if (n <= 0) return SIML_OK;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    k_fill<int32_t><<<dim3((unsigned)blocks), dim3(256), 0, ctx->stream>>>(p, n, 0);
    return launch_check("k_fill");
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

const char* siml_last_error(void) { return siml::g_err.c_str(); }

int siml_version(void) { return 1; }

int siml_device_count(int* count) {
    if (!count) return fail(SIML_EINVAL, "null pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return SIML_OK;
}

int siml_create(int device, siml_ctx** out) {
    if (!out) return fail(SIML_EINVAL, "siml_create: null output pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(SIML_EDEVICE, "siml_create: no HIP device is visible");
    }
    if (device < 0 || device >= n) return fail(SIML_EINVAL, "siml_create: device %d outside [0, %d)", device, n);
    HIP_TRY(hipSetDevice(device));
    siml_ctx* ctx = new siml_ctx();
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(SIML_EDEVICE, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return SIML_OK;
}

int siml_destroy(siml_ctx* ctx) {
    if (!ctx) return SIML_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->ref.release();
    ctx->qry.release();
    ctx->res.release();
    ctx->res2.release();
    ctx->scratch.release();
    ctx->rowlist.release();
    for (auto& pr : ctx->pending_ev) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SIML_OK;
}

int siml_set_ref(siml_ctx* ctx, const int32_t* lingos, const int32_t* counts, const int32_t* mags,
                 const int32_t* lengths, int64_t n, int64_t width, int64_t pitch) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    HostSet s;
    s.ids = lingos;
    s.counts = counts;
    s.mags = mags;
    s.lengths = lengths;
    s.n = n;
    s.width = width;
    s.pitch = pitch;
    clear_result(ctx);
    return upload_rows(ctx, s);
}

int siml_set_query(siml_ctx* ctx, const int32_t* lingos, const int32_t* counts, const int32_t* mags,
                   const int32_t* lengths, int64_t n, int64_t width, int64_t pitch) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    HostSet s;
    s.ids = lingos;
    s.counts = counts;
    s.mags = mags;
    s.lengths = lengths;
    s.n = n;
    s.width = width;
    s.pitch = pitch;
    clear_result(ctx);
    return upload_cols(ctx, s);
}

int siml_set_ref_ragged(siml_ctx* ctx, const int32_t* ids, const int32_t* counts, const int32_t* mags,
                        const int64_t* offsets, int64_t n) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (n > 0 && !offsets) return fail(SIML_EINVAL, "set_ref_ragged: null offsets");
    HostSet s;
    s.ids = ids;
    s.counts = counts;
    s.mags = mags;
    s.offsets = offsets;
    s.n = n;
    clear_result(ctx);
    return upload_rows(ctx, s);
}

int siml_set_query_ragged(siml_ctx* ctx, const int32_t* ids, const int32_t* counts, const int32_t* mags,
                          const int64_t* offsets, int64_t n) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (n > 0 && !offsets) return fail(SIML_EINVAL, "set_query_ragged: null offsets");
    HostSet s;
    s.ids = ids;
    s.counts = counts;
    s.mags = mags;
    s.offsets = offsets;
    s.n = n;
    clear_result(ctx);
    return upload_cols(ctx, s);
}

int siml_set_sizes(siml_ctx* ctx, int64_t* nref, int64_t* nquery) {
    if (!ctx || !nref || !nquery) return fail(SIML_EINVAL, "null pointer");
    *nref = ctx->ref.n;
    *nquery = ctx->qry.n;
    return SIML_OK;
}

int siml_multirow_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit) {
    TRY(check_rows(ctx, rowbase, rowlimit, "getMultipleRows"));
    TRY(set_device(ctx));
    clear_result(ctx);
    const int64_t nr = rowlimit - rowbase, nq = ctx->qry.n;
    TRY(ctx->res.reserve(sizeof(float) * (size_t)std::max<int64_t>(nr * nq, 1)));
    ctx->kind = SIML_RESULT_F32;
    ctx->rows = nr;
    ctx->cols = nq;
    if (nr == 0 || nq == 0) return SIML_OK;
    RowPlan plan;
    plan_rows(ctx, rowbase, rowlimit, plan);
    if (!plan.merge.empty()) TRY(upload_rowlist(ctx, plan.merge));
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, rowbase, rowlimit));
    TimedRegion tr(ctx, nr * nq);
    if (plan.all_merge) {
        k_rows_f32<false><<<row_grid(nr, ctx->qry.ntiles), dim3(kBlock), lds, ctx->stream>>>(
            ctx->ref.args(), ctx->qry.args(), nullptr, rowbase, ctx->res.as<float>(), nq);
        TRY(launch_check("k_rows_f32"));
        ctx->last_kernel = "rows_merge";
    } else {
        const int64_t ng = plan.g1 - plan.g0;
        if (ng > 0) {
            k_hash_rows_f32<<<row_grid(ng, ctx->qry.ntiles, 2048), dim3(kBlock), 0, ctx->stream>>>(
                ctx->ref.gargs(), plan.g0, ctx->ref.mag.as<int32_t>(), ctx->qry.args(), rowbase, rowlimit,
                ctx->res.as<float>(), nq);
            TRY(launch_check("k_hash_rows_f32"));
        }
        if (!plan.merge.empty()) {
            k_rows_f32<false><<<row_grid((int64_t)plan.merge.size(), ctx->qry.ntiles), dim3(kBlock), lds,
                                ctx->stream>>>(ctx->ref.args(), ctx->qry.args(), ctx->rowlist.as<int32_t>(), rowbase,
                                               ctx->res.as<float>(), nq);
            TRY(launch_check("k_rows_f32<rowlist>"));
        }
        ctx->last_kernel = plan.merge.empty() ? "rows_hash" : (ng > 0 ? "rows_hash+merge" : "rows_merge");
    }
    tr.done();
    return SIML_OK;
}

int siml_single_row_async(siml_ctx* ctx, int64_t row) {
    TRY(check_rows(ctx, row, row + 1, "getTanimotoRow"));
    TRY(set_device(ctx));
    clear_result(ctx);
    const int64_t nq = ctx->qry.n;
    TRY(ctx->res.reserve(sizeof(float) * (size_t)std::max<int64_t>(nq, 1)));
    ctx->kind = SIML_RESULT_F32;
    ctx->rows = 1;
    ctx->cols = nq;
    if (nq == 0) return SIML_OK;
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, row, row + 1));
    TimedRegion tr(ctx, nq);
    k_rows_f32<true><<<row_grid(1, ctx->qry.ntiles, 2048), dim3(kBlock), lds, ctx->stream>>>(
        ctx->ref.args(), ctx->qry.args(), nullptr, row, ctx->res.as<float>(), nq);
    TRY(launch_check("k_rows_f32<single>"));
    tr.done();
    ctx->last_kernel = "single_row_merge";
    return SIML_OK;
}

int siml_hist_rows_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit) {
    TRY(check_rows(ctx, rowbase, rowlimit, "getMultipleHistogrammedRows"));
    TRY(set_device(ctx));
    clear_result(ctx);
    const int64_t nr = rowlimit - rowbase, nq = ctx->qry.n;
    TRY(ctx->res.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(nr * 101, 1)));
    ctx->kind = SIML_RESULT_I32;
    ctx->rows = nr;
    ctx->cols = 101;
    TRY(zero_fill_i32(ctx, ctx->res.as<int32_t>(), nr * 101));
    if (nr == 0 || nq == 0) return SIML_OK;
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, rowbase, rowlimit));
    TimedRegion tr(ctx, nr * nq);
    k_hist_rows<<<row_grid(nr, ctx->qry.ntiles), dim3(kBlock), lds, ctx->stream>>>(
        ctx->ref.args(), ctx->qry.args(), rowbase, ctx->res.as<int32_t>());
    TRY(launch_check("k_hist_rows"));
    tr.done();
    ctx->last_kernel = "hist_rows_merge";
    return SIML_OK;
}

int siml_neighbors_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, float lo, float hi) {
    TRY(check_rows(ctx, rowbase, rowlimit, "getNeighbors"));
    TRY(set_device(ctx));
    clear_result(ctx);
    const int64_t nr = rowlimit - rowbase, nq = ctx->qry.n;
    ctx->kind = SIML_RESULT_CSR;
    ctx->rows = nr;
    ctx->cols = nq;
    ctx->csr_rows.assign((size_t)nr + 1, 0);
    if (nr == 0 || nq == 0) {
        TRY(ctx->res.reserve(256));
        return SIML_OK;
    }
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, rowbase, rowlimit));
    const dim3 grid = row_grid(nr, ctx->qry.ntiles, 8192, int64_t(1) << 24);
    const int64_t nw = (int64_t)grid.y * (kBlock / kWave);
    const int64_t nslots = nr * nw;
    const size_t counts_bytes = (sizeof(int32_t) * (size_t)nslots + 255) & ~size_t(255);
    TRY(ctx->scratch.reserve(counts_bytes + sizeof(int64_t) * (size_t)nslots));
    int32_t* d_counts = ctx->scratch.as<int32_t>();
    int64_t* d_offs = reinterpret_cast<int64_t*>(ctx->scratch.as<char>() + counts_bytes);
    TimedRegion tr(ctx, nr * nq);
    k_neighbors<false><<<grid, dim3(kBlock), lds, ctx->stream>>>(ctx->ref.args(), ctx->qry.args(), rowbase, lo, hi,
                                                                   d_counts, nullptr, nullptr);
    TRY(launch_check("k_neighbors<count>"));
    std::vector<int32_t> h_counts((size_t)nslots);
    HIP_TRY(hipMemcpyAsync(h_counts.data(), d_counts, sizeof(int32_t) * nslots, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<int64_t> h_offs((size_t)nslots);
    int64_t acc = 0;
    for (int64_t r = 0; r < nr; ++r) {
        ctx->csr_rows[(size_t)r] = acc;
        for (int64_t w = 0; w < nw; ++w) {
            h_offs[(size_t)(r * nw + w)] = acc;
            acc += h_counts[(size_t)(r * nw + w)];
        }
    }
    ctx->csr_rows[(size_t)nr] = acc;
    ctx->nnz = acc;
    HIP_TRY(hipMemcpyAsync(d_offs, h_offs.data(), sizeof(int64_t) * nslots, hipMemcpyHostToDevice, ctx->stream));
    TRY(ctx->res.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(acc, 1)));
    k_neighbors<true><<<grid, dim3(kBlock), lds, ctx->stream>>>(ctx->ref.args(), ctx->qry.args(), rowbase, lo, hi,
                                                                  nullptr, d_offs, ctx->res.as<int32_t>());
    TRY(launch_check("k_neighbors<fill>"));
    tr.done();
    // h_offs must outlive the async copy above.
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->last_kernel = "neighbors_merge";
    return SIML_OK;
}

int siml_topk_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, int32_t k) {
    (void)ctx;
    (void)rowbase;
    (void)rowlimit;
    (void)k;
    return fail(SIML_ESTATE, "siml_topk_async: not built yet");
}

int siml_hist1000_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit) {
    (void)ctx;
    (void)rowbase;
    (void)rowlimit;
    return fail(SIML_ESTATE, "siml_hist1000_async: not built yet");
}

int siml_query_done(siml_ctx* ctx, int* done) {
    if (!ctx || !done) return fail(SIML_EINVAL, "null pointer");
    TRY(set_device(ctx));
    hipError_t e = hipStreamQuery(ctx->stream);
    if (e == hipSuccess) {
        *done = 1;
    } else if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        *done = 0;
    } else {
        return fail(SIML_EDEVICE, "hipStreamQuery: %s", hipGetErrorString(e));
    }
    return SIML_OK;
}

int siml_synchronize(siml_ctx* ctx) {
    TRY(set_device(ctx));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SIML_OK;
}

int siml_result_info(siml_ctx* ctx, int* kind, int64_t* rows, int64_t* cols, int64_t* nnz) {
    if (!ctx || !kind || !rows || !cols || !nnz) return fail(SIML_EINVAL, "null pointer");
    *kind = ctx->kind;
    *rows = ctx->rows;
    *cols = ctx->cols;
    *nnz = ctx->nnz;
    return SIML_OK;
}

static int retrieve_matrix(siml_ctx* ctx, void* host, int64_t pitch, int want, size_t esz) {
    if (ctx->kind != want) return fail(SIML_ESTATE, "no pending result of this kind");
    if (pitch < ctx->cols) return fail(SIML_EINVAL, "host pitch %lld < columns %lld", (long long)pitch, (long long)ctx->cols);
    TRY(set_device(ctx));
    if (ctx->rows > 0 && ctx->cols > 0) {
        if (!host) return fail(SIML_EINVAL, "null host buffer");
        HIP_TRY(hipMemcpy2DAsync(host, esz * pitch, ctx->res.p, esz * ctx->cols, esz * ctx->cols, ctx->rows,
                                 hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    clear_result(ctx);
    return SIML_OK;
}

int siml_retrieve_f32(siml_ctx* ctx, float* host, int64_t pitch) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    return retrieve_matrix(ctx, host, pitch, SIML_RESULT_F32, sizeof(float));
}

int siml_retrieve_i32(siml_ctx* ctx, int32_t* host, int64_t pitch) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    return retrieve_matrix(ctx, host, pitch, SIML_RESULT_I32, sizeof(int32_t));
}

int siml_peek_rows_f32(siml_ctx* ctx, int64_t r0, int64_t r1, float* host, int64_t pitch) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (ctx->kind != SIML_RESULT_F32) return fail(SIML_ESTATE, "no pending float32 result");
    if (r0 < 0 || r1 > ctx->rows || r1 < r0) return fail(SIML_EINVAL, "rows outside the pending result");
    if (pitch < ctx->cols) return fail(SIML_EINVAL, "host pitch smaller than columns");
    TRY(set_device(ctx));
    if (r1 > r0 && ctx->cols > 0)
        HIP_TRY(hipMemcpy2DAsync(host, sizeof(float) * pitch, ctx->res.as<float>() + r0 * ctx->cols,
                                 sizeof(float) * ctx->cols, sizeof(float) * ctx->cols, r1 - r0, hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SIML_OK;
}

int siml_retrieve_csr(siml_ctx* ctx, int64_t* offsets, int32_t* indices) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (ctx->kind != SIML_RESULT_CSR) return fail(SIML_ESTATE, "no pending neighbour result");
    if (!offsets) return fail(SIML_EINVAL, "null offsets buffer");
    TRY(set_device(ctx));
    std::memcpy(offsets, ctx->csr_rows.data(), sizeof(int64_t) * ctx->csr_rows.size());
    if (ctx->nnz > 0) {
        if (!indices) return fail(SIML_EINVAL, "null indices buffer");
        HIP_TRY(hipMemcpyAsync(indices, ctx->res.p, sizeof(int32_t) * ctx->nnz, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    clear_result(ctx);
    return SIML_OK;
}

int siml_retrieve_topk(siml_ctx* ctx, int32_t* idx, float* t) {
    (void)idx;
    (void)t;
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    return fail(SIML_ESTATE, "no pending top-k result");
}

int siml_retrieve_i64(siml_ctx* ctx, int64_t* host) {
    (void)host;
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    return fail(SIML_ESTATE, "no pending int64 result");
}

int siml_set_policy(siml_ctx* ctx, int policy) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (policy != SIML_POLICY_AUTO && policy != SIML_POLICY_MERGE) return fail(SIML_EINVAL, "unknown policy %d", policy);
    ctx->policy = policy;
    return SIML_OK;
}

int siml_set_timing(siml_ctx* ctx, int on) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    ctx->timing = on != 0;
    return SIML_OK;
}

int siml_kernel_stats(siml_ctx* ctx, double* total_ms, int64_t* launches, int64_t* pairs, int reset) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    TRY(set_device(ctx));
    TRY(collect_timing(ctx));
    if (total_ms) *total_ms = ctx->kernel_ms;
    if (launches) *launches = ctx->launches;
    if (pairs) *pairs = ctx->pairs;
    if (reset) {
        ctx->kernel_ms = 0.0;
        ctx->launches = 0;
        ctx->pairs = 0;
    }
    return SIML_OK;
}

const char* siml_last_kernel(siml_ctx* ctx) { return ctx ? ctx->last_kernel.c_str() : "none"; }

// ---------------------------------------------------------------------------
// Stateless entry points
// ---------------------------------------------------------------------------

static std::mutex g_shared_mu;
static std::vector<siml_ctx*> g_shared;

static int shared_ctx(int device, siml_ctx** out) {
    std::lock_guard<std::mutex> lk(g_shared_mu);
    if (device < 0) return fail(SIML_EINVAL, "negative device");
    if ((size_t)device >= g_shared.size()) g_shared.resize((size_t)device + 1, nullptr);
    if (!g_shared[(size_t)device]) TRY(siml_create(device, &g_shared[(size_t)device]));
    *out = g_shared[(size_t)device];
    return SIML_OK;
}

int siml_tanimoto_block(int device, const int32_t* rlingos, const int32_t* rcounts, const int32_t* rmags,
                        const int32_t* rlengths, int64_t nr, int64_t rwidth, int64_t rpitch,
                        const int32_t* qlingos, const int32_t* qcounts, const int32_t* qmags,
                        const int32_t* qlengths, int64_t nq, int64_t qwidth, int64_t qpitch, float* out,
                        int64_t opitch) {
    if (!rmags || !qmags) return fail(SIML_EINVAL, "tanimoto_block: magnitude vectors are required");
    siml_ctx* ctx = nullptr;
    TRY(shared_ctx(device, &ctx));
    TRY(siml_set_ref(ctx, rlingos, rcounts, rmags, rlengths, nr, rwidth, rpitch));
    TRY(siml_set_query(ctx, qlingos, qcounts, qmags, qlengths, nq, qwidth, qpitch));
    TRY(siml_multirow_async(ctx, 0, nr));
    return siml_retrieve_f32(ctx, out, opitch);
}

static int row_hist_common(int device, const void* data, size_t esz, int64_t rows, int64_t cols, int64_t pitch,
                           uint32_t* hist) {
    if (rows < 0 || cols < 0 || pitch < cols) return fail(SIML_EINVAL, "rowHistogram101: bad shape or pitch");
    if (rows == 0) return SIML_OK;
    if (!data || !hist) return fail(SIML_EINVAL, "rowHistogram101: null buffer");
    siml_ctx* ctx = nullptr;
    TRY(shared_ctx(device, &ctx));
    TRY(set_device(ctx));
    TRY(ctx->scratch.reserve(esz * (size_t)std::max<int64_t>(rows * pitch, 1)));
    TRY(ctx->res2.reserve(sizeof(uint32_t) * (size_t)rows * 101));
    HIP_TRY(hipMemcpyAsync(ctx->scratch.p, data, esz * (size_t)(rows * pitch), hipMemcpyHostToDevice, ctx->stream));
    if (esz == 4)
        k_row_hist101<float><<<dim3((unsigned)rows), dim3(kBlock), 0, ctx->stream>>>(
            ctx->scratch.as<float>(), cols, pitch, ctx->res2.as<uint32_t>());
    else
        k_row_hist101<double><<<dim3((unsigned)rows), dim3(kBlock), 0, ctx->stream>>>(
            ctx->scratch.as<double>(), cols, pitch, ctx->res2.as<uint32_t>());
    TRY(launch_check("k_row_hist101"));
    HIP_TRY(hipMemcpyAsync(hist, ctx->res2.p, sizeof(uint32_t) * (size_t)rows * 101, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SIML_OK;
}

int siml_row_hist101_f32(int device, const float* data, int64_t rows, int64_t cols, int64_t pitch, uint32_t* hist) {
    return row_hist_common(device, data, sizeof(float), rows, cols, pitch, hist);
}

int siml_row_hist101_f64(int device, const double* data, int64_t rows, int64_t cols, int64_t pitch,
                         uint32_t* hist) {
    return row_hist_common(device, data, sizeof(double), rows, cols, pitch, hist);
}

}  // extern "C"
