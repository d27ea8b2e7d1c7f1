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
#include <atomic>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include "bitgroups.h"
#include "host_threads.h"
#include "csr.h"
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

// Query molecules are stored at "positions" (lane slots of the tiles);
// perm[p] is the caller's column index of position p (null: identity).
// len/mag/csum are indexed by position.
struct ColArgs {
    const int2* tiles;
    const int64_t* toff;
    const int32_t* tlen;
    const int32_t* len;
    const int32_t* mag;
    const int32_t* csum;
    const int32_t* perm;
    const int2* btiles;    // level stream (bit-sliced kernels)
    const int64_t* btoff;
    const int32_t* btlen;
    const int2* tmag;      // (min, max) magnitude of the molecules of each tile
    int64_t n;
    int64_t ntiles;        // tiles visited by the tile-mode bit-sliced kernel ...
    int64_t tstride;       // ... namely tiles 0, tstride, 2 tstride, ... (a sample when > 1)
};

__device__ __forceinline__ int64_t col_of(const ColArgs& C, int64_t p) { return C.perm ? (int64_t)C.perm[p] : p; }

// Candidate short divisions for an integer numerator n in [0, 2^16) and an
// integer denominator d in [1, 2^17), the operands of every Tanimoto ratio
// with small magnitudes.  Method 1: q0 = n * rcp(d), then one residual
// correction (Markstein's quotient step, q0 + (n - q0 d) rcp(d) with fused
// multiply-adds).  Method 2 first refines the reciprocal by one
// Newton-Raphson step.  siml_selftest_division compares a method with the
// IEEE quotient on every (n, d) of that domain.
__device__ __forceinline__ float short_div(float n, float d, int method) {
    float r = __builtin_amdgcn_rcpf(d);
    if (method == 2) r = __builtin_fmaf(__builtin_fmaf(-d, r, 1.0f), r, r);
    const float q0 = n * r;
    const float e = __builtin_fmaf(-q0, d, n);
    return __builtin_fmaf(e, r, q0);
}

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

// Streaming store of 16 bytes of a float32 matrix: the nontemporal hint
// keeps the written lines from displacing the level stream, tables and
// postings the kernels read from the caches.
__device__ __forceinline__ void store_stream(float* p, float4 v) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<f4v*>(p));
}

// Element i of an array no kernel writes while it runs, read through the
// constant address space: with a wave-uniform index the load is a scalar
// load (scalar cache, scalar registers) instead of a vector load.
template <class T>
__device__ __forceinline__ T load_uniform(const T* p, int64_t i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}
__device__ __forceinline__ int2 load_uniform(const int2* p, int64_t i) {
    const int32_t* q = reinterpret_cast<const int32_t*>(p);
    return make_int2(load_uniform(q, 2 * i), load_uniform(q, 2 * i + 1));
}

// Number of set bits of `m` below this lane.
__device__ __forceinline__ int rank_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Sum of v over the wave (every lane gets it).
__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// OR of v over the wave, returned wave-uniform: DPP row shifts 1, 2, 4, 8
// OR each row of 16 lanes into its lane 15, row broadcasts 15 and 31 carry
// the rows into lane 63.  Every lane must be active.
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), off);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

// Evaluated-work counter: one pair of global atomics per wave.
__device__ __forceinline__ void add_work(unsigned long long* work, unsigned long long pairs, unsigned long long lsum) {
    pairs = wave_sum_u64(pairs);
    lsum = wave_sum_u64(lsum);
    if (lane_id() == 0 && pairs) {
        atomicAdd(work, pairs);
        atomicAdd(work + 1, lsum);
    }
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
        orow[col_of(C, c)] = RECOMPUTE ? tanimoto_recomputed(isct, La, Lb, ma, C.csum[c]) : tanimoto_mags(isct, ma, C.mag[c]);
    }
}

// ---------------------------------------------------------------------------
// 1000-bin global histogram and top-k (no counterpart in the reference)
// ---------------------------------------------------------------------------

// Edges fl32(k/1000), k = 0..1000, as one correctly rounded division each.
__device__ __forceinline__ void stage_edges(float* e) {
    for (int k = threadIdx.x; k <= 1000; k += blockDim.x) e[k] = (float)k / 1000.0f;
}

// max{k : t >= e[k]} clipped to [0, 999]; t < 0 or NaN -> 0.  The guess
// from t*1000 is corrected against the float32 edges.
__device__ __forceinline__ int bin1000(float t, const float* e) {
    if (!(t >= 0.0f)) return 0;
    const float p = t * 1000.0f;
    int k = p >= 999.0f ? 999 : (int)p;
    while (k < 999 && t >= e[k + 1]) ++k;
    while (k > 0 && t < e[k]) --k;
    return k;
}

// Ordering key of (T, column): larger T first, equal T by smaller column.
__device__ __forceinline__ uint64_t topk_key(float t, int64_t c) {
    const uint32_t u = __float_as_uint(t);
    const uint32_t o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((uint64_t)o << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)c);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Candidate buffer of one (row, wave): k + kTopkSlack keys rounded up to a
// multiple of 64, so a re-selection happens at most every kTopkSlack / 64
// pushes; a lane holds kTopkRegs keys while the wave re-selects.
constexpr int kTopkSlack = 128;
constexpr int kTopkRegs = (SIML_MAX_TOPK + kTopkSlack + 63) / 64;

// Replaces the n > k distinct keys of buf[0..n) by the k largest (in any
// order) and returns the k-th largest.  One wave: the keys are read into
// registers once, an MSB-first 8-bit radix select over them finds the k-th
// largest (hist: 256 wave-private LDS counters), and the keys not below it
// are written back.  Called in wave-uniform control flow.
__device__ uint64_t wave_keep_topk(uint64_t* buf, int n, int k, uint32_t* hist) {
    const int lane = lane_id();
    uint64_t v[kTopkRegs];
#pragma unroll
    for (int j = 0; j < kTopkRegs; ++j) {
        const int i = j * kWave + lane;
        v[j] = i < n ? buf[i] : 0ull;  // 0 is below every key
    }
    uint64_t prefix = 0, pmask = 0;
    uint32_t need = (uint32_t)k;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int b = lane; b < 256; b += kWave) hist[b] = 0u;
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < kTopkRegs; ++j)
            if (v[j] != 0ull && (v[j] & pmask) == prefix) atomicAdd(&hist[(uint32_t)(v[j] >> shift) & 255u], 1u);
        wave_lds_sync();
        const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2], h3 = hist[4 * lane + 3];
        const uint32_t s = h0 + h1 + h2 + h3;
        uint32_t S = s;  // keys in bins of lanes >= this lane
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t t = __shfl_down(S, off);
            if (lane + off < kWave) S += t;
        }
        const uint64_t ge = __ballot(S >= need);
        const int L = 63 - __builtin_clzll(ge);
        uint32_t above = __shfl(S - s, L);
        const uint32_t hb[4] = {(uint32_t)__shfl(h0, L), (uint32_t)__shfl(h1, L), (uint32_t)__shfl(h2, L),
                                (uint32_t)__shfl(h3, L)};
        int d = 4 * L;
        uint32_t cnt_d = hb[0];
        for (int q = 3; q >= 0; --q) {
            if (above + hb[q] >= need) {
                d = 4 * L + q;
                cnt_d = hb[q];
                break;
            }
            above += hb[q];
        }
        need -= above;
        prefix |= (uint64_t)d << shift;
        pmask |= (uint64_t)255 << shift;
        wave_lds_sync();
        if (cnt_d == need) {
            // every key of the chosen bucket is needed: the k-th largest is
            // the bucket's smallest key
            uint64_t mn = ~0ull;
#pragma unroll
            for (int j = 0; j < kTopkRegs; ++j)
                if (v[j] != 0ull && (v[j] & pmask) == prefix && v[j] < mn) mn = v[j];
            for (int off = 1; off < kWave; off <<= 1) {
                const uint64_t o = ((uint64_t)(uint32_t)__shfl_xor((int)(mn >> 32), off) << 32) |
                                   (uint32_t)__shfl_xor((int)(uint32_t)mn, off);
                mn = o < mn ? o : mn;
            }
            prefix = mn;
            break;
        }
    }
    // the keys >= prefix are exactly the k largest (keys are distinct)
    int w = 0;
#pragma unroll
    for (int j = 0; j < kTopkRegs; ++j) {
        const bool keep = v[j] >= prefix && v[j] != 0ull;
        const uint64_t m = __ballot(keep);
        if (keep) buf[w + rank_below(m)] = v[j];
        w += __popcll(m);
    }
    return prefix;
}

// Candidate buffer of one (row, wave): up to cap keys; once armed, only
// keys above theta (the k-th largest seen so far) are added.
struct TopkBuf {
    uint64_t* buf;
    int cnt;
    bool armed;
    uint64_t theta;
};

__device__ __forceinline__ void topk_push(TopkBuf& B, bool valid, uint64_t key, int k, int cap, uint32_t* hist) {
    if (B.cnt + kWave > cap) {
        B.theta = wave_keep_topk(B.buf, B.cnt, k, hist);
        B.cnt = k;
        B.armed = true;
    }
    const bool pred = valid && (!B.armed || key > B.theta);
    const uint64_t m = __ballot(pred);
    if (pred) B.buf[B.cnt + rank_below(m)] = key;
    B.cnt += __popcll(m);
}

__global__ __launch_bounds__(kBlock) void k_hist1000(RowArgs R, ColArgs C, const int32_t* __restrict__ rowlist,
                                                     int64_t rowbase, unsigned long long* __restrict__ hist) {
    extern __shared__ int2 lds_row[];
    __shared__ uint32_t bins[1000];
    __shared__ float edges[1001];
    for (int k = threadIdx.x; k < 1000; k += blockDim.x) bins[k] = 0u;
    stage_edges(edges);
    const int64_t row = rowlist ? (int64_t)rowlist[blockIdx.x] : rowbase + blockIdx.x;
    const int La = R.len[row];
    const int ma = R.mag[row];
    stage_row(lds_row, R.elems + R.off[row], La);
    __syncthreads();
    int64_t t0, t1, gw, nw;
    wave_tiles(C.ntiles, t0, t1, gw, nw);
    const int lane = lane_id();
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t c = t * kWave + lane;
        if (c >= C.n) continue;
        const int isct = merge_isct(lds_row, La, C.tiles + C.toff[t] + lane, C.len[c]);
        atomicAdd(&bins[bin1000(tanimoto_mags(isct, ma, C.mag[c]), edges)], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 1000; k += blockDim.x)
        if (bins[k]) atomicAdd(&hist[k], (unsigned long long)bins[k]);
}

__global__ __launch_bounds__(kBlock) void k_topk(RowArgs R, ColArgs C, const int32_t* __restrict__ rowlist,
                                                 int64_t rowbase, int k, int cap, uint64_t* __restrict__ buf,
                                                 int32_t* __restrict__ bufcnt) {
    extern __shared__ int2 lds_row[];
    __shared__ uint32_t rhist[kBlock / kWave][256];
    const int64_t row = rowlist ? (int64_t)rowlist[blockIdx.x] : rowbase + blockIdx.x;
    const int La = R.len[row];
    const int ma = R.mag[row];
    stage_row(lds_row, R.elems + R.off[row], La);
    __syncthreads();
    int64_t t0, t1, gw, nw;
    wave_tiles(C.ntiles, t0, t1, gw, nw);
    const int lane = lane_id();
    uint32_t* hist = rhist[threadIdx.x / kWave];
    TopkBuf B{buf + ((row - rowbase) * nw + gw) * (int64_t)cap, 0, false, 0ull};
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t c = t * kWave + lane;
        const bool valid = c < C.n;
        float tv = 0.0f;
        if (valid) tv = tanimoto_mags(merge_isct(lds_row, La, C.tiles + C.toff[t] + lane, C.len[c]), ma, C.mag[c]);
        topk_push(B, valid, topk_key(tv, valid ? col_of(C, c) : 0), k, cap, hist);
    }
    if (lane == 0) bufcnt[(row - rowbase) * nw + gw] = B.cnt;
}

// Final top-k of each row from its nw candidate buffers: a workgroup radix
// select of the k-th best key, then a bitonic sort of the (at most 256)
// keys at or above it.  Output rows are padded with (-1, -inf).
// Candidate buffers one selection workgroup takes (their keys are staged in
// LDS).
constexpr int kTopkFan = 16;
// Buffers per selection workgroup when they are sparse (seeded passes).
constexpr int kTopkStreamFan = 256;
// Sampled tiles per wave of a top-k seeding pass, and the fewest
// workgroups it runs on while its waves keep at least one tile each.
constexpr int64_t kSampleTiles = 8;
constexpr int64_t kSampleWorkgroups = 1024;

// Dynamic LDS of a selection launch whose segments hold at most segcap
// keys (at most 16 x 384 keys = 48 KiB, for k = 256): sized to the launch,
// so more selection workgroups fit per CU for small k.
inline size_t topk_lds(int segcap) { return sizeof(uint64_t) * (size_t)kTopkFan * (size_t)segcap; }

// The k-th largest of the keys cand[0, total) (total > k; keys distinct and
// non-zero), by an MSB-first 8-bit radix select in LDS.  Every thread of the
// workgroup calls it; the result is uniform.
__device__ uint64_t wg_kth_largest(const uint64_t* cand, uint32_t total, uint32_t k) {
    __shared__ uint32_t hist[256];
    __shared__ uint64_t s_prefix, s_pmask;
    __shared__ unsigned long long s_min;
    __shared__ uint32_t s_need, s_exact;
    if (threadIdx.x == 0) {
        s_prefix = 0ull;
        s_pmask = 0ull;
        s_need = k;
    }
    __syncthreads();
    for (int shift = 56; shift >= 0; shift -= 8) {
        hist[threadIdx.x] = 0u;
        __syncthreads();
        const uint64_t prefix = s_prefix, pmask = s_pmask;
        for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
            const uint64_t key = cand[i];
            const bool act = (key & pmask) == prefix;
            const uint32_t dig = (uint32_t)(key >> shift) & 255u;
            // the high digits of most keys are equal: when every active
            // lane has the same digit, one atomic adds them all (equal-
            // address atomics of one instruction serialise)
            const uint64_t am = __ballot(act);
            if (am == 0ull) continue;
            const int leader = __builtin_ctzll(am);
            const uint32_t d0 = (uint32_t)__shfl((int)dig, leader);
            if (__ballot(act && dig == d0) == am) {
                if (lane_id() == leader) atomicAdd(&hist[d0], (uint32_t)__popcll(am));
            } else if (act) {
                atomicAdd(&hist[dig], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < kWave) {
            // wave 0 finds the digit: suffix sums over 4 bins per lane
            const int lane = threadIdx.x;
            const uint32_t need = s_need;
            const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                           h3 = hist[4 * lane + 3];
            const uint32_t sl = h0 + h1 + h2 + h3;
            uint32_t S = sl;
            for (int off = 1; off < kWave; off <<= 1) {
                const uint32_t t = __shfl_down(S, off);
                if (lane + off < kWave) S += t;
            }
            const uint64_t ge = __ballot(S >= need);
            const int L = 63 - __builtin_clzll(ge);
            uint32_t above = __shfl(S - sl, L);
            const uint32_t hb[4] = {(uint32_t)__shfl(h0, L), (uint32_t)__shfl(h1, L), (uint32_t)__shfl(h2, L),
                                    (uint32_t)__shfl(h3, L)};
            int d = 4 * L;
            uint32_t cnt_d = hb[0];
            for (int q = 3; q >= 0; --q) {
                if (above + hb[q] >= need) {
                    d = 4 * L + q;
                    cnt_d = hb[q];
                    break;
                }
                above += hb[q];
            }
            if (lane == 0) {
                s_need = need - above;
                s_prefix = prefix | ((uint64_t)d << shift);
                s_pmask = pmask | ((uint64_t)255 << shift);
                s_exact = cnt_d == need - above ? 1u : 0u;
                s_min = ~0ull;
            }
        }
        __syncthreads();
        if (s_exact) {
            // every key of the chosen bucket is needed: the k-th largest
            // is the bucket's smallest key
            const uint64_t bp = s_prefix, bm = s_pmask;
            uint64_t mn = ~0ull;
            for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
                const uint64_t key = cand[i];
                if ((key & bm) == bp && key < mn) mn = key;
            }
            atomicMin(&s_min, (unsigned long long)mn);
            __syncthreads();
            break;
        }
    }
    const uint64_t K = s_exact ? (uint64_t)s_min : s_prefix;
    __syncthreads();
    return K;
}

// Workgroup selection over the candidate segments [s0, s1) of one row
// (segment w holds cnt[w] keys at buf + w * cap; keys distinct and
// non-zero).  Segments are copied to LDS (`cand`, kTopkFan * cap keys) as
// many at a time as fit; whenever more remain and the keys gathered exceed
// k, they are cut to their k largest (radix select, then the keys >= the
// k-th move to the front), so any number of segments streams through one
// workgroup: one pass when they fit, as in a selection level of kTopkFan
// segments.  `sorted` (kBlock keys) receives the min(k, total) largest keys
// (unordered, rest 0); returns their number.  With kth, only the k-th
// largest key is returned there (0 when there are fewer than k).
__device__ uint32_t wg_select_topk(const uint64_t* __restrict__ buf, const int32_t* __restrict__ cnt, int s0,
                                   int s1, int cap, int k, uint64_t* sorted, uint64_t* cand,
                                   uint64_t* kth = nullptr) {
    __shared__ int32_t s_base[kWave + 1];
    __shared__ uint32_t s_take, s_fill;
    const uint32_t room_all = (uint32_t)kTopkFan * (uint32_t)cap;
    uint32_t carry = 0, total = 0;  // keys kept in cand[0, carry); uniform
    for (int w0 = s0;;) {
        const int nwin = min(kWave, s1 - w0);
        if (threadIdx.x < kWave) {
            // window of up to 64 segments: offsets by one wave prefix sum;
            // the window's first segments whose keys fit the room left
            // (at least one: room >= kTopkFan * cap - k >= cap)
            const int lane = threadIdx.x;
            const int n = lane < nwin ? cnt[w0 + lane] : 0;
            int S = n;
            for (int off = 1; off < kWave; off <<= 1) {
                const int t = __shfl_up(S, off);
                if (lane >= off) S += t;
            }
            if (lane < nwin) s_base[lane] = S - n;
            if (lane == (nwin > 0 ? nwin - 1 : 0)) s_base[nwin] = nwin > 0 ? S : 0;
            const uint64_t fits = __ballot(lane < nwin && (uint32_t)S <= room_all - carry);
            if (lane == 0) {
                s_take = (uint32_t)__popcll(fits);
                s_fill = 0u;
            }
        }
        __syncthreads();
        // m >= 1 whenever a segment is left (a segment holds at most cap keys);
        // the loop ends after s1 - s0 segments whatever the counts say
        const int m = nwin > 0 ? max((int)s_take, 1) : 0;
        const uint32_t add = min((uint32_t)s_base[m], room_all - carry);
        // keys to LDS: every element load is independent of the others, and
        // a thread has four in flight
        constexpr int kUnroll = 4;
        for (uint32_t i0 = threadIdx.x; i0 < add; i0 += kUnroll * blockDim.x) {
            uint64_t v[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const uint32_t i = i0 + u * blockDim.x;
                if (i >= add) break;
                int lo = 0, hi = m - 1;  // the segment w with s_base[w] <= i < s_base[w + 1]
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if ((uint32_t)s_base[mid] <= i)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                v[u] = buf[(int64_t)(w0 + lo) * cap + (i - (uint32_t)s_base[lo])];
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const uint32_t i = i0 + u * blockDim.x;
                if (i < add) cand[carry + i] = v[u];
            }
        }
        __syncthreads();
        total = carry + add;
        w0 += m;
        if (w0 >= s1) break;
        if (total > (uint32_t)k) {
            // more segments follow: keep the k largest at the front
            const uint64_t K = wg_kth_largest(cand, total, (uint32_t)k);
            for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
                const uint64_t key = cand[i];
                if (key >= K) sorted[atomicAdd(&s_fill, 1u)] = key;
            }
            __syncthreads();
            for (int i = threadIdx.x; i < k; i += blockDim.x) cand[i] = sorted[i];
            total = (uint32_t)k;
            __syncthreads();
        }
        carry = total;
    }
    uint64_t K = 0ull;  // every key is > 0
    if (total > (uint32_t)k) K = wg_kth_largest(cand, total, (uint32_t)k);
    if (kth) {
        // only the k-th largest key is wanted (0 when there are fewer than k)
        if (total < (uint32_t)k) K = 0ull;
        if (total == (uint32_t)k) {
            // the smallest of exactly k keys
            __shared__ unsigned long long s_kmin;
            if (threadIdx.x == 0) s_kmin = ~0ull;
            __syncthreads();
            uint64_t mn = ~0ull;
            for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) mn = cand[i] < mn ? cand[i] : mn;
            atomicMin(&s_kmin, (unsigned long long)mn);
            __syncthreads();
            K = s_kmin;
        }
        *kth = K;
        return total;
    }
    if (threadIdx.x == 0) s_fill = 0u;
    sorted[threadIdx.x] = 0ull;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
        const uint64_t key = cand[i];
        if (key >= K) sorted[atomicAdd(&s_fill, 1u)] = key;
    }
    __syncthreads();
    return s_fill;
}

// One level of the per-row selection when a row has more candidate
// buffers than one workgroup selects from: workgroup (row, s) reduces
// buffers [s * fan, (s + 1) * fan) to their k largest keys (fan = kTopkFan
// for full buffers; more when the buffers are sparse, which the workgroup
// streams through its LDS).
__global__ __launch_bounds__(kBlock) void k_topk_stage(const uint64_t* __restrict__ buf,
                                                       const int32_t* __restrict__ bufcnt, int nw, int cap, int k,
                                                       int fan, uint64_t* __restrict__ out,
                                                       int32_t* __restrict__ outcnt) {
    extern __shared__ uint64_t topk_cand[];  // kTopkFan x cap keys (launch's dynamic LDS)
    __shared__ uint64_t sorted[kBlock];
    const int64_t r = blockIdx.x;
    const int s = blockIdx.y, ns = gridDim.y;
    const int s0 = s * fan, s1 = min(nw, s0 + fan);
    const uint32_t have = wg_select_topk(buf + r * nw * (int64_t)cap, bufcnt + r * nw, s0, s1, cap, k, sorted,
                                         topk_cand);
    uint64_t* o = out + (r * ns + s) * (int64_t)k;
    for (int i = threadIdx.x; i < (int)have; i += blockDim.x) o[i] = sorted[i];
    if (threadIdx.x == 0) outcnt[r * ns + s] = (int32_t)have;
}

// Final per-row selection: the k largest keys of the row's candidate
// segments, sorted (T descending, then smaller column).  With zero_fill
// the candidates hold only the pairs with T > 0 and every other pair of the
// row has T == 0: the slots after the positive candidates take the columns
// with T == 0 in ascending order -- the smallest column indices of [0, ncols)
// that are not among the positive candidates.
// Seed of each row: its k-th largest candidate key, 0 when it has fewer
// than k candidates (then every candidate must be kept).
__global__ __launch_bounds__(kBlock) void k_topk_kth(const uint64_t* __restrict__ buf,
                                                     const int32_t* __restrict__ bufcnt, int nw, int k, int cap,
                                                     uint64_t* __restrict__ seed) {
    extern __shared__ uint64_t topk_cand[];  // kTopkFan x cap keys (launch's dynamic LDS)
    __shared__ uint64_t sorted[kBlock];
    const int64_t r = blockIdx.x;
    uint64_t K = 0ull;
    wg_select_topk(buf + r * nw * (int64_t)cap, bufcnt + r * nw, 0, nw, cap, k, sorted, topk_cand, &K);
    if (threadIdx.x == 0) seed[r] = K;
}

__global__ __launch_bounds__(kBlock) void k_topk_final(const uint64_t* __restrict__ buf,
                                                       const int32_t* __restrict__ bufcnt, int nw, int k, int cap,
                                                       int zero_fill, int64_t ncols, int32_t* __restrict__ out_idx,
                                                       float* __restrict__ out_t) {
    extern __shared__ uint64_t topk_cand[];  // kTopkFan x cap keys (launch's dynamic LDS)
    __shared__ uint64_t sorted[kBlock];
    __shared__ uint32_t poscol[kBlock];
    const int64_t rloc = blockIdx.x;
    const uint32_t have =
        wg_select_topk(buf + rloc * nw * (int64_t)cap, bufcnt + rloc * nw, 0, nw, cap, k, sorted, topk_cand);
    // bitonic sort, descending, of kBlock keys (k <= kBlock)
    for (int size = 2; size <= kBlock; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int i = threadIdx.x, j = i ^ stride;
            if (j > i) {
                const bool desc = (i & size) == 0;
                const uint64_t a = sorted[i], b = sorted[j];
                if (desc ? (a < b) : (a > b)) {
                    sorted[i] = b;
                    sorted[j] = a;
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        int32_t ix = -1;
        float tv = -__builtin_inff();
        if ((uint32_t)i < have) {
            const uint64_t key = sorted[i];
            const uint32_t o = (uint32_t)(key >> 32);
            tv = __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
            ix = (int32_t)(0xFFFFFFFFu - (uint32_t)key);
        }
        out_idx[rloc * k + i] = ix;
        out_t[rloc * k + i] = tv;
    }
    if (!zero_fill || have >= (uint32_t)k) return;
    // columns of the positive candidates, ascending
    poscol[threadIdx.x] = (uint32_t)threadIdx.x < have ? 0xFFFFFFFFu - (uint32_t)sorted[threadIdx.x] : 0xFFFFFFFFu;
    __syncthreads();
    for (int size = 2; size <= kBlock; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int i = threadIdx.x, j = i ^ stride;
            if (j > i) {
                const bool asc = (i & size) == 0;
                const uint32_t a = poscol[i], b = poscol[j];
                if (asc ? (a > b) : (a < b)) {
                    poscol[i] = b;
                    poscol[j] = a;
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        uint32_t p = 0;
        int64_t c = 0;
        for (int i = (int)have; i < k && c < ncols; ++c) {
            while (p < have && poscol[p] < (uint32_t)c) ++p;
            if (p < have && poscol[p] == (uint32_t)c) continue;
            out_idx[rloc * k + i] = (int32_t)c;
            out_t[rloc * k + i] = 0.0f;
            ++i;
        }
    }
}

// ---------------------------------------------------------------------------
// Bit-sliced multi-row kernels (row groups of bitgroups.h)
//
// A workgroup stages its group's level table (kBitSlots x (level key, row
// mask)) in LDS.  Each lane streams one query molecule's levels as (key,
// slot word) pairs, two levels per 16-byte load; a level reads its two
// candidate slots at the precomputed byte offsets and keeps the mask of the
// slot whose key is equal.  The masks are added to 32 vertical binary
// counters held as bit planes P[0..NB): blocks of 8 masks go through a
// carry-save adder tree (xor3 / majority, one v_bitop3 each) and the
// block's carry ripples into the upper planes.  A bit-plane transpose then
// yields each row's intersection count as one byte (two for NB = 16).
// ---------------------------------------------------------------------------

struct BitGroupArgs {
    const siml::BitGroupHdr* hdr;
    const int32_t* rows;   // kBitRowStride row indices per group
    const uint4* ent;      // (key, mask, slot, -) per entry
    const int32_t* glist;  // the groups of this launch
    // persistent grid (k_bits only; null: one work item per workgroup):
    // kXcds counters, zero at launch, each handing out the nslot group
    // slots once; ny column ranges per group slot
    int32_t* queue = nullptr;
    int32_t nslot = 0, ny = 0;
};

// Row mask of the slot holding `key` (0 when neither candidate does); sw
// holds the byte offsets of both candidate slots.
__device__ __forceinline__ uint32_t bit_probe(const uint2* tab, uint32_t key, uint32_t sw) {
    const char* base = reinterpret_cast<const char*>(tab);
    const uint2 k1 = *reinterpret_cast<const uint2*>(base + (sw & 0xFFFFu));
    const uint2 k2 = *reinterpret_cast<const uint2*>(base + (sw >> 16));
    const uint32_t m2 = (k2.x == key) ? k2.y : 0u;
    return (k1.x == key) ? k1.y : m2;
}

// Full adder on 32 independent bit lanes: s += a + b, carry out in c.
__device__ __forceinline__ void full_add(uint32_t& s, uint32_t a, uint32_t b, uint32_t& c) {
    c = (s & a) | (s & b) | (a & b);
    s = s ^ a ^ b;
}

// Adds 8 masks to the counters P (carry-save tree, then ripple).
template <int NB>
__device__ __forceinline__ void csa8(uint32_t (&P)[NB], const uint32_t (&m)[8]) {
    uint32_t c0, c1, c2, c3, d0, d1, e;
    full_add(P[0], m[0], m[1], c0);
    full_add(P[0], m[2], m[3], c1);
    full_add(P[0], m[4], m[5], c2);
    full_add(P[0], m[6], m[7], c3);
    full_add(P[1], c0, c1, d0);
    full_add(P[1], c2, c3, d1);
    full_add(P[2], d0, d1, e);
#pragma unroll
    for (int k = 3; k < NB; ++k) {
        const uint32_t t = P[k] & e;
        P[k] ^= e;
        e = t;
    }
}

// Probes the 8 levels held in v[0..4) and adds their masks.
template <int NB>
__device__ __forceinline__ void probe8(const uint2* tab, const int4 (&v)[4], uint32_t (&P)[NB]) {
    uint32_t m[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        m[2 * q] = bit_probe(tab, (uint32_t)v[q].x, (uint32_t)v[q].y);
        m[2 * q + 1] = bit_probe(tab, (uint32_t)v[q].z, (uint32_t)v[q].w);
    }
    csa8<NB>(P, m);
}

// This lane's byte offset (16 x lane) in a tile of the level stream,
// computed where it is used: a value held across the tile loops is one the
// register allocator spills under the kernels' register limits, and the
// reload's wait then also waits for every store in flight.
__device__ __forceinline__ uint32_t lane_x16() {
    uint32_t l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0\n\tv_lshlrev_b32 %0, 4, %0"
                 : "=v"(l));
    return l;
}

// Loads the 8 levels of block k of a tile: `base` is the tile's stream
// (wave-uniform), voff the lane's byte offset.  Addresses are a scalar base
// plus 32-bit vector offsets (a tile holds at most 256 x 64 levels).
__device__ __forceinline__ void load8(int4 (&v)[4], const char* __restrict__ base, uint32_t voff, int k) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
        v[q] = *reinterpret_cast<const int4*>(base + (voff + (uint32_t)(k * 4 + q) * (uint32_t)(kWave * 16)));
}

// Streams the Lt (a multiple of 8) levels of one query molecule; base is
// the tile's stream.  Two register sets alternate, so the loads of the
// next 8 levels are in flight while the current 8 are probed.
template <int NB>
__device__ __forceinline__ void bits_accumulate(const uint2* tab, const char* __restrict__ base, int Lt,
                                                uint32_t (&P)[NB]) {
#pragma unroll
    for (int k = 0; k < NB; ++k) P[k] = 0u;
    const int nblk = Lt >> 3;
    if (nblk <= 0) return;
    const uint32_t voff = lane_x16();
    int4 a[4], b[4];
    load8(a, base, voff, 0);
    int k = 0;
    for (; k + 2 <= nblk; k += 2) {
        load8(b, base, voff, k + 1);
        probe8<NB>(tab, a, P);
        // may read the 8 levels after this molecule's last: the stream
        // carries one spare block at its end, so this stays in bounds
        load8(a, base, voff, k + 2);
        probe8<NB>(tab, b, P);
    }
    if (k < nblk) probe8<NB>(tab, a, P);
}

__device__ __forceinline__ void swap_bits(uint32_t& a, uint32_t& b, int sh, uint32_t mask) {
    const uint32_t t = ((a >> sh) ^ b) & mask;
    b ^= t;
    a ^= t << sh;
}

// In-place transpose of the 8x8 bit blocks of P[0..8): afterwards byte q of
// P[j] is the 8-bit value whose bit i was bit (8q + j) of the old P[i].
__device__ __forceinline__ void transpose8(uint32_t* P) {
    swap_bits(P[0], P[4], 4, 0x0F0F0F0Fu);
    swap_bits(P[1], P[5], 4, 0x0F0F0F0Fu);
    swap_bits(P[2], P[6], 4, 0x0F0F0F0Fu);
    swap_bits(P[3], P[7], 4, 0x0F0F0F0Fu);
    swap_bits(P[0], P[2], 2, 0x33333333u);
    swap_bits(P[1], P[3], 2, 0x33333333u);
    swap_bits(P[4], P[6], 2, 0x33333333u);
    swap_bits(P[5], P[7], 2, 0x33333333u);
    swap_bits(P[0], P[1], 1, 0x55555555u);
    swap_bits(P[2], P[3], 1, 0x55555555u);
    swap_bits(P[4], P[5], 1, 0x55555555u);
    swap_bits(P[6], P[7], 1, 0x55555555u);
}

// Row r's count (as float) after transpose8 of each 8-plane half.
template <int NB>
__device__ __forceinline__ float count_of(const uint32_t (&P)[NB], int r) {
    const int j = r & 7, sh = 8 * (r >> 3);
    if constexpr (NB == 8) {
        return (float)((P[j] >> sh) & 0xFFu);
    } else {
        return (float)(((P[j] >> sh) & 0xFFu) | (((P[8 + j] >> sh) & 0xFFu) << 8));
    }
}

// Rows of the G groups one workgroup serves (32 per group): row q * 32 + r
// is row r of the workgroup's q-th group.
template <int NR>
struct BitRows {
    int32_t row[NR];  // caller row index, -1 when not live
    int32_t ma[NR];
    float maf[NR];
    int32_t amin, amax;  // magnitude range of the live rows
};
using BitRowInfo = BitRows<siml::kBitRows>;

// Epilogue policies: init (LDS), begin (per wave), apply (per row, tile),
// end (per wave), finish (after the workgroup barrier).  G is the number of
// row groups (tables) a workgroup serves: 32 G rows, 4 G waves.
template <int G>
struct EpiF32 {
    static constexpr int kG = G;
    static constexpr bool kPrune = false;
    static constexpr bool kBlockMode = false;
    static constexpr bool kAdmit = false;
    using Rows = BitRows<32 * G>;
    float* out;
    int64_t opitch, rb;
    struct Lds {
        int unused;
    };
    struct Reg {};
    __device__ void init(Lds&) const {}
    __device__ void begin(Lds&, Reg&, const Rows&, int64_t, int64_t) const {}
    __device__ void apply(Lds&, Reg&, int, int32_t row, bool valid, int64_t col, int, int, float t, int64_t,
                          int64_t, int) const {
        if (valid) out[(row - rb) * opitch + col] = t;
    }
    // a row none of the tile's columns intersects: T = 0 for every pair
    __device__ bool zero_skips() const { return true; }
    __device__ void skip_row(Lds&, Reg&, int, int32_t row, bool valid, int64_t col) const {
        if (valid) out[(row - rb) * opitch + col] = 0.0f;
    }
    __device__ void end(Lds&, Reg&, const Rows&, int64_t, int64_t) const {}
    __device__ void finish(Lds&, const Rows&) const {}
};

// Float32 matrix written a 256-column block at a time: each wave computes
// the 4 tiles of a block, stages a row's 256 values in LDS at their column
// offsets inside the block and stores them as one contiguous 1 KiB run.
// Needs a layout that keeps columns inside their block, out and opitch
// 16-byte aligned.
// Waves per workgroup of the float32 block kernel: with 8 bit planes, 8
// waves share the workgroup's table (3 workgroups per CU by LDS, 6 waves per
// SIMD within 80 VGPRs); with 16 planes, 4.
template <int NB>
constexpr int block_waves() {
    return NB == 8 ? 8 : 4;
}

struct EpiF32Block {
    static constexpr int kG = 1;
    static constexpr bool kPrune = false;
    static constexpr bool kBlockMode = true;
    static constexpr bool kAdmit = false;
    float* out;
    int64_t opitch, rb;
    struct Lds {
        int unused;  // the kernel sizes its staging rows by its wave count
    };
    struct Reg {};
    __device__ void init(Lds&) const {}
    __device__ void finish(Lds&, const BitRowInfo&) const {}
};

// 101-bin rows: pairs in bin 0 (T * 100 < 1, the bulk of unrelated pairs)
// are counted in per-lane registers, the rest with LDS atomics.
struct EpiHist101 {
    static constexpr int kG = 1;
    static constexpr bool kPrune = false;
    static constexpr bool kBlockMode = false;
    static constexpr bool kAdmit = false;
    int32_t* hist;
    int64_t rb;
    struct Lds {
        int bins[siml::kBitRows][101];
    };
    struct Reg {
        uint32_t zero[siml::kBitRows];
    };
    __device__ void init(Lds& L) const {
        for (int k = threadIdx.x; k < siml::kBitRows * 101; k += blockDim.x) (&L.bins[0][0])[k] = 0;
    }
    __device__ void begin(Lds&, Reg& R, const BitRowInfo&, int64_t, int64_t) const {
#pragma unroll
        for (int r = 0; r < siml::kBitRows; ++r) R.zero[r] = 0u;
    }
    __device__ void apply(Lds& L, Reg& R, int r, int32_t, bool valid, int64_t, int, int, float t, int64_t,
                          int64_t, int) const {
        const bool z = !(t * 100.0f >= 1.0f);
        R.zero[r] += (valid && z) ? 1u : 0u;
        const bool nz = valid && !z;
        if (__ballot(nz) != 0ull && nz) atomicAdd(&L.bins[r][bin101_f32(t)], 1);
    }
    __device__ bool zero_skips() const { return true; }
    __device__ void skip_row(Lds&, Reg& R, int r, int32_t, bool valid, int64_t) const { R.zero[r] += valid ? 1u : 0u; }
    __device__ void end(Lds& L, Reg& R, const BitRowInfo& ri, int64_t, int64_t) const {
#pragma unroll
        for (int r = 0; r < siml::kBitRows; ++r)
            if (ri.row[r] >= 0 && R.zero[r]) atomicAdd(&L.bins[r][0], (int)R.zero[r]);
    }
    __device__ void finish(Lds& L, const BitRowInfo& ri) const {
        for (int k = threadIdx.x; k < siml::kBitRows * 101; k += blockDim.x) {
            const int r = k / 101, b = k - r * 101;
            if (ri.row[r] >= 0 && L.bins[r][b]) atomicAdd(&hist[((int64_t)ri.row[r] - rb) * 101 + b], L.bins[r][b]);
        }
    }
};

// 1000-bin histogram of the whole block: bin-0 pairs (T < fl32(1/1000)) in
// per-lane registers, the rest with LDS atomics.
template <int G>
struct EpiHist1000 {
    static constexpr int kG = G;
    static constexpr bool kPrune = false;
    static constexpr bool kBlockMode = false;
    static constexpr bool kAdmit = false;
    using Rows = BitRows<32 * G>;
    unsigned long long* hist;
    // Pairs with I in [1, kIMax] and U = ma + mb - I below kIU (most
    // non-zero pairs of unrelated molecules) are counted by (I, U), T being a
    // function of those two: the counters spread over many more LDS
    // addresses than the few bins such T fall in, so fewer atomics of one
    // instruction collide.  finish() adds them to their bins.
    // (kIU = 448 keeps two-table workgroups within half of a CU's LDS)
    static constexpr int kIU = 448;
    static constexpr int kIMax = 4;
    struct Lds {
        uint32_t bins[1000];
        float edges[1001];
        uint32_t iu[kIMax * kIU];
    };
    struct Reg {
        uint32_t zero;
    };
    __device__ void init(Lds& L) const {
        for (int k = threadIdx.x; k < 1000; k += blockDim.x) L.bins[k] = 0u;
        for (int k = threadIdx.x; k < kIMax * kIU; k += blockDim.x) L.iu[k] = 0u;
        stage_edges(L.edges);
    }
    __device__ void begin(Lds&, Reg& R, const Rows&, int64_t, int64_t) const { R.zero = 0u; }
    __device__ void apply(Lds& L, Reg& R, int, int32_t, bool valid, int64_t, int ma, int mb, float t, int64_t,
                          int64_t, int isct) const {
        const bool z = !(t >= 1.0f / 1000.0f);  // edge 1 of stage_edges, fl32(1/1000)
        R.zero += (valid && z) ? 1u : 0u;
        const bool nz = valid && !z;
        if (__ballot(nz) == 0ull || !nz) return;
        // t > 0 here, so isct >= 1; when u < kIU, ma + mb is small and t is
        // exactly fl32(isct / u) (u = 0 gives +inf, as (float)isct / 0.0f)
        const int64_t u = (int64_t)ma + mb - isct;
        if (isct <= kIMax && u >= 0 && u < kIU)
            atomicAdd(&L.iu[(isct - 1) * kIU + u], 1u);
        else
            atomicAdd(&L.bins[bin1000(t, L.edges)], 1u);
    }
    __device__ bool zero_skips() const { return true; }
    __device__ void skip_row(Lds&, Reg& R, int, int32_t, bool valid, int64_t) const { R.zero += valid ? 1u : 0u; }
    __device__ void end(Lds& L, Reg& R, const Rows&, int64_t, int64_t) const {
        if (R.zero) atomicAdd(&L.bins[0], R.zero);
    }
    __device__ void finish(Lds& L, const Rows&) const {
        // the float expression of pair_tanimoto: (float)I / (float)U, both
        // exact integers (short_div equals IEEE division there)
        for (int k = threadIdx.x; k < kIMax * kIU; k += blockDim.x)
            if (L.iu[k]) atomicAdd(&L.bins[bin1000((float)(k / kIU + 1) / (float)(k % kIU), L.edges)], L.iu[k]);
        __syncthreads();
        for (int k = threadIdx.x; k < 1000; k += blockDim.x)
            if (L.bins[k]) atomicAdd(&hist[k], (unsigned long long)L.bins[k]);
    }
};

// Neighbour count (FILL = false) / fill pass: one running position per
// (wave, row) in LDS, ballot compaction in ascending storage order.  The
// bound prefilter of gpuGetNeighbors_multirow (kernels.py:451-455) is
// evaluated only where lo <= T < hi already holds; the result is the same
// conjunction.
template <bool FILL, int G>
struct EpiNbr {
    static constexpr int kG = G;
    static constexpr bool kPrune = true;
    static constexpr bool kBlockMode = false;
    static constexpr bool kAdmit = false;
    static constexpr int kRows = 32 * G;
    using Rows = BitRows<kRows>;
    float lo, hi;
    int32_t* counts;
    const int64_t* offs;
    int32_t* idx;
    int64_t rb;
    // count pass only: the first bcap indices of each (row, wave) slot are
    // also kept in buf (slot-major), so that without overflow no fill pass
    // is needed
    int32_t* buf;
    int bcap;
    // evaluated-work counter (null: off): [0] += pairs of the tiles that the
    // magnitude band did not prune, [1] += their sum of La + Lb (unique
    // lingo counts, the merge work the pairs stand for); rlen gives La
    unsigned long long* work;
    const int32_t* rlen;
    struct Lds {
        int64_t pos[4 * G][kRows];
    };
    struct Reg {
        unsigned long long pairs = 0, lsum = 0;
        int nrows = 0, sum_la = 0;
    };
    // True when no pair of the group's rows (magnitudes [amin, amax]) with
    // the tile's columns (magnitudes tm) can pass the bound prefilter:
    // fl(min/max) is monotone in both magnitudes, so the closest pair of
    // the two ranges bounds it.
    __device__ bool skip_tile(int2 tm, int amin, int amax) const {
        if (!(lo > 0.0f) || amin < 1 || tm.x < 1) return false;
        if (tm.y < amin) return (float)tm.y / (float)amin < lo;
        if (amax < tm.x) return (float)amax / (float)tm.x < lo;
        return false;
    }
    __device__ void init(Lds&) const {}
    __device__ void begin(Lds& L, Reg& R, const Rows& ri, int64_t gw, int64_t nw) const {
        const int w = threadIdx.x / kWave, lane = lane_id();
        for (int i = lane; i < kRows; i += kWave)
            L.pos[w][i] = (FILL && ri.row[i] >= 0) ? offs[((int64_t)ri.row[i] - rb) * nw + gw] : 0;
        if (work) {
            int nrows = 0, sum_la = 0;
            for (int i0 = 0; i0 < kRows; i0 += kWave) {
                const int i = i0 + lane;
                const bool live = i < kRows && ri.row[i < kRows ? i : 0] >= 0;
                nrows += __popcll(__ballot(live));
                sum_la += wave_sum_i32(live ? rlen[ri.row[i]] : 0);
            }
            R.nrows = nrows;
            R.sum_la = sum_la;
        }
        wave_lds_sync();
    }
    // one tile survived the magnitude band: lane c's column meets every row
    __device__ void count_tile(Reg& R, const ColArgs& C, bool valid, int64_t c) const {
        if (work && valid) {
            R.pairs += (unsigned)R.nrows;
            R.lsum += (unsigned long long)R.nrows * (unsigned)C.len[c] + (unsigned)R.sum_la;
        }
    }
    __device__ void apply(Lds& L, Reg&, int r, int32_t row, bool valid, int64_t col, int ma, int mb, float t,
                          int64_t gw, int64_t nw, int) const {
        bool nb = valid && t >= lo && t < hi;
        if (__ballot(nb) == 0ull) return;
        if (nb) {
            const float bound = (float)min(ma, mb) / (float)max(ma, mb);
            nb = bound >= lo;
        }
        const uint64_t m = __ballot(nb);
        if (m == 0ull) return;
        const int w = threadIdx.x / kWave;
        const int64_t pos = L.pos[w][r];
        if (FILL && nb) idx[pos + rank_below(m)] = (int32_t)col;
        if (!FILL && buf && nb) {
            const int64_t j = pos + rank_below(m);
            if (j < bcap) buf[(((int64_t)row - rb) * nw + gw) * bcap + j] = (int32_t)col;
        }
        wave_lds_sync();
        if (lane_id() == 0) L.pos[w][r] = pos + __popcll(m);
        wave_lds_sync();
    }
    // T = 0 is a neighbour only when lo <= 0
    __device__ bool zero_skips() const { return lo > 0.0f; }
    __device__ void skip_row(Lds&, Reg&, int, int32_t, bool, int64_t) const {}
    __device__ void end(Lds& L, Reg& R, const Rows& ri, int64_t gw, int64_t nw) const {
        const int w = threadIdx.x / kWave, lane = lane_id();
        if (!FILL)
            for (int i = lane; i < kRows; i += kWave)
                if (ri.row[i] >= 0) counts[((int64_t)ri.row[i] - rb) * nw + gw] = (int32_t)L.pos[w][i];
        if (work) add_work(work, R.pairs, R.lsum);
    }
    __device__ void finish(Lds&, const Rows&) const {}
};

// Top-k candidates per (row, wave).  The per-row state lives in registers:
// row i's threshold key (0 until armed) is lane i % 64 of Reg::thr[i / 64]
// and the number of keys in its buffer lane i % 64 of Reg::cnt[i / 64].  A tile in which no key of a
// row passes costs one 64-bit compare and a ballot; otherwise the passing
// keys are appended, after the buffer has been cut to its k largest if the
// append could overflow it.
template <int G>
struct EpiTopk {
    static constexpr int kG = G;
    static constexpr bool kPrune = false;
    static constexpr bool kBlockMode = false;
    static constexpr bool kAdmit = true;
    static constexpr int kRows = 32 * G;
    static constexpr int kRegs = (kRows + kWave - 1) / kWave;
    using Rows = BitRows<kRows>;
    int k, cap;
    uint64_t* buf;
    int32_t* bufcnt;
    int64_t rb;
    bool pos_only;           // only pairs with T > 0 are candidates
    const uint64_t* theta0;  // per row: a key at most the row's k-th largest (0: none), or null
    struct Lds {
        uint32_t hist[4 * G][256];
        // per row: (1 + s, s, s * ma) of the admission test below
        float4 adm[kRows];
    };
    struct Reg {
        uint64_t thr[kRegs];
        int32_t cnt[kRegs];
    };
    __device__ void init(Lds&) const {}
    // Every wave calls begin once, after the workgroup prologue.
    __device__ void begin(Lds& L, Reg& R, const Rows& ri, int64_t, int64_t) const {
        const int lane = lane_id();
        // Admission constants.  A seeded row buffers only keys >= its seed,
        // so only pairs with T >= th, the seed's T (the key's high word is
        // T's bit pattern with the sign bit set for T >= 0).  With
        // s = th (1 - 2^-10) - 2^-10, every such pair passes
        // I (1 + s) >= s (ma + mb): T = fl(I / U) >= th gives
        // I - s U >= 2^-11 U >= 2^-12 (ma + mb) in exact arithmetic (U >= I,
        // so U >= (ma + mb) / 2), while the float evaluation in admits()
        // errs by less than 2^-20 (ma + mb).  Rows without a seed (or with
        // s <= 0) get (1, 0, 0): every pair passes.
        if (threadIdx.x < kRows) {
            const int i = threadIdx.x;
            float a = 1.0f, b = 0.0f, c = 0.0f;
            if (theta0 && ri.row[i] >= 0) {
                const uint64_t s0 = theta0[(int64_t)ri.row[i] - rb];
                const uint32_t hi = (uint32_t)(s0 >> 32);
                if (s0 != 0ull && (hi & 0x80000000u)) {
                    const float th = __uint_as_float(hi & 0x7FFFFFFFu);
                    const float sl = th * (1.0f - 0x1p-10f) - 0x1p-10f;
                    if (sl > 0.0f) {
                        a = 1.0f + sl;
                        b = sl;
                        c = sl * (float)ri.ma[i];
                    }
                }
            }
            L.adm[i] = make_float4(a, b, c, 0.0f);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kRegs; ++j) {
            // a seeded row admits only keys >= its seed: threshold = seed - 1
            const int i = j * kWave + lane;
            uint64_t th = 0ull;
            if (theta0 && i < kRows && ri.row[i < kRows ? i : 0] >= 0) {
                const uint64_t s0 = theta0[(int64_t)ri.row[i] - rb];
                th = s0 ? s0 - 1 : 0ull;
            }
            R.thr[j] = th;
            R.cnt[j] = 0;
        }
    }
    __device__ void apply(Lds& L, Reg& R, int r, int32_t row, bool valid, int64_t col, int, int, float t,
                          int64_t gw, int64_t nw, int) const {
        const int w = threadIdx.x / kWave, lane = lane_id();
        const int j = r / kWave, rl = r % kWave;
        const uint64_t key = topk_key(t, col);
        uint64_t thr = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(R.thr[j] >> 32), rl) << 32) |
                       (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)R.thr[j], rl);
        valid = valid && (!pos_only || t > 0.0f);
        uint64_t m = __ballot(valid && key > thr);
        if (m == 0ull) return;
        int cnt = __builtin_amdgcn_readlane(R.cnt[j], rl);
        uint64_t* b = buf + (((int64_t)row - rb) * nw + gw) * (int64_t)cap;
        if (cnt + kWave > cap) {
            // cnt > cap - 64 >= k distinct keys: keep the k largest
            thr = wave_keep_topk(b, cnt, k, L.hist[w]);
            cnt = k;
            m = __ballot(valid && key > thr);
        }
        if (valid && key > thr) b[cnt + rank_below(m)] = key;
        cnt += __popcll(m);
        if (lane == rl) {
            R.thr[j] = thr;
            R.cnt[j] = cnt;
        }
    }
    // Whether some lane's pair of row r (intersection fi, column magnitude
    // mbf) can pass the row's seed; wave-uniform.  Rows that cannot skip the
    // division and apply().
    __device__ bool admits(const Lds& L, int r, float fi, float mbf, bool valid) const {
        if (!theta0) return true;
        const float4 p = L.adm[r];
        return __ballot(valid && fi * p.x >= fmaf(p.y, mbf, p.z)) != 0ull;
    }
    // with pos_only a pair with T = 0 is never a candidate
    __device__ bool zero_skips() const { return pos_only; }
    __device__ void skip_row(Lds&, Reg&, int, int32_t, bool, int64_t) const {}
    __device__ void end(Lds&, Reg& R, const Rows& ri, int64_t gw, int64_t nw) const {
        const int lane = lane_id();
#pragma unroll
        for (int j = 0; j < kRegs; ++j) {
            const int i = j * kWave + lane;
            if (i < kRows && ri.row[i < kRows ? i : 0] >= 0) bufcnt[((int64_t)ri.row[i] - rb) * nw + gw] = R.cnt[j];
        }
    }
    __device__ void finish(Lds&, const Rows&) const {}
};

// Tanimoto of one pair from its intersection count fi, the row magnitude
// maf and the column magnitude mbf, as multisetTanimoto3_mags
// (_CPULingo.c:55,70): I / (((float)ma + mb) - I), 0 when a magnitude is 0.
// FAST (set by the host when every magnitude is in [0, 65536) and at least
// the molecule's count sum) takes short_div method 1: then fi < 2^16 and the
// denominator of every pair with both magnitudes non-zero is an integer in
// [1, 2^17), where short_div equals IEEE division (siml_selftest_division).
template <bool FAST>
__device__ __forceinline__ float pair_tanimoto(float fi, float maf, float mbf, bool zero) {
    const float u = (maf + mbf) - fi;
    const float d = FAST ? short_div(fi, u, 1) : fi / u;
    return zero ? 0.0f : d;
}

// Workgroup prologue of the bit-sliced kernels: the group's table and row
// information into LDS.
// Workgroup -> (group slot, column range) of a bit-sliced launch.  The
// dispatcher deals consecutive workgroups to the 8 XCDs in turn; when the
// number of column ranges Y is a multiple of 8, workgroup L goes to XCD
// L % 8 and takes a column range y == L % 8 (mod 8), so each XCD's L2 only
// ever sees an eighth of the column stream; within an XCD consecutive
// workgroups take consecutive groups of the same range, so the workgroups
// in flight share their column data.  Otherwise (x, y) = blockIdx.
constexpr int kXcds = 8;
__device__ __forceinline__ void bits_wg(int& gslot, int& y, int& ny) {
    ny = (int)gridDim.y;
    if (ny % kXcds == 0) {
        const unsigned L = blockIdx.x + gridDim.x * blockIdx.y;
        const unsigned xcd = L % kXcds, slot = L / kXcds;
        gslot = (int)(slot % gridDim.x);
        y = (int)(xcd + kXcds * (slot / gridDim.x));
    } else {
        y = (int)blockIdx.y;
        gslot = (int)blockIdx.x;
    }
}

// The workgroup's GT groups are glist[gslot * GT + q], q < GT; a negative
// entry (padding of the last slot) is an empty group: no live rows, an empty
// table that matches no level.
// Loads the rows and the tables of group slot `gslot` (GT groups) into
// LDS; ends with a workgroup barrier.  The caller has passed a barrier since
// the previous slot's tables were last read.
template <int GT, bool PRUNE>
__device__ __forceinline__ void bits_load_slot(const BitGroupArgs& G, const int32_t* __restrict__ rmag, int64_t rb,
                                               int64_t rl, int gslot, uint2 (*tab)[siml::kBitSlots],
                                               BitRows<32 * GT>& ri) {
    for (int s = threadIdx.x; s < GT * siml::kBitSlots; s += blockDim.x) (&tab[0][0])[s] = make_uint2(0u, 0u);
    if (threadIdx.x < 32 * GT) {
        const int q = threadIdx.x / 32, r = threadIdx.x % 32;
        const int g = G.glist[(int64_t)gslot * GT + q];
        const int nrows = g >= 0 ? G.hdr[g].nrows : 0;
        const int32_t row = r < nrows ? G.rows[(int64_t)g * siml::kBitRowStride + r] : -1;
        const bool live = row >= 0 && row >= rb && row < rl;
        ri.row[threadIdx.x] = live ? row : -1;
        const int32_t ma = live ? rmag[row] : 0;
        ri.ma[threadIdx.x] = ma;
        ri.maf[threadIdx.x] = (float)ma;
    }
    __syncthreads();
    if (PRUNE && threadIdx.x == 0) {
        int lo = INT32_MAX, hi = INT32_MIN;
        for (int r = 0; r < 32 * GT; ++r)
            if (ri.row[r] >= 0) {
                lo = min(lo, ri.ma[r]);
                hi = max(hi, ri.ma[r]);
            }
        ri.amin = lo;
        ri.amax = hi;
    }
#pragma unroll
    for (int q = 0; q < GT; ++q) {
        const int g = G.glist[(int64_t)gslot * GT + q];
        if (g < 0) continue;
        const siml::BitGroupHdr hd = G.hdr[g];
        for (int k = threadIdx.x; k < hd.nkeys; k += blockDim.x) {
            const uint4 e = G.ent[hd.entry_off + k];
            tab[q][e.z] = make_uint2(e.x, e.y);
        }
    }
    __syncthreads();
}

// Prologue of a workgroup of an ordinary grid: its group slot's rows and
// tables, and the epilogue's LDS.
template <int GT, class Epi>
__device__ __forceinline__ void bits_prologue(const BitGroupArgs& G, const int32_t* __restrict__ rmag, int64_t rb,
                                              int64_t rl, const Epi& epi, uint2 (*tab)[siml::kBitSlots],
                                              BitRows<32 * GT>& ri, typename Epi::Lds& el) {
    int gslot, y, ny;
    bits_wg(gslot, y, ny);
    epi.init(el);
    bits_load_slot<GT, Epi::kPrune>(G, rmag, rb, rl, gslot, tab, ri);
}

// Streams the Lt (a multiple of 8) levels of one query molecule against the
// GT tables of the workgroup: each block of 8 levels is loaded once and
// probed in every table, so the level stream is read once per GT groups.
template <int NB, int GT>
__device__ __forceinline__ void bits_accumulate_g(const uint2 (*tab)[siml::kBitSlots], const char* __restrict__ base,
                                                  int Lt, uint32_t (&P)[GT][NB]) {
#pragma unroll
    for (int q = 0; q < GT; ++q)
#pragma unroll
        for (int k = 0; k < NB; ++k) P[q][k] = 0u;
    const int nblk = Lt >> 3;
    if (nblk <= 0) return;
    // one table's 16 table reads are in flight at a time: the scheduling
    // barriers keep the compiler from hoisting every table's reads (and
    // their result registers) ahead of the first table's adder tree
    const uint32_t voff = lane_x16();
    int4 a[4], b[4];
    load8(a, base, voff, 0);
    int k = 0;
    for (; k + 2 <= nblk; k += 2) {
        load8(b, base, voff, k + 1);
#pragma unroll
        for (int q = 0; q < GT; ++q) {
            probe8<NB>(tab[q], a, P[q]);
            if constexpr (GT > 1) __builtin_amdgcn_sched_barrier(0);
        }
        // may read the 8 levels after this molecule's last: the stream
        // carries one spare block at its end, so this stays in bounds
        load8(a, base, voff, k + 2);
#pragma unroll
        for (int q = 0; q < GT; ++q) {
            probe8<NB>(tab[q], b, P[q]);
            if constexpr (GT > 1) __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (k < nblk)
#pragma unroll
        for (int q = 0; q < GT; ++q) probe8<NB>(tab[q], a, P[q]);
}

// First tile of wave gw out of nw: the waves take contiguous, ascending tile
// ranges holding equal shares of the level stream (binary search of the
// level offsets), so waves that share a workgroup's tables finish together
// whatever the storage order.  A sampled launch (tstride > 1) splits by
// tile count.
__device__ __forceinline__ int64_t wave_tile_start(const ColArgs& C, int64_t gw, int64_t nw) {
    if (gw <= 0) return 0;
    if (gw >= nw) return C.ntiles;
    if (C.tstride != 1) return (C.ntiles * gw) / nw;
    const int64_t want = load_uniform(C.btoff, C.ntiles) * gw / nw;
    int64_t lo = 0, hi = C.ntiles;  // smallest t with btoff[t] >= want
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (load_uniform(C.btoff, mid) >= want)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// Float32 matrix, a 256-column block per wave step (EpiF32Block), with
// block_waves<NB>() waves per workgroup held to the registers that let 6
// (8 planes) or 4 (16 planes) waves share each SIMD.
template <int NB, bool FAST>
__global__ __launch_bounds__(block_waves<NB>() * kWave, NB == 8 ? 6 : 4) void k_bits_block(
    BitGroupArgs G, const int32_t* __restrict__ rmag, ColArgs C, int64_t rb, int64_t rl, EpiF32Block epi) {
    __shared__ uint2 tabs[1][siml::kBitSlots];
    __shared__ BitRowInfo ri;
    __shared__ EpiF32Block::Lds el;
    __shared__ float stage_rows[block_waves<NB>()][5 * kWave];
    bits_prologue<1>(G, rmag, rb, rl, epi, tabs, ri, el);
    const uint2* tab = tabs[0];
    const int lane = lane_id();
    const int4* bt4 = reinterpret_cast<const int4*>(C.btiles);
    const int wpb = blockDim.x / kWave, w = threadIdx.x / kWave;
    int gslot, y, ny;
    bits_wg(gslot, y, ny);
    // wave-uniform block loop: tile offsets and lengths are scalar loads
    const int64_t gw = __builtin_amdgcn_readfirstlane(y * wpb + w), nw = (int64_t)ny * wpb;
    const int64_t nblocks = (C.ntiles + 3) / 4;
    const int64_t b0 = (nblocks * gw) / nw, b1 = (nblocks * (gw + 1)) / nw;
    float* stage = stage_rows[w];  // 256 column slots + 64 for lanes past the end
    // the rows' caller indices and magnitudes in lane registers (lane r holds
    // row r), and which rows are live, read once
    const int32_t row_lane = ri.row[lane & (siml::kBitRows - 1)];
    const int32_t ma_lane = ri.ma[lane & (siml::kBitRows - 1)];
    const uint32_t live = (uint32_t)__ballot(lane < siml::kBitRows && row_lane >= 0);
    // tile indices and in-tile column counts are 32-bit (tiles < 2^31): the
    // block loop's tests are then scalar compares, with no 64-bit values
    // copied to vector registers
    const int ntl = (int)C.ntiles;
    for (int64_t blk = b0; blk < b1; ++blk) {
        uint32_t Q[4][NB];
        float mbf[4];
        bool zb[4];
        int loc[4];
        int64_t toff[4];
        int32_t tlen[4];
        const int tb = (int)blk * 4;
        // the four tiles' level offsets and lengths: unconditional scalar
        // loads (clamped to the last tile) issued together, a missing tile
        // gets length 0
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = min(tb + q, ntl - 1);
            toff[q] = load_uniform(C.btoff, t);
            const int32_t len = load_uniform(C.btlen, t);
            tlen[q] = tb + q < ntl ? len : 0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            bits_accumulate<NB>(tab, reinterpret_cast<const char*>(bt4 + (toff[q] >> 1)), tlen[q], Q[q]);

            transpose8(Q[q]);
            if constexpr (NB == 16) transpose8(Q[q] + 8);
        }
        // the columns' magnitudes and in-block positions, read after the
        // accumulation so that they hold no registers during it
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t c0 = (int64_t)(tb + q) * kWave;
            const int in_tile = (int)std::max<int64_t>(0, std::min<int64_t>(C.n - c0, kWave));
            const bool valid = lane < in_tile;
            const int64_t c = c0 + lane;
            const int mb = valid ? C.mag[c] : 0;
            mbf[q] = (float)mb;
            zb[q] = mb == 0;
            loc[q] = valid ? (int)(col_of(C, c) - blk * 4 * kWave) : 4 * kWave + lane;
        }
        const int64_t cbase = blk * 4 * kWave;
        // columns of this block (4 kWave except in the last block), so that
        // the store test is one 32-bit compare per lane and keeps no 64-bit
        // column values live across the rows
        const int rem = (int)std::min<int64_t>(C.n - cbase, 4 * kWave);
        const int o4 = 4 * lane;
        // opaque to the compiler per block: otherwise it hoists the 32 row
        // tests below out of the block loop into 64 live SGPRs
        uint32_t lv = live;
        asm volatile("" : "+s"(lv));
#pragma unroll
        for (int r = 0; r < siml::kBitRows; ++r) {
            if (!((lv >> r) & 1u)) continue;
            const int32_t row = __builtin_amdgcn_readlane(row_lane, r);
            const int ma = __builtin_amdgcn_readlane(ma_lane, r);
            const bool za = ma == 0;
            const float maf = (float)ma;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                stage[loc[q]] = pair_tanimoto<FAST>(count_of<NB>(Q[q], r), maf, mbf[q], za || zb[q]);
            wave_lds_sync();
            float* orow = epi.out + (int64_t)(row - epi.rb) * epi.opitch + cbase;
            if (o4 + 4 <= rem) {
                // the lane index recomputed here (two VALU ops) rather than
                // held: under the register limit the compiler otherwise
                // spills it and reloads it for every row, and the reload's
                // wait also waits for the previous rows' stores
                int l;
                asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
                store_stream(orow + 4 * l, reinterpret_cast<const float4*>(stage)[l]);
            } else {
                for (int i = 0; i < 4 && o4 + i < rem; ++i) orow[o4 + i] = stage[o4 + i];
            }
            wave_lds_sync();
        }
    }
}

// One work item of a bit-sliced workgroup: column range y of ny against the
// rows and tables in LDS (loaded by bits_load_slot).  Initialises the
// epilogue's LDS, and ends with a barrier after the epilogue's finish.
template <int NB, bool FAST, class Epi>
__device__ __forceinline__ void bits_item(const ColArgs& C, const Epi& epi, const uint2 (*tab)[siml::kBitSlots],
                                          const BitRows<32 * Epi::kG>& ri, typename Epi::Lds& el, int y, int ny) {
    constexpr int GT = Epi::kG;
    epi.init(el);
    __syncthreads();
    const int lane = lane_id();
    const int4* bt4 = reinterpret_cast<const int4*>(C.btiles);
    typename Epi::Reg R;
    int64_t t0, t1, gw, nw;
    {
        const int wpb = blockDim.x / kWave;
        // wave-uniform: the tile loop, its offsets and lengths then live in
        // scalar registers and are read through the scalar cache
        gw = __builtin_amdgcn_readfirstlane(y * wpb + (int)(threadIdx.x / kWave));
        nw = (int64_t)ny * wpb;
        t0 = wave_tile_start(C, gw, nw);
        t1 = wave_tile_start(C, gw + 1, nw);
    }
    epi.begin(el, R, ri, gw, nw);
    // the rows' caller indices and magnitudes in lane registers (lane r of
    // table q: row q * 32 + r), and which rows are live, read once
    int32_t row_lane[GT], ma_lane[GT];
    uint32_t live[GT];
#pragma unroll
    for (int q = 0; q < GT; ++q) {
        const int i = q * siml::kBitRows + (lane & (siml::kBitRows - 1));
        row_lane[q] = ri.row[i];
        ma_lane[q] = ri.ma[i];
        live[q] = (uint32_t)__ballot(lane < siml::kBitRows && row_lane[q] >= 0);
    }
    // With consistent magnitudes (FAST) a pair whose intersection is 0 has
    // T = +0 exactly; an epilogue that can take such pairs in bulk skips the
    // rows none of a tile's columns intersects (one wave-wide OR per table
    // and tile).
    const bool skip = FAST && epi.zero_skips();
    // scalar loads of the next tile's level offset and length are issued
    // while the current tile is probed
    int64_t next_off = t0 < t1 ? load_uniform(C.btoff, t0 * C.tstride) : 0;
    int32_t next_len = t0 < t1 ? load_uniform(C.btlen, t0 * C.tstride) : 0;
    for (int64_t ti = t0; ti < t1; ++ti) {
        const int64_t t = ti * C.tstride;
        const int64_t off = next_off;
        const int32_t len = next_len;
        if (ti + 1 < t1) {
            next_off = load_uniform(C.btoff, t + C.tstride);
            next_len = load_uniform(C.btlen, t + C.tstride);
        }
        if constexpr (Epi::kPrune) {
            if (epi.skip_tile(load_uniform(C.tmag, t), ri.amin, ri.amax)) continue;
        }
        const int64_t c = t * kWave + lane;
        const bool valid = c < C.n;
        // the columns' magnitudes and positions, in flight during the probes
        const int mb = valid ? C.mag[c] : 0;
        const int64_t col = valid ? col_of(C, c) : 0;
        uint32_t P[GT][NB];
        bits_accumulate_g<NB, GT>(tab, reinterpret_cast<const char*>(bt4 + (off >> 1)), len, P);
        // bit r of rows_any[q]: row r of table q has a non-zero count in some
        // lane of the tile
        uint32_t rows_any[GT];
#pragma unroll
        for (int q = 0; q < GT; ++q) {
            rows_any[q] = 0xFFFFFFFFu;
            if (skip) {
                uint32_t nz = P[q][0];
#pragma unroll
                for (int k = 1; k < NB; ++k) nz |= P[q][k];
                rows_any[q] = wave_or_u32(valid ? nz : 0u);
            }
            transpose8(P[q]);
            if constexpr (NB == 16) transpose8(P[q] + 8);
        }
        if constexpr (Epi::kPrune) epi.count_tile(R, C, valid, c);
        const float mbf = (float)mb;
#pragma unroll
        for (int q = 0; q < GT; ++q) {
#pragma unroll
            for (int r = 0; r < siml::kBitRows; ++r) {
                const int rr = q * siml::kBitRows + r;
                if (!((live[q] >> r) & 1u)) continue;
                const int32_t row = __builtin_amdgcn_readlane(row_lane[q], r);
                if (!((rows_any[q] >> r) & 1u)) {
                    epi.skip_row(el, R, rr, row, valid, col);
                    continue;
                }
                const float fi = count_of<NB>(P[q], r);
                if constexpr (Epi::kAdmit) {
                    if (!epi.admits(el, rr, fi, mbf, valid)) continue;
                }
                const int ma = __builtin_amdgcn_readlane(ma_lane[q], r);
                const float tv = pair_tanimoto<FAST>(fi, (float)ma, mbf, ma == 0 || mb == 0);
                epi.apply(el, R, rr, row, valid, col, ma, mb, tv, gw, nw, (int)fi);
            }
        }
    }
    epi.end(el, R, ri, gw, nw);
    __syncthreads();
    epi.finish(el, ri);
    __syncthreads();
}

// Every other epilogue: each wave walks a contiguous range of tiles and
// probes the workgroup's Epi::kG tables (32 rows each) with every level it
// loads.  An ordinary grid has one work item per workgroup: (group slot,
// column range) from the block index (bits_wg).  A persistent grid (G.queue
// set, about as many workgroups as are resident at once) takes group slots
// from kXcds queues instead: queue x hands out each slot once, and the
// workgroup that takes it loads the slot's tables once and runs its column
// ranges x, x + kXcds, ... (the ranges bits_wg gives XCD x).  A workgroup
// starts at queue (its block index mod kXcds), the XCD the hardware
// dispatcher places it on, and when that queue is empty goes on to the
// next; it exits when all kXcds are empty, so every workgroup finishes
// after at most nslot + 1 takes per queue.
// Workgroups serving several tables take the LDS of 2 or 4 single-table
// ones, so they are held to 128 VGPRs: 4 waves per SIMD (the second bound
// is the minimum number of waves per SIMD).
template <int NB, bool FAST, class Epi>
__global__ __launch_bounds__(kBlock * Epi::kG, Epi::kG > 1 ? 4 : 1) void k_bits(
    BitGroupArgs G, const int32_t* __restrict__ rmag, ColArgs C, int64_t rb, int64_t rl, Epi epi) {
    constexpr int GT = Epi::kG;
    __shared__ uint2 tab[GT][siml::kBitSlots];
    __shared__ BitRows<32 * GT> ri;
    __shared__ typename Epi::Lds el;
    __shared__ int taken;
    if (!G.queue) {
        int gslot, y, ny;
        bits_wg(gslot, y, ny);
        bits_load_slot<GT, Epi::kPrune>(G, rmag, rb, rl, gslot, tab, ri);
        bits_item<NB, FAST, Epi>(C, epi, tab, ri, el, y, ny);
        return;
    }
    const int home = (int)(blockIdx.x % kXcds);
    for (int tried = 0; tried < kXcds;) {
        const int x = (home + tried) % kXcds;
        if (threadIdx.x == 0) taken = atomicAdd(G.queue + x, 1);
        __syncthreads();
        const int slot = taken;
        __syncthreads();
        if (slot >= G.nslot || x >= G.ny) {
            ++tried;
            continue;
        }
        bits_load_slot<GT, Epi::kPrune>(G, rmag, rb, rl, slot, tab, ri);
        for (int y = x; y < G.ny; y += kXcds) bits_item<NB, FAST, Epi>(C, epi, tab, ri, el, y, G.ny);
    }
}


// ---------------------------------------------------------------------------
// Sparse path of the float32 matrix: an inverted index of the query set.
//
// The level stream of a query set holds, for every query molecule, the keys
// of its levels (x, t) (bitgroups.h).  The index maps each distinct key to
// the query molecules (caller column indices) holding that level: an
// open-addressing table of S = 2^k slots (key, count, offset) over one
// posting array.  A grouped reference row's intersection with every query
// molecule is then the number of its levels' postings naming that molecule,
// counted in LDS; the row's float32 Tanimoto values follow from the counts
// for all columns, most of them 0.  When the query molecules share few
// levels (postings are short) the counting is a small fraction of the
// matrix's pairs and the kernel is bound by writing the matrix; the handle
// chooses this path per call from the index's posting statistics (the
// expected postings per level).
// ---------------------------------------------------------------------------

// Level key of (x, t), as siml::level_key (bitgroups.h) on the host.
__device__ __forceinline__ uint32_t dev_level_key(int32_t x, uint32_t t) {
    return (uint32_t)x | ((t & 1u) << 7) | ((t & 2u) << 14) | ((t & 4u) << 21) | ((t & 8u) << 28);
}

// Home slot of a key in the index table (a 32-bit mix, masked).
__device__ __forceinline__ uint32_t post_home(uint32_t key, uint32_t mask) {
    uint32_t h = key * 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h & mask;
}

// The index while it is built: separate key, count and offset arrays.
struct PostingArgs {
    uint32_t* key;        // slot keys, 0 = empty
    int32_t* cnt;         // postings per slot
    const int64_t* off;   // first posting of each slot (exclusive scan of cnt)
    int32_t* col;         // caller column indices, grouped by slot
    int32_t* cur;         // fill cursors (index build only)
    uint32_t mask;        // slots - 1
    int32_t* overflow;    // set when a key found no slot (index build only)
};

// Slot of `key`, or -1 when the table holds no such key.  Linear probing
// from the home slot ends at the key or at an empty slot (the table has at
// least twice as many slots as the stream has entries, unless capped; a
// full table stops after one round).
__device__ __forceinline__ int64_t post_find(const PostingArgs& P, uint32_t key) {
    uint32_t s = post_home(key, P.mask);
    for (uint32_t i = 0; i <= P.mask; ++i) {
        const uint32_t k = P.key[s];
        if (k == key) return s;
        if (k == 0u) return -1;
        s = (s + 1) & P.mask;
    }
    return -1;
}

// The index as the sparse kernel reads it: one 16-byte slot per key (key,
// posting count, first posting), so a probe is one load, and the postings
// as uint16 when the query set has at most 65536 molecules (int32 else).
struct PostingQuery {
    const uint4* slot;        // (key, count, first posting low, high); key 0 = empty
    const int32_t* col;       // postings (int32), when col16 is null
    const uint16_t* col16;    // postings (uint16), or null
    uint32_t mask;            // slots - 1
    const int64_t* drow;      // per slot: its row of the chunk directory
    const int32_t* dir;       // chunk directory: row r, entry y = first posting of chunk y (of ndir - 1)
    int32_t ndir;             // entries per directory row: chunks + 1
};

// Posting range of `key`: true with (first, count) when the index holds it
// (and its slot in *slot_out when that is not null).
__device__ __forceinline__ bool post_lookup(const PostingQuery& Q, uint32_t key, int64_t& first, int32_t& n,
                                            uint32_t* slot_out = nullptr) {
    uint32_t s = post_home(key, Q.mask);
    for (uint32_t i = 0; i <= Q.mask; ++i) {
        const uint4 v = Q.slot[s];
        if (v.x == key) {
            first = (int64_t)((uint64_t)v.z | ((uint64_t)v.w << 32));
            n = (int32_t)v.y;
            if (slot_out) *slot_out = s;
            return true;
        }
        if (v.x == 0u) return false;
        s = (s + 1) & Q.mask;
    }
    return false;
}

// 1 for each slot that holds a key (the directory rows are numbered by an
// exclusive scan of these).
__global__ __launch_bounds__(kBlock) void k_post_occupied(const int32_t* __restrict__ cnt, int64_t slots,
                                                          int32_t* __restrict__ occ) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += (int64_t)gridDim.x * blockDim.x)
        occ[i] = cnt[i] > 0 ? 1 : 0;
}

// Chunk directory after the fill of chunk y: each occupied slot's fill
// cursor (its postings of chunks 0..y) becomes entry y + 1 of its row
// (entry 0 is 0).
__global__ __launch_bounds__(kBlock) void k_post_dir(const int32_t* __restrict__ cnt, const int32_t* __restrict__ cur,
                                                     const int64_t* __restrict__ drow, int64_t slots,
                                                     int32_t* __restrict__ dir, int32_t ndir, int32_t y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += (int64_t)gridDim.x * blockDim.x) {
        if (cnt[i] <= 0) continue;
        int32_t* d = dir + drow[i] * ndir;
        if (y == 0) d[0] = 0;
        d[y + 1] = cur[i];
    }
}

// Index build, last pass: the build's separate key / count / offset arrays
// packed into 16-byte slots.
__global__ __launch_bounds__(kBlock) void k_post_pack(const uint32_t* __restrict__ key, const int32_t* __restrict__ cnt,
                                                      const int64_t* __restrict__ off, int64_t slots,
                                                      uint4* __restrict__ slot) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t o = (uint64_t)off[i];
        slot[i] = make_uint4(key[i], (uint32_t)cnt[i], (uint32_t)o, (uint32_t)(o >> 32));
    }
}

// Postings narrowed to uint16 (caller columns < 65536).
__global__ __launch_bounds__(kBlock) void k_post_narrow(const int32_t* __restrict__ col, int64_t n,
                                                        uint16_t* __restrict__ col16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        col16[i] = (uint16_t)col[i];
}

// Visits every level of the level stream: f(key, storage position).  One
// wave per tile, lane l holds the levels of the tile's molecule l; padding
// levels (key 0) are skipped.
template <class F>
__device__ __forceinline__ void for_stream_levels(const ColArgs& C, F&& f) {
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    for (int64_t t = wave; t < C.ntiles; t += nwaves) {
        const int64_t p = t * kWave + lane;
        if (p >= C.n) continue;
        const int64_t base = load_uniform(C.btoff, t);
        const int32_t lt = load_uniform(C.btlen, t);
        for (int32_t j = 0; j < lt; ++j) {
            const int2 e = C.btiles[base + (int64_t)(j >> 1) * 2 * kWave + 2 * lane + (j & 1)];
            if (e.x != 0) f((uint32_t)e.x, p);
        }
    }
}

// Index build, pass 1: every stream level's key gets a slot (inserted by
// compare-and-swap) and counts one posting there.
__global__ __launch_bounds__(kBlock) void k_post_insert(ColArgs C, PostingArgs P) {
    for_stream_levels(C, [&](uint32_t key, int64_t) {
        uint32_t s = post_home(key, P.mask);
        for (uint32_t i = 0; i <= P.mask; ++i) {
            uint32_t k = P.key[s];
            if (k == 0u) k = atomicCAS(P.key + s, 0u, key);
            if (k == 0u || k == key) {
                atomicAdd(P.cnt + s, 1);
                return;
            }
            s = (s + 1) & P.mask;
        }
        atomicExch(P.overflow, 1);
    });
}

// Index build, pass 2 (after the scan of the counts into P.off): every
// stream level writes its caller column into its slot's postings.
__global__ __launch_bounds__(kBlock) void k_post_fill(ColArgs C, PostingArgs P) {
    for_stream_levels(C, [&](uint32_t key, int64_t p) {
        const int64_t s = post_find(P, key);
        if (s < 0) return;
        const int32_t at = atomicAdd(P.cur + s, 1);
        P.col[P.off[s] + at] = (int32_t)col_of(C, p);
    });
}

// Sum over slots of count^2 (stats[0]), of count (stats[1]) and the
// occupied slots (stats[2]): the expected postings of a stream level drawn
// at random is stats[0] / stats[1].
__global__ __launch_bounds__(kBlock) void k_post_stats(const int32_t* __restrict__ cnt, int64_t slots,
                                                       unsigned long long* stats) {
    unsigned long long sq = 0, sum = 0, used = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long c = (unsigned long long)cnt[i];
        sq += c * c;
        sum += c;
        used += c > 0 ? 1ull : 0ull;
    }
    sq = wave_sum_u64(sq);
    sum = wave_sum_u64(sum);
    used = wave_sum_u64(used);
    if (lane_id() == 0) {
        atomicAdd(stats, sq);
        atomicAdd(stats + 1, sum);
        atomicAdd(stats + 2, used);
    }
}

// Storage position of every caller column (the inverse of the layout's
// permutation).
__global__ __launch_bounds__(kBlock) void k_inv_perm(const int32_t* __restrict__ perm, int64_t n,
                                                     int32_t* __restrict__ inv) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        inv[perm[p]] = (int32_t)p;
}

// Index build, pass 2 for the caller columns [c0, c1) only: a thread per
// column writes the column into the postings of its levels.  Run once per
// column chunk, in chunk order, it leaves every key's postings grouped by
// chunk (chunk 0's columns first, in no particular order within a chunk),
// so a walk over the chunks in order reads each posting list front to back.
__global__ __launch_bounds__(kBlock) void k_post_fill_cols(ColArgs C, PostingArgs P, const int32_t* __restrict__ inv,
                                                           int64_t c0, int64_t c1) {
    const int64_t c = c0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    const int64_t p = inv ? (int64_t)inv[c] : c;
    const int64_t t = p / kWave;
    const int lane = (int)(p % kWave);
    const int64_t base = C.btoff[t];
    const int32_t lt = C.btlen[t];
    for (int32_t j = 0; j < lt; ++j) {
        const int2 e = C.btiles[base + (int64_t)(j >> 1) * 2 * kWave + 2 * lane + (j & 1)];
        if (e.x == 0) continue;
        const int64_t s = post_find(P, (uint32_t)e.x);
        if (s < 0) continue;
        const int32_t at = atomicAdd(P.cur + s, 1);
        P.col[P.off[s] + at] = (int32_t)c;
    }
}

// Magnitudes in caller column order (4 trailing zeros: the sparse kernel
// reads them four at a time).
__global__ __launch_bounds__(kBlock) void k_caller_mags(ColArgs C, int32_t* __restrict__ cmag) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < C.n + 4; p += (int64_t)gridDim.x * blockDim.x)
        if (p < C.n)
            cmag[col_of(C, p)] = C.mag[p];
        else
            cmag[p] = 0;
}

// The caller-order magnitudes narrowed to one byte (valid when every
// magnitude is at most 255) and to two (below 65536), each followed by 32
// zeros, and their largest value.
__global__ __launch_bounds__(kBlock) void k_caller_mags_small(const int32_t* __restrict__ cmag, int64_t n,
                                                              uint8_t* __restrict__ m8, uint16_t* __restrict__ m16,
                                                              unsigned* maxmag) {
    unsigned most = 0u;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n + 32; p += (int64_t)gridDim.x * blockDim.x) {
        const unsigned v = p < n ? (unsigned)max(cmag[p], 0) : 0u;
        most = v > most ? v : most;
        m8[p] = (uint8_t)min(v, 255u);
        m16[p] = (uint16_t)min(v, 65535u);
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const unsigned o = (unsigned)__shfl_xor((int)most, off);
        most = o > most ? o : most;
    }
    if (lane_id() == 0) atomicMax(maxmag, most);
}

// Reference rows per workgroup and query molecules per chunk of the sparse
// kernel: one byte per pair in LDS, kSparseRows x kSparseCols = 32 KiB, so
// that four 8-wave workgroups share a CU.  Measured on the 32768^2 matrix:
// 1 row and 512 threads 0.80 ms; 2 rows (64 KiB) and 1024 threads 0.88, 512
// threads 1.10; 4 rows of 16384 columns 1.21; 1 row and 1024 threads 0.94,
// 256 threads 1.01.
constexpr int kSparseRows = 1;
constexpr int kSparseCols = 32768;
// most levels of a listed row (a byte holds every count)
constexpr int kSparseMaxLevels = 255;

// Float32 rows of the listed reference rows (grouped rows with at most 255
// levels) from the inverted index: workgroup (x, y) counts rows
// [kSparseRows x, kSparseRows (x + 1)) of the list against query molecules
// [y kSparseCols, (y + 1) kSparseCols), then writes their Tanimoto values
// with the same float expression as every other kernel.  Counting first
// lists the rows' levels with their posting ranges in LDS; each wave then
// walks whole levels, its lanes reading 64 consecutive postings at a time,
// so a level with long postings is spread over a wave.
constexpr int kSparseBlock = 512;  // 8 waves per workgroup, 32 per CU: posting reads in flight

template <bool NARROW>
__device__ __forceinline__ int64_t post_col(const PostingQuery& P, int64_t i) {
    if constexpr (NARROW)
        return (int64_t)P.col16[i];
    else
        return (int64_t)P.col[i];
}

template <bool FAST, bool NARROW>
__global__ __launch_bounds__(kSparseBlock) void k_sparse_f32(RowArgs R, const int32_t* __restrict__ rowlist,
                                                             int64_t nlist, PostingQuery P,
                                                             const int32_t* __restrict__ cmag, int64_t nq, int64_t rb,
                                                             float* __restrict__ out, int64_t pitch) {
    __shared__ uint32_t cnt[kSparseRows][kSparseCols / 4];
    __shared__ int64_t lv_start[kSparseRows * kSparseMaxLevels];
    __shared__ int32_t lv_len[kSparseRows * kSparseMaxLevels];
    __shared__ int32_t nlv;
    const int64_t c0 = (int64_t)blockIdx.y * kSparseCols;
    const int cw = (int)min((int64_t)kSparseCols, nq - c0);
    int32_t rows[kSparseRows];
#pragma unroll
    for (int i = 0; i < kSparseRows; ++i) {
        const int64_t k = (int64_t)blockIdx.x * kSparseRows + i;
        rows[i] = k < nlist ? load_uniform(rowlist, k) : -1;
    }
    for (int i = threadIdx.x; i < kSparseRows * kSparseCols / 4; i += blockDim.x) (&cnt[0][0])[i] = 0u;
    if (threadIdx.x == 0) nlv = 0;
    __syncthreads();
    // the rows' levels found in the index: (posting start, count, row)
#pragma unroll
    for (int i = 0; i < kSparseRows; ++i) {
        if (rows[i] < 0) continue;
        const int64_t o = R.off[rows[i]];
        const int32_t len = R.len[rows[i]];
        for (int e = threadIdx.x; e < len; e += blockDim.x) {
            const int2 el = R.elems[o + e];
            for (int t = 1; t <= el.y; ++t) {
                int64_t first;
                int32_t cnt_t;
                // no query molecule has level t, so none has t + 1
                if (!post_lookup(P, dev_level_key(el.x, (uint32_t)t), first, cnt_t)) break;
                const int at = atomicAdd(&nlv, 1);
                lv_start[at] = first;
                lv_len[at] = cnt_t | (i << 28);
            }
        }
    }
    __syncthreads();
    const int n = nlv;
    const int lane = lane_id(), w = (int)(threadIdx.x / kWave), nwv = (int)(blockDim.x / kWave);
    // each wave walks whole levels; a level with more than 192 postings is
    // read four passes at a time (four independent reads in flight per lane)
    for (int l = w; l < n; l += nwv) {
        const int64_t st = lv_start[l];
        const int32_t lr = lv_len[l];
        const int32_t ln = lr & 0x0FFFFFFF;
        uint32_t* crow = cnt[lr >> 28];
        int q = lane;
        for (; q + 3 * kWave < ln; q += 4 * kWave) {
            int64_t c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) c[u] = post_col<NARROW>(P, st + q + u * kWave) - c0;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (c[u] >= 0 && c[u] < cw) atomicAdd(crow + (c[u] >> 2), 1u << (8 * (c[u] & 3)));
        }
        for (; q < ln; q += kWave) {
            const int64_t c = post_col<NARROW>(P, st + q) - c0;
            if (c >= 0 && c < cw) atomicAdd(crow + (c >> 2), 1u << (8 * (c & 3)));
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSparseRows; ++i) {
        if (rows[i] < 0) continue;
        const int ma = R.mag[rows[i]];
        const float maf = (float)ma;
        float* orow = out + ((int64_t)rows[i] - rb) * pitch + c0;
        for (int j = threadIdx.x; 4 * j < cw; j += blockDim.x) {
            const uint32_t wd = cnt[i][j];
            const int4 mb = *reinterpret_cast<const int4*>(cmag + c0 + 4 * j);
            float4 v;
            v.x = pair_tanimoto<FAST>((float)(wd & 0xFFu), maf, (float)mb.x, ma == 0 || mb.x == 0);
            v.y = pair_tanimoto<FAST>((float)((wd >> 8) & 0xFFu), maf, (float)mb.y, ma == 0 || mb.y == 0);
            v.z = pair_tanimoto<FAST>((float)((wd >> 16) & 0xFFu), maf, (float)mb.z, ma == 0 || mb.z == 0);
            v.w = pair_tanimoto<FAST>((float)(wd >> 24), maf, (float)mb.w, ma == 0 || mb.w == 0);
            if (4 * j + 4 <= cw) {
                store_stream(orow + 4 * j, v);
            } else {
                const float vv[4] = {v.x, v.y, v.z, v.w};
                for (int k = 0; k < cw - 4 * j; ++k) orow[4 * j + k] = vv[k];
            }
        }
    }
}

// Sparse top-k: candidate keys a workgroup holds in LDS, and admitted pairs
// a wave queues before it computes their keys.
constexpr int kSparseTopkCap = 768;
constexpr int kSparseQueue = 2 * kWave;

// Bit 7 of each byte of w that is at least n (1 <= n <= 128): per byte,
// bit 7 of (b & 0x7F) + 128 - n (no carry leaves the byte) or of b itself.
__device__ __forceinline__ uint32_t bytes_ge(uint32_t w, uint32_t n) {
    return (((w & 0x7F7F7F7Fu) + (128u - n) * 0x01010101u) | w) & 0x80808080u;
}

// The intersection of the pair behind a threshold key: the key holds
// T0 = fl(I0 / (ma + mb0 - I0)) and T(I) = fl(I / (ma + mb0 - I)) grows
// strictly with I, so one I in [1, 255] gives T0; 0 when none does.
__device__ uint32_t sparse_recover_isct(float t0, float maf, float mbf) {
    const int i0 = (int)rintf(t0 * (maf + mbf) / (1.0f + t0));
    for (int d = -2; d <= 2; ++d) {
        const int ii = i0 + d;
        if (ii >= 1 && ii <= 255 && (float)ii <= maf && (float)ii <= mbf &&
            pair_tanimoto<true>((float)ii, maf, mbf, false) == t0)
            return (uint32_t)ii;
    }
    return 0u;
}

// The row's levels found in the index (Lingo level keys; a level t missing
// from the index ends its Lingo's levels, since a molecule with level t + 1
// has level t): per level its first posting and its row of the chunk
// directory; *nlv counts the levels.  Every thread of the workgroup calls,
// between barriers.
template <bool NARROW>
__device__ void sparse_row_levels(const RowArgs& R, int64_t row, const PostingQuery& P, int64_t* lv_first,
                                  int32_t* lv_drow, uint32_t* nlv) {
    const int64_t o = R.off[row];
    const int32_t len = R.len[row];
    for (int e = threadIdx.x; e < len; e += blockDim.x) {
        const int2 el = R.elems[o + e];
        for (int t = 1; t <= el.y; ++t) {
            int64_t first;
            int32_t n;
            uint32_t slot;
            if (!post_lookup(P, dev_level_key(el.x, (uint32_t)t), first, n, &slot)) break;
            const int at = (int)atomicAdd(nlv, 1u);
            lv_first[at] = first;
            lv_drow[at] = (int32_t)P.drow[slot];
        }
    }
}

// Per-chunk arrays of the walk: each level's first posting of the chunk and
// the exclusive prefix of the levels' posting counts in the chunk.
struct SparseChunkLds {
    int64_t st[kSparseMaxLevels];
    int32_t pre[kSparseMaxLevels + 1];
};

// The directory entries of chunk y for this thread's level (thread l takes
// level l < nl; nl <= 255 < the workgroup): loaded ahead of their use, so
// that the loads' latency overlaps other work; sparse_chunk_store puts
// them in W.
struct SparseRange {
    int32_t a = 0, b = 0;
};
__device__ __forceinline__ SparseRange sparse_chunk_load(const int32_t* lv_drow, const PostingQuery& P, int nl,
                                                         int64_t y) {
    SparseRange r;
    if ((int)threadIdx.x < nl) {
        const int32_t* d = P.dir + (int64_t)lv_drow[threadIdx.x] * P.ndir + y;
        r.a = d[0];
        r.b = d[1];
    }
    return r;
}
__device__ __forceinline__ void sparse_chunk_store(SparseChunkLds& W, const int64_t* lv_first, int nl,
                                                   const SparseRange& r) {
    if ((int)threadIdx.x < nl) {
        W.st[threadIdx.x] = lv_first[threadIdx.x] + r.a;
        W.pre[threadIdx.x] = r.b - r.a;
    }
}

// Counts of the row's pairs with the chunk of columns [c0, c0 + 32768)
// into the byte counters cnt (cleared by the caller), from W (each level's
// first posting and posting count of the chunk, stored before the barrier
// that precedes the call): each wave takes an eighth of the chunk's
// postings of all levels in order, 64 consecutive ones per read and four
// reads in flight, a lane's level advancing from the previous read's last
// one (a binary search of the prefix of the counts places the wave's first
// posting; one per posting measured 123 ms against 110 on the batched
// top-100).  Every thread of the workgroup calls, between barriers; it has
// barriers.
template <bool NARROW>
__device__ void sparse_walk_chunk(uint32_t* cnt, SparseChunkLds& W, int nl, const PostingQuery& P, int64_t c0) {
    const int lane = lane_id();
    if (threadIdx.x < kWave) {
        // exclusive prefix over at most 256 levels, 4 per lane
        int32_t v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = 4 * lane + i < nl ? W.pre[4 * lane + i] : 0;
        const int32_t sl = v[0] + v[1] + v[2] + v[3];
        int32_t S = sl;
        for (int off = 1; off < kWave; off <<= 1) {
            const int32_t t = __shfl_up(S, off);
            if (lane >= off) S += t;
        }
        int32_t run = S - sl;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (4 * lane + i <= nl) W.pre[4 * lane + i] = run;
            run += v[i];
        }
    }
    __syncthreads();
    const int32_t total = W.pre[nl];
    const int32_t c0i = (int32_t)c0;
    // wave w takes postings [s0, s1) of the chunk, 64 consecutive ones per
    // read; a lane's level advances from the previous read's last level
    constexpr int kWaves = kSparseBlock / kWave;
    const int w = (int)(threadIdx.x / kWave);
    const int32_t per = (total + kWaves - 1) / kWaves;
    const int32_t s0 = min(total, w * per), s1 = min(total, s0 + per);
    if (s0 >= s1) return;
    int lv = 0;
    {
        int lo = 0, hi = nl - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (W.pre[mid] <= s0)
                lo = mid;
            else
                hi = mid - 1;
        }
        lv = lo;
    }
    for (int32_t g0 = s0; g0 < s1; g0 += 4 * kWave) {
        int32_t c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int32_t g = g0 + u * kWave + lane;
            c[u] = -1;
            int l = lv;
            if (g < s1) {
                while (W.pre[l + 1] <= g) ++l;
                c[u] = (int32_t)post_col<NARROW>(P, W.st[l] + (g - W.pre[l]));
            }
            // the level of the read's last posting
            if (g0 + u * kWave < s1) {
                const int32_t glast = min(s1 - 1, g0 + u * kWave + kWave - 1);
                lv = __shfl(l, glast - g0 - u * kWave);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t cl = (uint32_t)(c[u] - c0i);
            if (c[u] >= 0 && cl < (uint32_t)kSparseCols) atomicAdd(&cnt[cl >> 2], 1u << (8 * (cl & 3)));
        }
    }
}

struct alignas(16) SparseTopkLds {
    alignas(16) uint32_t cnt[kSparseCols / 4];  // per-pair counts of the chunk, one byte each
    uint64_t cand[kSparseTopkCap];              // candidate keys of the row
    int64_t lv_first[kSparseMaxLevels];         // per level: first posting,
    int32_t lv_drow[kSparseMaxLevels];          // directory row
    SparseChunkLds W;
    uint32_t queue[kSparseBlock / kWave][kSparseQueue];  // per wave: (column in chunk << 8) | count
    uint32_t hist[256];
    uint64_t prefix, pmask, thr;
    unsigned long long kmin;
    uint32_t need, exact, nlv, ncand;
    uint32_t imin, ti, tu;                      // the threshold as a ratio ti / tu (ti = 0: none)
};

// Top-k of reference rows [rb, rb + gridDim.x) (grouped rows of at most 255
// levels, magnitudes at least the count sums) from the inverted index of
// the query set, whose postings are grouped by chunk of kSparseCols
// columns.  Workgroup (x, s) takes row rb + x and the chunks [s cps,
// (s + 1) cps): it looks up the row's levels once, then per chunk counts
// the pairs in LDS -- the chunk directory gives each level's postings of
// the chunk, spread over all threads; the next chunk's entries load while
// this chunk is scanned -- and finds the pairs whose key beats the row's
// threshold, the k-th best key held so far.  The chunks come in column order, so a pair of a later chunk
// beats the threshold key only with a larger T, hence a larger ratio I / U
// than the threshold pair's: that exact integer test, after a test of the
// count against the smallest intersection it allows (imin), passes over
// most pairs without a division.  The scan tests 16 counts per lane at a
// time against imin (a byte mask per word); the counts that reach it are
// queued per wave, and each 64 queued pairs get their magnitudes, the
// ratio test and, when they pass, their T and key.  Candidates collect in LDS; when a chunk's do not
// fit, the k-th best of those stored becomes the threshold and the chunk
// is scanned again against it.  Only pairs with T > 0 are candidates: the
// T = 0 columns fill a row with fewer than k of them, in ascending order.
// With one column range (gridDim.y == 1) the workgroup writes the row's
// sorted top-k (and that fill) to out_idx / out_t; otherwise its candidates
// (at most k keys) to buf / bufcnt, segment x * gridDim.y + s, for the
// final selection.  Magnitudes come from the caller-order copy of MB bytes
// per column.
template <bool NARROW, int MB>
__global__ __launch_bounds__(kSparseBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_sparse_topk(
    RowArgs R, int64_t rb, PostingQuery P, const void* __restrict__ cmagn, int64_t nq, int k, int64_t cps,
    uint64_t* __restrict__ buf, int32_t* __restrict__ bufcnt, int32_t* __restrict__ out_idx,
    float* __restrict__ out_t) {
    __shared__ SparseTopkLds L;
    const int64_t row = rb + blockIdx.x;
    const int64_t nchunks = (nq + kSparseCols - 1) / kSparseCols;
    const int64_t y0 = (int64_t)blockIdx.y * cps;
    const int64_t y1 = min(nchunks, y0 + cps);
    const int ma = R.mag[row];
    const float maf = (float)ma;
    const int lane = lane_id(), w = (int)(threadIdx.x / kWave);
    auto mag_at = [&](int64_t c) -> uint32_t {
        if constexpr (MB == 1)
            return static_cast<const uint8_t*>(cmagn)[c];
        else
            return static_cast<const uint16_t*>(cmagn)[c];
    };
    if (threadIdx.x == 0) {
        L.nlv = 0u;
        L.ncand = 0u;
        L.thr = 0ull;
        L.imin = 1u;
        L.ti = 0u;
        L.tu = 1u;
    }
    __syncthreads();
    sparse_row_levels<NARROW>(R, row, P, L.lv_first, L.lv_drow, &L.nlv);
    __syncthreads();
    const int nl = (int)L.nlv;
    int64_t c0 = 0;
    uint32_t* wq = L.queue[w];
    // valid lanes' keys to cand, one atomic per wave (every lane calls)
    auto push = [&](uint64_t key, bool valid) {
        const uint64_t m = __ballot(valid);
        if (m == 0ull) return;
        uint32_t base = 0u;
        if (lane == 0) base = atomicAdd(&L.ncand, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, 0);
        const uint32_t at = base + (uint32_t)rank_below(m);
        if (valid && at < (uint32_t)kSparseTopkCap) L.cand[at] = key;
    };
    // The chunk's pairs whose count reaches imin go to the wave's queue (a
    // lane tests 16 counts at a time, a byte mask per word packed into one
    // mask, and queues those that reach imin one at a time); each 64 queued
    // pairs get their magnitude, and those whose ratio beats ti / tu (every
    // one when ti = 0) their T and key; valid lanes (T > 0, key > thr) are
    // pushed.  Every lane of the workgroup calls.
    auto chunk_keys = [&](uint64_t thr, uint32_t imin, uint32_t ti, uint32_t tu) {
        int qn = 0;  // queued pairs (wave-uniform)
        auto flush = [&](int n) {
            wave_lds_sync();
            uint64_t key = 0ull;
            bool valid = false;
            if (lane < n) {
                const uint32_t e = wq[lane];
                const int64_t c = c0 + (int64_t)(e >> 8);
                const uint32_t isct = e & 0xFFu, mb = mag_at(c);
                // I tu > ti (ma + mb - I)
                if (ti == 0u || isct * tu > ti * ((uint32_t)ma + mb - isct)) {
                    const float t = pair_tanimoto<true>((float)isct, maf, (float)mb, ma == 0 || mb == 0u);
                    key = topk_key(t, c);
                    valid = t > 0.0f && key > thr;
                }
            }
            push(key, valid);
            const int rem = qn - n;
            const uint32_t keep = lane < rem ? wq[n + lane] : 0u;
            wave_lds_sync();
            if (lane < rem) wq[lane] = keep;
            qn = rem;
        };
#pragma unroll 1
        for (int i = 0; i < kSparseCols / 16 / kSparseBlock; ++i) {
            const int q = (int)threadIdx.x + i * kSparseBlock;
            const uint4 v = reinterpret_cast<const uint4*>(L.cnt)[q];
            // one bit per count of at least imin among the 16 (bit 4h + b
            // for byte b of word h; a count of 0 marks a column past the
            // chunk's end), then those counts one at a time
            auto pack = [](uint32_t m) {  // bits 7, 15, 23, 31 -> 0..3
                return ((m >> 7) & 1u) | ((m >> 14) & 2u) | ((m >> 21) & 4u) | ((m >> 28) & 8u);
            };
            uint32_t mk = pack(bytes_ge(v.x, imin)) | (pack(bytes_ge(v.y, imin)) << 4) |
                          (pack(bytes_ge(v.z, imin)) << 8) | (pack(bytes_ge(v.w, imin)) << 12);
            while (__ballot(mk != 0u) != 0ull) {
                const bool has = mk != 0u;
                uint32_t e = 0u;
                if (has) {
                    const uint32_t j = (uint32_t)__builtin_ctz(mk);
                    mk &= mk - 1u;
                    const uint32_t h = j >> 2, b = j & 3u;
                    const uint32_t wd = h == 0 ? v.x : h == 1 ? v.y : h == 2 ? v.z : v.w;
                    e = ((uint32_t)(16 * q) + j) << 8 | ((wd >> (8 * b)) & 0xFFu);
                }
                const uint64_t m = __ballot(has);
                if (has) wq[qn + rank_below(m)] = e;
                qn += __popcll(m);
                if (qn >= kWave) flush(kWave);
            }
        }
        if (qn > 0) flush(qn);
    };
    // The k-th largest of cand[0, n) (n > k; keys distinct): MSB-first 8-bit
    // radix select in LDS.  Called by every thread; the result is uniform.
    auto select_kth = [&](uint32_t n) -> uint64_t {
        __syncthreads();  // every thread has read ncand
        if (threadIdx.x == 0) {
            L.prefix = 0ull;
            L.pmask = 0ull;
            L.need = (uint32_t)k;
            L.exact = 0u;
        }
        for (int shift = 56; shift >= 0; shift -= 8) {
            if (threadIdx.x < 256) L.hist[threadIdx.x] = 0u;
            __syncthreads();
            const uint64_t prefix = L.prefix, pmask = L.pmask;
            for (uint32_t i = threadIdx.x; i < ((n + kWave - 1) & ~(uint32_t)(kWave - 1)); i += blockDim.x) {
                const uint64_t key = i < n ? L.cand[i] : 0ull;
                const bool act = i < n && (key & pmask) == prefix;
                const uint32_t dig = (uint32_t)(key >> shift) & 255u;
                // equal digits of one instruction add once (equal-address
                // LDS atomics serialise)
                const uint64_t am = __ballot(act);
                if (am == 0ull) continue;
                const int leader = __builtin_ctzll(am);
                const uint32_t d0 = (uint32_t)__shfl((int)dig, leader);
                if (__ballot(act && dig == d0) == am) {
                    if (lane == leader) atomicAdd(&L.hist[d0], (uint32_t)__popcll(am));
                } else if (act) {
                    atomicAdd(&L.hist[dig], 1u);
                }
            }
            __syncthreads();
            if (threadIdx.x < kWave) {
                // wave 0 finds the digit: suffix sums over 4 bins per lane
                const uint32_t need = L.need;
                const uint32_t h0 = L.hist[4 * lane], h1 = L.hist[4 * lane + 1], h2 = L.hist[4 * lane + 2],
                               h3 = L.hist[4 * lane + 3];
                const uint32_t sl = h0 + h1 + h2 + h3;
                uint32_t S = sl;
                for (int off = 1; off < kWave; off <<= 1) {
                    const uint32_t t = __shfl_down(S, off);
                    if (lane + off < kWave) S += t;
                }
                const uint64_t ge = __ballot(S >= need);
                const int Lb = 63 - __builtin_clzll(ge);
                uint32_t above = __shfl(S - sl, Lb);
                const uint32_t hb[4] = {(uint32_t)__shfl(h0, Lb), (uint32_t)__shfl(h1, Lb), (uint32_t)__shfl(h2, Lb),
                                        (uint32_t)__shfl(h3, Lb)};
                int d = 4 * Lb;
                uint32_t cnt_d = hb[0];
                for (int q = 3; q >= 0; --q) {
                    if (above + hb[q] >= need) {
                        d = 4 * Lb + q;
                        cnt_d = hb[q];
                        break;
                    }
                    above += hb[q];
                }
                if (lane == 0) {
                    L.need = need - above;
                    L.prefix = prefix | ((uint64_t)d << shift);
                    L.pmask = pmask | ((uint64_t)255 << shift);
                    L.exact = cnt_d == need - above ? 1u : 0u;
                    L.kmin = ~0ull;
                }
            }
            __syncthreads();
            if (L.exact) {
                // every key of the chosen bucket is kept: K is its smallest
                const uint64_t bp = L.prefix, bm = L.pmask;
                unsigned long long mn = ~0ull;
                for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                    const uint64_t key = L.cand[i];
                    if ((key & bm) == bp && key < mn) mn = key;
                }
                atomicMin(&L.kmin, mn);
                __syncthreads();
                break;
            }
        }
        const uint64_t K = L.exact ? (uint64_t)L.kmin : L.prefix;
        __syncthreads();
        return K;
    };
    // cand keeps its keys >= K among cand[0, n) (n <= kSparseTopkCap)
    auto keep_from = [&](uint32_t n, uint64_t K) {
        constexpr int kPer = (kSparseTopkCap + kSparseBlock - 1) / kSparseBlock;
        uint64_t mine[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t i = threadIdx.x + j * kSparseBlock;
            mine[j] = i < n ? L.cand[i] : 0ull;
        }
        __syncthreads();
        if (threadIdx.x == 0) L.ncand = 0u;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPer; ++j) push(mine[j], mine[j] != 0ull && mine[j] >= K);
    };
    // the threshold K as a ratio (its pair recovered from its column) and
    // the smallest intersection it allows: I tu > ti (ma + mb - I) >= ti ma
    // (mb >= I) needs I > ti ma / tu
    auto ratio_of = [&](uint64_t K, uint32_t& ti, uint32_t& tu, uint32_t& imin) {
        const uint32_t col = 0xFFFFFFFFu - (uint32_t)K;
        const uint32_t mb0 = mag_at((int64_t)col);
        const float t0 = __uint_as_float((uint32_t)(K >> 32) & 0x7FFFFFFFu);
        const uint32_t i0 = sparse_recover_isct(t0, maf, (float)mb0);
        if (i0 == 0u) {
            ti = 0u;
            tu = 1u;
            imin = 1u;
            return;
        }
        ti = i0;
        tu = (uint32_t)ma + mb0 - i0;
        const uint64_t lo = (uint64_t)i0 * (uint64_t)ma / tu + 1;
        imin = lo > 128 ? 128u : (uint32_t)lo;
    };
    // cuts the candidates to the k keys >= K, their k-th largest, which
    // becomes the threshold
    auto reduce_to_k = [&](uint32_t n) {
        const uint64_t K = select_kth(n);
        keep_from(n, K);
        if (threadIdx.x == 0) {
            L.thr = K;
            ratio_of(K, L.ti, L.tu, L.imin);
        }
        __syncthreads();
    };
    if (y0 < y1) sparse_chunk_store(L.W, L.lv_first, nl, sparse_chunk_load(L.lv_drow, P, nl, y0));
    for (int64_t y = y0; y < y1; ++y) {
        c0 = y * kSparseCols;
        for (int i = threadIdx.x; i < kSparseCols / 16; i += blockDim.x)
            reinterpret_cast<uint4*>(L.cnt)[i] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
        // the row's threshold and candidates as the chunk starts (read by
        // every thread before any of the chunk's keys is added)
        const uint64_t thr = L.thr;
        const uint32_t imin = L.imin, n0 = L.ncand, ti = L.ti, tu = L.tu;
        sparse_walk_chunk<NARROW>(L.cnt, L.W, nl, P, c0);
        __syncthreads();
        // the next chunk's directory entries load while this one is scanned
        const SparseRange nxt = y + 1 < y1 ? sparse_chunk_load(L.lv_drow, P, nl, y + 1) : SparseRange{};
        chunk_keys(thr, imin, ti, tu);
        __syncthreads();
        if (y + 1 < y1) sparse_chunk_store(L.W, L.lv_first, nl, nxt);
        uint32_t nc = L.ncand;
        if (nc > (uint32_t)kSparseTopkCap) {
            // the chunk's keys did not fit: the k-th best of the stored ones
            // (a lower bound of the k-th best of all, above thr) becomes the
            // threshold, the held keys below it go and the chunk is scanned
            // again (with its own admission test, which every key above thr
            // passes) for its keys at or above it -- until they fit
            uint32_t held = n0;
            uint64_t K = 0ull;
            while (nc > (uint32_t)kSparseTopkCap) {
                K = select_kth((uint32_t)kSparseTopkCap);
                keep_from(held, K);
                __syncthreads();
                held = L.ncand;
                chunk_keys(K - 1, imin, ti, tu);
                __syncthreads();
                nc = L.ncand;
            }
            if (threadIdx.x == 0) {
                L.thr = K;
                ratio_of(K, L.ti, L.tu, L.imin);
            }
            __syncthreads();
        }
        if (nc > (uint32_t)(kSparseTopkCap / 2) && nc > (uint32_t)k) reduce_to_k(nc);
    }
    {
        const uint32_t nc = L.ncand;
        if (nc > (uint32_t)k) reduce_to_k(nc);
    }
    const uint32_t have = L.ncand;
    if (gridDim.y > 1) {
        const int64_t seg = (int64_t)blockIdx.x * gridDim.y + blockIdx.y;
        for (uint32_t i = threadIdx.x; i < have; i += blockDim.x) buf[seg * k + i] = L.cand[i];
        if (threadIdx.x == 0) bufcnt[seg] = (int32_t)have;
        return;
    }
    // sorted (T descending, then smaller column): bitonic over kSparseBlock keys
    static_assert(SIML_MAX_TOPK <= kSparseBlock && kSparseBlock <= kSparseTopkCap, "sort width");
    uint64_t* srt = L.cand;
    if (threadIdx.x >= have) srt[threadIdx.x] = 0ull;
    __syncthreads();
    for (int size = 2; size <= kSparseBlock; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int i = threadIdx.x, j = i ^ stride;
            if (j > i) {
                const bool desc = (i & size) == 0;
                const uint64_t a = srt[i], b = srt[j];
                if (desc ? (a < b) : (a > b)) {
                    srt[i] = b;
                    srt[j] = a;
                }
            }
            __syncthreads();
        }
    }
    int32_t* oi = out_idx + (int64_t)blockIdx.x * k;
    float* ot = out_t + (int64_t)blockIdx.x * k;
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        int32_t ix = -1;
        float tv = -__builtin_inff();
        if ((uint32_t)i < have) {
            const uint64_t key = srt[i];
            const uint32_t o = (uint32_t)(key >> 32);
            tv = __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
            ix = (int32_t)(0xFFFFFFFFu - (uint32_t)key);
        }
        oi[i] = ix;
        ot[i] = tv;
    }
    if (have >= (uint32_t)k) return;
    // fewer than k positive pairs: the columns with T == 0 in ascending
    // order follow (the smallest columns that are not candidates)
    uint32_t* poscol = L.cnt;
    poscol[threadIdx.x] = threadIdx.x < have ? 0xFFFFFFFFu - (uint32_t)srt[threadIdx.x] : 0xFFFFFFFFu;
    __syncthreads();
    for (int size = 2; size <= kSparseBlock; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int i = threadIdx.x, j = i ^ stride;
            if (j > i) {
                const bool asc = (i & size) == 0;
                const uint32_t a = poscol[i], b = poscol[j];
                if (asc ? (a > b) : (a < b)) {
                    poscol[i] = b;
                    poscol[j] = a;
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        uint32_t p = 0;
        int64_t c = 0;
        for (int i = (int)have; i < k && c < nq; ++c) {
            while (p < have && poscol[p] < (uint32_t)c) ++p;
            if (p < have && poscol[p] == (uint32_t)c) continue;
            oi[i] = (int32_t)c;
            ot[i] = 0.0f;
            ++i;
        }
    }
}

// Sparse 1000-bin histogram: pairs with I <= kIMax and U < kIU are counted
// by (I, U) (their T is a function of the two), the rest into their bins.
struct alignas(16) SparseHistLds {
    alignas(16) uint32_t cnt[kSparseCols / 4];
    int64_t lv_first[kSparseMaxLevels];
    int32_t lv_drow[kSparseMaxLevels];
    SparseChunkLds W;
    uint32_t bins[1000];
    float edges[1001];
    uint32_t iu[EpiHist1000<1>::kIMax * EpiHist1000<1>::kIU];
    uint32_t nlv;
};

// The 1000-bin int64 histogram of rows [rb, rb + gridDim.x) (grouped rows
// of at most 255 levels, magnitudes at least the count sums) against every
// query molecule, from the inverted index: workgroup (x, s) takes row
// rb + x and the chunks [s cps, (s + 1) cps), counts each chunk's pairs in
// LDS as the sparse top-k does (10M x 1024: 7.4 ms; 8.0 with the waves
// walking each level's postings from where the previous chunk stopped),
// then bins every pair of the chunk: a count
// of 0 is T = 0 (bin 0), any other pair gets T from the same float
// expression as every kernel.  Counts by (I, U) as in the bit-sliced
// histogram epilogue keep most LDS atomics off the few bins such pairs
// share; the workgroup adds its bins to `hist` at the end.
template <bool NARROW, int MB>
__global__ __launch_bounds__(kSparseBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_sparse_hist1000(
    RowArgs R, int64_t rb, PostingQuery P, const void* __restrict__ cmagn, int64_t nq, int64_t cps,
    unsigned long long* __restrict__ hist) {
    constexpr int kIU = EpiHist1000<1>::kIU, kIMax = EpiHist1000<1>::kIMax;
    __shared__ SparseHistLds L;
    const int64_t row = rb + blockIdx.x;
    const int64_t nchunks = (nq + kSparseCols - 1) / kSparseCols;
    const int64_t y0 = (int64_t)blockIdx.y * cps;
    const int64_t y1 = min(nchunks, y0 + cps);
    const int ma = R.mag[row];
    const float maf = (float)ma;
    for (int i = threadIdx.x; i < 1000; i += blockDim.x) L.bins[i] = 0u;
    for (int i = threadIdx.x; i < kIMax * kIU; i += blockDim.x) L.iu[i] = 0u;
    stage_edges(L.edges);
    if (threadIdx.x == 0) L.nlv = 0u;
    __syncthreads();
    sparse_row_levels<NARROW>(R, row, P, L.lv_first, L.lv_drow, &L.nlv);
    __syncthreads();
    const int nl = (int)L.nlv;
    uint32_t zero = 0u;  // this thread's pairs with T < fl32(1/1000)
    if (y0 < y1) sparse_chunk_store(L.W, L.lv_first, nl, sparse_chunk_load(L.lv_drow, P, nl, y0));
    for (int64_t y = y0; y < y1; ++y) {
        const int64_t c0 = y * kSparseCols;
        const int cw = (int)min((int64_t)kSparseCols, nq - c0);
        for (int i = threadIdx.x; i < kSparseCols / 16; i += blockDim.x)
            reinterpret_cast<uint4*>(L.cnt)[i] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
        sparse_walk_chunk<NARROW>(L.cnt, L.W, nl, P, c0);
        __syncthreads();
        // the next chunk's directory entries load while this one is binned
        const SparseRange nxt = y + 1 < y1 ? sparse_chunk_load(L.lv_drow, P, nl, y + 1) : SparseRange{};
#pragma unroll 1
        for (int i = 0; i < kSparseCols / 16 / kSparseBlock; ++i) {
            const int q = (int)threadIdx.x + i * kSparseBlock;
            const int nv = min(16, cw - 16 * q);  // columns of the block in the chunk
            if (nv <= 0) continue;
            const uint4 v = reinterpret_cast<const uint4*>(L.cnt)[q];
            const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
            // pairs with a count of 0: T = 0
            uint32_t pos = 0u;
#pragma unroll
            for (int h = 0; h < 4; ++h) pos += (uint32_t)__popc(bytes_ge(wv[h], 1u));
            zero += (uint32_t)nv - pos;
            if (pos == 0u) continue;
            uint32_t mw[4 * MB];
            {
                const uint4* mp = reinterpret_cast<const uint4*>(static_cast<const char*>(cmagn) + MB * (c0 + 16 * q));
#pragma unroll
                for (int u = 0; u < MB; ++u) {
                    const uint4 m = mp[u];
                    mw[4 * u] = m.x;
                    mw[4 * u + 1] = m.y;
                    mw[4 * u + 2] = m.z;
                    mw[4 * u + 3] = m.w;
                }
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint32_t isct = (wv[j >> 2] >> (8 * (j & 3))) & 0xFFu;
                if (isct == 0u) continue;
                const uint32_t mb = MB == 1 ? (mw[j >> 2] >> (8 * (j & 3))) & 0xFFu
                                            : (mw[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
                const uint32_t u = (uint32_t)ma + mb - isct;
                if (isct <= (uint32_t)kIMax && u < (uint32_t)kIU) {
                    atomicAdd(&L.iu[(isct - 1) * kIU + u], 1u);
                } else {
                    const float t = pair_tanimoto<true>((float)isct, maf, (float)mb, ma == 0 || mb == 0u);
                    if (!(t >= 1.0f / 1000.0f))
                        ++zero;
                    else
                        atomicAdd(&L.bins[bin1000(t, L.edges)], 1u);
                }
            }
        }
        __syncthreads();
        if (y + 1 < y1) sparse_chunk_store(L.W, L.lv_first, nl, nxt);
    }
    // (I, U) counts to their bins: T = (float)I / (float)U, both exact
    // integers (short_div equals IEEE division there)
    zero = wave_sum_u64(zero);
    if (lane_id() == 0 && zero) atomicAdd(&L.bins[0], zero);
    for (int k = threadIdx.x; k < kIMax * kIU; k += blockDim.x)
        if (L.iu[k]) atomicAdd(&L.bins[bin1000((float)(k / kIU + 1) / (float)(k % kIU), L.edges)], L.iu[k]);
    __syncthreads();
    for (int k = threadIdx.x; k < 1000; k += blockDim.x)
        if (L.bins[k]) atomicAdd(&hist[k], (unsigned long long)L.bins[k]);
}

// gpuHistogrammedLingo_multirow (kernels.py:606-667): fused Tanimoto and
// per-row 101-bin histogram; per-workgroup LDS bins, flushed with global
// atomics because several workgroups share a row.
__global__ __launch_bounds__(kBlock) void k_hist_rows(RowArgs R, ColArgs C, const int32_t* __restrict__ rowlist,
                                                      int64_t rowbase, int32_t* __restrict__ hist) {
    extern __shared__ int2 lds_row[];
    __shared__ int bins[101];
    const int64_t row = rowlist ? (int64_t)rowlist[blockIdx.x] : rowbase + blockIdx.x;
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
        if (bins[k]) atomicAdd(&hist[(row - rowbase) * 101 + k], bins[k]);
}

// gpuGetNeighbors_multirow (kernels.py:393-475): magnitude-bound prefilter
// (:440-441), exact T only when it passes, lo <= T < hi (:448).  Each wave
// owns a contiguous tile range; neighbours are compacted with a wave ballot
// and written in ascending column order.  Pass 1 (FILL = false) counts per
// (row, wave); pass 2 writes at the scanned offsets.
template <bool FILL>
__global__ __launch_bounds__(kBlock) void k_neighbors(RowArgs R, ColArgs C, const int32_t* __restrict__ rowlist,
                                                      int64_t rowbase, float lo, float hi,
                                                      int32_t* __restrict__ counts, const int64_t* __restrict__ offs,
                                                      int32_t* __restrict__ idx, unsigned long long* __restrict__ work) {
    extern __shared__ int2 lds_row[];
    const int64_t row = rowlist ? (int64_t)rowlist[blockIdx.x] : rowbase + blockIdx.x;
    const int La = R.len[row];
    const int ma = R.mag[row];
    stage_row(lds_row, R.elems + R.off[row], La);
    __syncthreads();
    int64_t t0, t1, gw, nw;
    wave_tiles(C.ntiles, t0, t1, gw, nw);
    const int lane = lane_id();
    const int64_t slot = (row - rowbase) * nw + gw;
    int64_t pos = FILL ? offs[slot] : 0;
    unsigned long long wpairs = 0, wlsum = 0;  // merges run (evaluated-work counter)
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t c = t * kWave + lane;
        bool nb = false;
        if (c < C.n) {
            const int mb = C.mag[c];
            const float bound = (float)min(ma, mb) / (float)max(ma, mb);
            if (bound >= lo) {
                wpairs += 1;
                wlsum += (unsigned)(La + C.len[c]);
                const int isct = merge_isct(lds_row, La, C.tiles + C.toff[t] + lane, C.len[c]);
                const float tv = tanimoto_mags(isct, ma, mb);
                nb = (tv >= lo) && (tv < hi);
            }
        }
        const uint64_t m = __ballot(nb);
        if (FILL && nb) idx[pos + rank_below(m)] = (int32_t)col_of(C, c);
        pos += __popcll(m);
    }
    if (!FILL && lane == 0) counts[slot] = (int32_t)pos;
    if (work) add_work(work, wpairs, wlsum);
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

// Copies the buffered neighbour indices of every (row, wave) slot to its
// CSR position (no slot holds more than bcap).
__global__ __launch_bounds__(kBlock) void k_nbr_compact(const int32_t* __restrict__ buf,
                                                        const int32_t* __restrict__ counts,
                                                        const int64_t* __restrict__ offs, int64_t nslots, int bcap,
                                                        int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots * bcap) return;
    const int64_t slot = i / bcap;
    const int j = (int)(i - slot * bcap);
    if (j < counts[slot]) out[offs[slot] + j] = buf[i];
}

// Achievable int32 VALU rate: 8 independent v_add_u32 (op 0) or v_min_i32
// (op 1) chains per lane, `iters` x 16 steps each; the result is written so
// nothing is dead.  Lane-operations = threads * iters * 16 * 8.
__global__ __launch_bounds__(kBlock) void k_valu_peak(uint32_t* __restrict__ out, uint32_t seed, int iters, int op) {
    uint32_t x0 = seed + threadIdx.x, x1 = x0 * 3u, x2 = x0 * 5u, x3 = x0 * 7u, x4 = x0 * 11u, x5 = x0 * 13u,
             x6 = x0 * 17u, x7 = x0 * 19u;
    const uint32_t c = seed | 1u;
    if (op == 0) {
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(x0) : "v"(c));
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(x1) : "v"(c));
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(x2) : "v"(c));
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(x3) : "v"(c));
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(x4) : "v"(c));
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(x5) : "v"(c));
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(x6) : "v"(c));
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(x7) : "v"(c));
            }
        }
    } else {
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                asm volatile("v_min_i32 %0, %0, %1" : "+v"(x0) : "v"(c));
                asm volatile("v_min_i32 %0, %0, %1" : "+v"(x1) : "v"(c));
                asm volatile("v_min_i32 %0, %0, %1" : "+v"(x2) : "v"(c));
                asm volatile("v_min_i32 %0, %0, %1" : "+v"(x3) : "v"(c));
                asm volatile("v_min_i32 %0, %0, %1" : "+v"(x4) : "v"(c));
                asm volatile("v_min_i32 %0, %0, %1" : "+v"(x5) : "v"(c));
                asm volatile("v_min_i32 %0, %0, %1" : "+v"(x6) : "v"(c));
                asm volatile("v_min_i32 %0, %0, %1" : "+v"(x7) : "v"(c));
            }
        }
    }
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7;
}

// Achievable LDS rate of 8-byte reads (ds_read_b64, the table read of the
// bit-sliced kernels) from a 32 KiB table per workgroup of 4 waves, as in
// those kernels.  OP 0: a wave's 64 reads cover 512 contiguous bytes (no
// bank conflicts); OP 1: every lane reads a pseudo-random slot, the access
// pattern of the table probes.  Each thread issues iters x 8 independent
// reads; bytes read = threads * iters * 64.
template <int OP>
__global__ __launch_bounds__(kBlock) void k_lds_peak(uint32_t* __restrict__ out, int iters, uint32_t step) {
    __shared__ uint2 tab[siml::kBitSlots];
    for (int s = threadIdx.x; s < siml::kBitSlots; s += blockDim.x) tab[s] = make_uint2((uint32_t)s, (uint32_t)s * 3u);
    __syncthreads();
    uint32_t acc[8], idx[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        acc[q] = 0u;
        idx[q] = OP == 0 ? threadIdx.x + 256u * q : (threadIdx.x + 1u) * 2654435761u + 40503u * q;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t slot = OP == 0 ? (idx[q] & (siml::kBitSlots - 1)) : (idx[q] >> (32 - siml::kBitSlotBits));
            const uint2 v = tab[slot];
            acc[q] += v.x ^ v.y;
            idx[q] = OP == 0 ? idx[q] + step : idx[q] * 1664525u + 1013904223u;
        }
    }
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] =
        acc[0] ^ acc[1] ^ acc[2] ^ acc[3] ^ acc[4] ^ acc[5] ^ acc[6] ^ acc[7];
}

// Exhaustive comparison of short_div(method) with IEEE division over
// n in [0, 65536) x d in [1, 131072): one thread per d, grid.y splits n.
__global__ __launch_bounds__(kBlock) void k_div_check(int method, unsigned long long* bad,
                                                      unsigned long long* first) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x + 1u;
    const uint32_t n0 = blockIdx.y * 4096u, n1 = n0 + 4096u;
    uint32_t cnt = 0;
    unsigned long long f = ~0ull;
    if (d < 131072u) {
        const float df = (float)d;
        for (uint32_t n = n0; n < n1; ++n) {
            const float nf = (float)n;
            const float a = short_div(nf, df, method);
            const float b = nf / df;
            if (__float_as_uint(a) != __float_as_uint(b)) {
                ++cnt;
                const unsigned long long key = ((unsigned long long)d << 32) | n;
                f = key < f ? key : f;
            }
        }
    }
    if (cnt) {
        atomicAdd(bad, (unsigned long long)cnt);
        atomicMin(first, f);
    }
}

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

// Pinned host staging buffer (hipHostMalloc), grown on demand.
struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    int reserve(size_t bytes) {
        if (bytes <= cap && p) return SIML_OK;
        release();
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(SIML_ENOMEM, "hipHostMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        }
        cap = bytes;
        return SIML_OK;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
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
    // every magnitude in [0, 65536) and at least its molecule's count sum
    bool mags_small = true;
};

int host_threads() { return siml::default_threads(32); }

// Host threads of copies between pinned staging and the caller's memory.
int copy_threads() { return std::min(host_threads(), 16); }

// parts[i] (pointer, bytes) copied to dst + at[i], split by bytes over up
// to nt threads.
void parallel_copy(char* dst, const std::vector<std::pair<const char*, size_t>>& parts, const std::vector<size_t>& at,
                   int nt) {
    size_t total = 0;
    for (const auto& p : parts) total += p.second;
    nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)nt, total >> 20));
    auto work = [&](size_t lo, size_t hi) {  // bytes [lo, hi) of the concatenation
        size_t base = 0;
        for (size_t i = 0; i < parts.size() && base < hi; ++i) {
            const size_t a = std::max(lo, base), b = std::min(hi, base + parts[i].second);
            if (a < b) std::memcpy(dst + at[i] + (a - base), parts[i].first + (a - base), b - a);
            base += parts[i].second;
        }
    };
    if (nt == 1) {
        work(0, total);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, total * t / nt, total * (t + 1) / nt);
    work(0, total / nt);
    for (auto& x : th) x.join();
}

template <class F>
void parallel_for(int64_t n, F&& f) {
    const int nt = host_threads();
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
    std::vector<char> big((size_t)std::max<int64_t>(s.n, 1), 0);
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
            const int32_t mag = s.mags ? s.mags[r] : (int32_t)sum;
            info.mag[(size_t)r] = mag;
            if (!ok) unsorted[(size_t)r] = 1;
            if (neg || mag < 0 || mag >= 65536 || (int64_t)mag < (int64_t)sum) big[(size_t)r] = 1;
        }
    });
    info.maxlen = 0;
    for (int64_t r = 0; r < s.n; ++r) {
        info.maxlen = std::max(info.maxlen, info.len[(size_t)r]);
        if (unsorted[(size_t)r]) info.sorted_unique = false;
        if (negative[(size_t)r]) info.counts_nonneg = false;
        if (big[(size_t)r]) info.mags_small = false;
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
    bool mags_small = false;
    // row groups may serve this set (the policy was not MERGE at set_ref);
    // per row, whether the sparse paths can take it
    bool groups_ok = false;
    std::vector<uint8_t> h_sparse_ok;
    // bit-sliced row groups, built on first use: [0] consecutive rows, [1]
    // rows ordered by magnitude (the neighbour search, whose tile pruning
    // needs narrow magnitude bands)
    struct BitGrouping {
        bool built = false;
        DevBuf hdr, rows, ent;
        std::vector<siml::BitGroupHdr> h_hdr;
        std::vector<int32_t> h_rows, merge_rows;
        void release() {
            hdr.release();
            rows.release();
            ent.release();
            reset();
        }
        // forget the grouping but keep the device buffers for the next one:
        // no hipFree (which waits for the whole device) between sets
        void reset() {
            h_hdr.clear();
            h_rows.clear();
            merge_rows.clear();
            built = false;
        }
    } bits[2];
    // host copy of the ragged set, for groupings built after set_ref
    std::vector<int32_t> h_ids, h_counts, h_mag, h_csum;
    std::vector<int64_t> h_start;
    RowArgs args() const {
        return {elems.as<int2>(), off.as<int64_t>(), len.as<int32_t>(), mag.as<int32_t>(), csum.as<int32_t>()};
    }
    void release() {
        elems.release();
        off.release();
        len.release();
        mag.release();
        csum.release();
        bits[0].release();
        bits[1].release();
        h_ids.clear();
        h_counts.clear();
        h_mag.clear();
        h_csum.clear();
        h_start.clear();
        h_sparse_ok.clear();
        groups_ok = false;
        n = -1;
    }
};

struct ColSet {
    ColSet() = default;
    ColSet(const ColSet&) = delete;
    ColSet& operator=(const ColSet&) = delete;
    ~ColSet() { release(); }
    int64_t n = -1, ntiles = 0;
    int64_t tiles_total = 0, btiles_total = 0;  // int2 elements in tiles / btiles
    int32_t maxlen = 0;
    int layout = 0;
    bool sorted_unique = true;
    bool counts_nonneg = true;
    bool mags_small = false;
    DevBuf tiles, toff, tlen, len, mag, csum, perm, tmag;
    bool identity = true;
    bool blocked = true;  // perm keeps every column inside its 256-column block
    // level stream of the bit-sliced kernels (build_colset)
    DevBuf btiles, btoff, btlen;
    bool levels_ok = false;
    // inverted index of the level stream (the sparse path of the float32
    // matrix), built on first use by any handle using the set
    struct Postings {
        bool built = false, ok = false;
        bool narrow = false;     // postings stored as uint16
        int64_t slots = 0, total = 0;
        double per_level = 0.0;  // expected postings of a stream level
        DevBuf slot, col, cmag;
        DevBuf cmag8, cmag16;    // caller-order magnitudes as 1 and 2 bytes
        int mag_bytes = 0;       // the narrowest of those that holds every magnitude (0: neither)
        DevBuf drow, dir;        // chunk directory (PostingQuery)
        int32_t ndir = 0;
        void release() {
            slot.release();
            col.release();
            cmag.release();
            cmag8.release();
            cmag16.release();
            drow.release();
            dir.release();
            mag_bytes = 0;
            built = ok = narrow = false;
        }
        PostingQuery query() const {
            return {slot.as<uint4>(), narrow ? nullptr : col.as<int32_t>(), narrow ? col.as<uint16_t>() : nullptr,
                    (uint32_t)(slots - 1), drow.as<int64_t>(), dir.as<int32_t>(), ndir};
        }
    } post;
    std::mutex post_mu;
    ColArgs args() const {
        return {tiles.as<int2>(), toff.as<int64_t>(), tlen.as<int32_t>(), len.as<int32_t>(), mag.as<int32_t>(),
                csum.as<int32_t>(), identity ? nullptr : perm.as<int32_t>(), btiles.as<int2>(), btoff.as<int64_t>(),
                btlen.as<int32_t>(), tmag.as<int2>(), n, ntiles, 1};
    }
    void release() {
        tiles.release();
        toff.release();
        tlen.release();
        len.release();
        mag.release();
        csum.release();
        perm.release();
        tmag.release();
        btiles.release();
        btoff.release();
        btlen.release();
        post.release();
        n = -1;
    }
};

}  // namespace

struct siml_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    RowSet ref;
    std::shared_ptr<ColSet> qry = std::make_shared<ColSet>();  // may be shared with a siml_db
    // pending result
    int kind = SIML_RESULT_NONE;
    int64_t rows = 0, cols = 0, nnz = 0;
    int64_t dpitch = 0;  // row pitch (elements) of a matrix result in res
    DevBuf res, res2, scratch;
    bool csr_on_device = false;     // a CSR result's row offsets are in csr_dev
    // timing
    bool timing = false;
    // (start, selection start or null, end) of each timed operation
    struct Timed {
        hipEvent_t a, m, b;
    };
    std::vector<Timed> pending_ev;
    std::vector<hipEvent_t> ev_pool;
    double kernel_ms = 0.0;
    double select_ms = 0.0;  // the part of kernel_ms spent after selection marks
    int64_t launches = 0, pairs = 0, pending_pairs = 0;
    // evaluated-work counter of the neighbour kernels (while timing is on):
    // device [pairs, sum of La + Lb], and the values at the last reset
    DevBuf work;
    unsigned long long work_base[2] = {0, 0};
    std::string last_kernel = "none";
    int policy = SIML_POLICY_AUTO;
    int layout = SIML_LAYOUT_BLOCK;
    // launch shape of the bit-sliced kernels: workgroups aimed at per launch
    // and, for top-k, the most waves per row (chosen by the autotuner when
    // on).  The default, measured on the MI355X over 1..32768 rows of 1M- and
    // 10M-molecule databases, is within 15% of the fastest shape at every
    // size; 256 waves per row without seeding (the earlier default) left
    // calls of up to a few hundred rows 3x slower on too few workgroups.
    int64_t target_wg = 8192;
    unsigned topk_wave_cap = 1024;
    // top-k: seed row thresholds from every topk_seed-th query tile first
    // (0: no seeding, 1: stride from the query set's size, seed_stride())
    int topk_seed = 1;
    // most row groups (tables) one bit-sliced workgroup serves: 1, 2 or 4
    int max_tables = 2;
    // top-k and 1000-bin histogram launches on a persistent grid of the
    // bit-sliced kernel (work queue per XCD) instead of an ordinary grid
    bool persistent = false;
    int cus = 256;  // compute units of the device
    bool autotune = false;
    // graph mode (siml_set_graphs): the kernel sequence of a top-k or
    // 1000-bin histogram call is recorded by stream capture and launched as
    // one HIP graph.  Instantiated graphs are kept by the sequence's
    // signature (node count and kernels in order) and updated in place with
    // the next call's arguments; `done` marks the end of an instance's last
    // launch, which must be complete before its arguments are updated.
    bool graphs = false;
    struct GraphExec {
        hipGraphExec_t exec = nullptr;
        hipEvent_t done = nullptr;
    };
    std::map<uint64_t, GraphExec> graph_exec;
    int64_t graph_launches = 0, graph_builds = 0;
    DevBuf rowlist, csr_dev, glist;
    std::vector<int32_t> h_glist;
    std::vector<int32_t> h_rowlist;  // host source of the last rowlist upload
    std::vector<char> h_stage;       // staging of strided result copies
    PinnedBuf pinned[2];             // double-buffered staging of large copies
    siml::BitGroupTables bit_tables; // host tables of the last row-group build
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

// Brackets the compute launches of one operation with events when timing;
// select() marks where a top-k operation's selection kernels begin.
struct TimedRegion {
    siml_ctx* ctx;
    hipEvent_t a = nullptr, m = nullptr, b = nullptr;
    int64_t pairs;
    TimedRegion(siml_ctx* c, int64_t p, bool on = true) : ctx(c), pairs(p) {
        if (on && ctx->timing) {
            a = take_event(ctx);
            b = take_event(ctx);
            if (a) (void)hipEventRecord(a, ctx->stream);
        }
    }
    void select() {
        if (ctx->timing && a && b && !m) {
            m = take_event(ctx);
            if (m) (void)hipEventRecord(m, ctx->stream);
        }
    }
    void done() {
        if (ctx->timing && a && b) {
            (void)hipEventRecord(b, ctx->stream);
            ctx->pending_ev.push_back({a, m, b});
            ctx->pending_pairs += pairs;
            a = m = b = nullptr;
        }
    }
    ~TimedRegion() {
        if (a) ctx->ev_pool.push_back(a);
        if (m) ctx->ev_pool.push_back(m);
        if (b) ctx->ev_pool.push_back(b);
    }
};

// Records the kernel launches issued between begin() and end() on the
// handle's stream as a HIP graph, when graph mode is on; launch() then
// launches it.  Only kernels may be issued in between: host copies and
// memsets stay outside, before begin().  A region left without end() (an
// error return) ends the capture and drops the graph.
struct GraphRegion {
    siml_ctx* ctx;
    bool on = false;                          // capturing
    siml_ctx::GraphExec* ready = nullptr;     // instance to launch after end()
    explicit GraphRegion(siml_ctx* c) : ctx(c) {}
    int begin() {
        if (!ctx->graphs) return SIML_OK;
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        on = true;
        return SIML_OK;
    }
    // Ends the capture and brings the cached instance of the graph's
    // signature up to date (or instantiates one); launch() then launches it.
    int end();
    int launch();
    ~GraphRegion() {
        if (!on) return;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
    }
};

// Signature of a captured graph: its node count, node types and kernel
// functions in node order.  Graphs of equal signature differ only in
// arguments and launch dimensions, which hipGraphExecUpdate changes.
int graph_signature(hipGraph_t g, uint64_t* sig) {
    size_t n = 0;
    HIP_TRY(hipGraphGetNodes(g, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    if (n) HIP_TRY(hipGraphGetNodes(g, nodes.data(), &n));
    uint64_t h = 0xcbf29ce484222325ull ^ (uint64_t)n;
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType t = hipGraphNodeTypeEmpty;
        HIP_TRY(hipGraphNodeGetType(nd, &t));
        uint64_t v = (uint64_t)t;
        if (t == hipGraphNodeTypeKernel) {
            hipKernelNodeParams p{};
            HIP_TRY(hipGraphKernelNodeGetParams(nd, &p));
            v ^= (uint64_t)(uintptr_t)p.func << 8;
        }
        h = (h ^ v) * 0x100000001b3ull;
    }
    *sig = h;
    return SIML_OK;
}

int GraphRegion::end() {
    if (!on) return SIML_OK;
    on = false;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture(ctx->stream, &g));
    if (!g) return fail(SIML_EDEVICE, "stream capture produced no graph");
    uint64_t sig = 0;
    int rc = graph_signature(g, &sig);
    siml_ctx::GraphExec* ge = rc == SIML_OK ? &ctx->graph_exec[sig] : nullptr;
    if (ge && ge->exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult r = hipGraphExecUpdateError;
        hipError_t e = ge->done ? hipEventSynchronize(ge->done) : hipSuccess;
        if (e != hipSuccess) {
            rc = fail(SIML_EDEVICE, "waiting for the previous graph launch failed: %s", hipGetErrorString(e));
        } else if (hipGraphExecUpdate(ge->exec, g, &bad, &r) != hipSuccess || r != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(ge->exec);
            ge->exec = nullptr;
        }
    }
    if (rc == SIML_OK && !ge->exec) {
        hipError_t e = hipGraphInstantiate(&ge->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            ge->exec = nullptr;
            rc = fail(SIML_EDEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        } else {
            ++ctx->graph_builds;
        }
    }
    (void)hipGraphDestroy(g);
    if (rc != SIML_OK) return rc;
    ready = ge;
    return SIML_OK;
}

int GraphRegion::launch() {
    if (!ready) return SIML_OK;
    siml_ctx::GraphExec* ge = ready;
    ready = nullptr;
    HIP_TRY(hipGraphLaunch(ge->exec, ctx->stream));
    if (!ge->done) HIP_TRY(hipEventCreateWithFlags(&ge->done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(ge->done, ctx->stream));
    ++ctx->graph_launches;
    return SIML_OK;
}

int collect_timing(siml_ctx* ctx) {
    if (ctx->pending_ev.empty()) return SIML_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (auto& t : ctx->pending_ev) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
        ctx->kernel_ms += ms;
        if (t.m) {
            HIP_TRY(hipEventElapsedTime(&ms, t.m, t.b));
            ctx->select_ms += ms;
            ctx->ev_pool.push_back(t.m);
        }
        ctx->launches += 1;
        ctx->ev_pool.push_back(t.a);
        ctx->ev_pool.push_back(t.b);
    }
    ctx->pending_ev.clear();
    ctx->pairs += ctx->pending_pairs;
    ctx->pending_pairs = 0;
    return SIML_OK;
}

// Builds bit-sliced grouping `which` of the reference set (0: consecutive
// rows, 1: rows ordered by magnitude) unless it exists.
// Host spans copied back to back into device memory at dev, on the handle's
// stream; returns when the copy is complete.  Large copies go through the
// two pinned slabs: host threads fill one slab while the other is in
// flight (a copy from pageable memory runs at a fraction of the DMA rate).
struct HostSpan {
    const void* p;
    size_t bytes;
};

int h2d_spans(siml_ctx* ctx, void* dev, const std::vector<HostSpan>& spans) {
    size_t total = 0;
    for (const HostSpan& s : spans) total += s.bytes;
    char* d = static_cast<char*>(dev);
    if (total < (size_t(32) << 20)) {
        for (const HostSpan& s : spans) {
            if (s.bytes) HIP_TRY(hipMemcpyAsync(d, s.p, s.bytes, hipMemcpyHostToDevice, ctx->stream));
            d += s.bytes;
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return SIML_OK;
    }
    constexpr size_t kSlab = size_t(64) << 20;
    TRY(ctx->pinned[0].reserve(kSlab));
    TRY(ctx->pinned[1].reserve(kSlab));
    hipEvent_t ev[2] = {nullptr, nullptr};
    HIP_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    if (hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return fail(SIML_EDEVICE, "hipEventCreate failed");
    }
    const int nt = copy_threads();
    size_t si = 0, soff = 0, done = 0;
    hipError_t e = hipSuccess;
    for (int64_t k = 0; done < total && e == hipSuccess; ++k) {
        const int b = (int)(k % 2);
        if (k >= 2) e = hipEventSynchronize(ev[b]);  // slab b's previous copy has landed
        if (e != hipSuccess) break;
        // gather up to kSlab bytes of the spans into slab b
        char* dst = static_cast<char*>(ctx->pinned[b].p);
        size_t fill = 0;
        std::vector<std::pair<const char*, size_t>> parts;
        std::vector<size_t> at;
        while (fill < kSlab && si < spans.size()) {
            const size_t take = std::min(kSlab - fill, spans[si].bytes - soff);
            if (take) {
                parts.emplace_back(static_cast<const char*>(spans[si].p) + soff, take);
                at.push_back(fill);
            }
            fill += take;
            soff += take;
            if (soff == spans[si].bytes) {
                ++si;
                soff = 0;
            }
        }
        parallel_copy(dst, parts, at, nt);
        e = hipMemcpyAsync(d + done, dst, fill, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[b], ctx->stream);
        done += fill;
    }
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = es;
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (e != hipSuccess) return fail(SIML_EDEVICE, "host-to-device copy failed: %s", hipGetErrorString(e));
    return SIML_OK;
}

int ensure_bits(siml_ctx* ctx, int which) {
    RowSet& R = ctx->ref;
    RowSet::BitGrouping& B = R.bits[which];
    if (B.built) return SIML_OK;
    siml::BitGroupInput bin{R.h_ids.data(), R.h_counts.data(), R.h_start.data(), R.h_len.data(), R.h_mag.data(),
                            R.n};
    siml::BitGroupTables& bt = ctx->bit_tables;  // reused: no fresh pages per build
    siml::build_bit_groups(bin, which == 1, bt, 0);
    TRY(upload(B.hdr, bt.hdr.data(), bt.hdr.size(), ctx->stream));
    TRY(upload(B.rows, bt.rows.data(), bt.rows.size(), ctx->stream));
    TRY(B.ent.reserve(sizeof(uint4) * (size_t)std::max<int64_t>(bt.nentries(), 1)));
    std::vector<HostSpan> spans;
    for (const siml::BitEntryVec& c : bt.chunks) spans.push_back({c.data(), sizeof(uint4) * c.size()});
    TRY(h2d_spans(ctx, B.ent.p, spans));
    B.h_hdr = bt.hdr;
    B.h_rows = bt.rows;
    B.merge_rows = bt.merge_rows;
    B.built = true;
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
    TRY(R.elems.reserve(sizeof(int2) * (size_t)std::max<int64_t>(total, 1)));
    TRY(h2d_spans(ctx, R.elems.p, {HostSpan{el.data(), sizeof(int2) * (size_t)total}}));
    TRY(upload(R.off, off.data(), (size_t)s.n, ctx->stream));
    TRY(upload(R.len, info.len.data(), (size_t)s.n, ctx->stream));
    TRY(upload(R.mag, info.mag.data(), (size_t)s.n, ctx->stream));
    TRY(upload(R.csum, info.csum.data(), (size_t)s.n, ctx->stream));
    R.bits[0].reset();
    R.bits[1].reset();
    R.h_ids.resize((size_t)total);
    R.h_counts.resize((size_t)total);
    parallel_for(total, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i) {
            R.h_ids[(size_t)i] = el[(size_t)i].x;
            R.h_counts[(size_t)i] = el[(size_t)i].y;
        }
    });
    R.h_start.assign(off.begin(), off.end());
    R.h_mag = info.mag;
    R.h_csum = info.csum;
    R.mags_small = info.mags_small;
    R.n = s.n;
    R.maxlen = info.maxlen;
    R.sorted_unique = info.sorted_unique;
    R.h_len = std::move(info.len);
    // the sparse paths take a row with strictly ascending ASCII ids, counts
    // of at most 15 (the index's levels) and at most 255 levels (byte
    // counters); row groups are built on first use
    R.groups_ok = ctx->policy != SIML_POLICY_MERGE;
    R.h_sparse_ok.assign((size_t)s.n, 0);
    parallel_for(s.n, [&](int64_t a, int64_t b) {
        for (int64_t r = a; r < b; ++r) {
            const int32_t* id = R.h_ids.data() + R.h_start[(size_t)r];
            const int32_t* ct = R.h_counts.data() + R.h_start[(size_t)r];
            const int32_t L = R.h_len[(size_t)r];
            int64_t lv = 0;
            bool ok = true;
            for (int32_t j = 0; j < L && ok; ++j) {
                ok = siml::ascii_id(id[j]) && ct[j] >= 0 && ct[j] <= siml::kBitMaxLevel && (j == 0 || id[j] > id[j - 1]);
                lv += ct[j];
            }
            R.h_sparse_ok[(size_t)r] = ok && lv <= kSparseMaxLevels ? 1 : 0;
        }
    });
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SIML_OK;
}

// Storage order of the query molecules.  NATURAL keeps the caller's order;
// BLOCK sorts by level count inside each block of kSortBlock consecutive
// columns, so a tile's step count (its longest molecule) is close to its mean
// while a wave's float32 outputs stay inside one 1 KiB span of the row;
// SORTED sorts the whole set by level count (outputs that carry their column
// index -- neighbours, top-k, histograms -- do not care about order).
constexpr int64_t kSortBlock = 256;

std::vector<int32_t> column_order(const std::vector<int32_t>& weight, int64_t n, int layout) {
    std::vector<int32_t> pos((size_t)n);
    for (int64_t c = 0; c < n; ++c) pos[(size_t)c] = (int32_t)c;
    auto by_weight = [&](int32_t x, int32_t y) { return weight[(size_t)x] < weight[(size_t)y]; };
    if (layout == SIML_LAYOUT_BLOCK)
        for (int64_t b = 0; b < n; b += kSortBlock)
            std::stable_sort(pos.begin() + b, pos.begin() + std::min(n, b + kSortBlock), by_weight);
    if (layout == SIML_LAYOUT_SORTED) std::stable_sort(pos.begin(), pos.end(), by_weight);
    return pos;
}

// Levels of one query molecule that a grouped row can hold (bitgroups.h):
// ASCII ids, t <= 15.
inline int32_t stream_levels(const int32_t* id, const int32_t* ct, int32_t L) {
    int32_t lv = 0;
    for (int32_t j = 0; j < L; ++j)
        if (siml::ascii_id(id[j]) && ct[j] > 0) lv += std::min(ct[j], siml::kBitMaxLevel);
    return lv;
}

// Builds a query set in device memory of the current device.
int build_colset(hipStream_t stream, int layout, const HostSet& s, ColSet& C) {
    SetInfo info;
    TRY(validate(s, info, "set_qsmiles"));
    if (s.n > 0x7FFFFFFF) return fail(SIML_ELIMIT, "set_qsmiles: more than 2^31-1 molecules");
    const bool levels_ok = info.sorted_unique && info.counts_nonneg;
    std::vector<int32_t> clv((size_t)s.n, 0);
    parallel_for(s.n, [&](int64_t a, int64_t b) {
        for (int64_t c = a; c < b; ++c)
            clv[(size_t)c] = levels_ok ? stream_levels(s.ids + s.start(c), s.counts + s.start(c), info.len[(size_t)c])
                                       : info.len[(size_t)c];
    });
    const std::vector<int32_t> order = column_order(clv, s.n, layout);
    bool identity = true;
    for (int64_t p = 0; p < s.n && identity; ++p) identity = order[(size_t)p] == (int32_t)p;
    std::vector<int32_t> plen((size_t)s.n), pmag((size_t)s.n), pcsum((size_t)s.n), plv((size_t)s.n);
    for (int64_t p = 0; p < s.n; ++p) {
        const size_t c = (size_t)order[(size_t)p];
        plen[(size_t)p] = info.len[c];
        pmag[(size_t)p] = info.mag[c];
        pcsum[(size_t)p] = info.csum[c];
        plv[(size_t)p] = clv[c];
    }
    const int64_t ntiles = (s.n + kWave - 1) / kWave;
    std::vector<int32_t> tlen((size_t)std::max<int64_t>(ntiles, 1), 0);
    std::vector<int64_t> toff((size_t)ntiles + 1, 0);
    for (int64_t t = 0; t < ntiles; ++t) {
        int32_t m = 0;
        for (int64_t p = t * kWave; p < std::min(s.n, (t + 1) * kWave); ++p) m = std::max(m, plen[(size_t)p]);
        tlen[(size_t)t] = m;
        toff[(size_t)t + 1] = toff[(size_t)t] + (int64_t)m * kWave;
    }
    const int64_t total = toff[(size_t)ntiles];
    std::vector<int2> tiles((size_t)std::max<int64_t>(total, 1), make_int2(0, 0));
    parallel_for(ntiles, [&](int64_t a, int64_t b) {
        for (int64_t t = a; t < b; ++t)
            for (int64_t p = t * kWave; p < std::min(s.n, (t + 1) * kWave); ++p) {
                const int64_t c = order[(size_t)p];
                const int32_t* id = s.ids + s.start(c);
                const int32_t* ct = s.counts + s.start(c);
                int2* d = tiles.data() + toff[(size_t)t] + (p - t * kWave);
                for (int32_t j = 0; j < plen[(size_t)p]; ++j) d[(size_t)j * kWave] = make_int2(id[j], ct[j]);
            }
    });
    TRY(upload(C.tiles, tiles.data(), (size_t)total, stream));
    C.tiles_total = total;
    TRY(upload(C.toff, toff.data(), (size_t)ntiles + 1, stream));
    TRY(upload(C.tlen, tlen.data(), (size_t)ntiles, stream));
    TRY(upload(C.len, plen.data(), (size_t)s.n, stream));
    TRY(upload(C.mag, pmag.data(), (size_t)s.n, stream));
    TRY(upload(C.csum, pcsum.data(), (size_t)s.n, stream));
    TRY(upload(C.perm, order.data(), (size_t)s.n, stream));
    std::vector<int2> tmag((size_t)std::max<int64_t>(ntiles, 1), make_int2(0, 0));
    for (int64_t t = 0; t < ntiles; ++t) {
        int32_t lo = INT32_MAX, hi = INT32_MIN;
        for (int64_t p = t * kWave; p < std::min(s.n, (t + 1) * kWave); ++p) {
            lo = std::min(lo, pmag[(size_t)p]);
            hi = std::max(hi, pmag[(size_t)p]);
        }
        tmag[(size_t)t] = make_int2(lo, hi);
    }
    TRY(upload(C.tmag, tmag.data(), (size_t)ntiles, stream));
    // level stream (bitgroups.h): (key, slot word) of every level a grouped
    // row can hold; level j of lane l of tile t is int2 element
    // btoff[t] + (j / 2) * 128 + 2 * l + (j & 1), so one 16-byte load per
    // lane reads two levels; step counts are padded to a multiple of 8 with
    // (0, 0), which matches only empty slots.
    C.levels_ok = levels_ok;
    if (C.levels_ok) {
        std::vector<int32_t> btlen((size_t)std::max<int64_t>(ntiles, 1), 0);
        std::vector<int64_t> btoff((size_t)ntiles + 1, 0);
        for (int64_t t = 0; t < ntiles; ++t) {
            int32_t m = 0;
            for (int64_t p = t * kWave; p < std::min(s.n, (t + 1) * kWave); ++p) m = std::max(m, plv[(size_t)p]);
            m = (m + 7) & ~7;
            btlen[(size_t)t] = m;
            btoff[(size_t)t + 1] = btoff[(size_t)t] + (int64_t)m * kWave;
        }
        const int64_t btotal = btoff[(size_t)ntiles];
        // one spare block of 8 levels x 64 lanes after the last tile: the
        // kernels' prefetch of the next block may read it
        const int64_t bspare = 8 * kWave;
        std::vector<int2> bt((size_t)(btotal + bspare), make_int2(0, 0));
        parallel_for(ntiles, [&](int64_t a, int64_t b) {
            for (int64_t t = a; t < b; ++t)
                for (int64_t p = t * kWave; p < std::min(s.n, (t + 1) * kWave); ++p) {
                    const int64_t c = order[(size_t)p];
                    const int32_t* id = s.ids + s.start(c);
                    const int32_t* ct = s.counts + s.start(c);
                    int2* d = bt.data() + btoff[(size_t)t] + 2 * (p - t * kWave);
                    size_t j = 0;
                    for (int32_t k = 0; k < plen[(size_t)p]; ++k) {
                        if (!siml::ascii_id(id[k])) continue;
                        for (int32_t lvl = 1; lvl <= std::min(ct[k], siml::kBitMaxLevel); ++lvl, ++j) {
                            const uint32_t key = siml::level_key(id[k], (uint32_t)lvl);
                            d[(j >> 1) * 2 * kWave + (j & 1)] =
                                make_int2((int32_t)key, (int32_t)siml::slot_word(key));
                        }
                    }
                }
        });
        TRY(upload(C.btiles, bt.data(), (size_t)(btotal + bspare), stream));
        C.btiles_total = btotal + bspare;
        TRY(upload(C.btoff, btoff.data(), (size_t)ntiles + 1, stream));
        TRY(upload(C.btlen, btlen.data(), (size_t)ntiles, stream));
    }
    HIP_TRY(hipStreamSynchronize(stream));
    C.n = s.n;
    C.ntiles = ntiles;
    C.maxlen = info.maxlen;
    C.sorted_unique = info.sorted_unique;
    C.counts_nonneg = info.counts_nonneg;
    C.mags_small = info.mags_small;
    C.identity = identity;
    C.blocked = layout != SIML_LAYOUT_SORTED;
    C.layout = layout;
    return SIML_OK;
}

int upload_cols(siml_ctx* ctx, const HostSet& s) {
    TRY(set_device(ctx));
    auto C = std::make_shared<ColSet>();
    TRY(build_colset(ctx->stream, ctx->layout, s, *C));
    ctx->qry = std::move(C);
    return SIML_OK;
}

int check_rows(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, const char* what) {
    if (ctx->ref.n < 0 || ctx->qry->n < 0)
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

int32_t* d_counts_of(siml_ctx* ctx) { return ctx->scratch.as<int32_t>(); }

int upload_rowlist(siml_ctx* ctx, const std::vector<int32_t>& rows) {
    ctx->h_rowlist = rows;
    return upload(ctx->rowlist, ctx->h_rowlist.data(), ctx->h_rowlist.size(), ctx->stream);
}

// Most row groups one bit-sliced workgroup serves (bits_tables).
constexpr int kMaxTables = 4;

// Bit-sliced plan of rows [rb, rl): the groups holding any of those rows,
// the planes they need, and the rows left to the merge kernel.
struct BitPlan {
    bool use = false;
    int nb = 8;
    int which = 0;  // grouping used
    std::vector<int32_t> groups, merge;
    const int32_t* d_glist = nullptr;  // the group list in device memory
};

// `by_mag` asks for the magnitude-ordered grouping (built here on first
// use).  Nothing is planned (use = false) when the policy is not AUTO, the
// query set has no level stream, or the reference set was given under
// another policy.  The group list goes to the handle's buffer, or to
// glist_dst (room for every group of the grouping plus kMaxTables) when
// given.
// The bit-sliced plan of rows [rb, rl): the groups holding them (the
// grouping is built on first use, except with no_build, when a grouping not
// built yet leaves the call to the merge kernel) and the rows no group holds.
int plan_bits(siml_ctx* ctx, int64_t rb, int64_t rl, BitPlan& p, bool by_mag = false, int32_t* glist_dst = nullptr,
              bool no_build = false) {
    RowSet& R = ctx->ref;
    p = BitPlan();
    if (ctx->policy == SIML_POLICY_MERGE || !ctx->qry->levels_ok || !R.groups_ok) return SIML_OK;
    p.which = by_mag ? 1 : 0;
    if (no_build && !R.bits[p.which].built) return SIML_OK;
    TRY(ensure_bits(ctx, p.which));
    const RowSet::BitGrouping& B = R.bits[p.which];
    p.use = true;
    for (size_t g = 0; g < B.h_hdr.size(); ++g) {
        bool hit = false;
        for (int r = 0; r < B.h_hdr[g].nrows && !hit; ++r) {
            const int32_t row = B.h_rows[g * siml::kBitRowStride + (size_t)r];
            hit = row >= rb && row < rl;
        }
        if (!hit) continue;
        p.groups.push_back((int32_t)g);
        p.nb = std::max(p.nb, B.h_hdr[g].nbits);
    }
    auto a = std::lower_bound(B.merge_rows.begin(), B.merge_rows.end(), (int32_t)rb);
    auto b = std::lower_bound(B.merge_rows.begin(), B.merge_rows.end(), rl > INT32_MAX ? INT32_MAX : (int32_t)rl);
    p.merge.assign(a, b);
    // the launches' group list, uploaded before any timed region begins
    // padded with empty groups (-1) to a multiple of kMaxTables, the most
    // groups one workgroup serves
    if (!p.groups.empty()) {
        ctx->h_glist = p.groups;
        while (ctx->h_glist.size() % kMaxTables) ctx->h_glist.push_back(-1);
        if (glist_dst) {
            HIP_TRY(hipMemcpyAsync(glist_dst, ctx->h_glist.data(), sizeof(int32_t) * ctx->h_glist.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
            p.d_glist = glist_dst;
        } else {
            TRY(upload(ctx->glist, ctx->h_glist.data(), ctx->h_glist.size(), ctx->stream));
            p.d_glist = ctx->glist.as<int32_t>();
        }
    }
    return SIML_OK;
}

// Tables (row groups) per workgroup of a bit-sliced launch: each column
// level a workgroup loads is probed in all of its tables, so the level
// stream is read from memory once per that many groups.  Fewer tables when
// the launch has few groups (an empty table costs its probes) or needs 16
// bit planes (the planes of several tables exceed the registers).
// Measured on the MI355X: two tables per workgroup made the 1M-molecule
// batched top-k 23% and the 10M-molecule histogram 13% faster than one,
// four tables neither; the neighbour search prunes tiles by the
// workgroup's magnitude band, which several groups widen, so it keeps one.
int bits_tables(const siml_ctx* ctx, const BitPlan& p, bool pruned = false) {
    const size_t n = p.groups.size();
    if (pruned || p.nb != 8 || n < 2 || ctx->max_tables < 2) return 1;
    return (n < 8 || ctx->max_tables < 4) ? 2 : 4;
}

// Calls f with std::integral_constant<int, G> for G = gt (1, 2 or 4).
template <class F>
int with_tables(int gt, F&& f) {
    if (gt == 4) return f(std::integral_constant<int, 4>{});
    if (gt == 2) return f(std::integral_constant<int, 2>{});
    return f(std::integral_constant<int, 1>{});
}

int launch_check(const char* name);

// Top-k seed stride of a handle's setting (1: from the query set's size).
// The sample should hold about 512 tiles per row: a sparser one (stride
// 64 on 1M molecules) gives a low seed and more candidates in the full
// pass, a denser one (stride 16 on 10M) more sampled candidates to select
// from.  Measured on the MI355X: 64 on 10M molecules ran 5-15% faster than
// 16, 32 on 1M up to 9%.
int64_t seed_stride(int setting, int64_t ntiles) {
    if (setting != 1) return setting;
    return std::min<int64_t>(64, std::max<int64_t>(16, ntiles / 512));
}

// Top-k candidate buffers of one call are held to this many bytes by
// running fewer column ranges (at least one: k + 128 keys per row and wave).
constexpr int64_t kTopkBufferBudget = int64_t(16) << 30;

// Neighbour indices kept per (row, wave) slot by the count pass.
constexpr int kNbrCap = 16;

// Launches k_bits<NB, Epi> over the plan's groups, Epi::kG groups per
// workgroup, with grid.y = gy column ranges.
// With `queue` (kXcds int32 of device memory), the k_bits launch runs on a
// persistent grid: as many workgroups as are resident at once (at most
// kXcds per group slot), taking the grid.x group slots from per-XCD queues,
// gy (a multiple of kXcds) column ranges each; the counters are cleared by
// a kernel on the stream first.
template <class Epi>
int launch_bits(siml_ctx* ctx, const BitPlan& p, unsigned gy, int64_t rb, int64_t rl, const Epi& epi,
                const char* name, const ColArgs* cols = nullptr, int32_t* queue = nullptr) {
    if (p.groups.empty()) return SIML_OK;
    constexpr int GT = Epi::kG;
    if (GT > 1 && p.nb != 8) return fail(SIML_EINVAL, "%s: %d tables per workgroup need 8 bit planes", name, GT);
    const dim3 grid((unsigned)((p.groups.size() + GT - 1) / GT), gy);
    const RowSet::BitGrouping& B = ctx->ref.bits[p.which];
    const BitGroupArgs G{B.hdr.as<siml::BitGroupHdr>(), B.rows.as<int32_t>(), B.ent.as<uint4>(), p.d_glist};
    const bool fast = ctx->ref.mags_small && ctx->qry->mags_small;
    const int32_t* rmag = ctx->ref.mag.as<int32_t>();
    const ColArgs C = cols ? *cols : ctx->qry->args();
    if constexpr (Epi::kBlockMode) {
        const dim3 b8(block_waves<8>() * kWave), b16(block_waves<16>() * kWave);
        if (p.nb == 8 && fast)
            k_bits_block<8, true><<<grid, b8, 0, ctx->stream>>>(G, rmag, C, rb, rl, epi);
        else if (p.nb == 8)
            k_bits_block<8, false><<<grid, b8, 0, ctx->stream>>>(G, rmag, C, rb, rl, epi);
        else if (fast)
            k_bits_block<16, true><<<grid, b16, 0, ctx->stream>>>(G, rmag, C, rb, rl, epi);
        else
            k_bits_block<16, false><<<grid, b16, 0, ctx->stream>>>(G, rmag, C, rb, rl, epi);
    } else {
        const dim3 block(kBlock * GT);
        if (queue && gy % kXcds != 0)
            return fail(SIML_EINVAL, "%s: a persistent grid needs a multiple of %d column ranges", name, kXcds);
        auto go = [&](auto kern) -> int {
            dim3 g = grid;
            BitGroupArgs GA = G;
            if (queue) {
                int per_cu = 0;
                HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)block.x, 0));
                const int64_t resident =
                    std::max<int64_t>((int64_t)std::max(per_cu, 1) * ctx->cus / kXcds * kXcds, kXcds);
                g = dim3((unsigned)std::min<int64_t>(resident, (int64_t)kXcds * grid.x), 1, 1);
                GA.queue = queue;
                GA.nslot = (int32_t)grid.x;
                GA.ny = (int32_t)gy;
                k_fill<int32_t><<<dim3(1), dim3(kWave), 0, ctx->stream>>>(queue, kXcds, 0);
                TRY(launch_check("k_fill<queue>"));
            }
            kern<<<g, block, 0, ctx->stream>>>(GA, rmag, C, rb, rl, epi);
            return SIML_OK;
        };
        if (p.nb == 8 && fast)
            TRY(go(k_bits<8, true, Epi>));
        else if (p.nb == 8)
            TRY(go(k_bits<8, false, Epi>));
        else if constexpr (GT == 1) {
            if (fast)
                TRY(go(k_bits<16, true, Epi>));
            else
                TRY(go(k_bits<16, false, Epi>));
        }
    }
    return launch_check(name);
}

// Column ranges of a bit-sliced launch rounded to a multiple of the XCD
// count, so that bits_wg keeps each XCD on an eighth of the column stream,
// when that stays within cap (small counts stay as they are).
unsigned xcd_gy(unsigned gy, unsigned cap) {
    if (gy < 4) return gy;
    const unsigned up = (gy + kXcds - 1) / kXcds * kXcds;
    if (up <= cap) return up;
    const unsigned down = gy / kXcds * kXcds;
    return down >= (unsigned)kXcds ? down : gy;
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
    ctx->rows = ctx->cols = ctx->nnz = ctx->dpitch = 0;
    ctx->csr_on_device = false;
}

int zero_fill_i32(siml_ctx* ctx, int32_t* p, int64_t n) {
    if (n <= 0) return SIML_OK;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    k_fill<int32_t><<<dim3((unsigned)blocks), dim3(256), 0, ctx->stream>>>(p, n, 0);
    return launch_check("k_fill");
}

// On-disk form of a built database: a header, then the device arrays of the
// column set in a fixed order, each at a 64-byte aligned offset.  Loading
// uploads the arrays as they are, skipping the host build (ordering,
// tiling, level stream).
constexpr char kDbMagic[8] = {'S', 'I', 'M', 'L', 'D', 'B', 0, 1};
constexpr int32_t kDbVersion = 3;  // level keys and slot words of bitgroups.h, tile magnitude ranges

struct DbFileHeader {
    char magic[8];
    int32_t version, layout;
    int64_t n, ntiles, tiles_total, btiles_total;
    int32_t maxlen;
    uint8_t identity, blocked, levels_ok, sorted_unique, counts_nonneg, mags_small, pad[2];
};

struct DbArray {
    DevBuf* buf;
    size_t bytes;
};

std::vector<DbArray> db_arrays(ColSet& C) {
    const size_t n = (size_t)std::max<int64_t>(C.n, 0), nt = (size_t)C.ntiles;
    std::vector<DbArray> a = {{&C.tiles, sizeof(int2) * (size_t)C.tiles_total},
                              {&C.toff, sizeof(int64_t) * (nt + 1)},
                              {&C.tlen, sizeof(int32_t) * nt},
                              {&C.len, sizeof(int32_t) * n},
                              {&C.mag, sizeof(int32_t) * n},
                              {&C.csum, sizeof(int32_t) * n},
                              {&C.perm, sizeof(int32_t) * n},
                              {&C.tmag, sizeof(int2) * nt}};
    if (C.levels_ok) {
        a.push_back({&C.btiles, sizeof(int2) * (size_t)C.btiles_total});
        a.push_back({&C.btoff, sizeof(int64_t) * (nt + 1)});
        a.push_back({&C.btlen, sizeof(int32_t) * nt});
    }
    return a;
}

size_t align64(size_t x) { return (x + 63) & ~size_t(63); }

// Array order of db_arrays.
enum DbArrayIndex {
    kDbTiles, kDbToff, kDbTlen, kDbLen, kDbMag, kDbCsum, kDbPerm, kDbTmag, kDbBtiles, kDbBtoff, kDbBtlen, kDbArrays
};

void db_header_to_set(const DbFileHeader& h, ColSet& C) {
    C.n = h.n;
    C.ntiles = h.ntiles;
    C.tiles_total = h.tiles_total;
    C.btiles_total = h.btiles_total;
    C.maxlen = h.maxlen;
    C.layout = h.layout;
    C.identity = h.identity;
    C.blocked = h.blocked;
    C.levels_ok = h.levels_ok;
    C.sorted_unique = h.sorted_unique;
    C.counts_nonneg = h.counts_nonneg;
    C.mags_small = h.mags_small;
}

// The per-tile and per-molecule arrays of a database file.
struct DbHostArrays {
    std::vector<int64_t> toff, btoff;
    std::vector<int32_t> tlen, btlen, len, mag, csum, perm;
    std::vector<int2> tmag;
    void resize(const DbFileHeader& h) {
        const size_t n = (size_t)h.n, nt = (size_t)h.ntiles;
        toff.resize(nt + 1);
        tlen.resize(nt);
        len.resize(n);
        mag.resize(n);
        csum.resize(n);
        perm.resize(n);
        tmag.resize(nt);
        if (h.levels_ok) {
            btoff.resize(nt + 1);
            btlen.resize(nt);
        }
    }
};

// Sizes and flags of a header that the arrays do not need to confirm.
int check_db_header(const DbFileHeader& h, const char* path) {
    auto flag = [](uint8_t v) { return v <= 1; };
    const bool ok = h.n >= 0 && h.n <= 0x7FFFFFFF && h.ntiles == (h.n + kWave - 1) / kWave && h.tiles_total >= 0 &&
                    h.btiles_total >= 0 && h.maxlen >= 0 &&
                    (h.layout == SIML_LAYOUT_NATURAL || h.layout == SIML_LAYOUT_BLOCK ||
                     h.layout == SIML_LAYOUT_SORTED) &&
                    flag(h.identity) && flag(h.blocked) && flag(h.levels_ok) && flag(h.sorted_unique) &&
                    flag(h.counts_nonneg) && flag(h.mags_small) &&
                    (!h.levels_ok || (h.sorted_unique && h.counts_nonneg)) && (h.levels_ok || h.btiles_total == 0);
    if (!ok) return fail(SIML_EINVAL, "siml_db_load: %s has an inconsistent header", path);
    return SIML_OK;
}

// Tile offsets, lengths, magnitudes, permutation and tile magnitude ranges:
// every index the kernels derive from them stays inside its array, and the
// values match what the builder derives from the molecules.
int check_db_index(const DbFileHeader& h, const DbHostArrays& H, const char* path) {
    const int64_t n = h.n, nt = h.ntiles;
    if (H.toff[0] != 0) return fail(SIML_EINVAL, "siml_db_load: %s: tile offsets do not start at 0", path);
    for (int64_t t = 0; t < nt; ++t)
        if (H.tlen[(size_t)t] < 0 || H.toff[(size_t)t + 1] != H.toff[(size_t)t] + (int64_t)kWave * H.tlen[(size_t)t])
            return fail(SIML_EINVAL, "siml_db_load: %s: tile %lld offsets do not match its length", path,
                        (long long)t);
    if (H.toff[(size_t)nt] != h.tiles_total)
        return fail(SIML_EINVAL, "siml_db_load: %s: tile offsets do not end at the stream size", path);
    int32_t maxlen = 0;
    std::vector<char> seen((size_t)std::max<int64_t>(n, 1), 0);
    for (int64_t p = 0; p < n; ++p) {
        const int32_t L = H.len[(size_t)p];
        if (L < 0 || L > H.tlen[(size_t)(p / kWave)])
            return fail(SIML_EINVAL, "siml_db_load: %s: molecule %lld is longer than its tile", path, (long long)p);
        maxlen = std::max(maxlen, L);
        const int32_t c = H.perm[(size_t)p];
        if (c < 0 || c >= n || seen[(size_t)c])
            return fail(SIML_EINVAL, "siml_db_load: %s: the column permutation is not a permutation", path);
        seen[(size_t)c] = 1;
        if ((h.identity && c != p) || (h.blocked && c / kSortBlock != p / kSortBlock))
            return fail(SIML_EINVAL, "siml_db_load: %s: the column permutation breaks its layout flags", path);
        const int32_t m = H.mag[(size_t)p];
        if (h.mags_small && (m < 0 || m >= 65536 || m < H.csum[(size_t)p]))
            return fail(SIML_EINVAL, "siml_db_load: %s: molecule %lld's magnitude breaks the small-magnitude flag",
                        path, (long long)p);
    }
    if (maxlen != h.maxlen) return fail(SIML_EINVAL, "siml_db_load: %s: the maximum length does not match", path);
    for (int64_t t = 0; t < nt; ++t) {
        int32_t lo = INT32_MAX, hi = INT32_MIN;
        for (int64_t p = t * kWave; p < std::min(n, (t + 1) * kWave); ++p) {
            lo = std::min(lo, H.mag[(size_t)p]);
            hi = std::max(hi, H.mag[(size_t)p]);
        }
        if (H.tmag[(size_t)t].x != lo || H.tmag[(size_t)t].y != hi)
            return fail(SIML_EINVAL, "siml_db_load: %s: tile %lld's magnitude range does not match", path,
                        (long long)t);
    }
    if (h.levels_ok) {
        if (H.btoff[0] != 0) return fail(SIML_EINVAL, "siml_db_load: %s: level offsets do not start at 0", path);
        for (int64_t t = 0; t < nt; ++t) {
            const int32_t m = H.btlen[(size_t)t];
            if (m < 0 || m % 8 != 0 || H.btoff[(size_t)t + 1] != H.btoff[(size_t)t] + (int64_t)kWave * m)
                return fail(SIML_EINVAL, "siml_db_load: %s: tile %lld's level offsets do not match its length", path,
                            (long long)t);
        }
        // the stream ends with one spare block of 8 levels per lane
        if (H.btoff[(size_t)nt] + 8 * kWave != h.btiles_total)
            return fail(SIML_EINVAL, "siml_db_load: %s: level offsets do not end at the stream size", path);
    }
    return SIML_OK;
}

// The merge stream: count sums equal csum, and ids ascending / counts
// non-negative where the header says so.
int check_db_tiles(const DbFileHeader& h, const DbHostArrays& H, const std::vector<int2>& tiles, const char* path) {
    std::atomic<int64_t> bad{-1};
    parallel_for(h.n, [&](int64_t a, int64_t b) {
        for (int64_t p = a; p < b && bad.load(std::memory_order_relaxed) < 0; ++p) {
            const int64_t t = p / kWave;
            const int2* e = tiles.data() + H.toff[(size_t)t] + (p - t * kWave);
            uint32_t sum = 0;
            bool ok = true;
            for (int32_t j = 0; j < H.len[(size_t)p]; ++j) {
                const int2 v = e[(size_t)j * kWave];
                sum += (uint32_t)v.y;
                if (h.counts_nonneg && v.y < 0) ok = false;
                if (h.sorted_unique && j > 0 && v.x <= e[(size_t)(j - 1) * kWave].x) ok = false;
            }
            if (!ok || (int32_t)sum != H.csum[(size_t)p]) bad.store(p);
        }
    });
    if (bad.load() >= 0)
        return fail(SIML_EINVAL, "siml_db_load: %s: molecule %lld's entries do not match its flags or count sum", path,
                    (long long)bad.load());
    return SIML_OK;
}

// The level stream: every (key, slot word) is an empty level (0, 0) or a
// key with the slot word the probe rule derives from it, so every table
// read of the kernels stays inside the table.
int check_db_levels(const DbFileHeader& h, const DbHostArrays& H, const std::vector<int2>& levels, const char* path) {
    (void)H;
    std::atomic<int64_t> bad{-1};
    parallel_for(h.btiles_total, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b && bad.load(std::memory_order_relaxed) < 0; ++i) {
            const uint32_t key = (uint32_t)levels[(size_t)i].x, sw = (uint32_t)levels[(size_t)i].y;
            if (key == 0u ? sw != 0u : sw != siml::slot_word(key)) bad.store(i);
        }
    });
    if (bad.load() >= 0)
        return fail(SIML_EINVAL, "siml_db_load: %s: level %lld has a slot word its key does not give", path,
                    (long long)bad.load());
    return SIML_OK;
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
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->cus = cus;
    (void)hipGetLastError();
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
    ctx->qry.reset();
    ctx->res.release();
    ctx->res2.release();
    ctx->scratch.release();
    ctx->rowlist.release();
    ctx->glist.release();
    ctx->csr_dev.release();
    ctx->work.release();
    ctx->pinned[0].release();
    ctx->pinned[1].release();
    for (auto& t : ctx->pending_ev) {
        (void)hipEventDestroy(t.a);
        if (t.m) (void)hipEventDestroy(t.m);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto& kv : ctx->graph_exec) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.done) (void)hipEventDestroy(kv.second.done);
    }
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

// ---------------------------------------------------------------------------
// Resident databases shared by several handles
// ---------------------------------------------------------------------------

struct siml_db {
    int device = 0;
    std::shared_ptr<ColSet> set;
};

int siml_db_create(int device, const int32_t* ids, const int32_t* counts, const int32_t* mags, const int64_t* offsets,
                   int64_t n, int layout, siml_db** out) {
    if (!out) return fail(SIML_EINVAL, "siml_db_create: null output pointer");
    *out = nullptr;
    if (n > 0 && !offsets) return fail(SIML_EINVAL, "siml_db_create: null offsets");
    if (layout != SIML_LAYOUT_NATURAL && layout != SIML_LAYOUT_BLOCK && layout != SIML_LAYOUT_SORTED)
        return fail(SIML_EINVAL, "siml_db_create: unknown layout %d", layout);
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) {
        (void)hipGetLastError();
        return fail(SIML_EDEVICE, "siml_db_create: device %d is not visible", device);
    }
    HIP_TRY(hipSetDevice(device));
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HostSet s;
    s.ids = ids;
    s.counts = counts;
    s.mags = mags;
    s.offsets = offsets;
    s.n = n;
    auto set = std::make_shared<ColSet>();
    const int rc = build_colset(st, layout, s, *set);
    (void)hipStreamDestroy(st);
    if (rc != SIML_OK) return rc;
    siml_db* db = new siml_db();
    db->device = device;
    db->set = std::move(set);
    *out = db;
    return SIML_OK;
}

int siml_db_destroy(siml_db* db) {
    if (!db) return SIML_OK;
    (void)hipSetDevice(db->device);
    delete db;  // the device buffers live on while a handle still uses them
    return SIML_OK;
}

int siml_db_size(siml_db* db, int64_t* n) {
    if (!db || !n) return fail(SIML_EINVAL, "null pointer");
    *n = db->set->n;
    return SIML_OK;
}

int siml_db_save(siml_db* db, const char* path) {
    if (!db || !path) return fail(SIML_EINVAL, "siml_db_save: null argument");
    HIP_TRY(hipSetDevice(db->device));
    ColSet& C = *db->set;
    DbFileHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kDbMagic, 8);
    h.version = kDbVersion;
    h.layout = C.layout;
    h.n = C.n;
    h.ntiles = C.ntiles;
    h.tiles_total = C.tiles_total;
    h.btiles_total = C.btiles_total;
    h.maxlen = C.maxlen;
    h.identity = C.identity;
    h.blocked = C.blocked;
    h.levels_ok = C.levels_ok;
    h.sorted_unique = C.sorted_unique;
    h.counts_nonneg = C.counts_nonneg;
    h.mags_small = C.mags_small;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(SIML_EINVAL, "siml_db_save: cannot open %s for writing", path);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    auto put = [&](const void* p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes; };
    const char zeros[64] = {0};
    size_t pos = sizeof(h);
    if (!put(&h, sizeof(h))) return fail(SIML_EINVAL, "siml_db_save: write to %s failed", path);
    std::vector<char> host;
    for (const DbArray& a : db_arrays(C)) {
        const size_t pad = align64(pos) - pos;
        if (!put(zeros, pad)) return fail(SIML_EINVAL, "siml_db_save: write to %s failed", path);
        pos += pad;
        host.resize(a.bytes);
        if (a.bytes) HIP_TRY(hipMemcpy(host.data(), a.buf->p, a.bytes, hipMemcpyDeviceToHost));
        if (!put(host.data(), a.bytes)) return fail(SIML_EINVAL, "siml_db_save: write to %s failed", path);
        pos += a.bytes;
    }
    if (std::fflush(f) != 0) return fail(SIML_EINVAL, "siml_db_save: write to %s failed", path);
    return SIML_OK;
}

int siml_db_load(int device, const char* path, siml_db** out) {
    if (!out || !path) return fail(SIML_EINVAL, "siml_db_load: null argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(SIML_EINVAL, "siml_db_load: cannot open %s", path);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    DbFileHeader h;
    if (std::fread(&h, 1, sizeof(h), f) != sizeof(h) || std::memcmp(h.magic, kDbMagic, 8) != 0)
        return fail(SIML_EINVAL, "siml_db_load: %s is not a database file", path);
    if (h.version != kDbVersion)
        return fail(SIML_EINVAL, "siml_db_load: %s has format version %d, this build reads %d", path, h.version,
                    kDbVersion);
    TRY(check_db_header(h, path));
    // every array is read and checked on the host before anything reaches
    // the device: the small per-tile and per-molecule arrays first, then the
    // two entry streams against them
    auto C = std::make_shared<ColSet>();
    db_header_to_set(h, *C);
    const std::vector<DbArray> arrays = db_arrays(*C);
    std::vector<int64_t> at(arrays.size());
    {
        size_t pos = sizeof(h);
        for (size_t i = 0; i < arrays.size(); ++i) {
            pos = align64(pos);
            at[i] = (int64_t)pos;
            pos += arrays[i].bytes;
        }
    }
    // the arrays must lie inside the file before any host memory is sized
    // by the header's counts
    if (fseeko(f, 0, SEEK_END) != 0) return fail(SIML_EINVAL, "siml_db_load: cannot seek in %s", path);
    const off_t fsize = ftello(f);
    if (fsize < 0 || (uint64_t)at.back() + arrays.back().bytes > (uint64_t)fsize)
        return fail(SIML_EINVAL, "siml_db_load: %s is truncated", path);
    auto read_at = [&](size_t i, void* dst) -> int {
        if (arrays[i].bytes == 0) return SIML_OK;
        if (fseeko(f, (off_t)at[i], SEEK_SET) != 0 || std::fread(dst, 1, arrays[i].bytes, f) != arrays[i].bytes)
            return fail(SIML_EINVAL, "siml_db_load: %s is truncated", path);
        return SIML_OK;
    };
    DbHostArrays H;
    H.resize(h);
    TRY(read_at(kDbToff, H.toff.data()));
    TRY(read_at(kDbTlen, H.tlen.data()));
    TRY(read_at(kDbLen, H.len.data()));
    TRY(read_at(kDbMag, H.mag.data()));
    TRY(read_at(kDbCsum, H.csum.data()));
    TRY(read_at(kDbPerm, H.perm.data()));
    TRY(read_at(kDbTmag, H.tmag.data()));
    if (h.levels_ok) {
        TRY(read_at(kDbBtoff, H.btoff.data()));
        TRY(read_at(kDbBtlen, H.btlen.data()));
    }
    TRY(check_db_index(h, H, path));
    std::vector<int2> stream((size_t)std::max<int64_t>(h.tiles_total, 1));
    TRY(read_at(kDbTiles, stream.data()));
    TRY(check_db_tiles(h, H, stream, path));
    std::vector<int2> levels;
    if (h.levels_ok) {
        levels.resize((size_t)std::max<int64_t>(h.btiles_total, 1));
        TRY(read_at(kDbBtiles, levels.data()));
        TRY(check_db_levels(h, H, levels, path));
    }
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) {
        (void)hipGetLastError();
        return fail(SIML_EDEVICE, "siml_db_load: device %d is not visible", device);
    }
    HIP_TRY(hipSetDevice(device));
    const void* src[kDbArrays] = {stream.data(), H.toff.data(), H.tlen.data(), H.len.data(), H.mag.data(),
                                  H.csum.data(), H.perm.data(), H.tmag.data(), levels.data(), H.btoff.data(),
                                  H.btlen.data()};
    for (size_t i = 0; i < arrays.size(); ++i) {
        TRY(arrays[i].buf->reserve(arrays[i].bytes));
        if (arrays[i].bytes) HIP_TRY(hipMemcpy(arrays[i].buf->p, src[i], arrays[i].bytes, hipMemcpyHostToDevice));
    }
    siml_db* db = new siml_db();
    db->device = device;
    db->set = std::move(C);
    *out = db;
    return SIML_OK;
}

int siml_attach_db(siml_ctx* ctx, siml_db* db) {
    if (!ctx || !db) return fail(SIML_EINVAL, "null pointer");
    if (db->device != ctx->device)
        return fail(SIML_EINVAL, "siml_attach_db: database on device %d, handle on device %d", db->device, ctx->device);
    TRY(set_device(ctx));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    clear_result(ctx);
    ctx->qry = db->set;
    return SIML_OK;
}

int siml_set_sizes(siml_ctx* ctx, int64_t* nref, int64_t* nquery) {
    if (!ctx || !nref || !nquery) return fail(SIML_EINVAL, "null pointer");
    *nref = ctx->ref.n;
    *nquery = ctx->qry->n;
    return SIML_OK;
}

}  // extern "C"

namespace {

// Launch shape and scratch layout of a top-k call over nr rows.  `bits`:
// the rows are served by the bit-sliced kernels in `groups` row groups, gt
// of them per workgroup (the merge kernel for rows the groups do not take
// runs gt times as many column ranges, so both index the same (row, wave)
// candidate buffers); otherwise every row is on the merge kernel.
struct TopkShape {
    int gt = 1, cap = 0;
    bool persist = false;  // the full pass runs on a persistent grid
    dim3 grid;
    int64_t nw = 0;
    size_t buf = 0, cnt = 0, s1 = 0, s2 = 0, c1 = 0, c2 = 0, seed = 0, queue = 0;
    size_t scratch() const { return buf + cnt + s1 + s2 + c1 + c2 + seed + queue; }
};

// `persistent`: the bit-sliced full pass on a persistent grid, which needs a
// multiple of kXcds column ranges (at least kXcds; an ordinary grid when the
// wave cap or the buffer budget allows fewer).
TopkShape topk_shape(const siml_ctx* ctx, int64_t nr, bool bits, int64_t groups, int gt, int32_t k, int64_t target_wg,
                     unsigned wave_cap, int64_t budget = kTopkBufferBudget, bool persistent = false) {
    TopkShape t;
    t.gt = bits ? gt : 1;
    t.cap = (int)(((k + kTopkSlack) + kWave - 1) / kWave * kWave);
    t.grid = row_grid(bits ? std::max<int64_t>((groups + t.gt - 1) / t.gt, 1) : nr, ctx->qry->ntiles,
                      bits ? target_wg : 2048);
    // column ranges: at most the wave cap, and candidate buffers of at most
    // `budget` bytes (one range always)
    const int64_t range_bytes = std::max<int64_t>(nr, 1) * (kBlock / kWave) * t.gt * (int64_t)t.cap *
                                (int64_t)sizeof(uint64_t);
    const unsigned ycap = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)(wave_cap / (unsigned)t.gt), budget / range_bytes, 65535}));
    t.grid.y = std::min<unsigned>(t.grid.y, ycap);
    if (bits) t.grid.y = xcd_gy(t.grid.y, ycap);
    if (bits && persistent && ycap >= (unsigned)kXcds) {
        t.persist = true;
        t.grid.y = std::min<unsigned>(std::max<unsigned>((t.grid.y + kXcds - 1) / kXcds * kXcds, kXcds),
                                      ycap / kXcds * kXcds);
    }
    t.nw = (int64_t)t.grid.y * (kBlock / kWave) * t.gt;
    const int64_t nslots = nr * t.nw;
    auto r256 = [](size_t b) { return (b + 255) & ~size_t(255); };
    t.buf = sizeof(uint64_t) * (size_t)nslots * (size_t)t.cap;
    t.cnt = r256(sizeof(int32_t) * (size_t)nslots);
    // selection levels: segments of the first level and (ping-pong) of the
    // later, smaller ones
    const int64_t ns1 = t.nw > kTopkFan ? nr * ((t.nw + kTopkFan - 1) / kTopkFan) : 0;
    const int64_t ns2 =
        t.nw > kTopkFan * kTopkFan ? nr * ((t.nw + kTopkFan * kTopkFan - 1) / (kTopkFan * kTopkFan)) : 0;
    t.s1 = r256(sizeof(uint64_t) * (size_t)ns1 * (size_t)k);
    t.s2 = r256(sizeof(uint64_t) * (size_t)ns2 * (size_t)k);
    t.c1 = r256(sizeof(int32_t) * (size_t)ns1);
    t.c2 = r256(sizeof(int32_t) * (size_t)ns2);
    t.seed = r256(sizeof(uint64_t) * (size_t)nr);
    t.queue = t.persist ? 256 : 0;
    return t;
}

// Launch shape and scratch layout of a neighbour search over nr rows (see
// topk_shape; the bit-sliced neighbour kernels use one table per workgroup
// unless gt says otherwise).  `one_pass`: the count pass also buffers
// kNbrCap indices per (row, wave) slot.
struct NbrShape {
    int gt = 1;
    dim3 grid;
    int64_t nw = 0;
    size_t counts = 0, offs = 0, nbuf = 0, scan = 0;
    size_t scratch() const { return counts + offs + nbuf + scan; }
};

NbrShape nbr_shape(const siml_ctx* ctx, int64_t nr, bool bits, int64_t groups, int gt, bool one_pass,
                   int64_t target_wg) {
    NbrShape t;
    t.gt = bits ? gt : 1;
    t.grid = row_grid(nr, ctx->qry->ntiles, 8192, int64_t(1) << 24);
    if (bits) {
        // bit-sliced workgroups serve gt groups of up to 32 rows: size
        // grid.y by workgroups, at most 2^24 (row, wave) slots
        const int64_t slots = std::max<int64_t>((groups + t.gt - 1) / t.gt, 1);
        const unsigned gyb = row_grid(slots, ctx->qry->ntiles, target_wg).y;
        const int64_t cap =
            std::min<int64_t>(65535 / t.gt, (int64_t(1) << 24) / (std::max<int64_t>(nr, 1) * (kBlock / kWave) * t.gt));
        t.grid.y = xcd_gy((unsigned)std::max<int64_t>(1, std::min<int64_t>(gyb, cap)), (unsigned)std::max<int64_t>(cap, 1));
    }
    // waves per row: the merge kernel (4 waves per workgroup) runs gt times
    // as many column ranges so that both kernels index the same slots
    t.nw = (int64_t)t.grid.y * (kBlock / kWave) * t.gt;
    const int64_t nslots = nr * t.nw;
    t.counts = (sizeof(int32_t) * (size_t)nslots + 255) & ~size_t(255);
    t.offs = (sizeof(int64_t) * (size_t)nslots + 255) & ~size_t(255);
    t.nbuf = one_pass ? sizeof(int32_t) * (size_t)nslots * kNbrCap : 0;
    t.scan = siml::csr_scan_bytes(nslots);
    return t;
}

}  // namespace

extern "C" {

// Largest index table (slots) built for a query set: the tables take 20 B
// per slot; a set with more distinct level keys than fit at half load is
// not indexed and keeps the bit-sliced path.
constexpr int64_t kPostMaxSlots = int64_t(1) << 25;
constexpr int64_t kPostFirstSlots = int64_t(1) << 20;

// Builds the inverted index of the handle's query set unless it exists
// (ColSet::post), on the handle's stream; returns when it is complete.
// Nothing is indexed for a set without a level stream.
int ensure_postings(siml_ctx* ctx) {
    ColSet& C = *ctx->qry;
    std::lock_guard<std::mutex> lk(C.post_mu);
    if (C.post.built) return SIML_OK;
    ColSet::Postings& X = C.post;
    X.built = true;
    X.ok = false;
    if (!C.levels_ok || C.n <= 0) return SIML_OK;
    // slots: a power of two at least twice the distinct keys, found by
    // insertion: the first table holds up to 2^20 slots (or twice the
    // stream's entries, if fewer); a table more than half full (or full) is
    // built again with twice the distinct keys' slots, at most kPostMaxSlots
    int64_t slots = 1024;
    while (slots < 2 * C.btiles_total && slots < kPostFirstSlots) slots <<= 1;
    DevBuf key, cnt, off, col32, cur, tmp, inv;
    const ColArgs CA = C.args();
    const unsigned wb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((C.ntiles + 3) / 4, 16384));
    int64_t host[8] = {0};
    int64_t* tail = nullptr;
    size_t scan_bytes = 0;
    PostingArgs P{};
    for (;;) {
        TRY(key.reserve(sizeof(uint32_t) * (size_t)slots));
        TRY(cnt.reserve(sizeof(int32_t) * (size_t)slots));
        TRY(off.reserve(sizeof(int64_t) * (size_t)slots));
        TRY(cur.reserve(sizeof(int32_t) * (size_t)slots));
        scan_bytes = siml::csr_scan_bytes(slots);
        TRY(tmp.reserve(scan_bytes + 256));
        // after the scan's scratch: {0, total} row offsets, the overflow
        // flag, then {sum of count^2, sum of count, occupied slots}
        tail = reinterpret_cast<int64_t*>(tmp.as<char>() + scan_bytes);
        const unsigned fb = (unsigned)std::min<int64_t>((slots + 255) / 256, 65535);
        k_fill<uint32_t><<<dim3(fb), dim3(256), 0, ctx->stream>>>(key.as<uint32_t>(), slots, 0u);
        k_fill<int32_t><<<dim3(fb), dim3(256), 0, ctx->stream>>>(cnt.as<int32_t>(), slots, 0);
        k_fill<int32_t><<<dim3(fb), dim3(256), 0, ctx->stream>>>(cur.as<int32_t>(), slots, 0);
        k_fill<int64_t><<<dim3(1), dim3(64), 0, ctx->stream>>>(tail, 8, 0);
        TRY(launch_check("k_fill<postings>"));
        P = PostingArgs{key.as<uint32_t>(), cnt.as<int32_t>(), off.as<int64_t>(), nullptr, cur.as<int32_t>(),
                        (uint32_t)(slots - 1), reinterpret_cast<int32_t*>(tail + 2)};
        k_post_insert<<<dim3(wb), dim3(kBlock), 0, ctx->stream>>>(CA, P);
        TRY(launch_check("k_post_insert"));
        k_post_stats<<<dim3(std::min<unsigned>(fb, 2048)), dim3(256), 0, ctx->stream>>>(
            cnt.as<int32_t>(), slots, reinterpret_cast<unsigned long long*>(tail + 4));
        TRY(launch_check("k_post_stats"));
        HIP_TRY(hipMemcpyAsync(host, tail, sizeof(host), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        const bool overflow = (host[2] & 0xFFFFFFFF) != 0;
        const int64_t distinct = host[6];
        if (!overflow && 2 * distinct <= slots) break;
        if (slots >= kPostMaxSlots) {
            if (!overflow) break;  // more than half full at the largest size
            X.release();
            X.built = true;
            return SIML_OK;
        }
        int64_t next = slots * 2;
        while (next < 2 * distinct && next < kPostMaxSlots) next <<= 1;
        slots = std::min<int64_t>(next, kPostMaxSlots);
    }
    X.slots = slots;
    TRY(siml::csr_scan(cnt.as<int32_t>(), 1, slots, off.as<int64_t>(), tail, tmp.p, ctx->stream));
    HIP_TRY(hipMemcpyAsync(host, tail, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int64_t total = host[1];
    const unsigned long long sq = (unsigned long long)host[4], sum = (unsigned long long)host[5];
    if (total <= 0 || (unsigned long long)total != sum) {
        X.release();
        X.built = true;
        return SIML_OK;
    }
    TRY(col32.reserve(sizeof(int32_t) * (size_t)total));
    TRY(X.cmag.reserve(sizeof(int32_t) * (size_t)(C.n + 4)));
    TRY(X.slot.reserve(sizeof(uint4) * (size_t)slots));
    P.col = col32.as<int32_t>();
    // the chunk directory: a row per occupied slot (numbered by a scan of
    // the occupied flags), an entry per chunk boundary
    const int64_t nchunks = (C.n + kSparseCols - 1) / kSparseCols;
    X.ndir = (int32_t)(nchunks + 1);
    const unsigned sb = (unsigned)std::min<int64_t>((slots + 255) / 256, 65535);
    {
        DevBuf occ;
        TRY(occ.reserve(sizeof(int32_t) * (size_t)slots));
        TRY(X.drow.reserve(sizeof(int64_t) * (size_t)slots));
        k_post_occupied<<<dim3(sb), dim3(kBlock), 0, ctx->stream>>>(cnt.as<int32_t>(), slots, occ.as<int32_t>());
        TRY(launch_check("k_post_occupied"));
        TRY(siml::csr_scan(occ.as<int32_t>(), 1, slots, X.drow.as<int64_t>(), tail, tmp.p, ctx->stream));
        HIP_TRY(hipMemcpyAsync(host, tail, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    const int64_t keys = host[1];
    TRY(X.dir.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(keys, 1) * (size_t)X.ndir));
    if (C.n <= kSparseCols) {
        k_post_fill<<<dim3(wb), dim3(kBlock), 0, ctx->stream>>>(CA, P);
        TRY(launch_check("k_post_fill"));
        k_post_dir<<<dim3(sb), dim3(kBlock), 0, ctx->stream>>>(cnt.as<int32_t>(), cur.as<int32_t>(),
                                                             X.drow.as<int64_t>(), slots, X.dir.as<int32_t>(), X.ndir,
                                                             0);
        TRY(launch_check("k_post_dir"));
    } else {
        // one fill per chunk of kSparseCols caller columns, in order: each
        // key's postings come out grouped by chunk, as the sparse top-k walk
        // reads them
        const int32_t* d_inv = nullptr;
        if (CA.perm) {
            TRY(inv.reserve(sizeof(int32_t) * (size_t)C.n));
            k_inv_perm<<<dim3((unsigned)std::min<int64_t>((C.n + kBlock - 1) / kBlock, 65535)), dim3(kBlock), 0,
                         ctx->stream>>>(CA.perm, C.n, inv.as<int32_t>());
            TRY(launch_check("k_inv_perm"));
            d_inv = inv.as<int32_t>();
        }
        for (int64_t c0 = 0; c0 < C.n; c0 += kSparseCols) {
            const int64_t c1 = std::min<int64_t>(C.n, c0 + kSparseCols);
            k_post_fill_cols<<<dim3((unsigned)((c1 - c0 + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream>>>(
                CA, P, d_inv, c0, c1);
            TRY(launch_check("k_post_fill_cols"));
            k_post_dir<<<dim3(sb), dim3(kBlock), 0, ctx->stream>>>(cnt.as<int32_t>(), cur.as<int32_t>(),
                                                                 X.drow.as<int64_t>(), slots, X.dir.as<int32_t>(),
                                                                 X.ndir, (int32_t)(c0 / kSparseCols));
            TRY(launch_check("k_post_dir"));
        }
    }
    {
        const unsigned fb = (unsigned)std::min<int64_t>((slots + 255) / 256, 65535);
        k_post_pack<<<dim3(fb), dim3(kBlock), 0, ctx->stream>>>(key.as<uint32_t>(), cnt.as<int32_t>(),
                                                              off.as<int64_t>(), slots, X.slot.as<uint4>());
        TRY(launch_check("k_post_pack"));
    }
    X.narrow = C.n <= 65536;
    if (X.narrow) {
        TRY(X.col.reserve(sizeof(uint16_t) * (size_t)total));
        const unsigned nb = (unsigned)std::min<int64_t>((total + 255) / 256, 65535);
        k_post_narrow<<<dim3(nb), dim3(kBlock), 0, ctx->stream>>>(col32.as<int32_t>(), total, X.col.as<uint16_t>());
        TRY(launch_check("k_post_narrow"));
    } else {
        std::swap(X.col, col32);
    }
    k_caller_mags<<<dim3((unsigned)std::min<int64_t>((C.n + 4 + 255) / 256, 65535)), dim3(256), 0, ctx->stream>>>(
        CA, X.cmag.as<int32_t>());
    TRY(launch_check("k_caller_mags"));
    TRY(X.cmag8.reserve((size_t)C.n + 32));
    TRY(X.cmag16.reserve(sizeof(uint16_t) * ((size_t)C.n + 32)));
    k_caller_mags_small<<<dim3((unsigned)std::min<int64_t>((C.n + 32 + 255) / 256, 65535)), dim3(256), 0,
                          ctx->stream>>>(X.cmag.as<int32_t>(), C.n, X.cmag8.as<uint8_t>(), X.cmag16.as<uint16_t>(),
                                         reinterpret_cast<unsigned*>(tail + 7));
    TRY(launch_check("k_caller_mags_small"));
    HIP_TRY(hipMemcpyAsync(host + 7, tail + 7, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    {
        const uint64_t maxmag = (uint64_t)host[7] & 0xFFFFFFFFull;
        X.mag_bytes = maxmag <= 255 ? 1 : (maxmag <= 65535 ? 2 : 0);
    }
    X.total = total;
    X.per_level = (double)sq / (double)sum;
    X.ok = true;
    return SIML_OK;
}

// Cost model of AUTO's choice (float32 matrix), in picoseconds of one
// MI355X, measured on 32768 x 32768 matrices of synthetic strings (67
// levels per molecule, 0.2 postings met per pair) and of the reference's
// ZINC molecules cycled (about 28 levels, 6.25 postings per pair): the
// sparse kernel takes 0.69 per pair (writing the matrix) plus 0.30 per
// posting its rows' levels meet (0.80 and 2.73 ms); the bit-sliced kernel
// 0.72 per pair plus 0.0112 per pair and query level (1.57 and 1.11 ms).
constexpr double kSparsePsPerPair = 0.69, kSparsePsPerPosting = 0.30;
constexpr double kBitsPsPerPair = 0.72, kBitsPsPerPairLevel = 0.0112;
// Top-k and the 1000-bin histogram (no matrix to write), measured on the
// top-100 of 257684 lognormal strings over 1M (0.09 postings met per pair,
// 45 levels per molecule): the sparse kernel takes 0.40 per pair (clearing
// and scanning the counts, its per-chunk steps) plus 0.30 per posting met
// (110 ms), the seeded bit-sliced passes 0.84 per pair plus 0.0112 per
// pair and query level (347 ms); the histograms of 1024 rows over 10M take
// 7.4 and 13.7 ms.
constexpr double kSparseTopkPsPerPair = 0.40, kBitsTopkPsPerPair = 0.84;

// Rows of [rb, rl) for the sparse kernels (rows with ascending ASCII ids,
// counts of at most 15 and at most 255 levels) and the rest, when the sparse
// path applies to the call: policy SPARSE, or AUTO with few expected
// postings.  Decided before any row-group planning, so a call that takes
// the sparse path never builds the row groups.
bool plan_sparse(siml_ctx* ctx, int64_t rb, int64_t rl, std::vector<int32_t>& sparse_rows,
                 std::vector<int32_t>& merge_rows, int* rc, bool topk = false) {
    *rc = SIML_OK;
    const RowSet& R = ctx->ref;
    if (!R.groups_ok || !ctx->qry->levels_ok ||
        (ctx->policy != SIML_POLICY_AUTO && ctx->policy != SIML_POLICY_SPARSE))
        return false;
    *rc = ensure_postings(ctx);
    if (*rc != SIML_OK) return false;
    const ColSet::Postings& X = ctx->qry->post;
    // the sparse kernel packs a level's posting count in 28 bits
    if (!X.ok || ctx->qry->n >= (int64_t(1) << 28)) return false;
    sparse_rows.clear();
    merge_rows.clear();
    double levels = 0.0;
    for (int64_t r = rb; r < rl; ++r) {
        if (R.h_sparse_ok[(size_t)r]) {
            sparse_rows.push_back((int32_t)r);
            levels += (double)R.h_csum[(size_t)r];
        } else {
            merge_rows.push_back((int32_t)r);
        }
    }
    if (ctx->policy == SIML_POLICY_SPARSE) return true;
    // expected postings met: the rows' levels times the postings of a
    // random stream level (exact when the rows are drawn like the queries)
    const double pairs = (double)(rl - rb) * (double)ctx->qry->n;
    const double qlevels = (double)X.total / (double)ctx->qry->n;
    if (topk)
        return kSparseTopkPsPerPair * pairs + kSparsePsPerPosting * levels * X.per_level <
               (kBitsTopkPsPerPair + kBitsPsPerPairLevel * qlevels) * pairs;
    return kSparsePsPerPair * pairs + kSparsePsPerPosting * levels * X.per_level <
           (kBitsPsPerPair + kBitsPsPerPairLevel * qlevels) * pairs;
}

// Launch shape of the sparse top-k and histogram: column ranges per row
// (several when the rows alone give too few workgroups -- four rounds of
// three workgroups per CU --, for the top-k at most kTopkFan so that one
// final selection workgroup per row merges them) and chunks per range.
constexpr int64_t kSparseTopkWorkgroups = 3072;
struct SparseTopkShape {
    int64_t nsplit = 1, cps = 1;
    size_t buf = 0, cnt = 0;  // candidate segments (nsplit > 1)
    size_t scratch() const { return buf + cnt; }
};
static SparseTopkShape sparse_topk_shape(int64_t nr, int64_t nq, int k, bool topk = true) {
    SparseTopkShape s;
    const int64_t nchunks = std::max<int64_t>(1, (nq + kSparseCols - 1) / kSparseCols);
    const int64_t want = std::max<int64_t>(1, (kSparseTopkWorkgroups + nr - 1) / std::max<int64_t>(nr, 1));
    s.nsplit = std::min<int64_t>(std::min<int64_t>(nchunks, topk ? (int64_t)kTopkFan : nchunks), want);
    s.cps = (nchunks + s.nsplit - 1) / s.nsplit;
    s.nsplit = (nchunks + s.cps - 1) / s.cps;
    if (s.nsplit > 1 && topk) {
        auto r256 = [](size_t b) { return (b + 255) & ~size_t(255); };
        s.buf = r256(sizeof(uint64_t) * (size_t)nr * (size_t)s.nsplit * (size_t)k);
        s.cnt = r256(sizeof(int32_t) * (size_t)nr * (size_t)s.nsplit);
    }
    return s;
}

// Top-k of rows [rb, rl) on the sparse path (every row grouped with at most
// 255 levels, magnitudes at least the count sums); results into the
// handle's result buffers.
static int sparse_topk(siml_ctx* ctx, int64_t rb, int64_t rl, int32_t k) {
    const int64_t nr = rl - rb, nq = ctx->qry->n;
    const ColSet::Postings& X = ctx->qry->post;
    const PostingQuery P = X.query();
    const SparseTopkShape shp = sparse_topk_shape(nr, nq, k);
    uint64_t* d_buf = nullptr;
    int32_t* d_cnt = nullptr;
    if (shp.nsplit > 1) {
        TRY(ctx->scratch.reserve(shp.scratch()));
        d_buf = ctx->scratch.as<uint64_t>();
        d_cnt = reinterpret_cast<int32_t*>(ctx->scratch.as<char>() + shp.buf);
    }
    int32_t* out_idx = ctx->res.as<int32_t>();
    float* out_t = ctx->res2.as<float>();
    GraphRegion gr(ctx);
    TRY(gr.begin());
    TimedRegion tr(ctx, nr * nq, !gr.on);
    const dim3 grid((unsigned)nr, (unsigned)shp.nsplit);
    auto go = [&](auto kern, const void* mags) {
        kern<<<grid, dim3(kSparseBlock), 0, ctx->stream>>>(ctx->ref.args(), rb, P, mags, nq, k, shp.cps, d_buf, d_cnt,
                                                           out_idx, out_t);
    };
    const void* m8 = X.cmag8.p;
    const void* m16 = X.cmag16.p;
    if (X.narrow && X.mag_bytes == 1)
        go(k_sparse_topk<true, 1>, m8);
    else if (X.narrow)
        go(k_sparse_topk<true, 2>, m16);
    else if (X.mag_bytes == 1)
        go(k_sparse_topk<false, 1>, m8);
    else
        go(k_sparse_topk<false, 2>, m16);
    TRY(launch_check("k_sparse_topk"));
    tr.select();
    if (shp.nsplit > 1) {
        k_topk_final<<<dim3((unsigned)nr), dim3(kBlock), topk_lds(k), ctx->stream>>>(d_buf, d_cnt, (int)shp.nsplit, k,
                                                                                   k, 1, nq, out_idx, out_t);
        TRY(launch_check("k_topk_final"));
    }
    tr.done();
    if (gr.on) {
        TRY(gr.end());
        TimedRegion tg(ctx, nr * nq);
        TRY(gr.launch());
        tg.done();
    }
    ctx->last_kernel = "topk_sparse";
    return SIML_OK;
}

static int multirow_async_impl(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit) {
    TRY(check_rows(ctx, rowbase, rowlimit, "getMultipleRows"));
    TRY(set_device(ctx));
    clear_result(ctx);
    const int64_t nr = rowlimit - rowbase, nq = ctx->qry->n;
    // device row pitch: a multiple of 16 floats (16-byte aligned 256-column
    // blocks, and the row pitch of the reference's padded result,
    // GPULingo.py:378-379, so that GPULingo's retrieval is one linear copy)
    const int64_t dp = (nq + 15) & ~int64_t(15);
    TRY(ctx->res.reserve(sizeof(float) * (size_t)std::max<int64_t>(nr * dp, 1)));
    ctx->kind = SIML_RESULT_F32;
    ctx->rows = nr;
    ctx->cols = nq;
    ctx->dpitch = dp;
    if (nr == 0 || nq == 0) return SIML_OK;
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, rowbase, rowlimit));
    std::vector<int32_t> srows, mrows;
    int prc = SIML_OK;
    const bool sparse = plan_sparse(ctx, rowbase, rowlimit, srows, mrows, &prc);
    TRY(prc);
    if (sparse) {
        // one list: the sparse kernel's rows, then the merge kernel's
        std::vector<int32_t> both(srows);
        both.insert(both.end(), mrows.begin(), mrows.end());
        if (!both.empty()) TRY(upload_rowlist(ctx, both));
        const ColSet::Postings& X = ctx->qry->post;
        const PostingQuery P = X.query();
        const bool fast = ctx->ref.mags_small && ctx->qry->mags_small;
        TimedRegion tr(ctx, nr * nq);
        if (!srows.empty()) {
            const dim3 grid((unsigned)((srows.size() + kSparseRows - 1) / kSparseRows),
                            (unsigned)((nq + kSparseCols - 1) / kSparseCols));
            auto go = [&](auto kern) {
                kern<<<grid, dim3(kSparseBlock), 0, ctx->stream>>>(ctx->ref.args(), ctx->rowlist.as<int32_t>(),
                                                                    (int64_t)srows.size(), P, X.cmag.as<int32_t>(),
                                                                    nq, rowbase, ctx->res.as<float>(), dp);
            };
            if (fast && X.narrow)
                go(k_sparse_f32<true, true>);
            else if (fast)
                go(k_sparse_f32<true, false>);
            else if (X.narrow)
                go(k_sparse_f32<false, true>);
            else
                go(k_sparse_f32<false, false>);
            TRY(launch_check("k_sparse_f32"));
        }
        if (!mrows.empty()) {
            k_rows_f32<false><<<row_grid((int64_t)mrows.size(), ctx->qry->ntiles), dim3(kBlock), lds, ctx->stream>>>(
                ctx->ref.args(), ctx->qry->args(), ctx->rowlist.as<int32_t>() + srows.size(), rowbase,
                ctx->res.as<float>(), dp);
            TRY(launch_check("k_rows_f32<rowlist>"));
        }
        tr.done();
        ctx->last_kernel = mrows.empty() ? "rows_sparse" : (srows.empty() ? "rows_merge" : "rows_sparse+merge");
        return SIML_OK;
    }
    BitPlan bp;
    TRY(plan_bits(ctx, rowbase, rowlimit, bp));
    if (bp.use) {
        if (!bp.merge.empty()) TRY(upload_rowlist(ctx, bp.merge));
        TimedRegion tr(ctx, nr * nq);
        const unsigned gy = xcd_gy(

            row_grid(std::max<int64_t>((int64_t)bp.groups.size(), 1), ctx->qry->ntiles, ctx->target_wg).y, 65535);
        if (ctx->qry->blocked)
            TRY(launch_bits(ctx, bp, gy, rowbase, rowlimit, EpiF32Block{ctx->res.as<float>(), dp, rowbase},
                            "k_bits<f32 block>"));
        else
            TRY(launch_bits(ctx, bp, gy, rowbase, rowlimit, EpiF32<1>{ctx->res.as<float>(), dp, rowbase},
                            "k_bits<f32>"));
        if (!bp.merge.empty()) {
            k_rows_f32<false><<<row_grid((int64_t)bp.merge.size(), ctx->qry->ntiles), dim3(kBlock), lds,
                                ctx->stream>>>(ctx->ref.args(), ctx->qry->args(), ctx->rowlist.as<int32_t>(), rowbase,
                                               ctx->res.as<float>(), dp);
            TRY(launch_check("k_rows_f32<rowlist>"));
        }
        tr.done();
        ctx->last_kernel = bp.merge.empty() ? "rows_bits" : (bp.groups.empty() ? "rows_merge" : "rows_bits+merge");
        return SIML_OK;
    }
    TimedRegion tr(ctx, nr * nq);
    k_rows_f32<false><<<row_grid(nr, ctx->qry->ntiles), dim3(kBlock), lds, ctx->stream>>>(
        ctx->ref.args(), ctx->qry->args(), nullptr, rowbase, ctx->res.as<float>(), dp);
    TRY(launch_check("k_rows_f32"));
    ctx->last_kernel = "rows_merge";
    tr.done();
    return SIML_OK;
}

int siml_single_row_async(siml_ctx* ctx, int64_t row) {
    TRY(check_rows(ctx, row, row + 1, "getTanimotoRow"));
    TRY(set_device(ctx));
    clear_result(ctx);
    const int64_t nq = ctx->qry->n;
    TRY(ctx->res.reserve(sizeof(float) * (size_t)std::max<int64_t>(nq, 1)));
    ctx->kind = SIML_RESULT_F32;
    ctx->rows = 1;
    ctx->cols = nq;
    ctx->dpitch = nq;
    if (nq == 0) return SIML_OK;
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, row, row + 1));
    TimedRegion tr(ctx, nq);
    k_rows_f32<true><<<row_grid(1, ctx->qry->ntiles, 2048), dim3(kBlock), lds, ctx->stream>>>(
        ctx->ref.args(), ctx->qry->args(), nullptr, row, ctx->res.as<float>(), nq);
    TRY(launch_check("k_rows_f32<single>"));
    tr.done();
    ctx->last_kernel = "single_row_merge";
    return SIML_OK;
}

static int hist_rows_async_impl(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit) {
    TRY(check_rows(ctx, rowbase, rowlimit, "getMultipleHistogrammedRows"));
    TRY(set_device(ctx));
    clear_result(ctx);
    const int64_t nr = rowlimit - rowbase, nq = ctx->qry->n;
    TRY(ctx->res.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(nr * 101, 1)));
    ctx->kind = SIML_RESULT_I32;
    ctx->rows = nr;
    ctx->cols = 101;
    ctx->dpitch = 101;
    TRY(zero_fill_i32(ctx, ctx->res.as<int32_t>(), nr * 101));
    if (nr == 0 || nq == 0) return SIML_OK;
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, rowbase, rowlimit));
    int32_t* hist = ctx->res.as<int32_t>();
    BitPlan bp;
    TRY(plan_bits(ctx, rowbase, rowlimit, bp));
    if (bp.use) {
        if (!bp.merge.empty()) TRY(upload_rowlist(ctx, bp.merge));
        TimedRegion tr(ctx, nr * nq);
        const unsigned gy = xcd_gy(
            row_grid(std::max<int64_t>((int64_t)bp.groups.size(), 1), ctx->qry->ntiles, ctx->target_wg).y, 65535);
        TRY(launch_bits(ctx, bp, gy, rowbase, rowlimit, EpiHist101{hist, rowbase}, "k_bits<hist101>"));
        if (!bp.merge.empty()) {
            k_hist_rows<<<row_grid((int64_t)bp.merge.size(), ctx->qry->ntiles), dim3(kBlock), lds, ctx->stream>>>(
                ctx->ref.args(), ctx->qry->args(), ctx->rowlist.as<int32_t>(), rowbase, hist);
            TRY(launch_check("k_hist_rows<rowlist>"));
        }
        tr.done();
        ctx->last_kernel = "hist_rows_bits";
        return SIML_OK;
    }
    TimedRegion tr(ctx, nr * nq);
    k_hist_rows<<<row_grid(nr, ctx->qry->ntiles), dim3(kBlock), lds, ctx->stream>>>(
        ctx->ref.args(), ctx->qry->args(), nullptr, rowbase, hist);
    TRY(launch_check("k_hist_rows"));
    tr.done();
    ctx->last_kernel = "hist_rows_merge";
    return SIML_OK;
}

static int neighbors_async_impl(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, float lo, float hi) {
    TRY(check_rows(ctx, rowbase, rowlimit, "getNeighbors"));
    TRY(set_device(ctx));
    clear_result(ctx);
    const int64_t nr = rowlimit - rowbase, nq = ctx->qry->n;
    ctx->kind = SIML_RESULT_CSR;
    ctx->rows = nr;
    ctx->cols = nq;
    if (nr == 0 || nq == 0) {
        TRY(ctx->res.reserve(256));
        return SIML_OK;
    }
    BitPlan bp;
    TRY(plan_bits(ctx, rowbase, rowlimit, bp, true));
    const std::vector<int32_t>& mrows = bp.merge;
    if (!mrows.empty()) TRY(upload_rowlist(ctx, mrows));
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, rowbase, rowlimit));
    // single pass when every row is on the bit-sliced kernel: the count pass
    // keeps the first kNbrCap indices of each slot
    const bool one_pass = bp.use && mrows.empty();
    const NbrShape shp = nbr_shape(ctx, nr, bp.use, (int64_t)bp.groups.size(), bp.use ? bits_tables(ctx, bp, true) : 1,
                                   one_pass, ctx->target_wg);
    const dim3 grid = shp.grid;
    const int gt = shp.gt;
    const int64_t nw = shp.nw;
    int32_t* d_nbuf = nullptr;
    const dim3 mgrid((unsigned)std::max<size_t>(mrows.size(), 1), grid.y * gt);
    unsigned long long* work = ctx->timing ? ctx->work.as<unsigned long long>() : nullptr;
    const int32_t* rlen = ctx->ref.args().len;
    auto launch = [&](bool fill, const int64_t* offs, int32_t* idx) -> int {
        int32_t* counts = fill ? nullptr : d_counts_of(ctx);
        if (bp.use) {
            TRY(with_tables(gt, [&](auto gc) {
                constexpr int GC = decltype(gc)::value;
                if (fill)
                    return launch_bits(ctx, bp, grid.y, rowbase, rowlimit,
                                       EpiNbr<true, GC>{lo, hi, nullptr, offs, idx, rowbase, nullptr, 0, work, rlen},
                                       "k_bits<nbr fill>");
                return launch_bits(ctx, bp, grid.y, rowbase, rowlimit,
                                   EpiNbr<false, GC>{lo, hi, counts, nullptr, nullptr, rowbase, d_nbuf, kNbrCap, work,
                                                     rlen},
                                   "k_bits<nbr count>");
            }));
        } else {
            if (fill)
                k_neighbors<true><<<grid, dim3(kBlock), lds, ctx->stream>>>(ctx->ref.args(), ctx->qry->args(), nullptr,
                                                                            rowbase, lo, hi, nullptr, offs, idx, work);
            else
                k_neighbors<false><<<grid, dim3(kBlock), lds, ctx->stream>>>(ctx->ref.args(), ctx->qry->args(), nullptr,
                                                                             rowbase, lo, hi, counts, nullptr, nullptr,
                                                                             work);
            return launch_check("k_neighbors");
        }
        // rows the multi-row kernels do not serve: merge kernel over a row list
        if (!mrows.empty()) {
            if (fill)
                k_neighbors<true><<<mgrid, dim3(kBlock), lds, ctx->stream>>>(
                    ctx->ref.args(), ctx->qry->args(), ctx->rowlist.as<int32_t>(), rowbase, lo, hi, nullptr, offs, idx,
                    work);
            else
                k_neighbors<false><<<mgrid, dim3(kBlock), lds, ctx->stream>>>(
                    ctx->ref.args(), ctx->qry->args(), ctx->rowlist.as<int32_t>(), rowbase, lo, hi, counts, nullptr,
                    nullptr, work);
            TRY(launch_check("k_neighbors<rowlist>"));
        }
        return SIML_OK;
    };
    const int64_t nslots = nr * nw;
    TRY(ctx->scratch.reserve(shp.scratch()));
    char* sbase = ctx->scratch.as<char>();
    int32_t* d_counts = reinterpret_cast<int32_t*>(sbase);
    int64_t* d_offs = reinterpret_cast<int64_t*>(sbase + shp.counts);
    d_nbuf = one_pass ? reinterpret_cast<int32_t*>(sbase + shp.counts + shp.offs) : nullptr;
    void* d_scan = sbase + shp.counts + shp.offs + shp.nbuf;
    TRY(ctx->csr_dev.reserve(sizeof(int64_t) * (size_t)(nr + 1)));
    int64_t* d_rows = ctx->csr_dev.as<int64_t>();
    TimedRegion tr(ctx, nr * nq);
    TRY(launch(false, nullptr, nullptr));
    // slot and row offsets on the device; the host reads back only the
    // totals it needs to size the index array
    TRY(siml::csr_scan(d_counts, nr, nw, d_offs, d_rows, d_scan, ctx->stream));
    int64_t stats[3];
    HIP_TRY(hipMemcpyAsync(stats, d_scan, sizeof(stats), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int64_t acc = stats[0], most = stats[1], longest = stats[2];
    ctx->nnz = acc;
    ctx->csr_on_device = true;
    TRY(ctx->res.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(acc, 1)));
    if (one_pass && most <= kNbrCap) {
        if (acc > 0) {
            const int64_t nthreads = nslots * kNbrCap;
            k_nbr_compact<<<dim3((unsigned)((nthreads + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream>>>(
                d_nbuf, d_counts, d_offs, nslots, kNbrCap, ctx->res.as<int32_t>());
            TRY(launch_check("k_nbr_compact"));
        }
    } else {
        TRY(launch(true, d_offs, ctx->res.as<int32_t>()));
    }
    if (!ctx->qry->identity && acc > 1 && longest > 1) {
        // indices were written in storage order: sort each row ascending
        TRY(ctx->res2.reserve(sizeof(int32_t) * (size_t)acc));
        int32_t* sorted = nullptr;
        TRY(siml::csr_sort_rows(ctx->res.as<int32_t>(), ctx->res2.as<int32_t>(), acc, nr, d_rows, longest, &sorted,
                                ctx->stream));
        if (sorted != ctx->res.as<int32_t>()) std::swap(ctx->res, ctx->res2);
    }
    tr.done();
    ctx->last_kernel = bp.use ? "neighbors_bits" : "neighbors_merge";
    return SIML_OK;
}

// Scratch of siml_topk_noalloc: a static device array, one copy per GPU,
// held until the process ends.  In host code the array's name is only a
// stand-in for it; its device address on each GPU comes from
// hipGetSymbolAddress with that GPU current, taken on the entry point's
// first call there and kept per device.
static __device__ __align__(256) unsigned char noalloc_scratch[SIML_NOALLOC_SCRATCH_BYTES];

static std::mutex g_noalloc_mu;
static std::map<int, char*> g_noalloc_addr;

// Device address of noalloc_scratch on `device`, which must be current.
static int noalloc_scratch_addr(int device, char** out) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    auto it = g_noalloc_addr.find(device);
    if (it == g_noalloc_addr.end()) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
        if (!p) return fail(SIML_EDEVICE, "hipGetSymbolAddress gave no address for the scratch array");
        it = g_noalloc_addr.emplace(device, static_cast<char*>(p)).first;
    }
    *out = it->second;
    return SIML_OK;
}

// Device memory an allocation-free top-k call brings: a scratch area for
// the group list, the row list and every buffer of the kernels, and the
// result arrays.
struct CallerMem {
    char* scratch;
    size_t cap;
    int32_t* idx;
    float* t;
};

// Top-k of rows [rowbase, rowlimit).  Without `mem`, device buffers are the
// handle's (grown as needed) and the result is the handle's pending result;
// with `mem`, nothing is allocated: the launch is narrowed to fewer column
// ranges until its buffers fit mem->cap bytes, and the result goes to
// mem->idx / mem->t.
static int topk_impl(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, int32_t k, const CallerMem* mem) {
    TRY(check_rows(ctx, rowbase, rowlimit, "getTopK"));
    if (k < 1 || k > SIML_MAX_TOPK) return fail(SIML_EINVAL, "getTopK: k=%d outside [1, %d]", k, SIML_MAX_TOPK);
    if (ctx->qry->n > 0x7FFFFFFF) return fail(SIML_ELIMIT, "getTopK: more than 2^31-1 query molecules");
    TRY(set_device(ctx));
    const int64_t nr = rowlimit - rowbase, nq = ctx->qry->n;
    if (!mem) {
        clear_result(ctx);
        TRY(ctx->res.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(nr * k, 1)));
        TRY(ctx->res2.reserve(sizeof(float) * (size_t)std::max<int64_t>(nr * k, 1)));
        ctx->kind = SIML_RESULT_TOPK;
        ctx->rows = nr;
        ctx->cols = k;
    }
    if (nr == 0) return SIML_OK;
    if (!mem && nq > 0 && nq <= 0x7FFFFFFF && nr <= 0x7FFFFFFF && ctx->ref.mags_small && ctx->qry->mags_small) {
        // the sparse path when every row can take it and the model (or the
        // policy) prefers it
        std::vector<int32_t> srows, smerge;
        int prc = SIML_OK;
        const bool sp = plan_sparse(ctx, rowbase, rowlimit, srows, smerge, &prc, true);
        TRY(prc);
        if (sp && smerge.empty() && ctx->qry->post.mag_bytes != 0) return sparse_topk(ctx, rowbase, rowlimit, k);
    }
    auto r256 = [](size_t b) { return (b + 255) & ~size_t(255); };
    // caller memory: the group list and the row list first
    size_t used = 0;
    int32_t* glist_dst = nullptr;
    int32_t* rowlist_dst = nullptr;
    if (mem) {
        glist_dst = reinterpret_cast<int32_t*>(mem->scratch);
        used = r256(sizeof(int32_t) * (ctx->ref.bits[0].h_hdr.size() + kMaxTables));
        rowlist_dst = reinterpret_cast<int32_t*>(mem->scratch + used);
        used += r256(sizeof(int32_t) * (size_t)nr);
        if (used > mem->cap)
            return fail(SIML_ELIMIT, "getTopK: the row and group lists of %lld rows need %zu bytes of scratch, more "
                        "than the %zu available", (long long)nr, used, mem->cap);
    }
    // caller memory: row groups only when an earlier call built them (their
    // build allocates), else the merge kernel
    BitPlan bp;
    TRY(plan_bits(ctx, rowbase, rowlimit, bp, false, glist_dst, mem != nullptr));
    const std::vector<int32_t>& mrows = bp.merge;
    const int32_t* d_rowlist = nullptr;
    if (!mrows.empty()) {
        if (mem) {
            ctx->h_rowlist = mrows;
            HIP_TRY(hipMemcpyAsync(rowlist_dst, ctx->h_rowlist.data(), sizeof(int32_t) * mrows.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
            d_rowlist = rowlist_dst;
        } else {
            TRY(upload_rowlist(ctx, mrows));
            d_rowlist = ctx->rowlist.as<int32_t>();
        }
    }
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, rowbase, rowlimit));
    // every row on the bit-sliced kernel with consistent magnitudes: T is in
    // [0, 1] for every pair, so only T > 0 pairs need candidate buffers (the
    // final selection fills up with T == 0 columns in index order)
    const bool pos_only = bp.use && mrows.empty() && ctx->ref.mags_small && ctx->qry->mags_small;
    static_assert(((SIML_MAX_TOPK + kTopkSlack) + kWave - 1) / kWave <= kTopkRegs,
                  "top-k buffer exceeds wave registers");
    const int gt_plan = bp.use ? bits_tables(ctx, bp) : 1;
    TopkShape shp = topk_shape(ctx, nr, bp.use, (int64_t)bp.groups.size(), gt_plan, k, ctx->target_wg,
                               ctx->topk_wave_cap, kTopkBufferBudget, ctx->persistent);
    if (mem) {
        const size_t room = mem->cap - used;
        for (int64_t budget = (int64_t)room; shp.scratch() > room; budget /= 2) {
            if (shp.grid.y <= 1)
                return fail(SIML_ELIMIT, "getTopK: %lld rows with k=%d need %zu bytes of scratch at the narrowest "
                            "launch, more than the %zu available", (long long)nr, k, used + shp.scratch(), mem->cap);
            shp = topk_shape(ctx, nr, bp.use, (int64_t)bp.groups.size(), gt_plan, k, ctx->target_wg,
                             ctx->topk_wave_cap, budget, ctx->persistent);
        }
    }
    const int cap = shp.cap, gt = shp.gt;
    const dim3 grid = shp.grid;
    int64_t nw = shp.nw;
    const size_t buf_bytes = shp.buf, cnt_bytes = shp.cnt, s1_bytes = shp.s1, s2_bytes = shp.s2,
                 c1_bytes = shp.c1, c2_bytes = shp.c2;
    char* base = nullptr;
    if (mem) {
        base = mem->scratch + used;
    } else {
        TRY(ctx->scratch.reserve(shp.scratch()));
        base = ctx->scratch.as<char>();
    }
    uint64_t* d_seed = reinterpret_cast<uint64_t*>(base + buf_bytes + cnt_bytes + s1_bytes + s2_bytes + c1_bytes +
                                                   c2_bytes);
    int32_t* d_queue = shp.persist ? reinterpret_cast<int32_t*>(reinterpret_cast<char*>(d_seed) + shp.seed) : nullptr;
    uint64_t* d_buf = reinterpret_cast<uint64_t*>(base);
    int32_t* d_cnt = reinterpret_cast<int32_t*>(base + buf_bytes);
    uint64_t* d_sbuf = reinterpret_cast<uint64_t*>(base + buf_bytes + cnt_bytes);
    uint64_t* d_sbuf2 = reinterpret_cast<uint64_t*>(base + buf_bytes + cnt_bytes + s1_bytes);
    int32_t* d_scnt = reinterpret_cast<int32_t*>(base + buf_bytes + cnt_bytes + s1_bytes + s2_bytes);
    int32_t* d_scnt2 = reinterpret_cast<int32_t*>(base + buf_bytes + cnt_bytes + s1_bytes + s2_bytes + c1_bytes);
    // reduce kTopkFan buffers per workgroup until one workgroup per row can
    // take the rest, then select and sort
    // sparse: the buffers of a seeded pass hold few keys (about k times the
    // seed stride per row, over all of the row's waves), so a selection
    // workgroup streams kTopkStreamFan of them instead of kTopkFan
    auto select = [&](int32_t* out_idx, float* out_t, uint64_t* kth, bool sparse = false) -> int {
        const uint64_t* src = d_buf;
        const int32_t* srccnt = d_cnt;
        int64_t nseg = nw;
        int segcap = cap;
        const int fan = sparse ? kTopkStreamFan : kTopkFan;
        for (int level = 0; nseg > fan; ++level) {
            // a sparse level has fewer outputs than a full one: the level
            // buffers (sized for full levels) hold them
            const int64_t ns = (nseg + fan - 1) / fan;
            uint64_t* dst = level % 2 ? d_sbuf2 : d_sbuf;
            int32_t* dstcnt = level % 2 ? d_scnt2 : d_scnt;
            k_topk_stage<<<dim3((unsigned)nr, (unsigned)ns), dim3(kBlock), topk_lds(segcap), ctx->stream>>>(
                src, srccnt, (int)nseg, segcap, k, fan, dst, dstcnt);
            TRY(launch_check("k_topk_stage"));
            src = dst;
            srccnt = dstcnt;
            nseg = ns;
            segcap = k;
        }
        if (kth) {
            k_topk_kth<<<dim3((unsigned)nr), dim3(kBlock), topk_lds(segcap), ctx->stream>>>(src, srccnt, (int)nseg,
                                                                                            k, segcap, kth);
            return launch_check("k_topk_kth");
        }
        k_topk_final<<<dim3((unsigned)nr), dim3(kBlock), topk_lds(segcap), ctx->stream>>>(src, srccnt, (int)nseg, k, segcap,
                                                                         pos_only ? 1 : 0, nq, out_idx, out_t);
        return launch_check("k_topk_final");
    };
    GraphRegion gr(ctx);
    TRY(gr.begin());
    TimedRegion tr(ctx, nr * nq, mem == nullptr && !gr.on);
    const uint64_t* seed = nullptr;
    const int64_t stride = seed_stride(ctx->topk_seed, ctx->qry->ntiles);
    if (bp.use && pos_only && stride > 1 && ctx->qry->ntiles >= 4 * stride) {
        // seeding: the k-th largest candidate key over every stride-th tile
        // is reached by at least k keys of the full row (with pos_only,
        // a row with fewer than k sampled candidates gets seed 0: no limit);
        // the full pass then buffers only keys >= that seed
        ColArgs S = ctx->qry->args();
        S.tstride = stride;
        S.ntiles = (ctx->qry->ntiles + stride - 1) / stride;
        // about kSampleTiles sampled tiles per wave when the launch has row
        // groups enough to fill the device (each wave's buffer then keeps
        // its own k largest of several tiles, so the seed's selection reads
        // a few keys per sampled tile instead of nearly every pair); at least
        // kSampleWorkgroups workgroups otherwise; at least one tile per wave
        const int64_t per = 4 * gt * kSampleTiles;
        const int64_t most = std::min<int64_t>(grid.y, (S.ntiles + 4 * gt - 1) / (4 * gt));
        const int64_t want = std::max<int64_t>((S.ntiles + per - 1) / per,
                                               (kSampleWorkgroups + grid.x - 1) / std::max<int64_t>(grid.x, 1));
        const unsigned gys = (unsigned)std::max<int64_t>(1, std::min<int64_t>(most, want));
        const int64_t nws = (int64_t)gys * (kBlock / kWave) * gt;
        TRY(with_tables(gt, [&](auto gc) {
            return launch_bits(ctx, bp, gys, rowbase, rowlimit,
                               EpiTopk<decltype(gc)::value>{k, cap, d_buf, d_cnt, rowbase, pos_only, nullptr},
                               "k_bits<topk sample>", &S);
        }));
        const int64_t nw_full = nw;
        nw = nws;
        const int rc = select(nullptr, nullptr, d_seed);
        nw = nw_full;
        TRY(rc);
        seed = d_seed;
    }
    if (bp.use) {
        TRY(with_tables(gt, [&](auto gc) {
            return launch_bits(ctx, bp, grid.y, rowbase, rowlimit,
                               EpiTopk<decltype(gc)::value>{k, cap, d_buf, d_cnt, rowbase, pos_only, seed},
                               "k_bits<topk>", nullptr, d_queue);
        }));
        if (!mrows.empty()) {
            k_topk<<<dim3((unsigned)mrows.size(), grid.y * gt), dim3(kBlock), lds, ctx->stream>>>(
                ctx->ref.args(), ctx->qry->args(), d_rowlist, rowbase, k, cap, d_buf, d_cnt);
            TRY(launch_check("k_topk<rowlist>"));
        }
    } else {
        k_topk<<<grid, dim3(kBlock), lds, ctx->stream>>>(ctx->ref.args(), ctx->qry->args(), nullptr, rowbase, k, cap,
                                                         d_buf, d_cnt);
        TRY(launch_check("k_topk"));
    }
    tr.select();
    TRY(select(mem ? mem->idx : ctx->res.as<int32_t>(), mem ? mem->t : ctx->res2.as<float>(), nullptr,
               seed != nullptr));
    tr.done();
    if (gr.on) {
        // the capture ends before the timing event is recorded
        TRY(gr.end());
        TimedRegion tg(ctx, nr * nq, mem == nullptr);
        TRY(gr.launch());
        tg.done();
    }
    ctx->last_kernel = bp.use ? "topk_bits" : "topk_merge";
    return SIML_OK;
}

static int topk_async_impl(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, int32_t k) {
    return topk_impl(ctx, rowbase, rowlimit, k, nullptr);
}

// The 1000-bin histogram of rows [rb, rl) on the sparse path (every row
// grouped with at most 255 levels, magnitudes at least the count sums),
// column ranges split as the sparse top-k's; into the handle's result.
static int sparse_hist1000(siml_ctx* ctx, int64_t rb, int64_t rl, unsigned long long* hist) {
    const int64_t nr = rl - rb, nq = ctx->qry->n;
    const ColSet::Postings& X = ctx->qry->post;
    const PostingQuery P = X.query();
    const SparseTopkShape shp = sparse_topk_shape(nr, nq, 1, false);
    GraphRegion gr(ctx);
    TRY(gr.begin());
    k_fill<unsigned long long><<<dim3(4), dim3(256), 0, ctx->stream>>>(hist, 1000, 0ull);
    TRY(launch_check("k_fill"));
    TimedRegion tr(ctx, nr * nq, !gr.on);
    const dim3 grid((unsigned)nr, (unsigned)shp.nsplit);
    auto go = [&](auto kern, const void* mags) {
        kern<<<grid, dim3(kSparseBlock), 0, ctx->stream>>>(ctx->ref.args(), rb, P, mags, nq, shp.cps, hist);
    };
    const void* m8 = X.cmag8.p;
    const void* m16 = X.cmag16.p;
    if (X.narrow && X.mag_bytes == 1)
        go(k_sparse_hist1000<true, 1>, m8);
    else if (X.narrow)
        go(k_sparse_hist1000<true, 2>, m16);
    else if (X.mag_bytes == 1)
        go(k_sparse_hist1000<false, 1>, m8);
    else
        go(k_sparse_hist1000<false, 2>, m16);
    TRY(launch_check("k_sparse_hist1000"));
    tr.done();
    if (gr.on) {
        TRY(gr.end());
        TimedRegion tg(ctx, nr * nq);
        TRY(gr.launch());
        tg.done();
    }
    ctx->last_kernel = "hist1000_sparse";
    return SIML_OK;
}

static int hist1000_async_impl(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit) {
    TRY(check_rows(ctx, rowbase, rowlimit, "getGlobalHistogram"));
    TRY(set_device(ctx));
    clear_result(ctx);
    const int64_t nr = rowlimit - rowbase, nq = ctx->qry->n;
    TRY(ctx->res.reserve(sizeof(uint64_t) * 1000));
    ctx->kind = SIML_RESULT_I64;
    ctx->rows = 1;
    ctx->cols = 1000;
    unsigned long long* hist = ctx->res.as<unsigned long long>();
    if (nr == 0 || nq == 0) {
        k_fill<unsigned long long><<<dim3(4), dim3(256), 0, ctx->stream>>>(hist, 1000, 0ull);
        return launch_check("k_fill");
    }
    if (ctx->ref.mags_small && ctx->qry->mags_small && nq <= 0x7FFFFFFF && nr <= 0x7FFFFFFF) {
        // the sparse path when every row can take it and the model (or the
        // policy) prefers it
        std::vector<int32_t> srows, smerge;
        int prc = SIML_OK;
        const bool sp = plan_sparse(ctx, rowbase, rowlimit, srows, smerge, &prc, true);
        TRY(prc);
        if (sp && smerge.empty() && ctx->qry->post.mag_bytes != 0) return sparse_hist1000(ctx, rowbase, rowlimit, hist);
    }
    BitPlan bp;
    TRY(plan_bits(ctx, rowbase, rowlimit, bp));
    const std::vector<int32_t>& mrows = bp.merge;
    if (!mrows.empty()) TRY(upload_rowlist(ctx, mrows));
    if (bp.use && ctx->persistent) TRY(ctx->scratch.reserve(256));  // persistent grid's queue counters
    // the bins are cleared by a kernel, which a graph may hold
    GraphRegion gr(ctx);
    TRY(gr.begin());
    k_fill<unsigned long long><<<dim3(4), dim3(256), 0, ctx->stream>>>(hist, 1000, 0ull);
    TRY(launch_check("k_fill"));
    const size_t lds = row_lds_bytes(max_len_in(ctx->ref, rowbase, rowlimit));
    const int gt = bp.use ? bits_tables(ctx, bp) : 1;
    dim3 grid = row_grid(bp.use ? std::max<int64_t>(((int64_t)bp.groups.size() + gt - 1) / gt, 1) : nr,
                         ctx->qry->ntiles, bp.use ? ctx->target_wg : 2048);
    // one workgroup counts at most 32 gt rows x its columns in uint32 LDS bins
    const int64_t min_gy = (siml::kBitRows * gt * kWave * ctx->qry->ntiles) / (int64_t(1) << 31) + 1;
    grid.y = (unsigned)std::min<int64_t>(65535, std::max<int64_t>(grid.y, min_gy));
    if (bp.use) grid.y = std::max<unsigned>(xcd_gy(grid.y, 65535), (unsigned)std::min<int64_t>(min_gy, 65535));
    int32_t* d_queue = nullptr;
    if (bp.use && ctx->persistent) {
        grid.y = std::min<unsigned>(std::max<unsigned>((grid.y + kXcds - 1) / kXcds * kXcds, kXcds),
                                    65535 / kXcds * kXcds);
        d_queue = ctx->scratch.as<int32_t>();
    }
    TimedRegion tr(ctx, nr * nq, !gr.on);
    if (bp.use) {
        TRY(with_tables(gt, [&](auto gc) {
            return launch_bits(ctx, bp, grid.y, rowbase, rowlimit, EpiHist1000<decltype(gc)::value>{hist},
                               "k_bits<hist1000>", nullptr, d_queue);
        }));
        if (!mrows.empty()) {
            k_hist1000<<<dim3((unsigned)mrows.size(), grid.y), dim3(kBlock), lds, ctx->stream>>>(
                ctx->ref.args(), ctx->qry->args(), ctx->rowlist.as<int32_t>(), rowbase, hist);
            TRY(launch_check("k_hist1000<rowlist>"));
        }
    } else {
        k_hist1000<<<grid, dim3(kBlock), lds, ctx->stream>>>(ctx->ref.args(), ctx->qry->args(), nullptr, rowbase, hist);
        TRY(launch_check("k_hist1000"));
    }
    tr.done();
    if (gr.on) {
        TRY(gr.end());
        TimedRegion tg(ctx, nr * nq);
        TRY(gr.launch());
        tg.done();
    }
    ctx->last_kernel = bp.use ? "hist1000_bits" : "hist1000_merge";
    return SIML_OK;
}

// ---------------------------------------------------------------------------
// Launch-shape autotuner.  With siml_set_autotune(ctx, 1), the first call of
// an operation for a shape bucket (operation, log2 of the rows, log2 of the
// query tiles, k) runs the operation once per candidate shape, timed with
// events on the handle's stream, and keeps the fastest for the bucket (shared
// by every handle of the process on that device).  Results do not depend on
// the shape, so the last run's result is the call's result.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

enum TuneOp { kTuneMultirow = 1, kTuneHistRows, kTuneNeighbors, kTuneTopk, kTuneHist1000 };

struct TuneShape {
    int64_t target_wg;
    unsigned topk_wave_cap;
    int topk_seed;  // sample stride, 0: no seeding, 1: from the query set's size
    bool persistent = false;
};

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int>, TuneShape> g_tune;

// Candidate launch shapes: the first 6 apply to every operation, the next
// 7 to top-k only, then 2 persistent-grid shapes for top-k and 2 for the
// 1000-bin histogram.
const TuneShape kCands[] = {{8192, 256, 0},     {2048, 256, 0},      {4096, 256, 0},      {16384, 256, 0},
                            {32768, 1024, 0},   {65536, 1024, 0},    {8192, 1024, 1},     {2048, 1024, 1},
                            {8192, 256, 16},    {2048, 256, 16},     {32768, 1024, 16},   {8192, 256, 4},
                            {8192, 64, 0},      {8192, 1024, 1, true}, {32768, 1024, 1, true},
                            {8192, 256, 0, true}, {32768, 256, 0, true}};
constexpr int kNumCands = (int)(sizeof(kCands) / sizeof(kCands[0]));
bool tune_candidate(int op, int i) {
    if (i < 6) return true;
    if (i < 15) return op == kTuneTopk;
    return op == kTuneHist1000;
}

int log2_bucket(int64_t x) {
    int b = 0;
    while (x > 1) {
        x >>= 1;
        ++b;
    }
    return b;
}

template <class F>
int tuned_call(siml_ctx* ctx, int op, int64_t rows, int k, F&& impl) {
    if (!ctx->autotune || ctx->qry->n < 0) return impl();
    const auto key = std::make_tuple(ctx->device, op, log2_bucket(rows), log2_bucket(ctx->qry->ntiles), k);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            ctx->target_wg = it->second.target_wg;
            ctx->topk_wave_cap = it->second.topk_wave_cap;
            ctx->topk_seed = it->second.topk_seed;
            ctx->persistent = it->second.persistent;
            return impl();
        }
    }
    TRY(set_device(ctx));
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(SIML_EDEVICE, "hipEventCreate failed");
    }
    TuneShape best = kCands[0];
    float best_ms = 0.f;
    int rc = SIML_OK;
    bool first = true;
    for (int i = 0; i < kNumCands && rc == SIML_OK; ++i) {
        if (!tune_candidate(op, i)) continue;
        ctx->target_wg = kCands[i].target_wg;
        ctx->topk_wave_cap = kCands[i].topk_wave_cap;
        ctx->topk_seed = kCands[i].topk_seed;
        ctx->persistent = kCands[i].persistent;
        (void)hipEventRecord(a, ctx->stream);
        rc = impl();
        if (rc != SIML_OK) break;
        (void)hipEventRecord(b, ctx->stream);
        float ms = 0.f;
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
            rc = fail(SIML_EDEVICE, "timing a candidate launch shape failed");
            break;
        }
        if (first || ms < best_ms) {
            best_ms = ms;
            best = kCands[i];
            first = false;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != SIML_OK) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = best;
    // later calls of the bucket use `best`; this call's result (the last
    // candidate's) is the same
    ctx->target_wg = best.target_wg;
    ctx->topk_wave_cap = best.topk_wave_cap;
    ctx->topk_seed = best.topk_seed;
    ctx->persistent = best.persistent;
    return SIML_OK;
}

}  // namespace

extern "C" {

int siml_multirow_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    return tuned_call(ctx, kTuneMultirow, rowlimit - rowbase, 0, [&] { return multirow_async_impl(ctx, rowbase, rowlimit); });
}

int siml_hist_rows_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    return tuned_call(ctx, kTuneHistRows, rowlimit - rowbase, 0, [&] { return hist_rows_async_impl(ctx, rowbase, rowlimit); });
}

int siml_neighbors_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, float lo, float hi) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    return tuned_call(ctx, kTuneNeighbors, rowlimit - rowbase, 0,
                      [&] { return neighbors_async_impl(ctx, rowbase, rowlimit, lo, hi); });
}

int siml_topk_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, int32_t k) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    return tuned_call(ctx, kTuneTopk, rowlimit - rowbase, k, [&] { return topk_async_impl(ctx, rowbase, rowlimit, k); });
}

int siml_topk_noalloc(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, int32_t k, int32_t* idx, float* t) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (!idx || !t) return fail(SIML_EINVAL, "siml_topk_noalloc: null result array");
    TRY(set_device(ctx));
    char* scratch = nullptr;
    TRY(noalloc_scratch_addr(ctx->device, &scratch));
    const CallerMem mem{scratch, SIML_NOALLOC_SCRATCH_BYTES, idx, t};
    return topk_impl(ctx, rowbase, rowlimit, k, &mem);
}

int siml_hist1000_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    return tuned_call(ctx, kTuneHist1000, rowlimit - rowbase, 0, [&] { return hist1000_async_impl(ctx, rowbase, rowlimit); });
}

int siml_set_autotune(siml_ctx* ctx, int on) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    ctx->autotune = on != 0;
    if (!ctx->autotune) {
        ctx->target_wg = 8192;
        ctx->topk_wave_cap = 1024;
        ctx->topk_seed = 1;
        ctx->persistent = false;
    }
    return SIML_OK;
}

int siml_set_persistent(siml_ctx* ctx, int on) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    ctx->autotune = false;
    ctx->persistent = on != 0;
    return SIML_OK;
}

int siml_set_graphs(siml_ctx* ctx, int on) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    ctx->graphs = on != 0;
    return SIML_OK;
}

int siml_graph_stats(siml_ctx* ctx, int64_t* launches, int64_t* builds) {
    if (!ctx || !launches || !builds) return fail(SIML_EINVAL, "null pointer");
    *launches = ctx->graph_launches;
    *builds = ctx->graph_builds;
    return SIML_OK;
}

int siml_launch_shape(siml_ctx* ctx, int64_t* target_wg, int32_t* topk_wave_cap, int32_t* topk_seed) {
    if (!ctx || !target_wg || !topk_wave_cap || !topk_seed) return fail(SIML_EINVAL, "null pointer");
    *target_wg = ctx->target_wg;
    *topk_wave_cap = (int32_t)ctx->topk_wave_cap;
    *topk_seed = ctx->topk_seed;
    return SIML_OK;
}

int siml_set_group_tables(siml_ctx* ctx, int max_tables) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (max_tables != 1 && max_tables != 2 && max_tables != 4)
        return fail(SIML_EINVAL, "tables per workgroup must be 1, 2 or 4, not %d", max_tables);
    ctx->max_tables = max_tables;
    return SIML_OK;
}

int siml_set_launch_shape(siml_ctx* ctx, int64_t target_wg, int32_t topk_wave_cap, int32_t topk_seed) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (target_wg < 1 || topk_wave_cap < 1) return fail(SIML_EINVAL, "launch shape values must be positive");
    ctx->autotune = false;
    ctx->target_wg = target_wg;
    ctx->topk_wave_cap = (unsigned)topk_wave_cap;
    if (topk_seed < 0 || topk_seed > 64) return fail(SIML_EINVAL, "top-k seed stride %d outside [0, 64]", topk_seed);
    ctx->topk_seed = topk_seed;
    return SIML_OK;
}

int siml_call_bytes(siml_ctx* ctx, int op, int64_t rows, int32_t k, int64_t* bytes) {
    if (!ctx || !bytes) return fail(SIML_EINVAL, "null pointer");
    if (ctx->qry->n < 0) return fail(SIML_ESTATE, "siml_call_bytes: no query set");
    if (rows < 0) return fail(SIML_EINVAL, "siml_call_bytes: negative row count");
    if (op == SIML_OP_TOPK && (k < 1 || k > SIML_MAX_TOPK))
        return fail(SIML_EINVAL, "siml_call_bytes: k=%d outside [1, %d]", k, SIML_MAX_TOPK);
    if (op != SIML_OP_TOPK && op != SIML_OP_NEIGHBORS && op != SIML_OP_HIST1000)
        return fail(SIML_EINVAL, "siml_call_bytes: unknown operation %d", op);
    // launch shapes the call may take: the handle's, or every candidate of
    // the autotuner (a bucket's first call runs them all)
    std::vector<TuneShape> shapes{{ctx->target_wg, ctx->topk_wave_cap, ctx->topk_seed, ctx->persistent}};
    if (ctx->autotune) {
        const int top = op == SIML_OP_TOPK ? kTuneTopk : op == SIML_OP_HIST1000 ? kTuneHist1000 : kTuneNeighbors;
        for (int i = 0; i < kNumCands; ++i)
            if (tune_candidate(top, i)) shapes.push_back(kCands[i]);
    }
    // plans: the fewest row groups the rows can form (the most column
    // ranges per group), with each table count allowed, or every row on the
    // merge kernel; scratch only shrinks as groups grow
    const int64_t gmin = (rows + siml::kBitRows - 1) / siml::kBitRows;
    std::vector<std::pair<bool, int>> plans{{false, 1}, {true, 1}};
    if (ctx->max_tables >= 2 && gmin >= 2) plans.emplace_back(true, 2);
    if (ctx->max_tables >= 4 && gmin >= 8) plans.emplace_back(true, 4);
    size_t most = 0;
    for (const auto& sh : shapes)
        for (const auto& pl : plans) {
            size_t b = 0;
            if (op == SIML_OP_TOPK)
                b = std::max(topk_shape(ctx, rows, pl.first, gmin, pl.second, k, sh.target_wg, sh.topk_wave_cap,
                                        kTopkBufferBudget, sh.persistent)
                                 .scratch(),
                             sparse_topk_shape(rows, ctx->qry->n, k).scratch());
            else if (op == SIML_OP_NEIGHBORS)
                b = nbr_shape(ctx, rows, pl.first, gmin, 1, true, sh.target_wg).scratch();
            else if (sh.persistent)
                b = 256;  // the persistent grid's queue counters
            most = std::max(most, b);
        }
    // results on the device and the per-call lists (row list, group list)
    size_t res = sizeof(uint64_t) * 1000;
    if (op == SIML_OP_TOPK) res = (sizeof(int32_t) + sizeof(float)) * (size_t)rows * (size_t)k;
    if (op == SIML_OP_NEIGHBORS) res = sizeof(int64_t) * ((size_t)rows + 1);
    *bytes = (int64_t)(most + res + sizeof(int32_t) * ((size_t)rows + (size_t)gmin + kMaxTables) + 256);
    return SIML_OK;
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

// Rows [0, rows) of a device matrix (row pitch dpitch elements) into host
// memory with row pitch hpitch, cols elements per row; waits for the copy.
// Equal pitches take one linear copy (the last row stops at cols).  Unequal
// pitches go through a host staging buffer in slabs of rows: one linear
// copy per slab, then a row-by-row host copy.  (A strided device-to-host
// copy into pageable memory runs at a fraction of the linear rate.)

// Rows [r0, r0 + n) of a staged slab (row pitch dpitch) into host memory
// (row pitch hpitch), cols elements each, split over host threads.
static void scatter_rows(void* host, int64_t hpitch, const char* stage, int64_t dpitch, int64_t r0, int64_t n,
                         int64_t cols, size_t esz, int nthreads) {
    auto work = [&](int64_t a, int64_t b) {
        if (hpitch == dpitch) {
            const size_t bytes_a = esz * (size_t)(a * dpitch), bytes_b = esz * (size_t)((b - 1) * dpitch + cols);
            std::memcpy(static_cast<char*>(host) + esz * (size_t)(r0 * hpitch) + bytes_a, stage + bytes_a,
                        bytes_b - bytes_a);
            return;
        }
        for (int64_t r = a; r < b; ++r)
            std::memcpy(static_cast<char*>(host) + esz * (size_t)((r0 + r) * hpitch), stage + esz * (size_t)(r * dpitch),
                        esz * (size_t)cols);
    };
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, n));
    if (nt == 1) {
        work(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, n * t / nt, n * (t + 1) / nt);
    work(0, n / nt);
    for (auto& x : th) x.join();
}

static int copy_rows_d2h(siml_ctx* ctx, void* host, int64_t hpitch, const void* dev, int64_t dpitch, int64_t rows,
                         int64_t cols, size_t esz) {
    if (rows <= 0 || cols <= 0) return SIML_OK;
    const size_t row_bytes = esz * (size_t)dpitch;
    const size_t total = esz * (size_t)((rows - 1) * dpitch + cols);
    if (total >= (size_t(32) << 20)) {
        // large results: device -> pinned slab (at PCIe rate) while host
        // threads copy the previous slab into the caller's memory, which
        // also spreads its first-touch page faults over the threads
        const int64_t slab = std::max<int64_t>(1, (int64_t)((size_t(64) << 20) / row_bytes));
        const size_t slab_bytes = row_bytes * (size_t)std::min<int64_t>(slab, rows);
        TRY(ctx->pinned[0].reserve(slab_bytes));
        TRY(ctx->pinned[1].reserve(slab_bytes));
        hipEvent_t ev[2] = {nullptr, nullptr};
        HIP_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        if (hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
            (void)hipEventDestroy(ev[0]);
            return fail(SIML_EDEVICE, "hipEventCreate failed");
        }
        const int nthreads = copy_threads();
        const int64_t nslabs = (rows + slab - 1) / slab;
        auto issue = [&](int64_t s) -> hipError_t {
            const int64_t r0 = s * slab, n = std::min<int64_t>(slab, rows - r0);
            hipError_t e = hipMemcpyAsync(ctx->pinned[s % 2].p, static_cast<const char*>(dev) + row_bytes * (size_t)r0,
                                          esz * (size_t)((n - 1) * dpitch + cols), hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(ev[s % 2], ctx->stream);
            return e;
        };
        hipError_t e = issue(0);
        for (int64_t s = 0; s < nslabs && e == hipSuccess; ++s) {
            if (s + 1 < nslabs) e = issue(s + 1);
            if (e == hipSuccess) e = hipEventSynchronize(ev[s % 2]);
            if (e == hipSuccess) {
                const int64_t r0 = s * slab, n = std::min<int64_t>(slab, rows - r0);
                scatter_rows(host, hpitch, static_cast<const char*>(ctx->pinned[s % 2].p), dpitch, r0, n, cols, esz,
                             nthreads);
            }
        }
        if (e != hipSuccess) (void)hipStreamSynchronize(ctx->stream);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        if (e != hipSuccess) return fail(SIML_EDEVICE, "result copy failed: %s", hipGetErrorString(e));
        return SIML_OK;
    }
    if (hpitch == dpitch) {
        HIP_TRY(hipMemcpyAsync(host, dev, esz * (size_t)((rows - 1) * dpitch + cols), hipMemcpyDeviceToHost,
                               ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return SIML_OK;
    }
    const int64_t slab = std::max<int64_t>(1, (int64_t)((size_t(64) << 20) / row_bytes));
    std::vector<char>& stage = ctx->h_stage;
    stage.resize(row_bytes * (size_t)std::min<int64_t>(slab, rows));
    for (int64_t r0 = 0; r0 < rows; r0 += slab) {
        const int64_t n = std::min<int64_t>(slab, rows - r0);
        HIP_TRY(hipMemcpyAsync(stage.data(), static_cast<const char*>(dev) + row_bytes * (size_t)r0,
                               esz * (size_t)((n - 1) * dpitch + cols), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (int64_t r = 0; r < n; ++r)
            std::memcpy(static_cast<char*>(host) + esz * (size_t)((r0 + r) * hpitch), stage.data() + row_bytes * (size_t)r,
                        esz * (size_t)cols);
    }
    return SIML_OK;
}

// n contiguous elements of esz bytes from the device: the bulk as rows of
// 16384 elements through copy_rows_d2h (pinned slabs when large), the tail
// with one copy.
static int copy_d2h(siml_ctx* ctx, void* host, const void* dev, int64_t n, size_t esz) {
    constexpr int64_t W = 16384;
    const int64_t rows = n / W, done = rows * W;
    if (rows > 0) TRY(copy_rows_d2h(ctx, host, W, dev, W, rows, W, esz));
    if (n > done) {
        HIP_TRY(hipMemcpyAsync(static_cast<char*>(host) + esz * (size_t)done,
                               static_cast<const char*>(dev) + esz * (size_t)done, esz * (size_t)(n - done),
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return SIML_OK;
}

static int retrieve_matrix(siml_ctx* ctx, void* host, int64_t pitch, int want, size_t esz) {
    if (ctx->kind != want) return fail(SIML_ESTATE, "no pending result of this kind");
    if (pitch < ctx->cols) return fail(SIML_EINVAL, "host pitch %lld < columns %lld", (long long)pitch, (long long)ctx->cols);
    TRY(set_device(ctx));
    if (ctx->rows > 0 && ctx->cols > 0) {
        if (!host) return fail(SIML_EINVAL, "null host buffer");
        TRY(copy_rows_d2h(ctx, host, pitch, ctx->res.p, ctx->dpitch, ctx->rows, ctx->cols, esz));
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
    if (r1 > r0 && ctx->cols > 0) {
        if (!host) return fail(SIML_EINVAL, "null host buffer");
        TRY(copy_rows_d2h(ctx, host, pitch, ctx->res.as<float>() + r0 * ctx->dpitch, ctx->dpitch, r1 - r0, ctx->cols,
                          sizeof(float)));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SIML_OK;
}

int siml_retrieve_csr(siml_ctx* ctx, int64_t* offsets, int32_t* indices) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (ctx->kind != SIML_RESULT_CSR) return fail(SIML_ESTATE, "no pending neighbour result");
    if (!offsets) return fail(SIML_EINVAL, "null offsets buffer");
    TRY(set_device(ctx));
    if (ctx->csr_on_device)
        TRY(copy_d2h(ctx, offsets, ctx->csr_dev.p, ctx->rows + 1, sizeof(int64_t)));
    else
        std::fill(offsets, offsets + ctx->rows + 1, int64_t(0));
    if (ctx->nnz > 0) {
        if (!indices) return fail(SIML_EINVAL, "null indices buffer");
        TRY(copy_d2h(ctx, indices, ctx->res.p, ctx->nnz, sizeof(int32_t)));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    clear_result(ctx);
    return SIML_OK;
}

int siml_retrieve_topk(siml_ctx* ctx, int32_t* idx, float* t) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (ctx->kind != SIML_RESULT_TOPK) return fail(SIML_ESTATE, "no pending top-k result");
    TRY(set_device(ctx));
    const size_t n = (size_t)(ctx->rows * ctx->cols);
    if (n) {
        if (!idx || !t) return fail(SIML_EINVAL, "null output buffer");
        TRY(copy_d2h(ctx, idx, ctx->res.p, (int64_t)n, sizeof(int32_t)));
        TRY(copy_d2h(ctx, t, ctx->res2.p, (int64_t)n, sizeof(float)));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    clear_result(ctx);
    return SIML_OK;
}

int siml_retrieve_i64(siml_ctx* ctx, int64_t* host) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (ctx->kind != SIML_RESULT_I64) return fail(SIML_ESTATE, "no pending int64 result");
    if (!host) return fail(SIML_EINVAL, "null output buffer");
    TRY(set_device(ctx));
    HIP_TRY(hipMemcpyAsync(host, ctx->res.p, sizeof(int64_t) * (size_t)ctx->cols, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    clear_result(ctx);
    return SIML_OK;
}

int siml_set_policy(siml_ctx* ctx, int policy) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (policy != SIML_POLICY_AUTO && policy != SIML_POLICY_MERGE && policy != SIML_POLICY_BITS &&
        policy != SIML_POLICY_SPARSE)
        return fail(SIML_EINVAL, "unknown policy %d", policy);
    ctx->policy = policy;
    return SIML_OK;
}

int siml_set_layout(siml_ctx* ctx, int layout) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (layout != SIML_LAYOUT_NATURAL && layout != SIML_LAYOUT_BLOCK && layout != SIML_LAYOUT_SORTED) return fail(SIML_EINVAL, "unknown layout %d", layout);
    ctx->layout = layout;
    return SIML_OK;
}

int siml_set_timing(siml_ctx* ctx, int on) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    if (on && !ctx->work.p) {
        TRY(set_device(ctx));
        TRY(ctx->work.reserve(2 * sizeof(unsigned long long)));
        HIP_TRY(hipMemset(ctx->work.p, 0, 2 * sizeof(unsigned long long)));
        ctx->work_base[0] = ctx->work_base[1] = 0;
    }
    ctx->timing = on != 0;
    return SIML_OK;
}

int siml_work_stats(siml_ctx* ctx, int64_t* pairs, int64_t* merge_len, int reset) {
    if (!ctx) return fail(SIML_EINVAL, "null handle");
    unsigned long long v[2] = {0, 0};
    if (ctx->work.p) {
        TRY(set_device(ctx));
        HIP_TRY(hipMemcpyAsync(v, ctx->work.p, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    if (pairs) *pairs = (int64_t)(v[0] - ctx->work_base[0]);
    if (merge_len) *merge_len = (int64_t)(v[1] - ctx->work_base[1]);
    if (reset) {
        ctx->work_base[0] = v[0];
        ctx->work_base[1] = v[1];
    }
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
        ctx->select_ms = 0.0;
        ctx->launches = 0;
        ctx->pairs = 0;
    }
    return SIML_OK;
}

int siml_kernel_split(siml_ctx* ctx, double* similarity_ms, double* selection_ms) {
    if (!ctx || !similarity_ms || !selection_ms) return fail(SIML_EINVAL, "null pointer");
    TRY(set_device(ctx));
    TRY(collect_timing(ctx));
    *similarity_ms = ctx->kernel_ms - ctx->select_ms;
    *selection_ms = ctx->select_ms;
    return SIML_OK;
}

int siml_row_groups(siml_ctx* ctx, int64_t* groups, int64_t* merge_rows) {
    if (!ctx || !groups || !merge_rows) return fail(SIML_EINVAL, "null pointer");
    const RowSet::BitGrouping& B = ctx->ref.bits[0];
    *groups = B.built ? (int64_t)B.h_hdr.size() : 0;
    *merge_rows = B.built ? (int64_t)B.merge_rows.size() : 0;
    return SIML_OK;
}

const char* siml_last_kernel(siml_ctx* ctx) { return ctx ? ctx->last_kernel.c_str() : "none"; }

// ---------------------------------------------------------------------------
// Stateless entry points
// ---------------------------------------------------------------------------

// One handle per device serves the stateless entry points.  ctypes releases
// the GIL around C calls (the reference ran these calls under it), so calls
// from several host threads can overlap: each call holds its device's lock
// from set-up to the copy of its result, and the handle's buffers are never
// resized while another call's kernels or copies use them.
// The molecules of a padded set as the kernels read them: per row its
// length, magnitude and first `length` (id, count) entries.
struct SetSnapshot {
    bool valid = false;
    std::vector<int32_t> len, mag, ids, counts;
    // whether (lingos, counts, mags, lengths) holds exactly these molecules
    bool same(const int32_t* L, const int32_t* C, const int32_t* M, const int32_t* N, int64_t n, int64_t width,
              int64_t pitch) const {
        if (!valid || n != (int64_t)len.size() || !L || !C || !M || !N || pitch < width) return false;
        size_t k = 0;
        for (int64_t r = 0; r < n; ++r) {
            const int32_t l = N[r];
            if (l != len[(size_t)r] || M[r] != mag[(size_t)r] || l < 0 || l > width) return false;
            if (std::memcmp(L + r * pitch, ids.data() + k, sizeof(int32_t) * (size_t)l) != 0 ||
                std::memcmp(C + r * pitch, counts.data() + k, sizeof(int32_t) * (size_t)l) != 0)
                return false;
            k += (size_t)l;
        }
        return true;
    }
    void take(const int32_t* L, const int32_t* C, const int32_t* M, const int32_t* N, int64_t n, int64_t pitch) {
        len.assign(N, N + n);
        mag.assign(M, M + n);
        ids.clear();
        counts.clear();
        for (int64_t r = 0; r < n; ++r) {
            ids.insert(ids.end(), L + r * pitch, L + r * pitch + N[r]);
            counts.insert(counts.end(), C + r * pitch, C + r * pitch + N[r]);
        }
        valid = true;
    }
};

struct SharedCtx {
    std::mutex mu;
    siml_ctx* ctx = nullptr;
    // the query set last uploaded by siml_tanimoto_block: the reference's
    // callers pass the same query set with every block of rows
    // (CPULingo.py:126-138), so it is rebuilt only when its contents change
    SetSnapshot query;
};
static std::mutex g_shared_mu;
static std::map<int, std::unique_ptr<SharedCtx>> g_shared;

// Locks device `device`'s shared handle for the caller's lifetime of `lk`.
static int shared_ctx(int device, std::unique_lock<std::mutex>& lk, siml_ctx** out, SharedCtx** shared = nullptr) {
    if (device < 0) return fail(SIML_EINVAL, "negative device");
    SharedCtx* sc = nullptr;
    {
        std::lock_guard<std::mutex> g(g_shared_mu);
        std::unique_ptr<SharedCtx>& slot = g_shared[device];
        if (!slot) slot.reset(new SharedCtx());
        sc = slot.get();
    }
    lk = std::unique_lock<std::mutex>(sc->mu);
    if (!sc->ctx) TRY(siml_create(device, &sc->ctx));
    *out = sc->ctx;
    if (shared) *shared = sc;
    return SIML_OK;
}

int siml_tanimoto_block(int device, const int32_t* rlingos, const int32_t* rcounts, const int32_t* rmags,
                        const int32_t* rlengths, int64_t nr, int64_t rwidth, int64_t rpitch,
                        const int32_t* qlingos, const int32_t* qcounts, const int32_t* qmags,
                        const int32_t* qlengths, int64_t nq, int64_t qwidth, int64_t qpitch, float* out,
                        int64_t opitch) {
    if (!rmags || !qmags) return fail(SIML_EINVAL, "tanimoto_block: magnitude vectors are required");
    siml_ctx* ctx = nullptr;
    SharedCtx* sc = nullptr;
    std::unique_lock<std::mutex> lk;
    TRY(shared_ctx(device, lk, &ctx, &sc));
    TRY(siml_set_ref(ctx, rlingos, rcounts, rmags, rlengths, nr, rwidth, rpitch));
    if (!sc->query.same(qlingos, qcounts, qmags, qlengths, nq, qwidth, qpitch)) {
        sc->query.valid = false;
        TRY(siml_set_query(ctx, qlingos, qcounts, qmags, qlengths, nq, qwidth, qpitch));
        sc->query.take(qlingos, qcounts, qmags, qlengths, nq, qpitch);
    }
    TRY(siml_multirow_async(ctx, 0, nr));
    return siml_retrieve_f32(ctx, out, opitch);
}

static int row_hist_common(int device, const void* data, size_t esz, int64_t rows, int64_t cols, int64_t pitch,
                           uint32_t* hist) {
    if (rows < 0 || cols < 0 || pitch < cols) return fail(SIML_EINVAL, "rowHistogram101: bad shape or pitch");
    if (rows == 0) return SIML_OK;
    if (!data || !hist) return fail(SIML_EINVAL, "rowHistogram101: null buffer");
    siml_ctx* ctx = nullptr;
    std::unique_lock<std::mutex> lk;
    TRY(shared_ctx(device, lk, &ctx));
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

int siml_selftest_valu(int device, int op, double* lane_ops_per_s) {
    if (!lane_ops_per_s) return fail(SIML_EINVAL, "null pointer");
    if (op != 0 && op != 1) return fail(SIML_EINVAL, "unknown operation %d", op);
    siml_ctx* ctx = nullptr;
    std::unique_lock<std::mutex> lk;
    TRY(shared_ctx(device, lk, &ctx));
    TRY(set_device(ctx));
    const int blocks = 256 * 4 * 8;  // 8 workgroups of 4 waves per CU
    const int iters = 2048;
    TRY(ctx->scratch.reserve(sizeof(uint32_t) * (size_t)blocks * kBlock));
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(SIML_EDEVICE, "hipEventCreate failed");
    }
    float ms = 0.f;
    int rc = SIML_OK;
    for (int rep = 0; rep < 2 && rc == SIML_OK; ++rep) {  // the first launch warms up
        (void)hipEventRecord(a, ctx->stream);
        k_valu_peak<<<dim3(blocks), dim3(kBlock), 0, ctx->stream>>>(ctx->scratch.as<uint32_t>(), 12345u, iters, op);
        rc = launch_check("k_valu_peak");
        (void)hipEventRecord(b, ctx->stream);
        if (rc == SIML_OK && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess))
            rc = fail(SIML_EDEVICE, "timing k_valu_peak failed");
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    TRY(rc);
    *lane_ops_per_s = (double)blocks * kBlock * iters * 16.0 * 8.0 / (ms * 1e-3);
    return SIML_OK;
}

int siml_selftest_lds(int device, int op, double* bytes_per_s) {
    if (!bytes_per_s) return fail(SIML_EINVAL, "null pointer");
    if (op != 0 && op != 1) return fail(SIML_EINVAL, "unknown access pattern %d", op);
    siml_ctx* ctx = nullptr;
    std::unique_lock<std::mutex> lk;
    TRY(shared_ctx(device, lk, &ctx));
    TRY(set_device(ctx));
    const int blocks = 256 * 4 * 8;  // as many workgroups as 8 rounds of 4 per CU
    const int iters = 4096;
    TRY(ctx->scratch.reserve(sizeof(uint32_t) * (size_t)blocks * kBlock));
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(SIML_EDEVICE, "hipEventCreate failed");
    }
    float ms = 0.f;
    int rc = SIML_OK;
    for (int rep = 0; rep < 2 && rc == SIML_OK; ++rep) {  // the first launch warms up
        (void)hipEventRecord(a, ctx->stream);
        if (op == 0)
            k_lds_peak<0><<<dim3(blocks), dim3(kBlock), 0, ctx->stream>>>(ctx->scratch.as<uint32_t>(), iters, 1024u);
        else
            k_lds_peak<1><<<dim3(blocks), dim3(kBlock), 0, ctx->stream>>>(ctx->scratch.as<uint32_t>(), iters, 1024u);
        rc = launch_check("k_lds_peak");
        (void)hipEventRecord(b, ctx->stream);
        if (rc == SIML_OK && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess))
            rc = fail(SIML_EDEVICE, "timing k_lds_peak failed");
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    TRY(rc);
    *bytes_per_s = (double)blocks * kBlock * iters * 8.0 * 8.0 / (ms * 1e-3);
    return SIML_OK;
}

int siml_selftest_division(int device, int method, int64_t* mismatches, int64_t* first_n, int64_t* first_d) {
    if (!mismatches || !first_n || !first_d) return fail(SIML_EINVAL, "null pointer");
    if (method != 1 && method != 2) return fail(SIML_EINVAL, "unknown division method %d", method);
    siml_ctx* ctx = nullptr;
    std::unique_lock<std::mutex> lk;
    TRY(shared_ctx(device, lk, &ctx));
    TRY(set_device(ctx));
    TRY(ctx->scratch.reserve(2 * sizeof(unsigned long long)));
    unsigned long long* d = ctx->scratch.as<unsigned long long>();
    k_fill<unsigned long long><<<dim3(1), dim3(64), 0, ctx->stream>>>(d, 1, 0ull);
    k_fill<unsigned long long><<<dim3(1), dim3(64), 0, ctx->stream>>>(d + 1, 1, ~0ull);
    k_div_check<<<dim3(131072 / kBlock, 16), dim3(kBlock), 0, ctx->stream>>>(method, d, d + 1);
    TRY(launch_check("k_div_check"));
    unsigned long long h[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *mismatches = (int64_t)h[0];
    *first_n = h[0] ? (int64_t)(h[1] & 0xFFFFFFFFull) : -1;
    *first_d = h[0] ? (int64_t)(h[1] >> 32) : -1;
    return SIML_OK;
}

}  // extern "C"
