/*
 * siml_oracle.c -- CPU restatement of the pySIML 1.5 LINGO similarity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * in siml_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py load it; the product never links, imports or calls it.
 *
 * Each function names the reference file:line it restates (paths relative to
 * the ihaque/SIML source tree).  Where the reference has undefined behaviour
 * the build's definition is stated next to the function.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   oracle_normalize      hand-derived known-answer strings, one per state
 *                         transition of the state machine; the pure-Python
 *                         digit-only variant on strings with no special chars.
 *   oracle_lingo_bag /    collections.Counter over every 4-byte window of the
 *   oracle_pack           normalised string (independent algorithm); invariants
 *                         mag == sum(counts) == len-3; strictly ascending ids;
 *                         the 4096 strings of examples/demo4096.smi.
 *   oracle_tanimoto_*     fractions.Fraction brute force on tiny sets,
 *                         closed forms (T(A,A)=1, disjoint=0, zero-mag=0),
 *                         symmetry, the min/max bound, serial == OpenMP.
 *   oracle_single_row     recomputed-magnitude rule checked against the block
 *                         with magnitudes = sum(counts) and zero-length rows.
 *   oracle_hist_*         exhaustive (I,U) enumeration binned with numpy
 *                         float32 / float64 arithmetic; row sums == columns.
 *   oracle_neighbors      filter over the full oracle matrix; the min/max
 *                         prefilter never drops a pair that passes T >= lo.
 *   oracle_hist1000       numpy float32 edges k/1000 with searchsorted.
 *   oracle_topk           numpy lexsort on (-T, index).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* SMILES packer                                                             */
/* ------------------------------------------------------------------------ */

enum { ST_NORMAL, ST_CHOMP, ST_REPLACE, ST_SPECIALNUMS, ST_DONE };

static inline int is_digit(unsigned char c) { return c >= '0' && c <= '9'; }

/* Digit normalisation state machine, _ccompiler.cpp:178-264.
 * Reads at most inlen bytes of `in` (stopping at a NUL as the reference's
 * `while (*inp ...)` does) and writes the normalised characters to `out`
 * (capacity >= inlen).  Returns the normalised length. */
int64_t oracle_normalize(const char* in, int64_t inlen, char* out) {
    int state = ST_NORMAL;
    int64_t n = 0;
    for (int64_t p = 0; p < inlen && in[p] != 0 && state != ST_DONE; ++p) {
        unsigned char c = (unsigned char)in[p];
        switch (state) {
        case ST_CHOMP: /* :194-197 -- drop one character */
            state = ST_REPLACE;
            break;
        case ST_REPLACE: /* :198-221 */
            if (is_digit(c)) {
                out[n++] = '0';
            } else if (c == ' ' || c == '\n' || c == '\r') {
                state = ST_DONE;
            } else if (c == '%') {
                out[n++] = '%';
                state = ST_CHOMP;
            } else {
                out[n++] = (char)c;
                state = ST_NORMAL;
            }
            break;
        case ST_SPECIALNUMS: /* :222-228 -- copied verbatim, even blanks */
            out[n++] = (char)c;
            if (!is_digit(c)) state = ST_NORMAL;
            break;
        default: /* ST_NORMAL, :229-256 */
            if (is_digit(c)) {
                out[n++] = '0';
                state = ST_REPLACE;
            } else if (c == '%') {
                out[n++] = '%';
                state = ST_CHOMP;
            } else if (c == '+' || c == '-' || c == 'H' || c == '[') {
                out[n++] = (char)c;
                state = ST_SPECIALNUMS;
            } else if (c == ' ' || c == '\n' || c == '\r') {
                state = ST_DONE;
            } else {
                out[n++] = (char)c;
            }
            break;
        }
    }
    return n;
}

/* Sorted-unique Lingo multiset of one normalised string,
 * _ccompiler.cpp:83-165 with LingoBag::insert at :48-77.
 * Every window out[k..k+3], k = 0..n-4, is the little-endian int32
 * c0 | c1<<8 | c2<<16 | c3<<24 (the raw int32 load at :121).  Ids are kept in
 * ascending SIGNED order by insertion (:51-64).  The magnitude is n-3, or 0
 * with no Lingos when n < 4 (:90-97).  The reference's read past the zeroed
 * buffer when no terminator follows a string of length 0 mod 4 (:120-125) is
 * not reproduced: only the n-3 real windows are inserted.
 * ids/counts need capacity n-3.  Returns the number of unique Lingos. */
int32_t oracle_lingo_bag(const char* s, int64_t n, int32_t* ids, int32_t* counts, int32_t* mag) {
    if (n < 4) {
        *mag = 0;
        return 0;
    }
    int32_t total = 0;
    for (int64_t k = 0; k + 3 < n; ++k) {
        uint32_t u = (uint32_t)(unsigned char)s[k] | ((uint32_t)(unsigned char)s[k + 1] << 8) |
                     ((uint32_t)(unsigned char)s[k + 2] << 16) | ((uint32_t)(unsigned char)s[k + 3] << 24);
        int32_t lingo = (int32_t)u;
        int32_t i;
        for (i = 0; i < total; ++i)
            if (ids[i] >= lingo) break;
        if (i == total || ids[i] != lingo) {
            for (int32_t j = total; j > i; --j) {
                ids[j] = ids[j - 1];
                counts[j] = counts[j - 1];
            }
            ids[i] = lingo;
            counts[i] = 1;
            ++total;
        } else {
            counts[i]++;
        }
    }
    *mag = (int32_t)(n - 3);
    return total;
}

/* Whole-set packer, _ccompiler.cpp:276-381.  Strings are given as one byte
 * buffer with n+1 offsets.  Pass 1 (lingos == NULL) fills lengths and mags;
 * pass 2 fills zero-padded (n, width) row-major lingo and count matrices
 * (:354-366).  Returns 0, or -1 if a row does not fit in `width`. */
int oracle_pack(const char* buf, const int64_t* offs, int64_t n, int32_t* lengths, int32_t* mags,
                int32_t* lingos, int32_t* counts, int64_t width) {
    int bad = 0;
    for (int64_t r = 0; r < n; ++r) {
        int64_t len = offs[r + 1] - offs[r];
        char* norm = (char*)malloc((size_t)len + 1);
        int64_t m = oracle_normalize(buf + offs[r], len, norm);
        int32_t cap = m > 3 ? (int32_t)(m - 3) : 1;
        int32_t* ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
        int32_t* cnt = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
        int32_t mag = 0;
        int32_t L = oracle_lingo_bag(norm, m, ids, cnt, &mag);
        lengths[r] = L;
        mags[r] = mag;
        if (lingos) {
            if (L > width) {
                bad = 1;
            } else {
                memset(lingos + r * width, 0, sizeof(int32_t) * (size_t)width);
                memset(counts + r * width, 0, sizeof(int32_t) * (size_t)width);
                memcpy(lingos + r * width, ids, sizeof(int32_t) * (size_t)L);
                memcpy(counts + r * width, cnt, sizeof(int32_t) * (size_t)L);
            }
        }
        free(norm);
        free(ids);
        free(cnt);
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Tanimoto                                                                  */
/* ------------------------------------------------------------------------ */

/* multisetTanimoto3_mags, _CPULingo.c:53-71: three-way sorted merge summing
 * min(count) over equal ids, then the exact float expression
 * isct/((float)amag+bmag-isct). */
float oracle_tanimoto_pair(const int32_t* a, const int32_t* b, const int32_t* asize, const int32_t* bsize,
                           int32_t alen, int32_t blen, int32_t amag, int32_t bmag) {
    if (amag == 0 || bmag == 0) return 0.0f;
    int32_t i = 0, j = 0, isct = 0;
    while (i < alen && j < blen) {
        if (a[i] == b[j]) {
            isct += imin(asize[i], bsize[j]);
            ++i;
            ++j;
        } else if (a[i] < b[j]) {
            ++i;
        } else {
            ++j;
        }
    }
    return isct / ((float)amag + bmag - isct);
}

/* Same merge, returning the integer intersection only (for tests). */
int32_t oracle_intersection(const int32_t* a, const int32_t* b, const int32_t* asize, const int32_t* bsize,
                            int32_t alen, int32_t blen) {
    int32_t i = 0, j = 0, isct = 0;
    while (i < alen && j < blen) {
        if (a[i] == b[j]) {
            isct += imin(asize[i], bsize[j]);
            ++i;
            ++j;
        } else if (a[i] < b[j]) {
            ++i;
        } else {
            ++j;
        }
    }
    return isct;
}

/* _CPULingo_getTanimotoBlock (_CPULingo.c:239-271) and, for nthreads != 1,
 * _CPULingo_getTanimotoBlockParallel (:427-462): row-major fill of an
 * (nr, nq) float32 block, OpenMP static schedule over reference rows.
 * Pitches are in elements.  nthreads <= 0 uses the OpenMP default. */
void oracle_tanimoto_block(const int32_t* rl, const int32_t* rc, const int32_t* rmag, const int32_t* rlen,
                           int64_t nr, int64_t rpitch, const int32_t* ql, const int32_t* qc,
                           const int32_t* qmag, const int32_t* qlen, int64_t nq, int64_t qpitch, float* out,
                           int64_t opitch, int nthreads) {
    if (nthreads == 1) {
        for (int64_t row = 0; row < nr; ++row) {
            const int32_t* a = rl + row * rpitch;
            const int32_t* ac = rc + row * rpitch;
            float* orow = out + row * opitch;
            for (int64_t col = 0; col < nq; ++col)
                orow[col] = oracle_tanimoto_pair(a, ql + col * qpitch, ac, qc + col * qpitch, rlen[row], qlen[col],
                                                 rmag[row], qmag[col]);
        }
        return;
    }
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t row = 0; row < nr; ++row) {
        const int32_t* a = rl + row * rpitch;
        const int32_t* ac = rc + row * rpitch;
        float* orow = out + row * opitch;
        for (int64_t col = 0; col < nq; ++col)
            orow[col] = oracle_tanimoto_pair(a, ql + col * qpitch, ac, qc + col * qpitch, rlen[row], qlen[col],
                                             rmag[row], qmag[col]);
    }
}

/* Integer intersection block (for tests of the intersection counts). */
void oracle_intersection_block(const int32_t* rl, const int32_t* rc, const int32_t* rlen, int64_t nr,
                               int64_t rpitch, const int32_t* ql, const int32_t* qc, const int32_t* qlen,
                               int64_t nq, int64_t qpitch, int32_t* out, int nthreads) {
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t row = 0; row < nr; ++row)
        for (int64_t col = 0; col < nq; ++col)
            out[row * nq + col] = oracle_intersection(rl + row * rpitch, ql + col * qpitch, rc + row * rpitch,
                                                      qc + col * qpitch, rlen[row], qlen[col]);
    (void)nthreads;
}

/* Single-row GPU variant, kernels.py:234-270 with the device function at
 * :122-159: the row's magnitude is recomputed as the sum of its counts
 * (:258-261), each column's magnitude as the sum of its counts (:143), the
 * zero test is on LENGTHS (:136), and the union is formed in int32 before
 * the conversion: __fdiv_rn(isct, (float)(amag+bmag-isct)) (:157-158).
 * The j-outer merge (:138-155) is state-for-state the same as the
 * three-way merge. */
void oracle_single_row(const int32_t* a, const int32_t* ac, int32_t alen, const int32_t* ql, const int32_t* qc,
                       const int32_t* qlen, int64_t nq, int64_t qpitch, float* out) {
    int32_t amag = 0;
    for (int32_t i = 0; i < alen; ++i) amag += ac[i];
    for (int64_t col = 0; col < nq; ++col) {
        const int32_t* b = ql + col * qpitch;
        const int32_t* bc = qc + col * qpitch;
        int32_t blen = qlen[col];
        if (alen == 0 || blen == 0) {
            out[col] = 0.0f;
            continue;
        }
        int32_t bmag = 0;
        for (int32_t j = 0; j < blen; ++j) bmag += bc[j];
        int32_t isct = oracle_intersection(a, b, ac, bc, alen, blen);
        float u = (float)(amag + bmag - isct);
        out[col] = (float)isct / u;
    }
}

/* ------------------------------------------------------------------------ */
/* Histograms                                                                */
/* ------------------------------------------------------------------------ */

/* Bin of doHistogram<T>, _cpuutil.cpp:65-67: (unsigned)(x*100) clamped to
 * 100.  The conversion is undefined in C for negative, NaN and very large
 * products; the build defines those as the documented bins
 * (utilities.py:48-53): below 0 or NaN -> 0, at or above 100 -> 100. */
static inline uint32_t bin101_f32(float x) {
    float p = x * 100;
    if (!(p >= 0.0f)) return 0;
    if (p >= 100.0f) return 100;
    return (uint32_t)p;
}
static inline uint32_t bin101_f64(double x) {
    double p = x * 100;
    if (!(p >= 0.0)) return 0;
    if (p >= 100.0) return 100;
    return (uint32_t)p;
}

/* _cpuutil_rowHistogram101 (_cpuutil.cpp:74-126) over a float32 (R, C)
 * matrix with row pitch `pitch` elements -> uint32 (R, 101). */
void oracle_row_hist101_f32(const float* data, int64_t rows, int64_t cols, int64_t pitch, uint32_t* hist) {
    memset(hist, 0, sizeof(uint32_t) * 101 * (size_t)rows);
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t c = 0; c < cols; ++c) hist[r * 101 + bin101_f32(data[r * pitch + c])] += 1;
}
/* Same for float64 input, multiplied in double (_cpuutil.cpp:92,123). */
void oracle_row_hist101_f64(const double* data, int64_t rows, int64_t cols, int64_t pitch, uint32_t* hist) {
    memset(hist, 0, sizeof(uint32_t) * 101 * (size_t)rows);
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t c = 0; c < cols; ++c) hist[r * 101 + bin101_f64(data[r * pitch + c])] += 1;
}

/* Fused Tanimoto + per-row 101-bin histogram, gpuHistogrammedLingo_multirow
 * (kernels.py:606-667): bin (unsigned)(T*100) per pair (:657), int32 counts. */
void oracle_hist_rows101(const int32_t* rl, const int32_t* rc, const int32_t* rmag, const int32_t* rlen, int64_t nr,
                         int64_t rpitch, const int32_t* ql, const int32_t* qc, const int32_t* qmag,
                         const int32_t* qlen, int64_t nq, int64_t qpitch, int32_t* hist, int nthreads) {
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t row = 0; row < nr; ++row) {
        int32_t* h = hist + row * 101;
        memset(h, 0, sizeof(int32_t) * 101);
        for (int64_t col = 0; col < nq; ++col) {
            float t = oracle_tanimoto_pair(rl + row * rpitch, ql + col * qpitch, rc + row * rpitch, qc + col * qpitch,
                                           rlen[row], qlen[col], rmag[row], qmag[col]);
            h[bin101_f32(t)] += 1;
        }
    }
    (void)nthreads;
}

/* Global 1000-bin histogram with float32 edges e_k = fl32(k/1000),
 * k = 0..1000: bin = max{k : T >= e_k}, clipped to 999; T < 0 or NaN -> 0.
 * (No counterpart in the reference; int64 counts.) */
static inline int32_t bin1000(float t) {
    if (!(t >= 0.0f)) return 0;
    float p = t * 1000.0f;
    int32_t k = p >= 999.0f ? 999 : (int32_t)p;
    while (k < 999 && t >= (float)(k + 1) / 1000.0f) ++k;
    while (k > 0 && t < (float)k / 1000.0f) --k;
    return k;
}

void oracle_hist1000(const int32_t* rl, const int32_t* rc, const int32_t* rmag, const int32_t* rlen, int64_t nr,
                     int64_t rpitch, const int32_t* ql, const int32_t* qc, const int32_t* qmag, const int32_t* qlen,
                     int64_t nq, int64_t qpitch, int64_t* hist) {
    memset(hist, 0, sizeof(int64_t) * 1000);
    for (int64_t row = 0; row < nr; ++row)
        for (int64_t col = 0; col < nq; ++col) {
            float t = oracle_tanimoto_pair(rl + row * rpitch, ql + col * qpitch, rc + row * rpitch, qc + col * qpitch,
                                           rlen[row], qlen[col], rmag[row], qmag[col]);
            hist[bin1000(t)] += 1;
        }
}

/* Bin of one value (exported for tests). */
int32_t oracle_bin1000(float t) { return bin1000(t); }
uint32_t oracle_bin101_f32(float t) { return bin101_f32(t); }
uint32_t oracle_bin101_f64(double t) { return bin101_f64(t); }

/* ------------------------------------------------------------------------ */
/* Threshold neighbours                                                      */
/* ------------------------------------------------------------------------ */

/* gpuGetNeighbors_multirow, kernels.py:433-470.  A column j is a neighbour of
 * row r iff the magnitude bound fl32(min(ma,mb) / (float)max(ma,mb)) >= lo
 * (:440-441) and the Tanimoto T satisfies T >= lo && T < hi (:448); lo and hi
 * are float32.  Indices are written in ascending column order.  The list is
 * complete: idx must have room for nq entries per row (row-major, pitch nq);
 * counts[r] receives the full count. */
void oracle_neighbors(const int32_t* rl, const int32_t* rc, const int32_t* rmag, const int32_t* rlen, int64_t nr,
                      int64_t rpitch, const int32_t* ql, const int32_t* qc, const int32_t* qmag, const int32_t* qlen,
                      int64_t nq, int64_t qpitch, float lo, float hi, int32_t* idx, int64_t* counts, int nthreads) {
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t row = 0; row < nr; ++row) {
        int64_t cnt = 0;
        int32_t ma = rmag[row];
        for (int64_t col = 0; col < nq; ++col) {
            int32_t mb = qmag[col];
            float bound = imin(ma, mb) / ((float)imax(ma, mb));
            if (!(bound >= lo)) continue;
            float t = oracle_tanimoto_pair(rl + row * rpitch, ql + col * qpitch, rc + row * rpitch, qc + col * qpitch,
                                           rlen[row], qlen[col], ma, mb);
            if (t >= lo && t < hi) idx[row * nq + cnt++] = (int32_t)col;
        }
        counts[row] = cnt;
    }
    (void)nthreads;
}

/* ------------------------------------------------------------------------ */
/* Top-k (no counterpart in the reference)                                   */
/* ------------------------------------------------------------------------ */

/* Order: larger T first, equal T -> smaller column index first; T compared as
 * the IEEE total order of its bit pattern.  Rows with fewer than k columns are
 * padded with index -1 and T = -inf. */
static inline uint32_t ordered_bits(float t) {
    uint32_t u;
    memcpy(&u, &t, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

void oracle_topk(const int32_t* rl, const int32_t* rc, const int32_t* rmag, const int32_t* rlen, int64_t nr,
                 int64_t rpitch, const int32_t* ql, const int32_t* qc, const int32_t* qmag, const int32_t* qlen,
                 int64_t nq, int64_t qpitch, int32_t k, int32_t* out_idx, float* out_t, int nthreads) {
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t row = 0; row < nr; ++row) {
        /* keep a sorted list of the best k (insertion) */
        uint64_t* best = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(k > 0 ? k : 1));
        int32_t nb = 0;
        for (int64_t col = 0; col < nq; ++col) {
            float t = oracle_tanimoto_pair(rl + row * rpitch, ql + col * qpitch, rc + row * rpitch, qc + col * qpitch,
                                           rlen[row], qlen[col], rmag[row], qmag[col]);
            uint64_t key = ((uint64_t)ordered_bits(t) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)col);
            if (nb < k) {
                int32_t p = nb++;
                while (p > 0 && best[p - 1] < key) {
                    best[p] = best[p - 1];
                    --p;
                }
                best[p] = key;
            } else if (k > 0 && key > best[k - 1]) {
                int32_t p = k - 1;
                while (p > 0 && best[p - 1] < key) {
                    best[p] = best[p - 1];
                    --p;
                }
                best[p] = key;
            }
        }
        for (int32_t i = 0; i < k; ++i) {
            if (i < nb) {
                uint32_t ob = (uint32_t)(best[i] >> 32);
                uint32_t u = (ob & 0x80000000u) ? (ob & 0x7FFFFFFFu) : ~ob;
                float t;
                memcpy(&t, &u, 4);
                out_idx[row * k + i] = (int32_t)(0xFFFFFFFFu - (uint32_t)(best[i] & 0xFFFFFFFFu));
                out_t[row * k + i] = t;
            } else {
                out_idx[row * k + i] = -1;
                out_t[row * k + i] = -INFINITY;
            }
        }
        free(best);
    }
    (void)nthreads;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
