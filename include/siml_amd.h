/*
 * siml_amd.h -- C ABI of libsiml_amd.so, the MI355X (gfx950) LINGO
 * similarity library.
 *
 * Conventions
 *   - Every function returns 0 (SIML_OK) or a negative SIML_E* code; the
 *     message of the last failure on the calling thread is siml_last_error().
 *   - Sizes, offsets and pitches are int64 element counts (not bytes).
 *   - A "SMILES set" is four int32 arrays in the reference's argument order:
 *     Lingo matrix, count matrix, MAGNITUDES, LENGTHS (_CPULingo.c:103-105).
 *     Matrices are row-major (n, width) with row pitch `pitch` >= width;
 *     row r holds its lengths[r] Lingo ids ascending in the first columns.
 *   - Inputs are borrowed for the duration of the call; outputs are
 *     caller-owned host buffers.  Device buffers are owned by the handle.
 *   - Rows are the reference set (set_refsmiles), columns the query set
 *     (set_qsmiles), as in GPULingo.getMultipleRows (GPULingo.py:365-405).
 */
#ifndef SIML_AMD_H
#define SIML_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIML_OK 0
#define SIML_EINVAL (-1)  /* argument error (shape, dtype, range)          */
#define SIML_ENOMEM (-2)  /* host or device allocation failed               */
#define SIML_EDEVICE (-3) /* HIP runtime or kernel launch error             */
#define SIML_ESTATE (-4)  /* set not initialised / no pending result        */
#define SIML_ELIMIT (-5)  /* a size beyond a documented limit of this build */

/* Result kinds reported by siml_result_info(). */
#define SIML_RESULT_NONE 0
#define SIML_RESULT_F32 1   /* float32 (rows, cols) matrix                  */
#define SIML_RESULT_I32 2   /* int32 (rows, cols) matrix                    */
#define SIML_RESULT_CSR 3   /* int64 offsets (rows+1) + int32 indices (nnz) */
#define SIML_RESULT_TOPK 4  /* int32 (rows, k) indices + float32 (rows, k)  */
#define SIML_RESULT_I64 5   /* int64 vector of `cols` bins                  */

/* Largest number of unique Lingos in one reference row (a row is staged in
 * LDS).  The reference's limit came from 16 KiB of shared memory
 * (GPULingo.py:391): 2048 Lingos. */
#define SIML_MAX_ROW_LINGOS 8192

const char* siml_last_error(void);
int siml_version(void);
/* Number of visible HIP devices (0 without a GPU). */
int siml_device_count(int* count);

/* ---------------------------------------------------------------------- */
/* SMILES packer.  Replaces _ccompiler.SMILEStoMatrices                    */
/* (_ccompiler.cpp:276-381; digit state machine :178-264; LingoBag        */
/* :47-176).  Strings are one byte buffer with n+1 offsets.                */
/* ---------------------------------------------------------------------- */

/* Pass 1: unique-Lingo count (lengths) and magnitude (mags) per string. */
int siml_pack_measure(const char* buf, const int64_t* offsets, int64_t n, int32_t* lengths, int32_t* mags,
                      int nthreads);
/* Pass 2: zero-padded (n, width) Lingo and count matrices. */
int siml_pack_fill(const char* buf, const int64_t* offsets, int64_t n, int32_t* lingos, int32_t* counts,
                   int64_t width, int nthreads);
/* Ragged form: ids/counts of string r at [row_offsets[r], row_offsets[r+1]).
 * row_offsets is an exclusive scan of the lengths from siml_pack_measure. */
int siml_pack_fill_ragged(const char* buf, const int64_t* offsets, int64_t n, const int64_t* row_offsets,
                          int32_t* ids, int32_t* counts, int nthreads);

/* ---------------------------------------------------------------------- */
/* Stateless block on the GPU.  Replaces _CPULingo.getTanimotoBlock and    */
/* getTanimotoBlockParallel (_CPULingo.c:83-274, :277-465): float32       */
/* (nr, nq) Tanimoto block into a host buffer with row pitch `opitch`.     */
/* ---------------------------------------------------------------------- */
int siml_tanimoto_block(int device, const int32_t* rlingos, const int32_t* rcounts, const int32_t* rmags,
                        const int32_t* rlengths, int64_t nr, int64_t rwidth, int64_t rpitch,
                        const int32_t* qlingos, const int32_t* qcounts, const int32_t* qmags,
                        const int32_t* qlengths, int64_t nq, int64_t qwidth, int64_t qpitch, float* out,
                        int64_t opitch);

/* ---------------------------------------------------------------------- */
/* Row histograms of a host float matrix on the GPU.  Replaces             */
/* _cpuutil.rowHistogram101 (_cpuutil.cpp:74-126): bin                     */
/* min((unsigned)(x*100), 100), uint32 (rows, 101).  The float64 variant   */
/* multiplies in double (_cpuutil.cpp:92,123).                             */
/* ---------------------------------------------------------------------- */
int siml_row_hist101_f32(int device, const float* data, int64_t rows, int64_t cols, int64_t pitch,
                         uint32_t* hist);
int siml_row_hist101_f64(int device, const double* data, int64_t rows, int64_t cols, int64_t pitch,
                         uint32_t* hist);

/* ---------------------------------------------------------------------- */
/* Comparator handle.  Replaces GPULingo's device glue and launches        */
/* (_pycudaData, GPULingo.py:55-142; methods :186-648).  One handle owns   */
/* one HIP stream and at most one pending result (GPULingo.py:537-641).    */
/* ---------------------------------------------------------------------- */
typedef struct siml_ctx siml_ctx;

int siml_create(int device, siml_ctx** out);
int siml_destroy(siml_ctx* ctx);

/* set_refsmiles / set_qsmiles (GPULingo.py:186-314).  mags may be NULL:
 * the magnitudes are then the row sums of the counts, the values
 * accumulateRefMagnitudes / accumulateQueryMagnitudes produce
 * (kernels.py:45-79). */
int siml_set_ref(siml_ctx* ctx, const int32_t* lingos, const int32_t* counts, const int32_t* mags,
                 const int32_t* lengths, int64_t n, int64_t width, int64_t pitch);
int siml_set_query(siml_ctx* ctx, const int32_t* lingos, const int32_t* counts, const int32_t* mags,
                   const int32_t* lengths, int64_t n, int64_t width, int64_t pitch);
/* Ragged variants (no counterpart in the reference): the ids/counts of
 * molecule r are at [offsets[r], offsets[r+1]); n+1 offsets. */
int siml_set_ref_ragged(siml_ctx* ctx, const int32_t* ids, const int32_t* counts, const int32_t* mags,
                        const int64_t* offsets, int64_t n);
int siml_set_query_ragged(siml_ctx* ctx, const int32_t* ids, const int32_t* counts, const int32_t* mags,
                          const int64_t* offsets, int64_t n);
/* Resident query databases (no counterpart in the reference).  A database
 * is a ragged query set built once in device memory; any number of handles
 * on the same device attach it as their query set and run concurrently on
 * their own streams.  siml_db_destroy drops the caller's reference; the
 * device memory is freed when no handle uses it any more.  layout is one
 * of the SIML_LAYOUT_* storage orders below. */
typedef struct siml_db siml_db;
int siml_db_create(int device, const int32_t* ids, const int32_t* counts, const int32_t* mags,
                   const int64_t* offsets, int64_t n, int layout, siml_db** out);
int siml_db_destroy(siml_db* db);
int siml_db_size(siml_db* db, int64_t* n);
int siml_attach_db(siml_ctx* ctx, siml_db* db);
/* On-disk database (no counterpart in the reference, whose packed sets only
 * note that zero padding "will allow efficient on-disk compressed storage",
 * _ccompiler.cpp:354-355): siml_db_save writes the built device arrays
 * (tiles, level stream, permutation, magnitudes); siml_db_load uploads them
 * to `device` without rebuilding.  Files carry a format version; a file of
 * another version is rejected with SIML_EINVAL. */
int siml_db_save(siml_db* db, const char* path);
int siml_db_load(int device, const char* path, siml_db** out);
/* Sizes of the current sets (-1 when unset). */
int siml_set_sizes(siml_ctx* ctx, int64_t* nref, int64_t* nquery);

/* gpuLingoSim3_multirow via getMultipleRows_async (GPULingo.py:365-405,
 * kernels.py:315-352): float32 (rowlimit-rowbase, nquery). */
int siml_multirow_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit);
/* gpuLingoSim3_srgf_T_tex2ns via getTanimotoRow_async (GPULingo.py:327-352,
 * kernels.py:234-270, :122-159): float32 (1, nquery); magnitudes are
 * recomputed from the counts and the zero test is on lengths. */
int siml_single_row_async(siml_ctx* ctx, int64_t row);
/* gpuHistogrammedLingo_multirow via getMultipleHistogrammedRows_async
 * (GPULingo.py:419-464, kernels.py:606-667): int32 (rows, 101). */
int siml_hist_rows_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit);
/* gpuGetNeighbors_multirow via getNeighbors_async (GPULingo.py:466-514,
 * kernels.py:393-475): for each row, every query column with
 * fl32(min(ma,mb)/(float)max(ma,mb)) >= lo and lo <= T < hi, ascending.
 * The result is complete CSR (the reference truncates at its pitch).  The
 * call waits for the count pass (the host reads back three totals to size
 * the index array); offsets are scanned on the device, and the fill,
 * compaction and per-row sort are queued on the stream. */
int siml_neighbors_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, float lo, float hi);
/* Top-k per row (no counterpart in the reference): larger T first, equal T
 * ordered by smaller column index; rows with fewer than k columns padded
 * with index -1 and T = -inf.  1 <= k <= SIML_MAX_TOPK. */
#define SIML_MAX_TOPK 256
int siml_topk_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, int32_t k);
/* Global 1000-bin histogram (no counterpart in the reference) of every T in
 * the block, float32 edges fl32(k/1000), bin = max{k : T >= edge_k} clipped
 * to 999; exact int64 counts. */
int siml_hist1000_async(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit);
/* Top-k with no device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * result of siml_topk_async over the same rows, written on the handle's
 * stream to caller-owned device arrays idx (int32, rows x k) and t
 * (float32, rows x k); the call returns once the work is queued
 * (siml_synchronize waits) and leaves the handle's pending result alone.
 * Candidate buffers, selection levels and the row and group lists live in
 * a static device array of SIML_NOALLOC_SCRATCH_BYTES, one per GPU, held
 * until the process ends; the launch runs fewer column ranges until they
 * fit, and a call that does not fit at one range fails with SIML_ELIMIT.
 * Every call on a GPU shares that GPU's array: calls that overlap on one
 * GPU (from any handle or stream) race.  The autotuner and kernel timing
 * are not used.  The sets must already be on the device (set_ref /
 * set_query / attach_db allocate); the bit-sliced row groups serve the call
 * when an earlier call built them (building allocates), else the merge
 * kernel does, with the same results. */
#define SIML_NOALLOC_SCRATCH_BYTES (128u << 20)
int siml_topk_noalloc(siml_ctx* ctx, int64_t rowbase, int64_t rowlimit, int32_t k, int32_t* idx, float* t);

/* Device memory of one call (no counterpart in the reference): an upper
 * bound of the bytes a call of operation op over `rows` reference rows
 * against the current query set allocates on the device beyond the
 * reference set itself -- candidate buffers, neighbour slots, selection
 * levels and the result -- from the launch shapes the call may take (every
 * candidate shape while the autotuner is on).  The indices of neighbour
 * lists depend on the data and are not included (4 bytes per neighbour, 8
 * when the query set is stored in a permuted order).
 * A handle keeps its largest buffers between calls.  k is the top-k k
 * (ignored otherwise). */
#define SIML_OP_TOPK 1
#define SIML_OP_NEIGHBORS 2
#define SIML_OP_HIST1000 3
int siml_call_bytes(siml_ctx* ctx, int op, int64_t rows, int32_t k, int64_t* bytes);

/* Pending-result protocol (GPULingo.py:537-648). */
int siml_query_done(siml_ctx* ctx, int* done);   /* asyncOperationsDone */
int siml_synchronize(siml_ctx* ctx);
int siml_result_info(siml_ctx* ctx, int* kind, int64_t* rows, int64_t* cols, int64_t* nnz);
/* Copy the pending result to host buffers and clear it.  Matrix results use
 * the host row pitch `pitch`; when it equals the device row pitch
 * (16 * ceil(cols / 16), as GPULingo.py:378-379 allocates) the rows are
 * copied as one span, so elements [cols, pitch) of each row but the last
 * are overwritten with unspecified values, as in the reference's padded
 * result array; any other pitch leaves them untouched.
 * CSR fills offsets (rows+1) and indices (nnz);
 * top-k fills idx and t with pitch k; I64 fills `cols` values. */
int siml_retrieve_f32(siml_ctx* ctx, float* host, int64_t pitch);
int siml_retrieve_i32(siml_ctx* ctx, int32_t* host, int64_t pitch);
int siml_retrieve_csr(siml_ctx* ctx, int64_t* offsets, int32_t* indices);
int siml_retrieve_topk(siml_ctx* ctx, int32_t* idx, float* t);
int siml_retrieve_i64(siml_ctx* ctx, int64_t* host);
/* Copy rows [r0, r1) of a pending F32 result without clearing it. */
int siml_peek_rows_f32(siml_ctx* ctx, int64_t r0, int64_t r1, float* host, int64_t pitch);

/* ---------------------------------------------------------------------- */
/* Measurement (no counterpart in the reference).                          */
/* ---------------------------------------------------------------------- */
/* When on, every compute kernel launch is bracketed by HIP events on the
 * handle's stream; siml_kernel_stats returns the summed kernel time (ms),
 * the number of launches and the number of pairs those launches evaluated
 * since the last reset. */
int siml_set_timing(siml_ctx* ctx, int on);
int siml_kernel_stats(siml_ctx* ctx, double* total_ms, int64_t* launches, int64_t* pairs, int reset);
/* The time of siml_kernel_stats split into the similarity kernels and the
 * top-k selection kernels that follow them (zero for other operations). */
int siml_kernel_split(siml_ctx* ctx, double* similarity_ms, double* selection_ms);
/* Work the neighbour kernels actually did while timing was on, counted on
 * the device: the pairs they evaluated (those the magnitude bound did not
 * prune) and the sum of La + Lb (unique lingo counts) over those pairs, the
 * merge work the pairs stand for.  Values since the last reset. */
int siml_work_stats(siml_ctx* ctx, int64_t* pairs, int64_t* merge_len, int reset);
/* Kernel policy: AUTO serves sorted-unique sets with the bit-sliced
 * multi-row kernels (rows that do not qualify, and unsorted sets, use the
 * merge kernel) and, for the float32 matrix (multirow), the top-k and the
 * 1000-bin histogram, chooses per call between the bit-sliced kernels and
 * the sparse path (an inverted index of the query set's levels, built on
 * first use, and per-pair counts in LDS) from the index's expected postings
 * per level; the sparse top-k and histogram serve calls whose rows all
 * qualify and whose magnitudes are at least the count sums (the
 * allocation-free top-k never takes them).  BITS never takes the sparse
 * path; SPARSE takes it whenever the call can; MERGE uses the merge kernel
 * for everything.  All give the same bits.  The bit-sliced row
 * groups serve a reference set given while the policy is not MERGE; they
 * are built by the first call that needs them (a reference set given under
 * MERGE is served by the merge kernel). */
#define SIML_POLICY_AUTO 0
#define SIML_POLICY_MERGE 1
#define SIML_POLICY_BITS 2
#define SIML_POLICY_SPARSE 3
int siml_set_policy(siml_ctx* ctx, int policy);
/* Storage order used by the next set_query: NATURAL keeps the caller's
 * order; BLOCK (default) sorts the query molecules by size inside blocks
 * of 256 consecutive columns; SORTED sorts the whole set by size (best for
 * top-k, neighbours and histograms; the float32 matrix is written with
 * scattered stores).  Results are identical; only speed differs. */
#define SIML_LAYOUT_NATURAL 0
#define SIML_LAYOUT_BLOCK 1
#define SIML_LAYOUT_SORTED 2
int siml_set_layout(siml_ctx* ctx, int layout);
/* Name of the kernel variant the last similarity launch used. */
const char* siml_last_kernel(siml_ctx* ctx);
/* Launch-shape autotuner (no counterpart in the reference).  When on, the
 * first multirow / hist_rows / neighbors / topk / hist1000 call of a shape
 * bucket (operation, log2 of the rows, log2 of the query tiles, k) runs once
 * per candidate launch shape -- workgroups aimed at per launch and, for
 * top-k, the most waves per row -- times each with events and keeps the
 * fastest for the bucket, shared by every handle on the device.  Results do
 * not depend on the shape.  Off by default.  siml_launch_shape reports the
 * shape the handle used last; siml_set_launch_shape fixes one (and turns
 * the autotuner off): workgroups aimed at per launch, most waves per top-k
 * row, and the top-k seed stride: 0 = no seeding, s in [2, 64] = top-k first
 * seeds its row thresholds from every s-th query tile, 1 = a stride from
 * the query set's size (its 64-molecule tiles / 512, clamped to [16, 64]).
 * Default (and after siml_set_autotune(ctx, 0)): 8192 workgroups, 1024
 * waves per top-k row, seed stride 1 (from the size). */
int siml_set_autotune(siml_ctx* ctx, int on);
/* Graph mode (replaces the reference's per-call prepared_async_call
 * launches, GPULingo.py:391-397,501-509): when on, the kernel sequence of
 * each siml_topk_async / siml_topk_noalloc / siml_hist1000_async call
 * (similarity passes, selection stages, bin clearing) is recorded by stream
 * capture and launched as one HIP graph; host uploads of the call's row
 * lists stay outside the graph.  Instantiated graphs are cached per handle
 * by their kernel sequence and updated in place with the next call's
 * arguments.  Results are identical; off by default.  siml_graph_stats
 * reports the graph launches and the instantiations (cache misses) since
 * the handle was created. */
int siml_set_graphs(siml_ctx* ctx, int on);
/* Persistent grid (no counterpart in the reference, whose kernels launch
 * one block per row, kernels.py:315-352): when on, the bit-sliced similarity
 * pass of siml_topk_async / siml_topk_noalloc / siml_hist1000_async runs as
 * many workgroups as the device holds at once, which take (row-group slot)
 * work items from one queue per XCD and run each slot's column ranges of
 * that XCD with the slot's tables loaded once.  Results are identical.  Turns
 * the autotuner off (whose candidates include persistent shapes); off by
 * default. */
int siml_set_persistent(siml_ctx* ctx, int on);
int siml_graph_stats(siml_ctx* ctx, int64_t* launches, int64_t* builds);
int siml_launch_shape(siml_ctx* ctx, int64_t* target_wg, int32_t* topk_wave_cap, int32_t* topk_seed);
int siml_set_launch_shape(siml_ctx* ctx, int64_t target_wg, int32_t topk_wave_cap, int32_t topk_seed);
/* Bit-sliced row groups of the current reference set (no counterpart in
 * the reference): the number of groups (each probes every level of every
 * query molecule once per launch) and of rows left to the merge kernel;
 * both 0 while no call has needed the groups (they are built on first use:
 * a call that takes a sparse path builds none). */
int siml_row_groups(siml_ctx* ctx, int64_t* groups, int64_t* merge_rows);
/* Most row groups one bit-sliced workgroup serves (no counterpart in the
 * reference): 1, 2 or 4 (default 2).  A workgroup serving G groups holds G
 * tables in LDS and probes each column level it loads in all of them, so
 * the query stream is read from memory once per G groups.  Top-k and the
 * 1000-bin histogram use it; launches with few groups, rows of more than
 * 255 levels, and the magnitude-pruned neighbour search use one table.
 * Results do not depend on it. */
int siml_set_group_tables(siml_ctx* ctx, int max_tables);
/* Self-test of a short float32 division method of the kernels (no
 * counterpart in the reference): compares it on the device with IEEE
 * division for every integer numerator in [0, 65536) and denominator in
 * [1, 131072); reports the number of differing quotients and the first
 * (smallest denominator) pair, -1 when none. */
int siml_selftest_division(int device, int method, int64_t* mismatches, int64_t* first_n, int64_t* first_d);
/* Achievable int32 VALU rate of the device (no counterpart in the
 * reference): lane-operations per second of independent v_add_u32 (op 0) or
 * v_min_i32 (op 1) chains on every SIMD, the measured counterpart of the
 * 256 CU x 128 lanes x 2.4 GHz vector peak. */
int siml_selftest_valu(int device, int op, double* lane_ops_per_s);
/* Achievable LDS read rate of the device (no counterpart in the reference):
 * bytes per second of 8-byte reads from a 32 KiB table per workgroup on
 * every CU, without bank conflicts (op 0) or at pseudo-random slots (op 1,
 * the access pattern of the bit-sliced kernels' table probes). */
int siml_selftest_lds(int device, int op, double* bytes_per_s);

#ifdef __cplusplus
}
#endif

#endif /* SIML_AMD_H */
