/*
 * tjoin.h -- C ABI of libtjoin.so, the MI355X (gfx950) certified exact
 * epsilon-join.
 *
 * Contract of every join entry point: given row-major FP32 point sets R
 * [nR x d] and S [nS x d] in device memory and a squared threshold T
 * (binary64), emit every pair (i, j) whose sequential binary64 difference
 * form d2_seq(R_i, S_j) = sum_k fl(fl(R_ik - S_jk)^2), k ascending, no FMA,
 * satisfies d2_seq <= T, sorted ascending by (i, j).  A self-join (S == NULL)
 * emits i < j only.  The predicate is the one of
 * CEJ/src/run_g2a_tensorjoin.py:193-232 (refine_ambiguous_fp64_i64) in the
 * summation order of CEJ/src/audit_g2b_accepted_fp64.py:51-57.
 *
 * Conventions: all pointers are device pointers unless noted; every call
 * enqueues on `stream` and returns without waiting, except where a comment
 * says "synchronous".  Return value: TJ_OK or a negative TJ_ERR_* code;
 * tj_last_error_string() describes the last failure of the calling thread.
 * Capacity overflow is not an error of the call: the stats block reports the
 * counts that were needed and the caller retries with larger buffers (the
 * retry protocol of CEJ/src/run_g2a_tensorjoin.py:329-378).
 */
#ifndef TJOIN_H
#define TJOIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TJ_OK 0
#define TJ_ERR_INVALID (-1)     /* bad argument */
#define TJ_ERR_HIP (-2)         /* a HIP runtime call failed */
#define TJ_ERR_UNSUPPORTED (-3) /* shape/precision this build does not serve */
#define TJ_ERR_WORKSPACE (-4)   /* workspace smaller than tj_plan_workspace_bytes */

/* Screen precisions (the low-precision MFMA pass that certifies most pairs). */
#define TJ_PREC_NONE 0   /* no screen: plain FP64 brute force */
#define TJ_PREC_FP16 1   /* FP16 MFMA, one term */
#define TJ_PREC_FP16X3 2 /* FP16 MFMA, hi/lo split: hi.hi + hi.lo + lo.hi */
#define TJ_PREC_BF16 3   /* BF16 MFMA, one term */
#define TJ_PREC_BF16X3 4 /* BF16 MFMA, hi/lo split */
#define TJ_PREC_FP8 5    /* OCP e4m3 MFMA with per-row power-of-two scales */
#define TJ_PREC_FP8X3 6  /* e4m3 hi/lo split with per-row scales */

/* Join methods. */
#define TJ_METHOD_CERTIFIED 0  /* MFMA screen + FP64 refinement of the ambiguous band */
#define TJ_METHOD_PLAIN_FP64 1 /* FP64 brute-force tile, no certificate band */
#define TJ_METHOD_GRID 2       /* uniform grid (d <= 3), FP64 predicate per candidate */
#define TJ_METHOD_PLAIN_FP32 3 /* control: FP32 GEMM-form tile over every pair, rigorous binary64 interval,
                                  FP64 for the pairs it cannot decide (G16/src/fp32_gpu_norm_operator.py:107-161) */

/* Output pair formats. */
#define TJ_OUT_I32 0 /* int32 [P, 2] */
#define TJ_OUT_I64 1 /* int64 [P, 2] */
#define TJ_OUT_U64_CANONICAL 2 /* tj_host_join of a self-join: directed ids i*n+j, both directions and the n self pairs, ascending */

/* Flags. */
#define TJ_FLAG_AUDIT 1u  /* route certified accepts through FP64 too and count certificate violations */
#define TJ_FLAG_TIMING 2u /* record HIP events between pipeline stages (read with tj_plan_stage_times) */
#define TJ_FLAG_PRUNE 4u  /* batches: order each join's rows spatially and skip panel pairs whose bounding balls are certified apart */

/* Slots of the uint64 stats block (TJ_STATS_SLOTS entries, device memory).
 * Slot meanings follow the reference counters [accept, ambiguous, overflow,
 * fp64, reject, bitwise-equal] of CEJ/src/run_g3c_b_r1_gpu_cascade.py:108-129,
 * extended with the screen error statistics. */
#define TJ_ST_OUT 0           /* pairs written to the output = CERT_ACCEPT + FP32_ACCEPT + FP64_ACCEPT */
#define TJ_ST_AMB 1           /* ambiguous pairs of the per-pair interval (decided by the FP32 filter or the FP64 form) */
#define TJ_ST_CERT_ACCEPT 2   /* pairs certified inside (screen fast accept + per-pair interval) */
#define TJ_ST_FP64_ACCEPT 3   /* ambiguous pairs the FP64 refinement accepted (after the FP32 filter) */
#define TJ_ST_CANDIDATES 4    /* pairs the screen queued for the per-pair interval test (the amb_capacity queue):
                                 the band between the fast reject and the fast accept (every survivor in audit mode) */
#define TJ_ST_MAX_ERR 5       /* max |d2_screen - d2_fp64| over refined pairs (binary64 bits, real units) */
#define TJ_ST_MAX_RATIO 6     /* max |d2_screen - d2_fp64| / certified half-width (binary64 bits) */
#define TJ_ST_VIOLATIONS 7    /* refined pairs outside their certified interval (must be 0) */
#define TJ_ST_WORK 8          /* work-queue cursor of the persistent screen */
#define TJ_ST_INVALID 9       /* non-finite input coordinates seen by the row metadata pass */
#define TJ_ST_GRID_CAND 10    /* grid join: candidates, the points in the cells adjacent to each query */
#define TJ_ST_PAIRS 11        /* pair evaluations the call covers */
#define TJ_ST_SCALE_EXP 12    /* join-unit exponent G (x_join = x * 2^-G), as int64 */
#define TJ_ST_VIOL_KEY 13      /* diagnostic: key of a refined pair outside its interval (flag bit 63 = certified) */
#define TJ_ST_VIOL_D2 14       /* diagnostic: its binary64 d2 (join units, binary64 bits) */
#define TJ_ST_VIOL_EST 15      /* diagnostic: its screen estimate and half-width (two binary32) */
#define TJ_ST_FP32_ACCEPT 16   /* ambiguous pairs the certified FP32 filter decided inside */
#define TJ_ST_FP32_REJECT 17   /* ambiguous pairs the certified FP32 filter decided outside */
#define TJ_ST_FP64_REFINED 18  /* pairs evaluated in the sequential FP64 form (all ambiguous ones in audit mode) */
#define TJ_ST_PRUNED 19        /* batches: pairs certified outside by panel bounding balls (never screened) */
#define TJ_ST_PRUNED_WORK 20   /* batches: sum over those pairs of the join width d (screen work not done) */
#define TJ_STATS_SLOTS 24

/* ---------------------------------------------------------------- library */
int tj_version(void);
const char* tj_last_error_string(void);
int tj_device_info(int* cu_count, int* arch_major, int* arch_minor); /* synchronous */

/* ------------------------------------------------------------------ plans
 * A plan fixes shapes, method, precision and buffer capacities and lays out
 * the caller-owned workspace.  Replaces the per-run buffer set of
 * TG/src/g2_operator_r1.py:42-61 (Matrix.run) and
 * CEJ/src/run_g5_tensorjoin_public.py:384-394. */
typedef struct tj_plan tj_plan;

int tj_plan_create(tj_plan** plan, int64_t nR, int64_t nS /* 0 = self-join */, int d,
                   int method, int precision, int64_t out_capacity, int64_t amb_capacity,
                   unsigned flags);
int tj_plan_destroy(tj_plan* plan);
/* Screen tile of a certified plan: row_blocks = 32-row blocks per wave (0 =
 * the body's default; the four-block R x S FP16 / BF16 bodies of width 64 and
 * 128 also run with 2), cols_per_split = columns of S per work item (0 = the
 * default rule, rounded up to 32).  Never changes the output, only the time;
 * drops a captured graph.  Replaces the fixed BLOCK_M = BLOCK_N = 64 tile of
 * CEJ/src/run_g2b_tensorjoin_public.py:41-42 (the reference does not tune). */
int tj_plan_set_tile(tj_plan* plan, int row_blocks, int64_t cols_per_split);
int tj_plan_workspace_bytes(const tj_plan* plan, size_t* bytes);
int tj_plan_describe(const tj_plan* plan, char* buf, size_t len); /* host only */
/* Milliseconds of the stages of the last tj_join on a TJ_FLAG_TIMING plan:
 * ms[0] metadata (scale scan, parameters, row metadata), ms[1] screen (or the
 * plain brute force), ms[2] FP64 refinement, ms[3] sort + unpack.
 * Synchronous (waits for the last stage event).  Counterpart of the optional
 * per-stage events of OA/src/common.py:91-106. */
int tj_plan_stage_times(tj_plan* plan, float* ms, int n);

/* Full join: enqueue row metadata -> screen -> FP64 refinement -> sort ->
 * unpack.  `pairs` receives min(P, out_capacity) rows in `out_format`;
 * `stats` (uint64[TJ_STATS_SLOTS], device) receives the counters.  Replaces
 * Matrix.run (TG/src/g2_operator_r1.py:42-61) and run_operator
 * (G16/src/fp32_gpu_norm_operator.py:107-161). */
int tj_join(tj_plan* plan, const float* R, const float* S, double T, void* workspace,
            void* pairs, int out_format, uint64_t* stats, void* stream);

/* tj_join without any allocation, for callers that cannot allocate device
 * memory: code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  No counterpart in the reference.  The arguments are
 * those of tj_plan_create and tj_join together (nS == 0: self-join of R,
 * S ignored); the caller passes only inputs and outputs.  The workspace a
 * tj_plan of the same arguments would need comes from a device array of
 * TJ_NOALLOC_SCRATCH_BYTES that the library holds on each GPU until the
 * process ends; a join needing more (tj_plan_workspace_bytes of that plan)
 * returns TJ_ERR_WORKSPACE.  Flags: TJ_FLAG_AUDIT only.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (two streams or
 * threads at once) race and must be serialised by the caller.  Capacity
 * overflow is reported in stats as for tj_join. */
#define TJ_NOALLOC_SCRATCH_BYTES (64u << 20)
int tj_join_noalloc(const float* R, int64_t nR, const float* S, int64_t nS, int d, int method, int precision,
                    double T, int64_t out_capacity, int64_t amb_capacity, unsigned flags, void* pairs,
                    int out_format, uint64_t* stats, void* stream);

/* hipGraph form of tj_join.  tj_plan_capture (synchronous) records the whole
 * tj_join sequence for these buffers into a graph owned by the plan;
 * tj_plan_replay launches it on `stream`.  No memset nodes: every clear is
 * a kernel.  The reference used CUDA Graphs only in a diagnostic
 * (CEJ/src/run_g9_graph_diagnostic.py:223-232); here it is the launch path
 * for callers that repeat a join on fixed buffers. */
int tj_plan_capture(tj_plan* plan, const float* R, const float* S, double T, void* workspace,
                    void* pairs, int out_format, uint64_t* stats);
int tj_plan_replay(tj_plan* plan, void* stream);

/* --------------------------------------------------------- batches of joins
 * Many independent self-joins of uneven sizes and widths in one pipeline
 * with one persistent screen launch over the panels of every join, largest
 * first (one launch per kernel family present: FP16 widths 16/64/256, FP16
 * widths 32/128, FP8).  Replaces the host loop over 4,096-tile batches with a blocking
 * count read after every stage (CEJ/src/run_g5_tensorjoin_public.py:
 * 377-381,410-535; TG/src/g2_operator_r1.py:45-57).
 * tj_batch_create (synchronous) fixes the joins' sizes n[t] (host array),
 * widths d[t] (1..256) and screen precisions (TJ_PREC_FP16 or TJ_PREC_FP8
 * per join) and uploads the work list.  tj_batch_join takes x, a device
 * array of `joins` device pointers (join t: row-major FP32 [n[t] x d[t]]),
 * and T, a device array of the squared thresholds; it writes the local
 * (i, j), i < j, of every join as int32 rows into `pairs`, join after join,
 * each join's rows sorted, and offsets[t] (device int64[joins + 1]) = first
 * row of join t, offsets[joins] = total.  stats as tj_join, summed over the
 * joins; capacities and retry as tj_plan. */
typedef struct tj_batch tj_batch;

int tj_batch_create(tj_batch** batch, int joins, const int64_t* n, const int* d,
                    const int* precision, int64_t out_capacity, int64_t cand_capacity,
                    unsigned flags);
int tj_batch_destroy(tj_batch* batch);
int tj_batch_workspace_bytes(const tj_batch* batch, size_t* bytes);
int tj_batch_describe(const tj_batch* batch, char* buf, size_t len); /* host only */
int tj_batch_stage_times(tj_batch* batch, float* ms, int n);         /* as tj_plan_stage_times */
int tj_batch_join(tj_batch* batch, const float* const* x, const double* T, void* workspace,
                  int32_t* pairs, int64_t* offsets, uint64_t* stats, void* stream);
/* hipGraph form of tj_batch_join, as tj_plan_capture / tj_plan_replay. */
int tj_batch_capture(tj_batch* batch, const float* const* x, const double* T, void* workspace,
                     int32_t* pairs, int64_t* offsets, uint64_t* stats);
int tj_batch_replay(tj_batch* batch, void* stream);

/* ------------------------------------------------------- pipeline stages
 * The stages tj_join runs, exported for callers that compose their own
 * pipeline and for tests.  All counts are read from device memory so that
 * the sequence is capturable in a hipGraph. */

/* Row metadata for one side: quantized copies and the certificate terms.
 * Replaces half_metadata (TG/src/gpu_kernels.py:33-51), gpu_quantize_metadata
 * (G16/src/gpu_preparation.py:12-49) and gpu_norm_metadata
 * (G16/src/gpu_norms.py:11-25).  `params` is the device parameter block
 * written by tj_params. */
int tj_scale_scan(const float* x, int64_t n, int d, uint32_t* maxabs_bits, uint64_t* stats,
                  void* stream);
int tj_params(const uint32_t* maxabs_bits, double T, int d, int precision, void* params,
              uint64_t* stats, void* stream);
/* meta: 12 floats per row {nlo, nhi, H, E, L, t, v, e, w, a, 0, 0}: the
 * bracket of ||x||^2, the bounds of ||hi||, ||residual|| and ||lo||, the
 * screen's row level t (its MFMA accumulator seed, or added to the dot in
 * binary32 by the four-block bodies), reject level v and accept terms w, a,
 * and the FP8 scale exponents e (join units; see tj_params). */
int tj_rowmeta(const float* x, int64_t n, int d, int precision, const void* params,
               void* q_hi, void* q_lo, float* meta, uint64_t* stats, void* stream);

/* Sizes of one side's row metadata buffers: q_bytes for each code array
 * (q_hi, and q_lo for the hi/lo precisions) and meta_bytes for `meta`.
 * Host only. */
int tj_rowmeta_sizes(int64_t n, int d, int precision, size_t* q_bytes, size_t* meta_bytes);

/* The MFMA screen over A x B (nb == 0: the self-join of A, i < j), the
 * certified classification of every pair from its MFMA dot and the row
 * terms: pairs certified inside by the fast accept are appended to
 * out_keys (count stats[TJ_ST_OUT]); the band the screen cannot decide
 * (every survivor of the fast reject with TJ_FLAG_AUDIT) is appended to the
 * candidate queue: cand_keys[k] = i << jbits | j (jbits = the bits of the
 * number of B rows, the smallest b with 2^b >= max(nB, 2)) and cand_s[k] =
 * its binary32 screen value, count stats[TJ_ST_CANDIDATES].  The queue is in
 * no particular order.  Entries past a capacity are
 * counted, not written.  Inputs come from tj_rowmeta of each side with the
 * same params; stats must hold the counts of earlier stages (zero at the
 * start of a join); the screen's work cursor stats[TJ_ST_WORK] is reset
 * here.  Replaces scan16 / scan8 (TG/src/g2_kernels.py:7-82) with their
 * compaction (half_classify, TG/src/gpu_kernels.py:55-88), as launched
 * stage by stage by TG/src/frozen_driver.py:69-77. */
int tj_screen(int precision, int d, const void* qa_hi, const void* qa_lo, const float* meta_a, int64_t na,
              const void* qb_hi, const void* qb_lo, const float* meta_b, int64_t nb, const void* params,
              uint64_t* out_keys, int64_t out_capacity, uint64_t* cand_keys, float* cand_s, int64_t cand_capacity,
              unsigned flags, uint64_t* stats, void* stream);

/* The candidate queue of tj_screen -> decisions: per-pair certified
 * interval, then the certified FP32 filter and the sequential FP64 form for
 * what stays ambiguous; accepted keys appended to out_keys at
 * stats[TJ_ST_OUT].  Reads min(stats[TJ_ST_CANDIDATES], cand_capacity)
 * entries.  S == NULL for a self-join; n_cols = the number of S rows (of R
 * for a self-join), which fixes the key layout.  TJ_FLAG_AUDIT recomputes
 * certified pairs in FP64 too and counts violations.  Replaces the accept /
 * ambiguous queue / FP64 terminal cascade of TG/src/g2_operator_r1.py:52-57
 * with certified_fp32_filter_i64 (CEJ/src/run_g3c_b_r1_gpu_cascade.py:57-129)
 * and refine_ambiguous_fp64_i64 (CEJ/src/run_g2a_tensorjoin.py:193-232). */
int tj_certify(const float* R, const float* S, int d, int64_t n_cols, const uint64_t* cand_keys, const float* cand_s,
               int64_t cand_capacity, const float* meta_R, const float* meta_S, const void* params, uint64_t* out_keys,
               int64_t out_capacity, unsigned flags, uint64_t* stats, void* stream);

/* FP64 refinement of a queue of pair keys i << jbits | j (bit 63 ignored):
 * every key whose sequential binary64 difference form d2_seq(R_i, S_j) <= T
 * (real units) is appended to out_keys at stats[TJ_ST_OUT] (positions past
 * out_capacity counted, not written).  The queue length is read from the
 * device scalar *count, clamped to capacity -- e.g. stats + TJ_ST_CANDIDATES
 * for tj_screen's queue.  S == NULL: S = R.  Adds to stats[TJ_ST_FP64_REFINED]
 * and stats[TJ_ST_FP64_ACCEPT].  Replaces refine_ambiguous_fp64_i64
 * (CEJ/src/run_g2a_tensorjoin.py:193-232): ids in, accepted ids out. */
int tj_refine_fp64(const float* R, const float* S, int d, int jbits, double T, const uint64_t* keys,
                   const uint64_t* count, int64_t capacity, uint64_t* out_keys, int64_t out_capacity, uint64_t* stats,
                   void* stream);

/* Plain FP64 brute force (the comparison method): every pair evaluated in
 * the sequential FP64 difference form, accepted keys appended to out_keys.
 * Replaces the GDS brute-force kernel (GDS/kernel.cu:493-522) and the FP64
 * control of G16/src/fp32_gpu_norm_operator.py:107-161. */
int tj_plain_fp64(const float* R, int64_t nR, const float* S, int64_t nS, int d, int jbits,
                  double T, uint64_t* out_keys, int64_t out_capacity, uint64_t* stats,
                  void* stream);

/* The uniform-grid join (d <= 3) as one stage, with its own workspace
 * (tj_grid_workspace_bytes; capacity = rows of the output; the workspace
 * holds a candidate buffer of 9 capacity + 2 nR + 1024 ids, so any output
 * that fits the capacity is complete in one call).  tj_grid_join:
 * R x S (nS == 0: self-join of R, i < j) -> sorted (i, j) rows in
 * out_format, count stats[TJ_ST_OUT] (rows past capacity counted, not
 * written); stats is cleared first and receives TJ_ST_OUT, TJ_ST_GRID_CAND,
 * TJ_ST_INVALID, TJ_ST_PAIRS.  Replaces GDS-Join's GPU join
 * (GDS/main.cu:690-950, GDS/kernel.cu:132-486, GDS/GPU.cu:775-872) with the
 * predicate d2_seq <= T of this library. */
int tj_grid_workspace_bytes(int64_t nR, int64_t nS, int64_t capacity, size_t* bytes);
int tj_grid_join(const float* R, int64_t nR, const float* S, int64_t nS, int d, double T, void* workspace,
                 size_t workspace_bytes, int64_t capacity, void* pairs, int out_format, uint64_t* stats, void* stream);
/* GDS-Join's output (GDSJoinPy + copyResultIntoPythonArray, GDS/main.cu:360,
 * 507): for every point i of X, counts[i] = the number of points j (i
 * itself included) with d2_seq(X_i, X_j) <= T, and table = their ids grouped
 * by i in ascending i, each group ascending; total = stats[TJ_ST_OUT] (ids
 * past capacity counted, not written).  Workspace: tj_grid_workspace_bytes
 * (n, 0, capacity).  (GDS-Join decides sqrt(d2) <= eps in its own summation
 * order and leaves each group unordered.) */
int tj_grid_neighbors(const float* X, int64_t n, int d, double T, void* workspace, size_t workspace_bytes,
                      int64_t capacity, uint32_t* counts, uint32_t* table, uint64_t* stats, void* stream);

/* Ascending LSD radix sort of `count` (device scalar) uint64 keys over the
 * low `bits` bits.  Two-phase scratch query as output_sort
 * (OA/src/output.cu:26-31): call with temp == NULL to get temp_bytes.
 * The sorted keys end in `keys` (ping-pong through temp). */
int tj_sort_keys(void* temp, size_t* temp_bytes, uint64_t* keys, const uint64_t* count,
                 int64_t capacity, int bits, void* stream);

/* Sorted keys -> (i, j) rows of int32 or int64. */
int tj_unpack_pairs(const uint64_t* keys, const uint64_t* count, int64_t capacity, int jbits,
                    void* pairs, int out_format, void* stream);

/* Reference output format: upper pairs of an n-point self-join -> directed
 * u64 IDs i*n+j, both directions plus the n self pairs, sorted.
 * Replaces output_mirror + output_sort (OA/src/output.cu:8-31) and the
 * host canonical() of TG/src/retained_ops.py:17-19.  `ids` needs room for
 * 2*count + n entries; temp as tj_sort_keys (query with temp == NULL). */
int tj_canonical_u64(void* temp, size_t* temp_bytes, const int64_t* pairs_i64,
                     const uint64_t* count, int64_t capacity, int64_t n, uint64_t* ids,
                     void* stream);

/* ------------------------------------------------ host-to-host pipeline
 * The join from pageable host memory to a sorted host result: the public
 * runs of the reference time exactly this path (pageable FP32 source ->
 * device stages -> sorted canonical ids on the host,
 * CEJ/src/run_g5_tensorjoin_public.py:372-547).  Inputs go through pinned
 * staging buffers (parallel host copies + DMA on a copy stream); an R x S
 * join runs over chunks of R rows (chunk_rows, 0 = about nR/8) so that
 * staging and copying out overlap the device work; capacities grow and a
 * chunk re-runs when it overflows.  `threads` host copy threads (0 = up to 8).
 * The library allocates device buffers and pinned staging for the handle
 * (unlike tj_join, whose caller owns all memory). */
typedef struct tj_host tj_host;

int tj_host_create(tj_host** host, int64_t nR, int64_t nS /* 0 = self-join */, int d, int method,
                   int precision, unsigned flags /* TJ_FLAG_AUDIT */, int64_t chunk_rows, int threads);
int tj_host_destroy(tj_host* host);
/* Synchronous.  R, S: host row-major FP32 (pageable is fine).  out_format
 * TJ_OUT_I32 / TJ_OUT_I64 -> *result = P sorted (i, j) rows, *count = P;
 * TJ_OUT_U64_CANONICAL (self-joins) -> *result = the directed ids, *count =
 * 2P + n.  *result is malloc'ed by the library and released with
 * tj_host_free.  stats (host, TJ_STATS_SLOTS, may be NULL): counters summed
 * over chunks, maxima for TJ_ST_MAX_ERR / TJ_ST_MAX_RATIO. */
int tj_host_join(tj_host* host, const float* R, const float* S, double T, int out_format, void** result,
                 int64_t* count, uint64_t* stats);
void tj_host_free(void* result);
/* Host wall-clock milliseconds of the last tj_host_join: ms[0] total,
 * ms[1] staging inputs and enqueueing chunks, ms[2] waiting for the device,
 * ms[3] copying results out. */
int tj_host_times(const tj_host* host, double* ms, int n);

/* ------------------------------------------------------------ diagnostics
 * MFMA probe: nblocks independent 32x32 tiles, each the product of rows
 * A[blk*32 + i] and B[blk*32 + j] (row-major, ksteps MFMA steps of 16
 * (16-bit) or 64 (fp8) elements), accumulated exactly as the screen does;
 * fp8 rows carry power-of-two scale exponents scale_a / scale_b; seed (or
 * NULL) gives the initial accumulator of row blk*32 + i, as the screen seeds
 * its chains.  out[blk][i][j] (binary32).  Used to validate the accumulation model of the
 * certificate and the output layout on the device.  No counterpart in the
 * reference (its premise was inherited, F16C/NUMERICAL_SCOPE.md:11-43). */
int tj_mfma_probe(int precision, const void* A, const void* B, const int* scale_a,
                  const int* scale_b, const float* seed, int ksteps, int nblocks, float* out,
                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TJOIN_H */
