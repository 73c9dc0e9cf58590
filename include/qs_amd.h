/*
 * qs_amd.h -- C ABI of the MI355X (gfx950) block-scatter tensor-contraction
 * library.  Plain C types only: integers, doubles, pointers and sizes.
 *
 * The host planner (QuasiStrided.jl's `plan_contract`, or the Python mirror in
 * this package) resolves labels into M/N/K axis groups exactly as
 * src/planning/plan.jl:197-262 does, then hands the groups, the operand
 * storages and the folded conjugation flags to `qs_plan_create`.  The library
 * owns everything device-side: offset tables, tile choice, scratch.
 *
 * Error convention (replaces Julia exceptions, src/layout/axis_group.jl:13-41,
 * src/planning/labels.jl:28-102, src/microkernels/interface.jl:164-206): every
 * entry point returns a qs_status; a non-zero status leaves every output
 * untouched ("before any write") and `qs_last_error()` holds a message.
 * Nothing throws across the ABI.
 *
 * Threading: a plan is used by one thread at a time (src/planning/plan.jl:12-13);
 * library globals (device context, autotune cache) are thread-safe.  Every
 * execute takes an explicit HIP stream (hipStream_t passed as void*; NULL is
 * the default stream).
 */
#ifndef QS_AMD_H
#define QS_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: one per reference exception class ------------------- */
typedef enum qs_status {
  QS_OK = 0,
  QS_ERR_ARGUMENT = 1,    /* ArgumentError    */
  QS_ERR_DIMENSION = 2,   /* DimensionMismatch */
  QS_ERR_BOUNDS = 3,      /* BoundsError      */
  QS_ERR_OVERFLOW = 4,    /* OverflowError    */
  QS_ERR_DEVICE = 5,      /* a HIP runtime error */
  QS_ERR_MEMORY = 6,      /* device allocation failed or budget exceeded */
  QS_ERR_UNSUPPORTED = 7  /* valid input this build does not implement */
} qs_status;

/* ---- element types (src/planning/plan.jl:264 SUPPORTED_ELTYPES) -------- */
typedef enum qs_dtype {
  QS_F32 = 0,   /* Float32    */
  QS_F64 = 1,   /* Float64    */
  QS_C32 = 2,   /* ComplexF32 (interleaved re, im) */
  QS_C64 = 3    /* ComplexF64 (interleaved re, im) */
} qs_dtype;

#define QS_MAX_DIMS 16

/* One grouped axis with two stride maps: AxisGroup{D,2}
 * (src/layout/axis_group.jl:5-41).  `ndim` dims, enumerated by one zero-based
 * logical coordinate, first dim fastest; strides in elements, zero-based, no
 * base.  M groups carry maps (A, C), N groups (B, C), K groups (A, B) -- after
 * the planner's M/N swap, "A" is whichever operand feeds M. */
typedef struct qs_group {
  int32_t ndim;
  int32_t _pad;
  int64_t lengths[QS_MAX_DIMS];
  int64_t strides[2][QS_MAX_DIMS];
} qs_group;

/* Where an operand's storage lives. */
typedef enum qs_memory {
  QS_MEM_DEVICE = 0,  /* device pointer: the library reads/writes it in place   */
  QS_MEM_HOST = 1     /* host pointer: the library stages the touched span H2D,
                         and C's span back D2H, on each execute */
} qs_memory;

/* Precision modes. */
typedef enum qs_precision {
  QS_PLAIN = 0,       /* native FP64/FP32 MFMA in the compute type            */
  QS_SPLIT_INT8 = 1   /* FP64/FP32 from INT8 residues of scaled integer
                         operands (Ozaki scheme II, CRT) on the INT8 MFMA units */
} qs_precision;

/* Everything `ContractPlan` (src/planning/plan.jl:16-45) carries that the
 * device needs.  Operands: storage pointer, zero-based element offset of the
 * view (StridedViews `offset`), storage length in elements (for the exact span
 * check, src/layout/tiles.jl:123-134). */
typedef struct qs_contract_desc {
  int32_t a_type, b_type, c_type;   /* qs_dtype of each storage              */
  int32_t compute_type;             /* qs_dtype T (plan.jl:278-285)          */
  int32_t conj_a, conj_b;           /* atransform/btransform == conj         */
  int32_t memory;                   /* qs_memory, shared by A, B and C       */
  int32_t precision;                /* qs_precision                          */
  qs_group m, n, k;
  const void *a; int64_t a_base; int64_t a_len;
  const void *b; int64_t b_base; int64_t b_len;
  void *c;       int64_t c_base; int64_t c_len;
  double target_rel_err;            /* split precision: error to meet (0: default) */
  int32_t tile;                     /* tile-config override, -1 = automatic  */
  int32_t path_mode;                /* 0 auto; 1 force the tiled MFMA GEMM path */
  int32_t k_splits;                 /* split-K override, 0 = automatic       */
  int32_t _reserved;
} qs_contract_desc;

typedef struct qs_plan qs_plan;

/* The execution path a plan runs (src/execution/paths.jl:1-74 counterparts). */
typedef enum qs_path {
  QS_PATH_EMPTY = 0,   /* M or N empty: nothing to do                    */
  QS_PATH_SCALE = 1,   /* K empty: C = beta * C                          */
  QS_PATH_GEMM = 2,    /* tiled MFMA block-scatter GEMM, alpha/beta epilogue */
  QS_PATH_DOT = 3,     /* M == 1 or N == 1: bandwidth-bound reduction     */
  QS_PATH_OUTER = 4,   /* K == 1: bandwidth-bound streaming outer product */
  QS_PATH_SKINNY = 5,  /* small M: B streamed once, read in place         */
  QS_PATH_SPLIT = 6    /* split-precision (INT8 residue) GEMMs            */
} qs_path;

/* What a plan decided (for tests, logs and benchmarks). */
typedef struct qs_plan_info {
  int32_t path;          /* qs_path */
  int32_t tile_m, tile_n, tile_k;
  int32_t k_splits;      /* K partitions reduced in a fixed order (1 = none) */
  int32_t a_kfast, b_kfast;  /* operand gathers coalesced along K (1) or along M/N (0) */
  int32_t moduli;        /* split precision: CRT moduli (INT8 GEMMs)      */
  int64_t m, n, k;       /* axis lengths */
  int64_t workspace_bytes;
  int32_t paired_loads;  /* 1: the gathers load two adjacent elements per lane (real,
                            one storage type, tiles up to 64 x 128, spans below 2 GiB) */
  int32_t mixed_domain;  /* 1: complex A x real B, 2: real A x complex B, two MFMAs per
                            block (ComplexRealKernel / RealComplexKernel); 0 otherwise */
} qs_plan_info;

/* The gfx950 TargetProfile (counterpart of src/hardware/target.jl:26-47). */
typedef struct qs_target_profile {
  char name[64];            /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
  int32_t compute_units;
  int32_t wavefront_size;
  int32_t lds_bytes_per_cu;
  int32_t lds_bytes_per_block;
  int32_t clock_khz;
  int32_t memory_clock_khz;
  int32_t memory_bus_bits;
  int32_t l2_bytes;
  int64_t hbm_bytes;
  int32_t xcds;
  int32_t mfma_f64_m, mfma_f64_n, mfma_f64_k;   /* 16x16x4  */
  int32_t mfma_f32_m, mfma_f32_n, mfma_f32_k;   /* 16x16x4  */
  int32_t mfma_i8_m, mfma_i8_n, mfma_i8_k;      /* 16x16x64 */
  double peak_fp64_tflops, peak_fp32_tflops, peak_int8_tops, peak_hbm_gbs;
} qs_target_profile;

/* ---- workspace allocator --------------------------------------------------
 * Counterpart of plan_contract's `allocator` (src/planning/plan.jl:125-131)
 * and of ContractWorkspace's tensoralloc / release! (src/execution/
 * workspace.jl:83-123): the plan's device workspace -- split-K partial sums
 * and the split-precision residues, scales and samples -- comes from the
 * caller's allocator and goes back to it in reverse acquisition order, on
 * qs_plan_release or qs_plan_destroy.  The offset tables stay the library's,
 * as the reference's offset vectors are plain Vectors.  A Julia caller wraps
 * a TensorOperations allocator over device arrays in two @cfunction
 * callbacks; the Python mirror bridges torch's caching allocator. */
typedef struct qs_allocator {
  /* `bytes` of device memory, aligned to 256 bytes, or NULL. */
  void *(*alloc)(void *ctx, size_t bytes);
  void (*free)(void *ctx, void *ptr, size_t bytes);
  void *ctx;
} qs_allocator;

/* qs_plan_create with the workspace from `allocator` (NULL: the library's). */
qs_status qs_plan_create_with_allocator(const qs_contract_desc *desc, const qs_allocator *allocator,
                                        qs_plan **plan);

/* A device memory pool (hipMemPool_t) on `device` as a qs_allocator, for
 * callers with no allocator of their own; no counterpart in the reference,
 * whose DefaultAllocator is Julia's GC.  Blocks are taken with
 * hipMallocFromPoolAsync on the null stream and waited for; they go back with
 * hipFreeAsync when a plan releases its workspace (after the library's device
 * synchronization).  The pool keeps up to `release_threshold` bytes reserved
 * (UINT64_MAX: everything it has grown to), so plans built and destroyed in a
 * loop reuse its memory.  `out->ctx` owns the pool: destroy it once no plan
 * holds a block (QS_ERR_ARGUMENT otherwise). */
qs_status qs_pool_allocator_create(int device, uint64_t release_threshold, qs_allocator *out);
/* Bytes the pool holds from the device and bytes handed out now, and the
 * number of blocks plans hold. */
qs_status qs_pool_allocator_stats(const qs_allocator *allocator, uint64_t *reserved_bytes, uint64_t *used_bytes,
                                  int64_t *live_blocks);
qs_status qs_pool_allocator_destroy(qs_allocator *allocator);

/* release!(plan, allocator): hand the workspace back now, split-precision
 * buffer first, then the split-K partials.  The plan's info stays readable;
 * qs_execute and qs_autotune on it then fail with QS_ERR_ARGUMENT, and
 * qs_plan_destroy frees the rest. */
qs_status qs_plan_release(qs_plan *plan);

/* Last error message of the calling thread ("" if none). */
const char *qs_last_error(void);

/* Library version string. */
const char *qs_version(void);

/* Fill `out` from hipGetDeviceProperties of the current device plus the
 * gfx950 MFMA table.  Counterpart of `target_profile()`
 * (src/hardware/target.jl:181-191); never fails on a gfx950 device. */
qs_status qs_target_profile_get(qs_target_profile *out);

/* Measured peaks of the current device, reported beside the datasheet ones
 * (the roofline is priced against both): the FP64 MFMA rate that independent
 * chains of v_mfma_f64_16x16x4_f64 sustain on every SIMD, in TFLOP/s, and the
 * rate of a device-to-device copy of 1 GiB, in GB/s (read + write bytes).
 * Allocates 2 GiB of device memory for the copy while it runs.  Counterpart
 * of the facts `target_profile()` detects (src/hardware/target.jl:181-191). */
qs_status qs_target_measure(double *fp64_tflops, double *copy_gbs, void *stream);

/* Validate `desc` and build a plan: the axis-group overflow policy
 * (axis_group.jl:27-41), eltype/compute-type rules (plan.jl:278-295), offset
 * tables (fill_offsets!, axis_group.jl:87-145) uploaded once, the exact span of
 * every operand (tiles.jl:123-134), and the path/tile choice.
 * Counterpart of `plan_contract` after label resolution (plan.jl:115-133). */
qs_status qs_plan_create(const qs_contract_desc *desc, qs_plan **out);

/* C = alpha * sum_K op(A) op(B) + beta * C.  alpha and beta point to values of
 * the compute type (complex: two doubles or floats, re then im).  beta == 0
 * never reads C; alpha == 0 or K == 0 never reads A or B; beta is applied
 * exactly once.  Operand spans are checked before any launch (BoundsError
 * with nothing written).  Counterpart of `execute!` (execute.jl:11-20). */
qs_status qs_execute(qs_plan *plan, const void *alpha, const void *beta, void *stream);

/* Plan introspection. */
qs_status qs_plan_get_info(const qs_plan *plan, qs_plan_info *out);

/* Point an existing plan at other storages of the same layout (device
 * memory only).  The new spans are checked exactly by the next execute
 * (BoundsError before any write; alpha == 0 never reads A or B, so their
 * spans are checked only when alpha != 0).  Lets one plan serve many
 * requests. */
qs_status qs_plan_rebind(qs_plan *plan, const void *a, int64_t a_base, int64_t a_len,
                         const void *b, int64_t b_base, int64_t b_len,
                         void *c, int64_t c_base, int64_t c_len);

/* Free a plan and its device tables.  Counterpart of `release!`
 * (src/execution/workspace.jl:110-123).  NULL is a no-op. */
qs_status qs_plan_destroy(qs_plan *plan);

/* ---- permuted add and partial trace ----------------------------------------
 * Device counterparts of TensorOperations' `tensoradd!` / `tensortrace!`,
 * which the reference forwards to `StridedNative`
 * (src/integrations/tensoroperations.jl:90-107):
 *
 *     C[m, n] = alpha * sum_k op(A)[m, n, k] + beta * C[m, n]
 *
 * over the groups M and N (two maps each: A, then C) and K, the traced pairs
 * (map 0: the sum of the pair's two strides in A; map 1 unused).  The host
 * puts C's unit-stride dim first in M and A's first in N; a transposing
 * permutation is then coalesced on both sides.  op(A) = conj(A) when
 * `conj_a`.  Device memory only; alpha == 0 or K == 0 never reads A, beta ==
 * 0 never reads C; exact span checks before any write, as for contractions. */
typedef struct qs_add_desc {
  int32_t a_type, c_type, compute_type, conj_a;
  qs_group m, n, k;
  const void *a; int64_t a_base; int64_t a_len;
  void *c;       int64_t c_base; int64_t c_len;
} qs_add_desc;

typedef struct qs_add_plan qs_add_plan;

qs_status qs_add_plan_create(const qs_add_desc *desc, qs_add_plan **out);
/* alpha and beta hold one value of the compute type (complex: 2 reals). */
qs_status qs_add_execute(qs_add_plan *plan, const void *alpha, const void *beta, void *stream);
qs_status qs_add_plan_destroy(qs_add_plan *plan);
/* Point an add plan at other storages of the same layout (spans are checked
 * at execute).  No reference counterpart (StridedNative plans nothing):
 * repeated permutes and traces of one layout reuse the device tables. */
qs_status qs_add_plan_rebind(qs_add_plan *plan, const void *a, int64_t a_base, int64_t a_len,
                             void *c, int64_t c_base, int64_t c_len);

/* ---- split precision ------------------------------------------------------
 * No counterpart in the reference.  A plan created with precision
 * QS_SPLIT_INT8 (real compute types) runs Ozaki scheme II on the INT8 MFMA
 * units: A and B scaled row / column-wise to integers whose norms bound every
 * entry of the integer product below P / 2, one exact INT8 GEMM per modulus
 * of P (pairwise coprime moduli <= 256, their number chosen from
 * `target_rel_err` -- default 1e-13 for F64, 1e-6 for F32 -- and K), and
 * FP64 reconstruction by the Chinese remainder theorem.  With the report
 * switched on, every execute also evaluates sampled entries with Dot2
 * (compensated dot product) on the same operands and records the delivered
 * error. */
typedef struct qs_error_report {
  int32_t moduli;     /* CRT moduli (INT8 GEMMs)               */
  int32_t samples;    /* entries evaluated in the last execute  */
  double rel_fro;     /* sampled relative Frobenius error       */
  double max_rel;     /* largest sampled relative error         */
  double target;      /* the requested error (0 = default)      */
  int32_t exceeded;   /* 1: the last execute's sampled error (of the result
                         delivered) is above the target               */
  int32_t escalations;/* guarded mode: reruns with more moduli in the last execute */
  int32_t fallback;   /* guarded mode: 1 if the last execute was finished by
                         the plain FP64 / FP32 MFMA kernel              */
  int32_t _pad;
} qs_error_report;

/* The error bound.  The moduli count makes every entry satisfy
 * |C_ij - C_exact,ij| <= target * ||A_i||_2 ||B_j||_2 (row i of op(A) and
 * column j of op(B) over K; plus the final rounding), a normwise bound, like
 * the reference's norm-wise tolerance of mixed-precision results
 * (test/execution/test_mixed.jl:50).  For entries of random sign that is an
 * error about target / (2 sqrt(K)) of ||A_i|| ||B_j||, i.e. about target
 * relative to C.  Under cancellation (|C_ij| << ||A_i|| ||B_j||) the error
 * relative to C grows by the cancellation factor, as plain FP64's does (by
 * about 7x less at K = 4096).
 * `on`: 0 off; 1 every execute evaluates the report; 2 guarded -- an execute
 * whose sampled error relative to C exceeds the target is redone (from C as
 * it was, kept in the workspace when beta != 0) with more moduli, up to 15,
 * and then with the plain MFMA kernel; the raised moduli count stays for
 * later executes. */
qs_status qs_plan_set_error_report(qs_plan *plan, int32_t on);
qs_status qs_plan_get_error_report(const qs_plan *plan, qs_error_report *out);

/* ---- autotuner ----------------------------------------------------------
 * No counterpart in the reference (whose blocking is an analytic cache model,
 * src/planning/blocking.jl:66-77): the GPU tile and split-K factor are chosen
 * by timing candidates.  Results are kept per shape class (compute and
 * storage types, gather directions, half-octave buckets of M, N, K) in a
 * process-wide cache that `qs_plan_create` consults, and that can be saved
 * to and loaded from a text file. */
typedef struct qs_tune_result {
  int32_t tile, k_splits;        /* chosen configuration            */
  double best_ms;                /* its mean time per execute       */
  double default_ms;             /* the untuned plan's time         */
  int32_t candidates;            /* configurations timed            */
  int32_t _pad;
} qs_tune_result;

/* Time every candidate (tile, split-K) on the plan's own operands (device
 * plans only; results go to a scratch copy of C's span, so C is not
 * modified), adopt the fastest for this plan and record it for its shape
 * class. `reps` timed executes per candidate (0 = 5). */
qs_status qs_autotune(qs_plan *plan, void *stream, int32_t reps, qs_tune_result *out);

/* Save / load / clear the shape-class cache (text, one entry per line). */
qs_status qs_autotune_save(const char *path);
qs_status qs_autotune_load(const char *path);
qs_status qs_autotune_clear(void);

/* ---- grouped execution ----------------------------------------------------
 * No counterpart in the reference, which runs one contraction per call
 * (src/execution/execute.jl:11-20): tensor-network sweeps issue many small
 * contractions of uneven shape, and a group runs all of them in one
 * persistent launch per compute type and gather direction.  Members are
 * existing device-memory plans on the tiled GEMM path.  They must outlive the
 * group and every graph captured from it.  A member rebound after the group
 * was built (qs_plan_rebind) has the group's device tables rebuilt at the
 * next execute or capture; a graph replays the storages it was captured
 * with.  Every member's spans are checked exactly before an execute or a
 * capture launches anything (QS_ERR_BOUNDS).  Each member computes all of K
 * in registers, so its result does not depend on the other members. */
typedef struct qs_group_plan qs_group_plan;
typedef struct qs_graph qs_graph;

/* `tiles`: per-member workgroup tile (qs TileId 0..5, -1 = the grouped
 * heuristic), or NULL for all heuristic.  The grouped kernel runs 128x64,
 * 64x64, 32x64 and (Float64) 32x32 tiles; the other ids map onto the nearest
 * of these. */
qs_status qs_group_create(qs_plan *const *plans, int32_t n, const int32_t *tiles, qs_group_plan **out);

/* A batched contraction as a group: n members that all run `plan`, member i
 * with a_offsets[i] / b_offsets[i] / c_offsets[i] (elements) added to the
 * plan's A / B / C bases (every member's spans are checked at create:
 * QS_ERR_BOUNDS).  This is the device form of batch labels (a label
 * in A, B and C), which the reference rejects (src/planning/labels.jl:48-54,
 * the check `plan_contract` keeps); it is an extension with no counterpart
 * there.  Members' C spans must not overlap.  `tile` as in qs_group_create. */
qs_status qs_group_create_batched(qs_plan *plan, int32_t n, const int64_t *a_offsets, const int64_t *b_offsets,
                                  const int64_t *c_offsets, int32_t tile, qs_group_plan **out);

/* Run every member: C_i = alpha_i A_i B_i + beta_i C_i.  alphas / betas hold
 * two doubles per member (re, im; im ignored for real members); NULL means
 * alpha = 1 / beta = 0.  Spans of every member are checked before launch. */
qs_status qs_group_execute(qs_group_plan *g, const double *alphas, const double *betas, void *stream);

/* Per-shape-class tiles for grouped launches.  No counterpart in the
 * reference (whose register tile is chosen per shape from its profile,
 * src/planning/kernel_selection.jl:232-256).  The whole group runs (into a
 * scratch C: the members' C is not touched) once with every member on each
 * grouped tile class, the kernel timing each member's tiles inside that mixed
 * launch; per shape class (types, gather directions, octave buckets of M, N
 * and K) the tile whose members took the least time wins.  The tuned
 * assignment is kept only if the whole group runs faster with it than with
 * the padding heuristic.  Chosen tiles go to the autotune cache, which
 * qs_group_create consults and qs_autotune_save / qs_autotune_load persist,
 * and the group's tables are rebuilt.  Results do not change (every tile
 * sums K in the same order).  `reps` timed runs per whole-group timing
 * (0 = 3). */
typedef struct qs_group_tune_result {
  int32_t classes;          /* shape classes tuned                          */
  int32_t candidates;       /* tile classes tried (whole-group runs)        */
  int32_t members_changed;  /* members whose tile differs from before       */
  int32_t _pad;
  double before_ms;         /* the whole group with the padding heuristic's tiles */
  double after_ms;          /* the whole group with the tiles kept          */
} qs_group_tune_result;

qs_status qs_group_autotune(qs_group_plan *g, void *stream, int32_t reps, qs_group_tune_result *out);

/* Algorithmic flops of one execute (2 per real MAC, 8 per complex MAC). */
qs_status qs_group_flops(const qs_group_plan *g, int64_t *flops);

/* How the group runs (its device tables rebuilt first if a member was
 * rebound): `launches` grouped persistent launches (one per compute type,
 * gather direction and domain) and `standalone` members that take a launch
 * of their own (storage types other than the compute type's, except complex
 * times real with every type native, or operand spans past 2 GiB).  No
 * counterpart in the reference, which runs every contraction on its own. */
qs_status qs_group_stats(qs_group_plan *g, int32_t *launches, int32_t *standalone);

qs_status qs_group_destroy(qs_group_plan *g);

/* Capture `repeats` back-to-back executes of the group (with the given
 * coefficients, uploaded before capture) into a HIP graph. */
qs_status qs_group_capture(qs_group_plan *g, const double *alphas, const double *betas, int32_t repeats,
                           void *stream, qs_graph **out);
qs_status qs_graph_launch(qs_graph *gr, void *stream);
qs_status qs_graph_destroy(qs_graph *gr);

/* ---- serving -------------------------------------------------------------
 * No counterpart in the reference.  A continuous-batching request loop over
 * operands and plans resident in HBM: requests are admitted first-come
 * first-served while the bytes they hold (their operand spans) stay within
 * `budget_bytes`; admitted requests are packed into grouped launches on one of
 * four streams, so new work overlaps work already on the device.  Each request
 * computes its K sum unsplit with the tile of its own shape, so its result is
 * the same alone and in any mix.  `grouped == 0` runs every request as its
 * own launch, in arrival order on one stream (the per-request baseline). */
typedef struct qs_server qs_server;
typedef struct qs_server_stats {
  int64_t submitted, completed, rejected, batches;
  double mean_batch;
  double p50_ms, p90_ms, p99_ms, max_ms, mean_ms;   /* submit -> completion observed */
  double span_s;            /* first submit -> last completion */
  double throughput_rps, tflops;
  int64_t peak_bytes, budget_bytes;
} qs_server_stats;

qs_status qs_server_create(int64_t budget_bytes, int32_t max_batch, int32_t grouped, qs_server **out);

/* Submit `plan` against the given storages (NULL: the plan's own) with
 * alpha / beta as two doubles each (re, im).  The plan must outlive the
 * request; the storages must stay resident until it completes.  A request
 * whose spans alone exceed the budget is rejected with QS_ERR_MEMORY. */
qs_status qs_server_submit(qs_server *s, qs_plan *plan, const void *a, int64_t a_base, int64_t a_len,
                           const void *b, int64_t b_base, int64_t b_len, void *c, int64_t c_base, int64_t c_len,
                           const double *alpha, const double *beta, uint64_t *ticket);
qs_status qs_server_poll(qs_server *s, uint64_t ticket, int32_t *done, double *latency_ms);
qs_status qs_server_wait(qs_server *s, uint64_t ticket, double *latency_ms);
qs_status qs_server_drain(qs_server *s);
qs_status qs_server_stats_get(qs_server *s, qs_server_stats *out);
qs_status qs_server_reset_stats(qs_server *s);
qs_status qs_server_destroy(qs_server *s);

/* ---- a contraction that allocates nothing --------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc.  Plans and runs `desc` in one call (qs_plan_create +
 * qs_execute + qs_plan_destroy) with no device or host allocation on the
 * success path: the offset tables are written by a kernel on `stream` into a
 * fixed array of 4 MiB per GPU, so M + N + K <= 261824 (larger groups:
 * QS_ERR_UNSUPPORTED), and K runs unsplit.  Device memory only; split
 * precision is refused (QS_ERR_UNSUPPORTED).  Same validation, span checks
 * and kernels as qs_execute: results are bitwise those of an unsplit plan.
 * Every call on one GPU shares that array and leaves its tables there: calls
 * that overlap on one GPU (two streams, two threads) race on it. */
qs_status qs_contract_noalloc(const qs_contract_desc *desc, const void *alpha, const void *beta, void *stream);

/* ---- allocation counts ----------------------------------------------------
 * No counterpart in the reference, whose steady-state `execute!` allocates
 * nothing (test/execution/test_execute.jl:72-94).  Every device and pinned
 * host allocation the library makes, and every workspace block it takes from
 * a caller's qs_allocator, is counted since the library was loaded; a test
 * reads the counts around repeated executes of one plan and expects them to
 * stay put. */
typedef struct qs_alloc_stats {
  int64_t device_allocs, device_frees;   /* hipMalloc / hipFree calls that succeeded */
  int64_t device_bytes;                  /* bytes of those hipMallocs              */
  int64_t pinned_allocs, pinned_frees;   /* hipHostMalloc / hipHostFree            */
  int64_t bridge_allocs, bridge_frees;   /* qs_allocator alloc / free callbacks    */
} qs_alloc_stats;

qs_status qs_alloc_stats_get(qs_alloc_stats *out);

/* ---- fault injection --------------------------------------------------------
 * No counterpart in the reference, which has no allocation that can fail
 * partway (its errors are argument errors raised before any write).  For
 * tests of the library's own error paths: arms a one-shot failure of the
 * (after + 1)-th device allocation (QS_FAULT_DEVICE_ALLOC) or pinned host
 * allocation (QS_FAULT_PINNED_ALLOC) the library makes from now on.  That
 * allocation returns out-of-memory without calling the runtime, so the call
 * that made it returns QS_ERR_MEMORY (or QS_ERR_DEVICE where the allocation
 * is not the caller's memory) with nothing written to C and nothing leaked.
 * after < 0 disarms.  One counter per kind, shared by every thread. */
typedef enum qs_fault_kind { QS_FAULT_DEVICE_ALLOC = 0, QS_FAULT_PINNED_ALLOC = 1 } qs_fault_kind;

qs_status qs_fault_inject(int32_t kind, int64_t after);

/* Host-side offset enumeration, exactly `fill_offsets!`
 * (src/layout/axis_group.jl:87-145) for a two-map group: writes `count`
 * offsets of coordinates [first, first+count) into out0/out1.  Exposed so a
 * caller can check a group against its own indexing. */
qs_status qs_fill_offsets(const qs_group *g, int64_t first, int64_t count,
                          int64_t *out0, int64_t *out1);

#ifdef __cplusplus
}
#endif

#endif /* QS_AMD_H */
