/*
 * rocblas_amd_ext.h -- entry points beside the rocBLAS trsm ABI.
 *
 * These have no counterpart in the reference library.  They use rocBLAS's
 * types, enums and status codes (include/rocblas_amd.h) and its argument
 * conventions, so they sit next to rocblas_{s,d}trsm_batched without new
 * concepts.
 */
#ifndef ROCBLAS_AMD_EXT_H
#define ROCBLAS_AMD_EXT_H

#include "rocblas_amd.h"

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Distribution of a synthetic triangle (rocblas_amd_{s,d}synth_tri). */
typedef enum rocblas_amd_synth_dist_
{
    /* diagonal U[1,2], off-diagonal U[-1,1]/k */
    rocblas_amd_synth_uniform = 0,
    /* diagonal U[1,2], off-diagonal N(0,1)/sqrt(k) (Box-Muller) */
    rocblas_amd_synth_normal = 1
} rocblas_amd_synth_dist;

/* Variable-size batched triangular solve: instance i solves
 * op(A_i) X_i = alpha B_i (side left) or X_i op(A_i) = alpha B_i (side right)
 * with its own m[i], n[i], lda[i], ldb[i].  m, n, lda, ldb, A and B are device
 * arrays of batch_count entries; alpha is one scalar (host or device per the
 * pointer mode).  A one-workgroup planning launch orders the work by cost and
 * two persistent launches solve it (k <= 512 with the MFMA panel solve, larger
 * triangles one wavefront each); the call allocates nothing beyond the handle
 * workspace (36 bytes per instance + 64; batch_count < 2^28) and can be
 * captured in a HIP graph.  No counterpart in the reference: rocBLAS callers
 * loop over rocblas_dtrsm.  The per-instance sizes are device data, so they
 * are checked on the device: an instance with m[i] <= 0, n[i] <= 0,
 * lda[i] < k (k = m[i] for side left, n[i] for side right) or ldb[i] < m[i] is
 * skipped and its B left untouched. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_strsm_vbatched(rocblas_handle     handle,
                                                             rocblas_side       side,
                                                             rocblas_fill       uplo,
                                                             rocblas_operation  transA,
                                                             rocblas_diagonal   diag,
                                                             const rocblas_int* m,
                                                             const rocblas_int* n,
                                                             const float*       alpha,
                                                             const float* const A[],
                                                             const rocblas_int* lda,
                                                             float* const       B[],
                                                             const rocblas_int* ldb,
                                                             rocblas_int        batch_count);

ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_dtrsm_vbatched(rocblas_handle      handle,
                                                             rocblas_side        side,
                                                             rocblas_fill        uplo,
                                                             rocblas_operation   transA,
                                                             rocblas_diagonal    diag,
                                                             const rocblas_int*  m,
                                                             const rocblas_int*  n,
                                                             const double*       alpha,
                                                             const double* const A[],
                                                             const rocblas_int*  lda,
                                                             double* const       B[],
                                                             const rocblas_int*  ldb,
                                                             rocblas_int         batch_count);

/* rocblas_amd_dtrsm_vbatched that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, checks and results (bit for bit) as rocblas_amd_dtrsm_vbatched,
 * but the planning scratch (64 + 32 * batch_count bytes) lives in a
 * library-owned device array of 2,097,216 bytes per GPU, held until the
 * process ends, instead of the handle workspace: batch_count may be at most
 * 65536 (else rocblas_status_invalid_size), and a size query answers
 * rocblas_status_size_unchanged.  The first call on a device looks up the
 * array's address there; no call allocates, frees or touches the handle
 * workspace.  Every call on one GPU shares that array and each call rewrites
 * it before use, so calls whose executions overlap on one GPU (other
 * streams, other handles, concurrent graph replays) race and must be
 * serialised by the caller. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_dtrsm_vbatched_noalloc(rocblas_handle      handle,
                                                                     rocblas_side        side,
                                                                     rocblas_fill        uplo,
                                                                     rocblas_operation   transA,
                                                                     rocblas_diagonal    diag,
                                                                     const rocblas_int*  m,
                                                                     const rocblas_int*  n,
                                                                     const double*       alpha,
                                                                     const double* const A[],
                                                                     const rocblas_int*  lda,
                                                                     double* const       B[],
                                                                     const rocblas_int*  ldb,
                                                                     rocblas_int batch_count);

/* C := alpha op(A) op(B) + beta C with the MFMA kernel that carries the
 * trailing updates of the blocked solve.  Same signature and argument rules
 * as rocblas_{s,d}gemm (reference rocblas-functions.h:20047-20060); beta == 0
 * does not read C.  Prefixed because only the GEMM shapes the solve uses are
 * tuned. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_sgemm(rocblas_handle    handle,
                                                    rocblas_operation transA,
                                                    rocblas_operation transB,
                                                    rocblas_int       m,
                                                    rocblas_int       n,
                                                    rocblas_int       k,
                                                    const float*      alpha,
                                                    const float*      A,
                                                    rocblas_int       lda,
                                                    const float*      B,
                                                    rocblas_int       ldb,
                                                    const float*      beta,
                                                    float*            C,
                                                    rocblas_int       ldc);
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_dgemm(rocblas_handle    handle,
                                                    rocblas_operation transA,
                                                    rocblas_operation transB,
                                                    rocblas_int       m,
                                                    rocblas_int       n,
                                                    rocblas_int       k,
                                                    const double*     alpha,
                                                    const double*     A,
                                                    rocblas_int       lda,
                                                    const double*     B,
                                                    rocblas_int       ldb,
                                                    const double*     beta,
                                                    double*           C,
                                                    rocblas_int       ldc);

/* Low-precision form of the trailing updates of rocblas_amd_dtrsm_mixed. */
typedef enum rocblas_amd_precision_form_
{
    /* FP32 copy of A, FP32 MFMA products accumulated in FP32 */
    rocblas_amd_form_fp32 = 0,
    /* split-BF16x3: each operand value split into BF16 halves hi + lo, the
     * product hi*hi + hi*lo + lo*hi on BF16 MFMA accumulated in FP32 */
    rocblas_amd_form_bf16x3 = 1
} rocblas_amd_precision_form;

/* Outcome of rocblas_amd_dtrsm_mixed (host struct, optional). */
typedef struct rocblas_amd_refine_info_
{
    rocblas_int iterations;     /* refinement steps applied after the first solve */
    double      backward_error; /* |alpha B - op(A) X|_1 / (|op(A)|_1 |X|_1 + |alpha| |B|_1)
                                   of the returned X; -1 when tol <= 0 (not measured) */
    rocblas_int converged;      /* 1 when backward_error <= tol (or tol <= 0) */
} rocblas_amd_refine_info;

/* FP64 triangular solve, same problem and argument rules as rocblas_dtrsm,
 * computed with low-precision trailing updates and FP64 iterative refinement:
 * FP64 inverses of the block x block diagonal blocks of A (trtri), trailing
 * updates on FP32 MFMA (form rocblas_amd_form_fp32) or split-BF16x3 MFMA
 * (rocblas_amd_form_bf16x3), FP64 residual alpha B - op(A) X, and corrections
 * until the normwise backward error is <= tol or max_refine steps are taken.
 * block = 0 picks the block size tuned for the form.  tol <= 0 runs exactly max_refine
 * steps without measuring (no host synchronisation); tol > 0 synchronises
 * the stream once per step to test convergence.  Workspace (library-managed
 * handles grow on demand): 2 m n + k^2/2 + nblocks block^2 doubles.
 * No counterpart in the reference. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_dtrsm_mixed(rocblas_handle             handle,
                                                          rocblas_side               side,
                                                          rocblas_fill               uplo,
                                                          rocblas_operation          transA,
                                                          rocblas_diagonal           diag,
                                                          rocblas_int                m,
                                                          rocblas_int                n,
                                                          const double*              alpha,
                                                          const double*              A,
                                                          rocblas_int                lda,
                                                          double*                    B,
                                                          rocblas_int                ldb,
                                                          rocblas_amd_precision_form form,
                                                          rocblas_int                block,
                                                          rocblas_int                max_refine,
                                                          double                     tol,
                                                          rocblas_amd_refine_info*   info);

/* Seeded synthetic inputs generated in place on the handle's stream, with the
 * counter-based SplitMix64 streams of rocblas_amd/synth.py: the k x k
 * triangle `uplo` of A (every entry times `scale`; the other triangle and rows
 * k..lda-1 set to `poison`), and an m x n right-hand side U[-1,1] (rows
 * m..ldb-1 set to zero) drawn from stream `tag` (2 for B).  The uniform
 * distribution is bit-identical to the host generator.  No counterpart in the
 * reference (its clients fill matrices on the host, rocblas_init.hpp). */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_dsynth_tri(rocblas_handle         handle,
                                                         rocblas_fill           uplo,
                                                         rocblas_amd_synth_dist dist,
                                                         rocblas_int            k,
                                                         uint64_t               seed,
                                                         double                 scale,
                                                         double                 poison,
                                                         double*                A,
                                                         rocblas_int            lda);
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_ssynth_tri(rocblas_handle         handle,
                                                         rocblas_fill           uplo,
                                                         rocblas_amd_synth_dist dist,
                                                         rocblas_int            k,
                                                         uint64_t               seed,
                                                         double                 scale,
                                                         double                 poison,
                                                         float*                 A,
                                                         rocblas_int            lda);
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_dsynth_rhs(rocblas_handle handle,
                                                         rocblas_int    m,
                                                         rocblas_int    n,
                                                         uint64_t       seed,
                                                         rocblas_int    tag,
                                                         double*        B,
                                                         rocblas_int    ldb);
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_ssynth_rhs(rocblas_handle handle,
                                                         rocblas_int    m,
                                                         rocblas_int    n,
                                                         uint64_t       seed,
                                                         rocblas_int    tag,
                                                         float*         B,
                                                         rocblas_int    ldb);

/* ---- serving loop ---------------------------------------------------------
 * A server owns one device arena of hbm_budget_bytes (its admission budget)
 * and `streams` streams with a handle each.  rocblas_amd_server_run replays a
 * trace of requests.  Request i solves op(A) X = alpha B (side left) or
 * X op(A) = alpha B (side right) with its own side, uplo, transA, diag and
 * alpha, a synthetic triangle A of order m (the uniform distribution of
 * rocblas_amd_{s,d}synth_tri under factor_seed, in the request's uplo; kept
 * resident per factor_id and evicted least recently used when the arena is
 * full) and n synthetic right-hand-side vectors (rocblas_amd_{s,d}synth_rhs
 * under rhs_seed; B is m x n for side left, n x m for side right).
 * Requests are admitted FIFO once they have arrived and their memory fits,
 * up to max_inflight at once; those with m <= small_k are solved together in
 * one variable-size batched launch per precision (every case and alpha in
 * the same launch) on the first stream, larger ones by rocblas_{s,d}trsm
 * over the other streams in turn, so small requests never wait behind large
 * solves and new work overlaps work in flight.  With verify != 0 every
 * request's normwise backward error
 *   ||alpha B - op(A) X||_1 / (||op(A)||_1 ||X||_1 + |alpha| ||B||_1)
 * is computed on the device in FP64 right after its solve, as part of the
 * request (counted in its latency), and returned per request.  latency[i] is
 * the host time from arrival to observed completion (-1 if rejected).  No
 * counterpart in the reference. */
typedef struct rocblas_amd_server_* rocblas_amd_server;

typedef struct rocblas_amd_request_
{
    int64_t  factor_id;   /* requests with one id share one resident factor */
    uint64_t factor_seed; /* seed of the factor when it is generated */
    uint64_t rhs_seed;    /* seed of B */
    double   arrival;     /* seconds after the start of the run; non-decreasing */
    int32_t  m;           /* triangle order */
    int32_t  n;           /* right-hand-side vectors */
    int32_t  precision;   /* 0: FP64, 1: FP32 */
    int32_t  keep;        /* keep X on the device for rocblas_amd_server_result */
    int32_t  side;        /* rocblas_side (0: left) */
    int32_t  uplo;        /* rocblas_fill (0: lower); one factor id, one uplo */
    int32_t  transA;      /* rocblas_operation (0: none) */
    int32_t  diag;        /* rocblas_diagonal (0: non-unit) */
    double   alpha;       /* 0 gives X = 0 without reading A */
} rocblas_amd_request;

typedef struct rocblas_amd_serve_stats_
{
    int64_t completed;
    int64_t rejected; /* larger than the whole arena */
    int64_t launches;
    int64_t batched_launches;
    int64_t factor_hits;
    int64_t factor_misses;
    int64_t evictions;
    int64_t checked;        /* requests whose backward error was computed */
    int64_t over_tolerance; /* of those, above 1e-12 (FP64) or 1e-5 (FP32) */
    double  wall_seconds;
    double  p50_latency; /* seconds */
    double  p99_latency;
    double  max_latency;
    double  mean_latency;
    double  peak_arena_bytes;
    double  max_backward_error_fp64;
    double  max_backward_error_fp32;
} rocblas_amd_serve_stats;

ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_server_create(rocblas_amd_server* server,
                                                            size_t      hbm_budget_bytes,
                                                            rocblas_int streams,
                                                            rocblas_int max_inflight,
                                                            rocblas_int small_k,
                                                            rocblas_int verify);
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_server_destroy(rocblas_amd_server server);
/* backward_error (host, count entries, may be null): per request, -1 when not
 * computed (verify off, or rejected). */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_server_run(rocblas_amd_server         server,
                                                         const rocblas_amd_request* requests,
                                                         int64_t                    count,
                                                         double*                    latency,
                                                         double*                    backward_error,
                                                         rocblas_amd_serve_stats*   stats);
/* Copies the kept X of request `index` of the last run to host memory X
 * (B's shape: m x n for side left, n x m for side right; leading dimension
 * ldx; in the request's precision). */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_server_result(rocblas_amd_server server,
                                                            int64_t            index,
                                                            void*              X,
                                                            rocblas_int        ldx);

/* Overrides the tuning of the _vbatched solves for this process: every
 * instance is split into up to panel_groups work items (1..8) of the panel
 * solve; 0 restores the built-in table.  Used by the autotuner
 * (rocblas_amd/autotune_vbatched.py); not thread-safe against concurrent
 * _vbatched calls.  No counterpart in the reference. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_vbatched_set_tuning(rocblas_int panel_groups);

/* Per-size-bucket override of the same table: instances with k in bucket
 * (0: k <= 64, 1: <= 128, 2: <= 256, 3: larger) are split into panel_groups
 * work items (1..8; 0 restores the built-in value).  rocblas_amd_vbatched_set_tuning
 * with panel_groups != 0 sets every bucket.  Used by the autotuner; not
 * thread-safe against concurrent _vbatched calls.  No counterpart in the
 * reference. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_vbatched_set_bucket_tuning(rocblas_int bucket,
                                                                         rocblas_int panel_groups);

/* Completes the handle's stream (skipped while the stream is being captured)
 * and reports whether a kernel of an earlier call on this handle gave up one
 * of its bounded waits, which leaves that call's result wrong:
 * rocblas_status_internal_error if so (the report is then cleared), else
 * success.  Two kernels wait on other workgroups of their own launch, each
 * wait bounded to about a second: the chained single-instance solve, and the
 * triangular products of rocblas_?trtri whose K ranges are split over several
 * workgroups.  The next call on the handle that launches a solve or an
 * inversion also reports and clears it, returning
 * rocblas_status_internal_error without doing any work.  No counterpart in
 * the reference, whose waits are unbounded. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_get_async_status(rocblas_handle handle);

/* Test hook for the failure path above: with enable != 0, every chained
 * launch and every split trtri product on this handle gives up each wait
 * after a few polls (its result is wrong and the error is raised; every
 * kernel still runs to its end).  0 restores normal launches.  No
 * counterpart in the reference. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_set_chain_test_timeout(rocblas_handle handle,
                                                                     int            enable);

/* Measures the HBM rate a memory-bound kernel can reach on this device for a
 * given mix of reads and writes: a streaming kernel (16 bytes per lane per
 * access, several loads in flight per wavefront) reads read_bytes from src
 * and writes write_bytes to dst, iters times after one untimed launch, on the
 * handle's stream, for each of 4, 8 and 16 wavefronts per SIMD with plain and
 * with non-temporal loads and stores; *seconds is the best average time of
 * one pass (the call synchronises on its events).  read_bytes : write_bytes
 * must be one of 1:0, 0:1, 1:1, 2:1, 3:1 or 5:2, and each a whole number of
 * units (rocblas_amd_hbm_stream_step_bytes() bytes per access unit), else
 * rocblas_status_invalid_size.  With write_bytes == 0, dst must still hold
 * one access unit (the kernel may write there to keep its loads live).  The
 * triangular solves quote their bandwidth against this rate beside the
 * published peak.  No counterpart in the reference. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_hbm_stream(rocblas_handle handle,
                                                         const void*    src,
                                                         size_t         read_bytes,
                                                         void*          dst,
                                                         size_t         write_bytes,
                                                         rocblas_int    iters,
                                                         double*        seconds);

/* Bytes one access unit of rocblas_amd_hbm_stream moves per read or write
 * (16 bytes for each thread of a grid of 16 wavefronts per SIMD). */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_amd_hbm_stream_step_bytes(rocblas_handle handle,
                                                                    size_t*        bytes);

#ifdef __cplusplus
}
#endif

#endif /* ROCBLAS_AMD_EXT_H */
