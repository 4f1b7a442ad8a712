/*
 * Extension C ABI of libspai_hip.so: the forms and entry points that the
 * reference has no counterpart for in its C++ library, plus the PCG evaluator
 * that the reference implements in Cython.
 *
 * All structs are plain data passed by value; pointers are host pointers unless
 * the name says "dev".  Functions that return int64_t return a negative value
 * on failure, and spai_last_error() then holds the message for the calling
 * thread.
 */
#pragma once

#include <stdint.h>

#include "methods_c_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Method selector for the extension entry points.  The five values map one to
 * one onto the reference's entry points (include/methods_c_api.h:48-56). */
enum SpaiMethod {
  SPAI_GLOBAL_CG = 0,
  SPAI_GLOBAL_MR = 1,
  SPAI_GLOBAL_LOMR = 2,
  SPAI_INNER_OUTER_MR = 3,
  SPAI_INNER_OUTER_LOMR = 4
};

/* Forms beside the reference semantics.  All-zero selects exactly the
 * reference behaviour of spai_global_* / inner_outer_*. */
struct SpaiExtParams {
  /* 0: the pattern of M grows as in the reference.
   * 1: M stays on S = pattern(A) | pattern(M0) | I; the search directions are
   *    sampled on S and every step size is the exact minimiser of the method's
   *    objective over S (projected line search).  Dropping must be off. */
  int64_t fixed_pattern;
  /* 0: FP64 throughout.  1: FP32 bulk: M, the directions, the residual and the
   *    sparse products stream FP32 values; step sizes, trace inner products
   *    and ||I - A M||_F are accumulated in FP64 (fixed_pattern must be 1). */
  int64_t precision;
  /* 0: the reference's global nnz budget.  k > 0 (dropping on): each column of
   *    M and of the direction keeps its diagonal and its k largest |values|
   *    off the diagonal (ties kept by the smaller row) instead. */
  int64_t column_top_k;
  /* 0: trace inner products and norms are tree reductions (fast; results
   *    differ from the reference in the last bits).
   * 1: they are summed in the reference's order -- sequentially within each of
   *    the num_threads column batches, then over the batches in order -- so
   *    the reference-semantics builders reproduce its arithmetic bit for bit
   *    (slower for few batches: one device thread per batch). */
  int64_t batch_ordered_reductions;
  int64_t reserved[4];
};

/* CSR matrix with int64 offsets (rows + 1) and column indices. */
struct CsrData {
  int64_t rows;
  int64_t cols;
  int64_t nnz;
  const int64_t* row_offsets;
  const int64_t* col_indices;
  const double* values;
};

#ifndef __cplusplus
typedef struct SpaiExtParams SpaiExtParams;
typedef struct CsrData CsrData;
#endif

/* Any of the five methods in any form; the result is released by free_mem().
 * seconds_out (nullable) receives the wall time of the device work, measured
 * with HIP events from the first upload to the last download. */
struct Result spai_run_ex(int64_t method, struct MatrixData A, struct MatrixData M0, struct MatrixData Pr,
                          struct MethodParams params, struct SpaiExtParams ext, double* seconds_out);

/* Preconditioned CG on the device; replaces methods_cython/backward_error.pyx:69-144
 * (pcg).  history receives up to max_iterations relative residuals (entry 0 is
 * 1.0, the last is the true residual), *count their number; x_out (nullable)
 * receives the iterate.  m_fp32 != 0 streams M's values in FP32 (vectors,
 * A and every reduction stay FP64).  Returns the reference's status code: 0 =
 * iteration limit, 1 = converged, 2 = zero or non-finite recurrence scalar,
 * 3 = non-finite residual; negative on error. */
int64_t spai_pcg(struct CsrData A, struct CsrData M, const double* b, int64_t max_iterations, double tolerance,
                 double* history, int64_t* count, double* x_out, int64_t m_fp32);

/* y = A x on the device; replaces methods_cython/backward_error.pyx:56-66 (matvec). */
int64_t spai_matvec(struct CsrData A, const double* x, double* y);

/* One product of the reference-semantics path on the device: C = X Y, or
 * R = I - X Y when residual != 0 (CalculateResidualForBatch), with exact zeros
 * pruned when prune_zero != 0.  Replaces the per-batch product of
 * src/common/state_manager.cpp:47-66 (MultiplyBatches) and
 * src/common/dropping.cpp:97-104: the pattern is the one Eigen's conservative
 * product gives and the values sum in the same order.  X and Y are CSC with
 * sorted rows; the result, in the same form, is released by free_mem().
 * seconds_out (nullable) receives the mean wall time of one product over
 * `repeats` runs after one warm-up, operands resident on the device. */
struct Result spai_spgemm(struct MatrixData X, struct MatrixData Y, int64_t residual, int64_t prune_zero,
                          int64_t repeats, double* seconds_out);

/* Per-column top-k by |v| on the device: each column keeps its k off-diagonal
 * entries of largest magnitude (ties: smaller row first) and its diagonal
 * entry.  The column form of the magnitude budget of
 * src/common/dropping.cpp:177-224 (DropForMatrixType), used by
 * column_top_k in spai_run_ex.  X is CSC with sorted rows; the result, in the
 * same form, is released by free_mem(). */
struct Result spai_column_topk(struct MatrixData X, int64_t k);

/* Per-column top-k that allocates nothing: the same result as
 * spai_column_topk, on caller-owned device arrays.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * stream capture, serving loops, harnesses that forbid hipMalloc).
 *   X: CSC on the device, int32 column offsets[cols + 1] and row indices[nnz]
 *      (rows sorted in each column), FP64 values[nnz].
 *   out_*: device arrays, offsets[cols + 1] and room for nnz entries (the
 *      result never has more entries than X); the result's nnz is
 *      out_col_offsets[cols] once the work completes.
 * Limits: nnz <= 16,777,216 and cols <= 4,194,304 (larger sizes are rejected
 * with an error); the work uses a 16 MiB device array that each GPU holds
 * from its first call until the process ends.  The work is enqueued on
 * `stream` (a hipStream_t, 0 = null stream) without synchronisation, so
 * errors inside kernels surface at the caller's next synchronisation.  Calls
 * that overlap on one GPU share that array and race.  Returns 0, or -1 with
 * spai_last_error() set. */
int64_t spai_column_topk_noalloc(int64_t rows, int64_t cols, int64_t nnz, const int32_t* col_offsets_dev,
                                 const int32_t* row_indices_dev, const double* values_dev, int64_t k,
                                 int32_t* out_col_offsets_dev, int32_t* out_row_indices_dev, double* out_values_dev,
                                 uint64_t stream);

/* Achievable HBM bandwidth of this device, for stating the sparse kernels'
 * rates as a fraction of it: median GB/s over `repeats` launches of a
 * streaming read (sum), copy (y = x) and triad (a = b + s c) on FP64 arrays of
 * bytes_per_array bytes each (use sizes well beyond the caches, e.g. 2 GiB),
 * written to gbs_out[0..2].  No counterpart in the reference.  Returns 0, or
 * -1 with spai_last_error() set. */
int64_t spai_stream_bandwidth(int64_t bytes_per_array, int64_t repeats, double* gbs_out);

/* Message of the last failed call on this thread ("" if none). */
const char* spai_last_error(void);

/* Name of the HIP device the library runs on, for reports. */
int64_t spai_device_name(char* buffer, int64_t size);

/* ------------------------------------------------------------------------
 * Prepared PCG system: A and M uploaded once, then any number of right-hand
 * sides solved with every operand in HBM.  Replaces the harness classes
 * PreparedPCGSystem / PreparedPCGSolver (scripts/utils/backward_error.py:63-96),
 * which call backward_error.pcg once per right-hand side on host copies.
 * ------------------------------------------------------------------------ */
typedef struct SpaiPcgSystem SpaiPcgSystem;

/* Validates and uploads A and M (same checks as spai_pcg); NULL on error. */
SpaiPcgSystem* spai_pcg_prepare(struct CsrData A, struct CsrData M, int64_t m_fp32);
void spai_pcg_release(SpaiPcgSystem* system);
/* As spai_pcg for one right-hand side b (host, length n). */
int64_t spai_pcg_solve(SpaiPcgSystem* system, const double* b, int64_t max_iterations, double tolerance,
                       double* history, int64_t* count, double* x_out);
/* y = A x with the prepared A (row dots in column order, as spai_matvec). */
int64_t spai_pcg_matvec(SpaiPcgSystem* system, const double* x, double* y);
/* Row layouts of the prepared system's PCG iteration kernels (as
 * spai_session_pcg_layouts). */
int64_t spai_pcg_layouts(SpaiPcgSystem* system, int64_t* out);

/* ------------------------------------------------------------------------
 * Device-resident session: A, M0, Pr and the right-hand sides are uploaded once,
 * then builds and solves run with every operand already in HBM.  Used by the
 * benchmark and by callers that build several preconditioners for one A.
 * ------------------------------------------------------------------------ */
typedef struct SpaiSession SpaiSession;

SpaiSession* spai_session_create(int64_t method, struct MatrixData A, struct MatrixData M0, struct MatrixData Pr,
                                 struct MethodParams params, struct SpaiExtParams ext);
void spai_session_destroy(SpaiSession* session);
/* Builds M on the device; returns the number of iterations. */
int64_t spai_session_build(SpaiSession* session);
/* Uploads a right-hand side of length n into slot 0. */
int64_t spai_session_set_rhs(SpaiSession* session, const double* b);
/* PCG with the built M from x0 = 0; returns the status as spai_pcg. */
int64_t spai_session_pcg(SpaiSession* session, int64_t max_iterations, double tolerance, int64_t* count);
/* Copies the last PCG history (count entries) and iterate to the host (nullable). */
int64_t spai_session_pcg_output(SpaiSession* session, double* history, double* x);
/* Copies the built M and its history to a host Result (release with free_mem). */
struct Result spai_session_result(SpaiSession* session);
/* Average microseconds of one SpMV with A (plus <p, Ap>) in the layout the
 * session's PCG reads: one event pair around 20 back-to-back launches. */
double spai_session_pcg_iteration_us(SpaiSession* session);
/* Row layouts of the session's PCG iteration kernels, for tests and
 * measurement: out[0] for A, out[1] for M; 0 = CSR, 1 = sliced ELLPACK,
 * 2 = diagonals.  The environment variable SPAI_PCG_LAYOUT ("csr", "sell",
 * "auto") caps the choice when the solver is created.  No counterpart in the
 * reference (its PCG is a CSR loop, methods_cython/backward_error.pyx:16-34). */
int64_t spai_session_pcg_layouts(SpaiSession* session, int64_t* out);
/* on != 0: later builds record HIP events, on the session's stream, around
 * each per-iteration pass of the fused LOMR or MR build. */
int64_t spai_session_set_kernel_timing(SpaiSession* session, int64_t on);
/* Average microseconds of that pass in the last build and (launches) how many
 * passes were averaged; negative when the last build was not timed. */
double spai_session_build_kernel_us(SpaiSession* session, int64_t* launches);
/* Launch shape of the session's fixed-pattern build, for tests and
 * measurement: out[0] = 1 when LOMR/MR run the fused per-iteration pass, 2
 * when LOMR runs the diagonal pass (A, S and S2 on few diagonals), 0
 * otherwise; then lanes per column, columns per tile, tiles, resident-grid
 * blocks, wavefronts per block, nnz of the iterate pattern S, nnz of the
 * product pattern S2 = pattern(A S) | S | I, and the diagonals of A, S and
 * S2 on the diagonal pass (0 without it); len entries written, at most 11.
 * Returns 0, or -1. */
int64_t spai_session_info(SpaiSession* session, int64_t* out, int64_t len);

/* ------------------------------------------------------------------------
 * Differentiation form: f(M, A) = ||I - A M||_F^2 on fixed patterns.
 * df/dM = -2 (A^T R) sampled on pattern(M), df/dA = -2 (R M^T) sampled on
 * pattern(A), R = I - A M.  Values are device pointers in the CSC order of the
 * patterns given at creation (FP64, or FP32 when precision = 1, with FP64
 * reductions).  R is formed per block of block_cols columns (0 = all at once)
 * and never stored whole.  Work runs on `stream` (a hipStream_t, 0 = null
 * stream).  No counterpart in the reference, whose closest quantity is the
 * sampled A R of its drop score (src/common/dropping.cpp:19-29).
 * ------------------------------------------------------------------------ */
typedef struct SpaiLossPlan SpaiLossPlan;

SpaiLossPlan* spai_loss_plan_create(struct MatrixData A, struct MatrixData M, int64_t precision, int64_t block_cols,
                                    uint64_t stream);
void spai_loss_plan_destroy(SpaiLossPlan* plan);
/* Later forward and backward calls run on `stream` (a hipStream_t, 0 = null
 * stream), e.g. the caller's current stream at each call. */
int64_t spai_loss_set_stream(SpaiLossPlan* plan, uint64_t stream);
/* out4: nnz(pattern(A M) | I), largest block of R, number of blocks, nnz(M). */
int64_t spai_loss_plan_info(SpaiLossPlan* plan, int64_t* out4);
int64_t spai_loss_forward(SpaiLossPlan* plan, const void* m_values_dev, const void* a_values_dev, double* loss_out);
/* grad_m = grad_scale * df/dM, grad_a = grad_scale * df/dA (device outputs). */
int64_t spai_loss_backward(SpaiLossPlan* plan, const void* m_values_dev, const void* a_values_dev, double grad_scale,
                           void* grad_m_dev, void* grad_a_dev);
const char* spai_loss_last_error(void);

#ifdef __cplusplus
}
#endif
