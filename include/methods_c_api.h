/*
 * C ABI of the Global SPAI / inner-outer preconditioner builders.
 *
 * Drop-in for the reference interface include/methods_c_api.h:5-60
 * (implemented there by src/common/methods_c_api.cpp:130-172).  The four
 * structs keep the reference's field order, field types and by-value calling
 * convention, so a caller compiled against the reference header links against
 * libspai_hip.so unchanged.  Every entry point runs on the current HIP device.
 *
 * Matrices are CSC with int64 column offsets and row indices and FP64 values,
 * rows sorted inside each column.  Inputs are borrowed; the returned arrays are
 * owned by the library and released only by free_mem().
 */
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/methods_c_api.h:7-14 (MatrixData). */
struct MatrixData {
  int64_t rows;
  int64_t cols;
  int64_t nnz;
  const int64_t* col_offsets; /* cols + 1 entries */
  const int64_t* row_indices; /* nnz entries, sorted per column */
  const double* values;       /* nnz entries */
};

/* Replaces include/methods_c_api.h:16-25 (MethodParams). */
struct MethodParams {
  int64_t max_iterations;
  double tolerance;        /* absolute ||I - A M||_F stop */
  double max_density;      /* nnz budget max(n, floor(max_density * n^2)) */
  int64_t enable_dropping;
  int64_t num_threads;     /* number of contiguous column batches */
  int64_t use_initial_m;
  int64_t use_preconditioner;
  int64_t inner_iterations;
};

/* Replaces include/methods_c_api.h:27-32 (SingleIterationData). */
struct SingleIterationData {
  int64_t iteration;
  double residual_norm;
  double density_m;
  double density_direction;
};

/* Replaces include/methods_c_api.h:34-46 (Result). */
struct Result {
  int64_t rows;
  int64_t cols;
  int64_t nnz;
  int64_t iterations;
  int64_t converged;
  int64_t history_size;
  int64_t* col_offsets;
  int64_t* row_indices;
  double* values;
  struct SingleIterationData* history;
  char* error_message; /* non-NULL on failure; then every other field is zero */
};

#ifndef __cplusplus
typedef struct MatrixData MatrixData;
typedef struct MethodParams MethodParams;
typedef struct SingleIterationData SingleIterationData;
typedef struct Result Result;
#endif

/* Replaces include/methods_c_api.h:48 (Global SPAI conjugate gradient). */
struct Result spai_global_cg(struct MatrixData A, struct MatrixData M0, struct MatrixData Pr, struct MethodParams params);

/* Replaces include/methods_c_api.h:50 (Global SPAI minimal residual). */
struct Result spai_global_mr(struct MatrixData A, struct MatrixData M0, struct MatrixData Pr, struct MethodParams params);

/* Replaces include/methods_c_api.h:52 (Global SPAI locally optimal minimal residual). */
struct Result spai_global_lomr(struct MatrixData A, struct MatrixData M0, struct MatrixData Pr, struct MethodParams params);

/* Replaces include/methods_c_api.h:54 (inner-outer minimal residual). */
struct Result inner_outer_mr(struct MatrixData A, struct MatrixData M0, struct MatrixData Pr, struct MethodParams params);

/* Replaces include/methods_c_api.h:56 (inner-outer locally optimal minimal residual). */
struct Result inner_outer_lomr(struct MatrixData A, struct MatrixData M0, struct MatrixData Pr, struct MethodParams params);

/* Replaces include/methods_c_api.h:58; releases the arrays of a Result and nulls them. */
void free_mem(struct Result* result);

#ifdef __cplusplus
}
#endif
