# Declarations of include/methods_c_api.h and include/spai_ext.h for Cython.
from libc.stdint cimport int64_t

cdef extern from "methods_c_api.h":
    cdef struct MatrixData:
        int64_t rows
        int64_t cols
        int64_t nnz
        const int64_t* col_offsets
        const int64_t* row_indices
        const double* values

    cdef struct MethodParams:
        int64_t max_iterations
        double tolerance
        double max_density
        int64_t enable_dropping
        int64_t num_threads
        int64_t use_initial_m
        int64_t use_preconditioner
        int64_t inner_iterations

    cdef struct SingleIterationData:
        int64_t iteration
        double residual_norm
        double density_m
        double density_direction

    cdef struct Result:
        int64_t rows
        int64_t cols
        int64_t nnz
        int64_t iterations
        int64_t converged
        int64_t history_size
        int64_t* col_offsets
        int64_t* row_indices
        double* values
        SingleIterationData* history
        char* error_message

    Result spai_global_cg(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) noexcept nogil
    Result spai_global_mr(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) noexcept nogil
    Result spai_global_lomr(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) noexcept nogil
    Result inner_outer_mr(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) noexcept nogil
    Result inner_outer_lomr(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) noexcept nogil
    void free_mem(Result* result) noexcept nogil

cdef extern from "spai_ext.h":
    cdef struct SpaiExtParams:
        int64_t fixed_pattern
        int64_t precision
        int64_t column_top_k
        int64_t batch_ordered_reductions
        int64_t reserved[4]

    cdef struct CsrData:
        int64_t rows
        int64_t cols
        int64_t nnz
        const int64_t* row_offsets
        const int64_t* col_indices
        const double* values

    Result spai_run_ex(int64_t method, MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params,
                       SpaiExtParams ext, double* seconds_out) noexcept nogil
    int64_t spai_pcg(CsrData A, CsrData M, const double* b, int64_t max_iterations, double tolerance,
                     double* history, int64_t* count, double* x_out, int64_t m_fp32) noexcept nogil
    int64_t spai_matvec(CsrData A, const double* x, double* y) noexcept nogil
    const char* spai_last_error() noexcept nogil
