// C++ interface of libspai_hip.so with the reference's namespaces, types and
// signatures (reference include/methods.hpp:11-155), so that C++ callers of
// methods::global_spai::Run* and methods::inner_outer::Run* switch over by
// recompiling against this header and linking -lspai_hip.
//
// What differs: the reference's SparseMatrix is
// Eigen::SparseMatrix<double, Eigen::ColMajor, int64_t> (methods.hpp:17).
// Eigen is not a dependency here, so methods::common::SparseMatrix is this
// library's own compressed-sparse-column class with the same storage (int64
// column offsets and row indices, double values, rows sorted inside each
// column) and the subset of Eigen's interface that callers of the methods use
// to build inputs and read results: rows/cols/outerSize/innerSize/size/
// nonZeros, outerIndexPtr/innerIndexPtr/valuePtr, makeCompressed/isCompressed,
// setFromTriplets (duplicates summed), setIdentity, coeff, norm/squaredNorm,
// transpose, and +, -, scalar * and sparse * sparse.  The builders run on the
// current HIP device; the algebra operators run on the host (they are for
// checking results, as the reference's tests/test.cpp does).
//
// Errors are thrown as std::invalid_argument with the reference's messages
// (src/common/methods_common.cpp:12-60, src/inner_outer/inner_outer_common.hpp:16),
// or std::runtime_error for device failures.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace methods {

namespace common {

inline constexpr double EPS = 1e-9;

// (row, column, value) entry for SparseMatrix::setFromTriplets, with the
// accessors of Eigen::Triplet.
template <class Scalar = double, class StorageIndex = int64_t>
class Triplet {
 public:
  Triplet() = default;
  Triplet(StorageIndex row, StorageIndex col, Scalar value) : row_(row), col_(col), value_(value) {}
  StorageIndex row() const { return row_; }
  StorageIndex col() const { return col_; }
  Scalar value() const { return value_; }

 private:
  StorageIndex row_ = 0, col_ = 0;
  Scalar value_ = 0;
};

class SparseMatrix {
 public:
  using Scalar = double;
  using StorageIndex = int64_t;
  using Index = int64_t;

  SparseMatrix() : outer_(1, 0) {}
  SparseMatrix(int64_t rows, int64_t cols);
  // From CSC arrays (outer: cols + 1 offsets, inner: row indices, values);
  // rows are sorted and duplicates summed inside each column.
  SparseMatrix(int64_t rows, int64_t cols, std::vector<int64_t> outer, std::vector<int64_t> inner,
               std::vector<double> values);

  int64_t rows() const { return rows_; }
  int64_t cols() const { return cols_; }
  int64_t outerSize() const { return cols_; }
  int64_t innerSize() const { return rows_; }
  int64_t size() const { return rows_ * cols_; }
  int64_t nonZeros() const { return static_cast<int64_t>(values_.size()); }

  // Always compressed: offsets, rows and values are contiguous arrays.
  bool isCompressed() const { return true; }
  void makeCompressed() {}

  const int64_t* outerIndexPtr() const { return outer_.data(); }
  int64_t* outerIndexPtr() { return outer_.data(); }
  const int64_t* innerIndexPtr() const { return inner_.data(); }
  int64_t* innerIndexPtr() { return inner_.data(); }
  const double* valuePtr() const { return values_.data(); }
  double* valuePtr() { return values_.data(); }

  // Replaces the contents by the triplets' entries (duplicates summed); the
  // shape stays.  Throws std::invalid_argument for an entry outside it.
  template <class InputIterator>
  void setFromTriplets(InputIterator begin, InputIterator end) {
    std::vector<int64_t> r, c;
    std::vector<double> v;
    for (InputIterator it = begin; it != end; ++it) {
      r.push_back(static_cast<int64_t>(it->row()));
      c.push_back(static_cast<int64_t>(it->col()));
      v.push_back(static_cast<double>(it->value()));
    }
    AssignFromEntries(r, c, v);
  }

  // Ones on the main diagonal, nothing else (any shape).
  void setIdentity();
  void setZero();
  void resize(int64_t rows, int64_t cols);

  // Value at (row, col), 0 when the entry is absent.
  double coeff(int64_t row, int64_t col) const;
  double squaredNorm() const;
  double norm() const { return std::sqrt(squaredNorm()); }
  SparseMatrix transpose() const;

  friend SparseMatrix operator+(const SparseMatrix& x, const SparseMatrix& y);
  friend SparseMatrix operator-(const SparseMatrix& x, const SparseMatrix& y);
  friend SparseMatrix operator*(double s, const SparseMatrix& x);
  friend SparseMatrix operator*(const SparseMatrix& x, double s) { return s * x; }
  // Host product (column by column; each column's rows come out sorted).
  friend SparseMatrix operator*(const SparseMatrix& x, const SparseMatrix& y);

 private:
  void AssignFromEntries(const std::vector<int64_t>& r, const std::vector<int64_t>& c, const std::vector<double>& v);

  int64_t rows_ = 0, cols_ = 0;
  std::vector<int64_t> outer_;
  std::vector<int64_t> inner_;
  std::vector<double> values_;
};

inline SparseMatrix EnsureCompressedSparseMatrix(SparseMatrix matrix = SparseMatrix()) {
  matrix.makeCompressed();
  return matrix;
}

// Inputs of a builder (reference methods.hpp:28-62).  An empty M0 (size() ==
// 0) means M0 = I; an empty Pr means Pr = diag(A)^-1.  max_non_zeros is
// informational, as in the reference, where no method reads it (the density
// budget is max(n, floor(max_density n^2)), computed by the builders); here it
// holds that budget for the given A.
struct InputParams {
  SparseMatrix A;
  SparseMatrix M0;
  SparseMatrix Pr;

  int64_t max_iterations = 100;
  // Absolute Frobenius residual tolerance, with or without dropping.
  double tolerance = 1e-9;
  double max_density = 0.03;
  bool enable_dropping = true;
  int64_t max_non_zeros = -1;

  InputParams(SparseMatrix A = SparseMatrix(), SparseMatrix M0 = SparseMatrix(), SparseMatrix Pr = SparseMatrix(),
              const int64_t max_iterations = 100, const double tolerance = 1e-9, const double max_density = 0.03,
              const bool enable_dropping = true)
      : A(EnsureCompressedSparseMatrix(std::move(A))),
        M0(EnsureCompressedSparseMatrix(std::move(M0))),
        Pr(EnsureCompressedSparseMatrix(std::move(Pr))),
        max_iterations(max_iterations),
        tolerance(tolerance),
        max_density(max_density),
        enable_dropping(enable_dropping),
        max_non_zeros(std::max(this->A.rows(), static_cast<int64_t>(std::floor(
                                                   max_density * double(this->A.rows()) * double(this->A.rows()))))) {}
};

// One history row (reference methods.hpp:64-80).
struct SingleIterationData {
  int64_t iteration = 0;
  double residual_norm = 0.0;
  double density_m = 0.0;
  double density_direction = 0.0;

  SingleIterationData(int64_t iteration = 0, double residual_norm = 0.0, double density_m = 0.0,
                      double density_direction = 0.0)
      : iteration(iteration),
        residual_norm(residual_norm),
        density_m(density_m),
        density_direction(density_direction) {}
};

using IterationHistory = std::vector<SingleIterationData>;

// Result of a builder (reference methods.hpp:84-133; the reference's private
// bookkeeping of AppendIteration is internal to the library here).
struct OutputResult {
 private:
  double initial_residual_norm_ = 0.0;

 public:
  std::string method;
  SparseMatrix M;
  int64_t iterations = 0;
  bool converged = false;
  IterationHistory history;

  OutputResult() = default;
  OutputResult(std::string method, const double initial_residual_norm = 0.0)
      : initial_residual_norm_(initial_residual_norm), method(std::move(method)) {}
  OutputResult(SparseMatrix M, const int64_t iterations = 0, const bool converged = false, std::string method = "",
               IterationHistory history = IterationHistory())
      : initial_residual_norm_(history.empty() ? 0.0 : history.front().residual_norm),
        method(std::move(method)),
        M(EnsureCompressedSparseMatrix(std::move(M))),
        iterations(iterations),
        converged(converged),
        history(std::move(history)) {}
};

}  // namespace common

// num_threads is the number of contiguous column batches [n i/b, n (i+1)/b)
// that the reference's per-batch quantities (thresholds, reduction order)
// follow; the work itself runs on the device.
namespace global_spai {

common::OutputResult RunConjugateGradient(const common::InputParams& params, const int64_t num_threads = 1);

common::OutputResult RunMinimalResidual(const common::InputParams& params, const int64_t num_threads = 1);

common::OutputResult RunLocallyOptimalMinimalResidual(const common::InputParams& params,
                                                      const int64_t num_threads = 1);

}  // namespace global_spai

namespace inner_outer {

common::OutputResult RunMinimalResidual(const common::InputParams& params, const int64_t num_threads = 1,
                                        const int64_t inner_iterations = 2);

common::OutputResult RunLocallyOptimalMinimalResidual(const common::InputParams& params,
                                                      const int64_t num_threads = 1,
                                                      const int64_t inner_iterations = 2);

}  // namespace inner_outer

}  // namespace methods
