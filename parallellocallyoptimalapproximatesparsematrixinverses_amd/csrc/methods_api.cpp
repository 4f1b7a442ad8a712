// C++ interface of include/methods.hpp: the host SparseMatrix class and the
// builders methods::global_spai::Run* / methods::inner_outer::Run*
// (reference include/methods.hpp:137-153), which run on the device through
// the same engine as the C ABI.

#include "methods.hpp"

#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>

#include "engine.hpp"

namespace methods {
namespace common {

SparseMatrix::SparseMatrix(int64_t rows, int64_t cols) : rows_(rows), cols_(cols) {
  if (rows < 0 || cols < 0) throw std::invalid_argument("Sparse matrix dimensions must be non-negative");
  outer_.assign(size_t(cols) + 1, 0);
}

SparseMatrix::SparseMatrix(int64_t rows, int64_t cols, std::vector<int64_t> outer, std::vector<int64_t> inner,
                           std::vector<double> values)
    : SparseMatrix(rows, cols) {
  if ((int64_t)outer.size() != cols + 1) throw std::invalid_argument("Sparse matrix needs cols + 1 column offsets");
  if (outer[0] != 0 || outer[cols] != (int64_t)inner.size() || inner.size() != values.size())
    throw std::invalid_argument("Sparse matrix offsets do not match its entries");
  std::vector<int64_t> r, c;
  r.reserve(inner.size());
  c.reserve(inner.size());
  for (int64_t j = 0; j < cols; ++j) {
    if (outer[j] > outer[j + 1]) throw std::invalid_argument("Sparse matrix column offsets must be non-decreasing");
    for (int64_t p = outer[j]; p < outer[j + 1]; ++p) {
      r.push_back(inner[p]);
      c.push_back(j);
    }
  }
  AssignFromEntries(r, c, values);
}

void SparseMatrix::AssignFromEntries(const std::vector<int64_t>& r, const std::vector<int64_t>& c,
                                     const std::vector<double>& v) {
  const size_t m = v.size();
  for (size_t e = 0; e < m; ++e)
    if (r[e] < 0 || r[e] >= rows_ || c[e] < 0 || c[e] >= cols_)
      throw std::invalid_argument("Sparse matrix entry is out of range");
  // counting sort by column, then each column by row with duplicates summed
  std::vector<int64_t> count(size_t(cols_) + 1, 0);
  for (size_t e = 0; e < m; ++e) ++count[size_t(c[e]) + 1];
  for (int64_t j = 0; j < cols_; ++j) count[j + 1] += count[j];
  std::vector<size_t> order(m);
  {
    std::vector<int64_t> next(count.begin(), count.end() - 1);
    for (size_t e = 0; e < m; ++e) order[size_t(next[c[e]]++)] = e;
  }
  outer_.assign(size_t(cols_) + 1, 0);
  inner_.clear();
  values_.clear();
  inner_.reserve(m);
  values_.reserve(m);
  for (int64_t j = 0; j < cols_; ++j) {
    auto first = order.begin() + count[j], last = order.begin() + count[j + 1];
    // stable: duplicates add in input order
    std::stable_sort(first, last, [&](size_t a, size_t b) { return r[a] < r[b]; });
    for (auto it = first; it != last; ++it) {
      if (!inner_.empty() && (int64_t)inner_.size() > outer_[j] && inner_.back() == r[*it])
        values_.back() += v[*it];
      else {
        inner_.push_back(r[*it]);
        values_.push_back(v[*it]);
      }
    }
    outer_[j + 1] = (int64_t)inner_.size();
  }
}

void SparseMatrix::setIdentity() {
  const int64_t d = std::min(rows_, cols_);
  outer_.assign(size_t(cols_) + 1, 0);
  inner_.resize(size_t(d));
  values_.assign(size_t(d), 1.0);
  for (int64_t j = 0; j < cols_; ++j) outer_[j + 1] = std::min(j + 1, d);
  std::iota(inner_.begin(), inner_.end(), int64_t(0));
}

void SparseMatrix::setZero() {
  outer_.assign(size_t(cols_) + 1, 0);
  inner_.clear();
  values_.clear();
}

void SparseMatrix::resize(int64_t rows, int64_t cols) {
  if (rows < 0 || cols < 0) throw std::invalid_argument("Sparse matrix dimensions must be non-negative");
  rows_ = rows;
  cols_ = cols;
  setZero();
}

double SparseMatrix::coeff(int64_t row, int64_t col) const {
  if (row < 0 || row >= rows_ || col < 0 || col >= cols_) throw std::invalid_argument("coeff index out of range");
  const auto first = inner_.begin() + outer_[col], last = inner_.begin() + outer_[col + 1];
  const auto it = std::lower_bound(first, last, row);
  return (it != last && *it == row) ? values_[size_t(it - inner_.begin())] : 0.0;
}

double SparseMatrix::squaredNorm() const {
  double s = 0.0;
  for (double v : values_) s += v * v;
  return s;
}

SparseMatrix SparseMatrix::transpose() const {
  SparseMatrix t(cols_, rows_);
  std::vector<int64_t> count(size_t(rows_) + 1, 0);
  for (int64_t r : inner_) ++count[size_t(r) + 1];
  for (int64_t i = 0; i < rows_; ++i) count[i + 1] += count[i];
  t.outer_ = count;
  t.inner_.resize(inner_.size());
  t.values_.resize(values_.size());
  // columns in ascending order give each row of t ascending columns
  for (int64_t j = 0; j < cols_; ++j)
    for (int64_t p = outer_[j]; p < outer_[j + 1]; ++p) {
      const int64_t q = count[size_t(inner_[p])]++;
      t.inner_[q] = j;
      t.values_[q] = values_[p];
    }
  return t;
}

namespace {

// x + s y over the union of the patterns, column by column.
SparseMatrix AddScaled(const SparseMatrix& x, const SparseMatrix& y, double s) {
  if (x.rows() != y.rows() || x.cols() != y.cols()) throw std::invalid_argument("Sparse matrices must have the same shape");
  std::vector<int64_t> outer(size_t(x.cols()) + 1, 0), inner;
  std::vector<double> values;
  const int64_t *xp = x.outerIndexPtr(), *xi = x.innerIndexPtr(), *yp = y.outerIndexPtr(), *yi = y.innerIndexPtr();
  const double *xv = x.valuePtr(), *yv = y.valuePtr();
  for (int64_t j = 0; j < x.cols(); ++j) {
    int64_t a = xp[j], b = yp[j];
    while (a < xp[j + 1] || b < yp[j + 1]) {
      if (b >= yp[j + 1] || (a < xp[j + 1] && xi[a] < yi[b])) {
        inner.push_back(xi[a]);
        values.push_back(xv[a++]);
      } else if (a >= xp[j + 1] || yi[b] < xi[a]) {
        inner.push_back(yi[b]);
        values.push_back(s * yv[b++]);
      } else {
        inner.push_back(xi[a]);
        values.push_back(xv[a++] + s * yv[b++]);
      }
    }
    outer[j + 1] = (int64_t)inner.size();
  }
  return SparseMatrix(x.rows(), x.cols(), std::move(outer), std::move(inner), std::move(values));
}

}  // namespace

SparseMatrix operator+(const SparseMatrix& x, const SparseMatrix& y) { return AddScaled(x, y, 1.0); }

SparseMatrix operator-(const SparseMatrix& x, const SparseMatrix& y) { return AddScaled(x, y, -1.0); }

SparseMatrix operator*(double s, const SparseMatrix& x) {
  SparseMatrix out = x;
  for (int64_t p = 0; p < out.nonZeros(); ++p) out.valuePtr()[p] *= s;
  return out;
}

SparseMatrix operator*(const SparseMatrix& x, const SparseMatrix& y) {
  if (x.cols() != y.rows()) throw std::invalid_argument("Sparse matrix dimensions do not match");
  const int64_t rows = x.rows();
  std::vector<double> acc(size_t(rows), 0.0);
  std::vector<int64_t> mark(size_t(rows), -1), rows_j;
  std::vector<int64_t> outer(size_t(y.cols()) + 1, 0), inner;
  std::vector<double> values;
  const int64_t *xp = x.outerIndexPtr(), *xi = x.innerIndexPtr(), *yp = y.outerIndexPtr(), *yi = y.innerIndexPtr();
  const double *xv = x.valuePtr(), *yv = y.valuePtr();
  for (int64_t j = 0; j < y.cols(); ++j) {
    rows_j.clear();
    for (int64_t q = yp[j]; q < yp[j + 1]; ++q) {
      const int64_t k = yi[q];
      for (int64_t p = xp[k]; p < xp[k + 1]; ++p) {
        const int64_t i = xi[p];
        if (mark[i] != j) {
          mark[i] = j;
          acc[i] = 0.0;
          rows_j.push_back(i);
        }
        acc[i] += xv[p] * yv[q];
      }
    }
    std::sort(rows_j.begin(), rows_j.end());
    for (int64_t i : rows_j) {
      inner.push_back(i);
      values.push_back(acc[i]);
    }
    outer[j + 1] = (int64_t)inner.size();
  }
  return SparseMatrix(rows, y.cols(), std::move(outer), std::move(inner), std::move(values));
}

}  // namespace common

namespace {

namespace mc = methods::common;

MatrixData View(const mc::SparseMatrix& m) {
  MatrixData d;
  d.rows = m.rows();
  d.cols = m.cols();
  d.nnz = m.nonZeros();
  d.col_offsets = m.outerIndexPtr();
  d.row_indices = m.innerIndexPtr();
  d.values = m.valuePtr();
  return d;
}

// One builder call: an empty M0 means the identity and an empty Pr means
// diag(A)^-1 (methods_common.cpp:50-63,93-96); the problem checks and their
// messages follow methods_common.cpp:12-48 (in the engine).
mc::OutputResult Run(int64_t method, const char* name, const mc::InputParams& p, int64_t num_threads,
                     int64_t inner_iterations) {
  spai::EngineParams ep;
  ep.max_iterations = p.max_iterations;
  ep.tolerance = p.tolerance;
  ep.max_density = p.max_density;
  ep.enable_dropping = p.enable_dropping;
  ep.num_threads = num_threads;
  ep.inner_iterations = inner_iterations;
  const MatrixData a = View(p.A), m0 = View(p.M0), pr = View(p.Pr);
  const bool has_m0 = p.M0.size() != 0, has_pr = p.Pr.size() != 0;
  spai::EngineOutput out =
      spai::RunEngine(method, a, has_m0 ? &m0 : nullptr, has_pr ? &pr : nullptr, ep, /*check_csc=*/false);
  mc::IterationHistory history;
  history.reserve(out.history.size());
  for (const ::SingleIterationData& h : out.history)
    history.emplace_back(h.iteration, h.residual_norm, h.density_m, h.density_direction);
  const int64_t nnz = int64_t(out.M.idx.size());
  std::vector<int64_t> ptr(out.M.ptr.data(), out.M.ptr.data() + out.M.ptr.size());
  std::vector<int64_t> idx(out.M.idx.data(), out.M.idx.data() + nnz);
  std::vector<double> val(out.M.val.data(), out.M.val.data() + nnz);
  mc::SparseMatrix M(out.M.rows, out.M.cols, std::move(ptr), std::move(idx), std::move(val));
  return mc::OutputResult(std::move(M), out.iterations, out.converged, name, std::move(history));
}

}  // namespace

namespace global_spai {

common::OutputResult RunConjugateGradient(const common::InputParams& params, const int64_t num_threads) {
  return Run(SPAI_GLOBAL_CG, "cg", params, num_threads, 2);
}

common::OutputResult RunMinimalResidual(const common::InputParams& params, const int64_t num_threads) {
  return Run(SPAI_GLOBAL_MR, "mr", params, num_threads, 2);
}

common::OutputResult RunLocallyOptimalMinimalResidual(const common::InputParams& params, const int64_t num_threads) {
  return Run(SPAI_GLOBAL_LOMR, "lomr", params, num_threads, 2);
}

}  // namespace global_spai

namespace inner_outer {

common::OutputResult RunMinimalResidual(const common::InputParams& params, const int64_t num_threads,
                                        const int64_t inner_iterations) {
  return Run(SPAI_INNER_OUTER_MR, "inner_outer_mr", params, num_threads, inner_iterations);
}

common::OutputResult RunLocallyOptimalMinimalResidual(const common::InputParams& params, const int64_t num_threads,
                                                      const int64_t inner_iterations) {
  return Run(SPAI_INNER_OUTER_LOMR, "inner_outer_lomr", params, num_threads, inner_iterations);
}

}  // namespace inner_outer

}  // namespace methods
