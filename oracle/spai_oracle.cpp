// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement (C++17, OpenMP, no Eigen) of the reference's Global SPAI
// CG/MR/LOMR builders, its inner-outer MR/LOMR builders and its PCG evaluator.
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
// this library, and only as the checker or the timed CPU baseline: the product
// (libspai_hip.so) never links or calls it.
//
// Every function cites the reference file:line it restates.  Eigen 3 (unpinned;
// the reference's Dockerfile:12 installs Debian's libeigen3-dev, 3.4.x) is not
// present here, so the Eigen operations the reference calls are restated from
// Eigen's published algorithms:
//   * sparse * sparse      -> conservative column-wise Gustavson product
//                             (values[i] = x*y on first touch, += x*y after,
//                             inner index k ascending; explicit zeros kept)
//   * sparse + sparse      -> union of the patterns, functor(lhs, 0) / (0, rhs)
//   * cwiseProduct().sum() -> intersection, sequential column-major sum
//   * prune(ref, eps)      -> keep |v| > |ref| * eps (prune(0.0): keep v != 0)
//   * CompleteOrthogonalDecomposition on a 2x2 -> column-pivoted Householder QR
//                             plus the right Householder step when rank < 2.
// Floating-point contraction is disabled for this file (see Makefile): the
// reference is built for baseline x86-64, which has no FMA.
//
// Parity status per function (how each is pinned) is listed in the header of
// oracle/oracle.py.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <omp.h>

#include "methods_c_api.h"
#include "spai_ext.h"

namespace {

using i64 = int64_t;

// methods::common::kPruneValueThreshold (src/common/methods_common.hpp:32).
constexpr double kPruneEps = std::numeric_limits<double>::epsilon();
// methods::common::EPS (include/methods.hpp:15).
constexpr double kDensityEps = 1e-9;

struct Csc {
  i64 rows = 0, cols = 0;
  std::vector<i64> ptr;
  std::vector<i64> idx;
  std::vector<double> val;
  Csc() = default;
  Csc(i64 r, i64 c) : rows(r), cols(c), ptr(c + 1, 0) {}
  i64 nnz() const { return (i64)idx.size(); }
  bool empty_shape() const { return rows == 0 && cols == 0; }
};

struct Batch {
  i64 first = 0, width = 0;
};

// Whole-matrix builder: every batch produces its own CSC piece (width columns),
// pieces are concatenated in batch order.  Mirrors the per-batch matrices of
// ColumnBatch (src/common/methods_common.hpp:44-61) and RunBatchesParallel's
// OpenMP dynamic schedule (methods_common.hpp:16-28).
template <class F>
Csc BuildByBatches(i64 rows, i64 cols, const std::vector<Batch>& batches, int threads, F&& f) {
  const i64 nb = (i64)batches.size();
  std::vector<Csc> pieces(nb);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (threads > 1)
  for (i64 b = 0; b < nb; ++b) {
    pieces[b] = Csc(rows, batches[b].width);
    f(b, pieces[b]);
  }
  Csc out(rows, cols);
  i64 total = 0;
  for (auto& p : pieces) total += p.nnz();
  out.idx.resize(total);
  out.val.resize(total);
  i64 off = 0;
  for (i64 b = 0; b < nb; ++b) {
    const auto& p = pieces[b];
    for (i64 c = 0; c < batches[b].width; ++c) out.ptr[batches[b].first + c] = off + p.ptr[c];
    std::copy(p.idx.begin(), p.idx.end(), out.idx.begin() + off);
    std::copy(p.val.begin(), p.val.end(), out.val.begin() + off);
    off += p.nnz();
  }
  out.ptr[cols] = off;
  return out;
}

template <class F>
void ForBatches(const std::vector<Batch>& batches, int threads, F&& f) {
  const i64 nb = (i64)batches.size();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (threads > 1)
  for (i64 b = 0; b < nb; ++b) f(b);
}

// ---------------------------------------------------------------------------
// Column primitives (Eigen semantics restated)
// ---------------------------------------------------------------------------

// Conservative sparse product X * Y[:, cols of batch] (Eigen
// conservative_sparse_sparse_product; call sites src/common/state_manager.cpp:60,
// src/common/dropping.cpp:98, src/inner_outer/inner_outer_common.hpp:48).
// Per output column: accumulate over Y's entries k ascending, values[i] = x*y on
// the first touch of row i, values[i] += x*y afterwards; rows sorted.
// negate_lhs computes (-1.0 * X) * Y as in dropping.cpp:98.
void ProductColumns(const Csc& X, const Csc& Y, const Batch& batch, bool negate_lhs, std::vector<char>& mask,
                    std::vector<double>& values, std::vector<i64>& list, Csc& piece,
                    // residual epilogue: R = -(A M) then R(j,j) += 1 (dropping.cpp:99-102)
                    bool add_identity, bool prune_zero) {
  piece.idx.clear();
  piece.val.clear();
  for (i64 c = 0; c < batch.width; ++c) {
    const i64 j = batch.first + c;
    i64 nz = 0;
    for (i64 p = Y.ptr[j]; p < Y.ptr[j + 1]; ++p) {
      const i64 k = Y.idx[p];
      const double y = Y.val[p];
      for (i64 q = X.ptr[k]; q < X.ptr[k + 1]; ++q) {
        const i64 i = X.idx[q];
        const double x = negate_lhs ? -X.val[q] : X.val[q];
        if (!mask[i]) {
          mask[i] = 1;
          values[i] = x * y;
          list[nz++] = i;
        } else {
          values[i] += x * y;
        }
      }
    }
    if (add_identity) {
      // coeffRef(j, j) += 1.0: inserts a 0.0 when absent (dropping.cpp:101).
      if (!mask[j]) {
        mask[j] = 1;
        values[j] = 0.0;
        list[nz++] = j;
      }
      values[j] += 1.0;
    }
    std::sort(list.begin(), list.begin() + nz);
    for (i64 t = 0; t < nz; ++t) {
      const i64 i = list[t];
      mask[i] = 0;
      const double v = values[i];
      if (prune_zero && !(std::fabs(v) > 0.0)) continue;  // prune(0.0) keeps NaN
      piece.idx.push_back(i);
      piece.val.push_back(v);
    }
    piece.ptr[c + 1] = (i64)piece.idx.size();
  }
}

Csc Product(const Csc& X, const Csc& Y, const std::vector<Batch>& batches, int threads, bool prune_zero,
            bool residual = false) {
  if (X.cols != Y.rows) throw std::invalid_argument("Sparse matrix dimensions do not match");
  return BuildByBatches(X.rows, Y.cols, batches, threads, [&](i64 b, Csc& piece) {
    std::vector<char> mask(X.rows, 0);
    std::vector<double> values(X.rows, 0.0);
    std::vector<i64> list(X.rows + 1, 0);
    ProductColumns(X, Y, batches[b], residual, mask, values, list, piece, residual, prune_zero);
  });
}

// Union merge of two sorted columns with per-side value maps (Eigen's sparse
// binary iterator: functor(lhs, rhs) when both, functor(lhs, 0) / functor(0, rhs)
// otherwise), then optional prune of exact zeros.
template <class Both, class OnlyL, class OnlyR>
Csc UnionMerge(const Csc& L, const Csc& R, const std::vector<Batch>& batches, int threads, bool prune_zero,
               Both both, OnlyL only_l, OnlyR only_r) {
  if (L.rows != R.rows || L.cols != R.cols) throw std::invalid_argument("Sparse matrices must have the same shape");
  return BuildByBatches(L.rows, L.cols, batches, threads, [&](i64 b, Csc& piece) {
    for (i64 c = 0; c < batches[b].width; ++c) {
      const i64 j = batches[b].first + c;
      i64 a = L.ptr[j], ae = L.ptr[j + 1], r = R.ptr[j], re = R.ptr[j + 1];
      auto emit = [&](i64 row, double v) {
        if (prune_zero && !(std::fabs(v) > 0.0)) return;
        piece.idx.push_back(row);
        piece.val.push_back(v);
      };
      while (a < ae || r < re) {
        if (a < ae && r < re && L.idx[a] == R.idx[r]) {
          emit(L.idx[a], both(L.val[a], R.val[r]));
          ++a;
          ++r;
        } else if (a < ae && (r >= re || L.idx[a] < R.idx[r])) {
          emit(L.idx[a], only_l(L.val[a]));
          ++a;
        } else {
          emit(R.idx[r], only_r(R.val[r]));
          ++r;
        }
      }
      piece.ptr[c + 1] = (i64)piece.idx.size();
    }
  });
}

// AddScaled (src/common/sparse_operations.cpp:41-49): target += c * dir when
// c != 0, then prune(0.0).
Csc AddScaled(const Csc& target, const Csc& dir, double c, const std::vector<Batch>& batches, int threads) {
  if (target.rows != dir.rows || target.cols != dir.cols)
    throw std::invalid_argument("Sparse matrices must have the same shape");
  if (c == 0.0) {
    return UnionMerge(target, Csc(target.rows, target.cols), batches, threads, true, [](double l, double) { return l; },
                      [](double l) { return l; }, [](double r) { return r; });
  }
  return UnionMerge(
      target, dir, batches, threads, true, [c](double l, double r) { return l + c * r; },
      [](double l) { return l + 0.0; }, [c](double r) { return 0.0 + c * r; });
}

// LinearCombination (sparse_operations.cpp:22-26) with prune=false:
// y *= b; y += a * x; prune(0.0).  Returns the new y.
Csc LinearCombination(const Csc& x, double a, const Csc& y, double b, const std::vector<Batch>& batches,
                      int threads) {
  return UnionMerge(
      y, x, batches, threads, true, [a, b](double yv, double xv) { return yv * b + a * xv; },
      [b](double yv) { return yv * b + 0.0; }, [a](double xv) { return 0.0 + a * xv; });
}

// MulScalarColumnwise (sparse_operations.cpp:28-39): x(:, j) *= a[j], prune(0.0).
void MulScalarColumnwise(Csc& x, const std::vector<double>& a, i64 first) {
  Csc out(x.rows, x.cols);
  for (i64 j = 0; j < x.cols; ++j) {
    for (i64 p = x.ptr[j]; p < x.ptr[j + 1]; ++p) {
      const double v = x.val[p] * a[first + j];
      if (std::fabs(v) > 0.0 || std::isnan(v)) {
        out.idx.push_back(x.idx[p]);
        out.val.push_back(v);
      }
    }
    out.ptr[j + 1] = out.nnz();
  }
  x = std::move(out);
}

// FrobeniusDot (sparse_operations.cpp:51-56) over the columns of one batch:
// intersection of the patterns, summed sequentially column-major.
double DotBatch(const Csc& X, const Csc& Y, const Batch& b) {
  double s = 0.0;
  for (i64 j = b.first; j < b.first + b.width; ++j) {
    i64 p = X.ptr[j], pe = X.ptr[j + 1], q = Y.ptr[j], qe = Y.ptr[j + 1];
    while (p < pe && q < qe) {
      if (X.idx[p] == Y.idx[q]) {
        s += X.val[p] * Y.val[q];
        ++p;
        ++q;
      } else if (X.idx[p] < Y.idx[q]) {
        ++p;
      } else {
        ++q;
      }
    }
  }
  return s;
}

// squaredNorm() of one batch: sequential sum of squares column-major.
double SqNormBatch(const Csc& X, const Batch& b) {
  double s = 0.0;
  for (i64 p = X.ptr[b.first]; p < X.ptr[b.first + b.width]; ++p) s += X.val[p] * X.val[p];
  return s;
}

// Column dot r.col(j).dot(y.col(j)) (inner_outer_common.hpp:54-69).
double ColDot(const Csc& X, const Csc& Y, i64 j) {
  double s = 0.0;
  i64 p = X.ptr[j], pe = X.ptr[j + 1], q = Y.ptr[j], qe = Y.ptr[j + 1];
  while (p < pe && q < qe) {
    if (X.idx[p] == Y.idx[q]) {
      s += X.val[p] * Y.val[q];
      ++p;
      ++q;
    } else if (X.idx[p] < Y.idx[q]) {
      ++p;
    } else {
      ++q;
    }
  }
  return s;
}

double ColSqNorm(const Csc& X, i64 j) {
  double s = 0.0;
  for (i64 p = X.ptr[j]; p < X.ptr[j + 1]; ++p) s += X.val[p] * X.val[p];
  return s;
}

Csc Transpose(const Csc& X) {
  Csc T(X.cols, X.rows);
  std::vector<i64> cnt(X.rows + 1, 0);
  for (i64 p = 0; p < X.nnz(); ++p) cnt[X.idx[p] + 1]++;
  for (i64 i = 0; i < X.rows; ++i) cnt[i + 1] += cnt[i];
  T.ptr = cnt;
  T.idx.resize(X.nnz());
  T.val.resize(X.nnz());
  std::vector<i64> cur(cnt.begin(), cnt.end() - 1);
  for (i64 j = 0; j < X.cols; ++j)
    for (i64 p = X.ptr[j]; p < X.ptr[j + 1]; ++p) {
      const i64 d = cur[X.idx[p]]++;
      T.idx[d] = j;
      T.val[d] = X.val[p];
    }
  return T;
}

Csc Identity(i64 n) {
  Csc I(n, n);
  I.idx.resize(n);
  I.val.assign(n, 1.0);
  for (i64 i = 0; i < n; ++i) {
    I.ptr[i + 1] = i + 1;
    I.idx[i] = i;
  }
  return I;
}

bool IsFiniteNonZero(double v, double threshold) { return std::isfinite(v) && std::fabs(v) > threshold; }

// GetSafeQuotient (sparse_operations.cpp:14-20).
double SafeQuotient(double num, double den) {
  if (!IsFiniteNonZero(den, 0.0)) return 0.0;
  const double q = num / den;
  return std::isfinite(q) ? q : 0.0;
}

// ---------------------------------------------------------------------------
// SolveLeastSquares2x2 (sparse_operations.cpp:58-73): symmetric diagonal
// scaling, then Eigen's CompleteOrthogonalDecomposition::solve restated for a
// 2x2 (ColPivHouseholderQR: column-norm pivoting, Householder reflections
// v = [1; essential], beta = -sign(c0) * ||v||, tau = (beta - c0) / beta; rank =
// #{|r_ii| > maxpivot * 2 * eps}; rank 1 adds one right Householder step).
// ---------------------------------------------------------------------------
struct House {
  double essential, tau, beta;
};

// makeHouseholder on a 2-vector [c0, t].
House MakeHouse2(double c0, double t) {
  const double tail_sq = t * t;
  const double tol = std::numeric_limits<double>::min();
  House h;
  if (tail_sq <= tol) {
    h.tau = 0.0;
    h.beta = c0;
    h.essential = 0.0;
  } else {
    double beta = std::sqrt(c0 * c0 + tail_sq);
    if (c0 >= 0.0) beta = -beta;
    h.essential = t / (c0 - beta);
    h.tau = (beta - c0) / beta;
    h.beta = beta;
  }
  return h;
}

// applyHouseholderOnTheLeft on a 2-vector (v0, v1) with essential e, tau.
void ApplyHouse2(double e, double tau, double& v0, double& v1) {
  if (tau == 0.0) return;
  double tmp = e * v1;
  tmp += v0;
  v0 -= tau * tmp;
  v1 -= (tau * e) * tmp;
}

// rank_eps replaces eps in the final rank decision only (|r_ii| > maxpivot *
// 2 * rank_eps); Eigen's value, DBL_EPSILON, is the default.  A larger value
// restates the FP32 inner-outer form of the device library, which takes
// pivots below sqrt(FLT_EPSILON) as zero.
void SolveLS2x2(const double m[2][2], const double x[2], double out[2],
                double rank_eps = std::numeric_limits<double>::epsilon()) {
  out[0] = out[1] = 0.0;
  for (int i = 0; i < 2; ++i) {
    if (!std::isfinite(x[i])) return;
    for (int j = 0; j < 2; ++j)
      if (!std::isfinite(m[i][j])) return;
  }
  double s[2];
  for (int i = 0; i < 2; ++i) {
    const double r = std::sqrt(std::fabs(m[i][i]));
    s[i] = r > 0.0 ? 1.0 / r : 1.0;
  }
  double a[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) a[i][j] = (s[i] * m[i][j]) * s[j];
  double b[2] = {s[0] * x[0], s[1] * x[1]};

  // ColPivHouseholderQR of a (column-major 2x2).
  const double eps = std::numeric_limits<double>::epsilon();
  double norms_direct[2], norms_upd[2];
  for (int j = 0; j < 2; ++j) {
    norms_direct[j] = std::sqrt(a[0][j] * a[0][j] + a[1][j] * a[1][j]);
    norms_upd[j] = norms_direct[j];
  }
  const double maxnorm = std::max(norms_upd[0], norms_upd[1]);
  const double threshold_helper = (maxnorm * eps) * (maxnorm * eps) / 2.0;
  const double downdate_threshold = std::sqrt(eps);
  int nonzero_pivots = 2;
  double maxpivot = 0.0;
  int perm[2] = {0, 1};
  // k = 0
  {
    int biggest = (norms_upd[1] > norms_upd[0]) ? 1 : 0;  // first index of the maximum
    const double big_sq = norms_upd[biggest] * norms_upd[biggest];
    if (big_sq < threshold_helper * 2.0) nonzero_pivots = 0;
    if (biggest != 0) {
      std::swap(a[0][0], a[0][1]);
      std::swap(a[1][0], a[1][1]);
      std::swap(norms_upd[0], norms_upd[1]);
      std::swap(norms_direct[0], norms_direct[1]);
      std::swap(perm[0], perm[1]);
    }
  }
  const House h0 = MakeHouse2(a[0][0], a[1][0]);
  a[0][0] = h0.beta;
  a[1][0] = h0.essential;
  maxpivot = std::max(maxpivot, std::fabs(h0.beta));
  ApplyHouse2(h0.essential, h0.tau, a[0][1], a[1][1]);
  if (norms_upd[1] != 0.0) {
    double t = std::fabs(a[0][1]) / norms_upd[1];
    t = (1.0 + t) * (1.0 - t);
    if (t < 0.0) t = 0.0;
    const double ratio = norms_upd[1] / norms_direct[1];
    const double t2 = t * ratio * ratio;
    if (t2 <= downdate_threshold) {
      norms_direct[1] = std::fabs(a[1][1]);
      norms_upd[1] = norms_direct[1];
    } else {
      norms_upd[1] *= std::sqrt(t);
    }
  }
  // k = 1: a length-1 Householder is the identity (tau = 0, beta = a11).
  {
    const double big_sq = norms_upd[1] * norms_upd[1];
    if (nonzero_pivots == 2 && big_sq < threshold_helper * 1.0) nonzero_pivots = 1;
    maxpivot = std::max(maxpivot, std::fabs(a[1][1]));
  }
  const double premultiplied = maxpivot * (rank_eps * 2.0);
  int rank = 0;
  for (int i = 0; i < nonzero_pivots; ++i) rank += (std::fabs(a[i][i]) > premultiplied) ? 1 : 0;
  if (rank == 0) {
    return;
  }
  // c = Q^T b (reflector 0 only; reflector 1 has tau = 0).
  double c0 = b[0], c1 = b[1];
  ApplyHouse2(h0.essential, h0.tau, c0, c1);
  double y[2] = {0.0, 0.0};
  if (rank == 2) {
    // Upper-triangular back substitution, skipping zero right-hand entries.
    double r1 = c1, r0 = c0;
    if (r1 != 0.0) {
      r1 /= a[1][1];
      r0 -= r1 * a[0][1];
    }
    if (r0 != 0.0) r0 /= a[0][0];
    y[0] = r0;
    y[1] = r1;
  } else {
    // rank 1: right Householder on row 0 [r00, r01] -> [beta, 0].
    const House z0 = MakeHouse2(a[0][0], a[0][1]);
    const double t00 = z0.beta;
    double r0 = c0;
    if (r0 != 0.0) r0 /= t00;
    double d0 = r0, d1 = 0.0;
    ApplyHouse2(z0.essential, z0.tau, d0, d1);
    y[0] = d0;
    y[1] = d1;
  }
  double sol[2];
  sol[perm[0]] = y[0];
  sol[perm[1]] = y[1];
  out[0] = s[0] * sol[0];
  out[1] = s[1] * sol[1];
  if (!std::isfinite(out[0]) || !std::isfinite(out[1])) out[0] = out[1] = 0.0;
}

// ---------------------------------------------------------------------------
// State manager (src/common/methods_common.cpp, state_manager.cpp, dropping.cpp)
// ---------------------------------------------------------------------------

struct Params {
  Csc A, M0, Pr;  // M0/Pr: empty shape when absent
  i64 max_iterations = 100;
  double tolerance = 1e-9;
  double max_density = 0.03;
  bool enable_dropping = true;
  // extensions
  bool fixed_pattern = false;
  i64 top_k = 0;
  double rank_eps = std::numeric_limits<double>::epsilon();  // inner-outer 2x2 rank threshold
};

struct History {
  i64 iteration;
  double residual_norm, density_m, density_direction;
};

struct Output {
  Csc M;
  i64 iterations = 0;
  bool converged = false;
  std::vector<History> history;
  double initial_residual_norm = 0.0;
};

enum class Mat { M, R, Z, P, AP, AZ, Tmp };

struct State {
  const Params& prm;
  i64 n = 0;
  int threads = 1;
  std::vector<Batch> batches;
  std::vector<double> a_col_sq;
  bool preserve_symmetry = true;
  Csc Pr;
  Csc M, R, Z, P, AP, AZ, Tmp;
  // fixed pattern: S (pattern of M and of the directions)
  Csc S;

  explicit State(const Params& p) : prm(p) {}

  Csc& Get(Mat t) {
    switch (t) {
      case Mat::M: return M;
      case Mat::R: return R;
      case Mat::Z: return Z;
      case Mat::P: return P;
      case Mat::AP: return AP;
      case Mat::AZ: return AZ;
      case Mat::Tmp: return Tmp;
    }
    throw std::invalid_argument("unknown column-batch matrix");
  }

  // MultiplyBatches (state_manager.cpp:47-66): prune(0.0) unless dropping.
  Csc Multiply(const Csc& X, const Csc& Y) const {
    const bool prune = !prm.enable_dropping && !prm.fixed_pattern;
    return Product(X, Y, batches, threads, prune);
  }

  // CalculateResidualForBatch (dropping.cpp:97-104): R = -(A M) + I, prune(0.0).
  // Fixed pattern keeps the structure of A*S | I.
  void Residual() { R = Product(prm.A, M, batches, threads, !prm.fixed_pattern, true); }

  // GetMaxNonZeros (state_manager.cpp:68-71).
  i64 MaxNonZeros() const {
    const double m = prm.max_density * (double)n * (double)n;
    return std::max(n, (i64)std::floor(m));
  }

  // GetResidualNorm / StabilizeResiduals norm: per-batch squaredNorm summed in
  // batch order (methods_common.cpp:148-154).
  double ResidualNorm() const {
    std::vector<double> sq(batches.size());
    for (size_t b = 0; b < batches.size(); ++b) sq[b] = SqNormBatch(R, batches[b]);
    double s = 0.0;
    for (double v : sq) s += v;
    return std::sqrt(s);
  }

  // Sum over batches of FrobeniusDot, accumulated from 0.0 in batch order.
  double Dot(const Csc& X, const Csc& Y) const {
    std::vector<double> part(batches.size());
    ForBatches(batches, threads, [&](i64 b) { part[b] = DotBatch(X, Y, batches[b]); });
    double s = 0.0;
    for (double v : part) s += v;
    return s;
  }
  double SqNorm(const Csc& X) const {
    std::vector<double> part(batches.size());
    ForBatches(batches, threads, [&](i64 b) { part[b] = SqNormBatch(X, batches[b]); });
    double s = 0.0;
    for (double v : part) s += v;
    return s;
  }

  // StabilizeResiduals (state_manager.cpp:81-95).
  void StabilizeResiduals() {
    const double norm = ResidualNorm();
    const double floor_ = 64.0 * std::numeric_limits<double>::epsilon() * std::sqrt((double)n);
    const double thr = std::max(10.0 * prm.tolerance, floor_);
    if (std::isfinite(norm) && norm > thr) return;
    Residual();
  }

  // EnsureSymmetry (dropping.cpp:106-120): M := 0.5 * (M + M^T).
  void EnsureSymmetry() {
    if (!preserve_symmetry) return;
    const Csc MT = Transpose(M);
    M = UnionMerge(
        M, MT, batches, threads, false, [](double a, double b) { return 0.5 * (a + b); },
        [](double a) { return 0.5 * (a + 0.0); }, [](double b) { return 0.5 * (0.0 + b); });
  }

  // Per-column top-k (extension): keep the diagonal and the k largest |v| off
  // the diagonal; ties keep the smaller row.
  static void KeepColumnTopK(Csc& X, i64 k, bool protect_diag) {
    Csc out(X.rows, X.cols);
    std::vector<i64> order;
    for (i64 j = 0; j < X.cols; ++j) {
      order.clear();
      for (i64 p = X.ptr[j]; p < X.ptr[j + 1]; ++p)
        if (!(protect_diag && X.idx[p] == j)) order.push_back(p);
      std::sort(order.begin(), order.end(), [&](i64 a, i64 b) {
        const double fa = std::fabs(X.val[a]), fb = std::fabs(X.val[b]);
        if (fa != fb) return fa > fb;
        return X.idx[a] < X.idx[b];
      });
      std::vector<char> keep(X.ptr[j + 1] - X.ptr[j], 0);
      for (i64 t = 0; t < (i64)order.size() && t < k; ++t) keep[order[t] - X.ptr[j]] = 1;
      for (i64 p = X.ptr[j]; p < X.ptr[j + 1]; ++p) {
        if ((protect_diag && X.idx[p] == j) || keep[p - X.ptr[j]]) {
          out.idx.push_back(X.idx[p]);
          out.val.push_back(X.val[p]);
        }
      }
      out.ptr[j + 1] = out.nnz();
    }
    X = std::move(out);
  }

  struct Cand {
    double score;
    i64 offset;
  };
  static bool Less(const Cand& l, const Cand& r) {
    return l.score < r.score || (l.score == r.score && l.offset < r.offset);
  }
  // GetDropCutoff (dropping.cpp:86-93).
  static Cand Cutoff(std::vector<Cand>& c, i64 drop) {
    if (drop <= 0 || drop > (i64)c.size()) throw std::invalid_argument("Invalid global sparse dropping count");
    std::nth_element(c.begin(), c.begin() + (drop - 1), c.end(), Less);
    return c[drop - 1];
  }

  // GetApproximationDropScore (dropping.cpp:19-29): m^2 ||A e_k||^2 + 2 m (A R)_kl.
  static double Score(double m, double col_sq, double tmp) {
    const double r = m * m * col_sq + 2.0 * m * tmp;
    return std::isfinite(r) ? r : -std::numeric_limits<double>::infinity();
  }

  static double Lookup(const Csc& X, i64 row, i64 col) {
    const auto b = X.idx.begin() + X.ptr[col], e = X.idx.begin() + X.ptr[col + 1];
    const auto it = std::lower_bound(b, e, row);
    if (it != e && *it == row) return X.val[it - X.idx.begin()];
    return 0.0;
  }

  // ApplyDroppingStrategy (dropping.cpp:122-175).
  void ApplyDroppingStrategy() {
    EnsureSymmetry();
    // per-column eps * ||m_col||_2 threshold; the diagonal is protected.
    {
      Csc out = BuildByBatches(n, n, batches, threads, [&](i64 b, Csc& piece) {
        const Batch& bt = batches[b];
        for (i64 c = 0; c < bt.width; ++c) {
          const i64 j = bt.first + c;
          const double thr = kPruneEps * std::sqrt(ColSqNorm(M, j));
          for (i64 p = M.ptr[j]; p < M.ptr[j + 1]; ++p) {
            const double v = M.val[p];
            if (v != 0.0 && (M.idx[p] == j || std::fabs(v) >= thr)) {
              piece.idx.push_back(M.idx[p]);
              piece.val.push_back(v);
            }
          }
          piece.ptr[c + 1] = piece.nnz();
        }
      });
      M = std::move(out);
    }
    if (prm.top_k > 0) {
      KeepColumnTopK(M, prm.top_k, true);
      Residual();
      return;
    }
    Residual();
    const i64 maxnz = MaxNonZeros();
    const i64 nz = M.nnz();
    if (nz <= maxnz) return;

    Tmp = Multiply(prm.A, R);
    // candidate offsets: column-major over the off-diagonal entries of M.
    std::vector<Cand> cand;
    cand.reserve(nz);
    for (i64 j = 0; j < n; ++j)
      for (i64 p = M.ptr[j]; p < M.ptr[j + 1]; ++p) {
        const i64 row = M.idx[p];
        if (row == j) continue;
        cand.push_back({Score(M.val[p], a_col_sq[row], Lookup(Tmp, row, j)), (i64)cand.size()});
      }
    const Cand cut = Cutoff(cand, nz - maxnz);
    Csc out(n, n);
    i64 off = 0;
    for (i64 j = 0; j < n; ++j) {
      for (i64 p = M.ptr[j]; p < M.ptr[j + 1]; ++p) {
        const i64 row = M.idx[p];
        bool keep = true;
        if (row != j) {
          const Cand c{Score(M.val[p], a_col_sq[row], Lookup(Tmp, row, j)), off++};
          keep = Less(cut, c);
        }
        if (keep) {
          out.idx.push_back(row);
          out.val.push_back(M.val[p]);
        }
      }
      out.ptr[j + 1] = out.nnz();
    }
    M = std::move(out);
    Residual();
  }

  // DropForMatrixType (dropping.cpp:177-224).
  bool DropForMatrixType(Mat t) {
    if (!prm.enable_dropping) return false;
    Csc& X = Get(t);
    const i64 original = X.nnz();
    {
      Csc out = BuildByBatches(n, n, batches, threads, [&](i64 b, Csc& piece) {
        const Batch& bt = batches[b];
        const double thr = std::sqrt(SqNormBatch(X, bt)) * kPruneEps;  // prune(norm, eps)
        for (i64 c = 0; c < bt.width; ++c) {
          const i64 j = bt.first + c;
          for (i64 p = X.ptr[j]; p < X.ptr[j + 1]; ++p) {
            const double v = X.val[p];
            if (!(std::fabs(v) <= thr)) {
              piece.idx.push_back(X.idx[p]);
              piece.val.push_back(v);
            }
          }
          piece.ptr[c + 1] = piece.nnz();
        }
      });
      X = std::move(out);
    }
    if (prm.top_k > 0) {
      KeepColumnTopK(X, prm.top_k, true);
      return X.nnz() != original;
    }
    const i64 maxnz = MaxNonZeros();
    const i64 nz = X.nnz();
    if (nz <= maxnz) return nz != original;
    std::vector<Cand> cand(nz);
    for (i64 p = 0; p < nz; ++p) cand[p] = {std::fabs(X.val[p]), p};
    const Cand cut = Cutoff(cand, nz - maxnz);
    Csc out(n, n);
    for (i64 j = 0; j < n; ++j) {
      for (i64 p = X.ptr[j]; p < X.ptr[j + 1]; ++p) {
        if (Less(cut, Cand{std::fabs(X.val[p]), p})) {
          out.idx.push_back(X.idx[p]);
          out.val.push_back(X.val[p]);
        }
      }
      out.ptr[j + 1] = out.nnz();
    }
    X = std::move(out);
    return true;
  }

  // Fixed-pattern projection onto S (extension): structural sample.
  Csc Sample(const Csc& X) const {
    Csc out(n, n);
    out.idx = S.idx;
    out.ptr = S.ptr;
    out.val.assign(S.nnz(), 0.0);
    for (i64 j = 0; j < n; ++j)
      for (i64 p = S.ptr[j]; p < S.ptr[j + 1]; ++p) out.val[p] = Lookup(X, S.idx[p], j);
    return out;
  }
  // Dense-structure copy of X onto S's pattern (X's pattern must be within S).
  Csc OnS(const Csc& X) const { return Sample(X); }
};

// ValidateProblem (methods_common.cpp:12-36).
void ValidateProblem(const Params& p) {
  const Csc& A = p.A;
  if (A.rows != A.cols || A.rows == 0) throw std::invalid_argument("A must be a non-empty square matrix");
  if (!p.M0.empty_shape() && (p.M0.rows != A.rows || p.M0.cols != A.cols))
    throw std::invalid_argument("Initial preconditioner M0 must have the same shape as A");
  if (!p.Pr.empty_shape() && (p.Pr.rows != A.rows || p.Pr.cols != A.cols))
    throw std::invalid_argument("Pr must have the same shape as A");
  if (p.max_iterations <= 0) throw std::invalid_argument("max_iterations must be positive");
  if (!std::isfinite(p.tolerance) || p.tolerance < 0.0)
    throw std::invalid_argument("tolerance must be finite and non-negative");
  if (!std::isfinite(p.max_density) || p.max_density <= kDensityEps || p.max_density > 1.0)
    throw std::invalid_argument("max_density must be finite and in the interval (0, 1]");
  const double max_entries = p.max_density * (double)A.rows * (double)A.rows;
  if (p.enable_dropping && max_entries < (double)A.rows)
    throw std::invalid_argument("max_density is too small to preserve the diagonal of M");
  if (p.fixed_pattern && p.enable_dropping)
    throw std::invalid_argument("fixed_pattern requires enable_dropping = 0");
}

// PrepareStateManager (methods_common.cpp:126-157).
void Prepare(State& s, const Params& p, i64 num_threads, bool preserve_symmetry) {
  ValidateProblem(p);
  const i64 n = p.A.cols;
  if (num_threads <= 0) throw std::invalid_argument("Invalid threads count value");
  const i64 tc = std::min(num_threads, n);
  s.n = n;
  s.threads = (int)std::max<i64>(1, std::min<i64>(tc, omp_get_max_threads()));
  s.preserve_symmetry = preserve_symmetry;
  for (i64 i = 0; i < tc; ++i) s.batches.push_back({n * i / tc, n * (i + 1) / tc - n * i / tc});
  // InitializeBatches: M0 or identity.
  s.M = p.M0.empty_shape() ? Identity(n) : p.M0;
  // GetColumnNormsSquared (methods_common.cpp:65-76), only when dropping.
  s.a_col_sq.assign(n, 0.0);
  if (p.enable_dropping)
    for (i64 k = 0; k < n; ++k) s.a_col_sq[k] = ColSqNorm(p.A, k);
  // GetPreconditioner / GetDiagonalInversePreconditioner (methods_common.cpp:50-63,105-110).
  if (!p.Pr.empty_shape()) {
    s.Pr = p.Pr;
  } else {
    std::vector<double> d(n, 0.0);
    for (i64 j = 0; j < n; ++j) d[j] = State::Lookup(p.A, j, j);
    for (double v : d)
      if (!std::isfinite(v) || v <= 0.0)
        throw std::invalid_argument("A has an invalid diagonal entry; provide Pr explicitly");
    Csc D(n, n);
    for (i64 j = 0; j < n; ++j) {
      const double inv = 1.0 / d[j];
      if (!std::isfinite(inv))
        throw std::invalid_argument("A diagonal cannot be inverted safely; provide Pr explicitly");
      D.idx.push_back(j);
      D.val.push_back(inv);
      D.ptr[j + 1] = j + 1;
    }
    s.Pr = D;
  }
  const Csc Zero(n, n);
  s.R = Zero;
  s.Z = Zero;
  s.P = Zero;
  s.AP = Zero;
  s.AZ = Zero;
  s.Tmp = Zero;
  if (p.fixed_pattern) {
    // S = pattern(A) | pattern(M0) | I, values of M moved onto S.
    Csc I = Identity(n);
    Csc SA = UnionMerge(
        p.A, I, s.batches, s.threads, false, [](double, double) { return 0.0; }, [](double) { return 0.0; },
        [](double) { return 0.0; });
    s.S = UnionMerge(
        SA, s.M, s.batches, s.threads, false, [](double, double) { return 0.0; }, [](double) { return 0.0; },
        [](double) { return 0.0; });
    s.M = s.OnS(s.M);
    s.P = s.OnS(Zero);
  }
  if (p.enable_dropping) {
    s.ApplyDroppingStrategy();
  } else {
    s.Residual();
  }
}

// AppendIteration (src/common/result_operations.cpp:7-67).
bool AppendIteration(Output& out, const State& s, const Csc& direction) {
  const i64 it = (i64)out.history.size();
  out.iterations = it;
  const double entries = (double)s.n * (double)s.n;
  const double rn = out.history.empty() ? out.initial_residual_norm : s.ResidualNorm();
  out.history.push_back({it, rn, (double)s.M.nnz() / entries, (double)direction.nnz() / entries});
  if (rn < s.prm.tolerance) return out.converged = true;
  if (s.prm.fixed_pattern) return false;
  return !s.prm.enable_dropping && s.prm.max_density < 1.0 && out.history.back().density_m >= s.prm.max_density;
}

// Z := Pr * R, sampled on S in the fixed-pattern form.
void UpdateZ(State& s) {
  if (s.prm.fixed_pattern) {
    s.Z = s.Sample(Product(s.Pr, s.R, s.batches, s.threads, false));
  } else {
    s.Z = s.Multiply(s.Pr, s.R);
  }
}

// RunConjugateGradient (src/global_spai/conjugate_gradient.cpp:15-83).
Output RunCG(const Params& prm, i64 threads) {
  State s(prm);
  Prepare(s, prm, threads, true);
  Output out;
  out.initial_residual_norm = s.ResidualNorm();
  UpdateZ(s);
  s.P = s.Z;
  double r_z = s.Dot(s.R, s.Z);
  s.DropForMatrixType(Mat::P);
  if (AppendIteration(out, s, s.P)) {
    out.M = s.M;
    return out;
  }
  s.AP = s.Multiply(prm.A, s.P);
  for (i64 i = 1; i <= prm.max_iterations; ++i) {
    const double den = s.Dot(s.P, s.AP);
    if (!IsFiniteNonZero(r_z, 0.0) || !IsFiniteNonZero(den, 0.0)) break;
    const double alpha = r_z / den;
    if (!std::isfinite(alpha)) break;
    const double prev = r_z;
    if (prm.fixed_pattern) {
      s.M = UnionMerge(
          s.M, s.P, s.batches, s.threads, false, [alpha](double l, double r) { return l + alpha * r; },
          [](double l) { return l; }, [alpha](double r) { return alpha * r; });
      s.R = UnionMerge(
          s.R, s.AP, s.batches, s.threads, false, [alpha](double l, double r) { return l + (-alpha) * r; },
          [](double l) { return l; }, [alpha](double r) { return (-alpha) * r; });
    } else {
      s.M = AddScaled(s.M, s.P, alpha, s.batches, s.threads);
      if (prm.enable_dropping) {
        s.ApplyDroppingStrategy();
      } else {
        s.R = AddScaled(s.R, s.AP, -alpha, s.batches, s.threads);
      }
    }
    if (AppendIteration(out, s, s.P)) break;
    UpdateZ(s);
    r_z = s.Dot(s.R, s.Z);
    const double beta = r_z / prev;
    if (!std::isfinite(beta)) break;
    if (prm.fixed_pattern) {
      s.P = UnionMerge(
          s.P, s.Z, s.batches, s.threads, false, [beta](double p, double z) { return p * beta + z; },
          [beta](double p) { return p * beta; }, [](double z) { return z; });
    } else {
      s.P = LinearCombination(s.Z, 1.0, s.P, beta, s.batches, s.threads);
    }
    s.DropForMatrixType(Mat::P);
    s.AP = s.Multiply(prm.A, s.P);
  }
  out.M = s.M;
  return out;
}

// Pr-weighted products used by the projected forms: <Pr R, Y> with Pr diagonal
// equals <R, Pr Y>; computed as the batch-ordered dot of (Pr R) with Y.
double DotPrR(State& s, const Csc& Y) {
  const Csc PrR = Product(s.Pr, s.R, s.batches, s.threads, false);
  return s.Dot(PrR, Y);
}

// RunMinimalResidual (src/global_spai/minimal_residual.cpp:31-93).
Output RunMR(const Params& prm, i64 threads) {
  State s(prm);
  Prepare(s, prm, threads, true);
  Output out;
  out.initial_residual_norm = s.ResidualNorm();
  UpdateZ(s);
  if (AppendIteration(out, s, s.Z)) {
    out.M = s.M;
    return out;
  }
  auto project = [&](double& num, double& den) {
    s.AZ = s.Multiply(prm.A, s.Z);
    s.Tmp = s.Multiply(s.Pr, s.AZ);
    num = prm.fixed_pattern ? DotPrR(s, s.Tmp) : s.Dot(s.Z, s.Tmp);
    den = s.SqNorm(s.Tmp);
  };
  double num = 0.0, den = 0.0;
  project(num, den);
  for (i64 i = 1; i <= prm.max_iterations; ++i) {
    if (!IsFiniteNonZero(den, 0.0)) break;
    const double alpha = num / den;
    if (!std::isfinite(alpha)) break;
    if (prm.fixed_pattern) {
      s.M = UnionMerge(
          s.M, s.Z, s.batches, s.threads, false, [alpha](double l, double r) { return l + alpha * r; },
          [](double l) { return l; }, [alpha](double r) { return alpha * r; });
    } else {
      s.M = AddScaled(s.M, s.Z, alpha, s.batches, s.threads);
    }
    if (prm.enable_dropping) {
      s.ApplyDroppingStrategy();
    } else {
      s.Residual();
    }
    if (AppendIteration(out, s, s.Z)) break;
    UpdateZ(s);
    project(num, den);
  }
  out.M = s.M;
  return out;
}

// GetDeltaGamma (src/global_spai/locally_minimal_residual.cpp:36-50).
struct Lomr {
  double az_praz = 0, ap_prap = 0, az_prap = 0, z_az = 0, z_ap = 0;
};
bool DeltaGamma(const Lomr& p, i64 i, double& delta, double& gamma) {
  if (i == 1) {
    if (!IsFiniteNonZero(p.az_praz, 0.0)) return true;
    delta = p.z_az / p.az_praz;
    gamma = 1.0;
    return false;
  }
  const double m[2][2] = {{p.az_praz, p.az_prap}, {p.az_prap, p.ap_prap}};
  const double x[2] = {p.z_az, p.z_ap};
  double c[2];
  SolveLS2x2(m, x, c);
  delta = c[0];
  gamma = c[1];
  return false;
}

// RunLocallyOptimalMinimalResidual (locally_minimal_residual.cpp:54-118).
Output RunLOMR(const Params& prm, i64 threads) {
  State s(prm);
  Prepare(s, prm, threads, true);
  Output out;
  out.initial_residual_norm = s.ResidualNorm();
  UpdateZ(s);
  if (AppendIteration(out, s, s.Z)) {
    out.M = s.M;
    return out;
  }
  for (i64 i = 1; i <= prm.max_iterations; ++i) {
    s.AP = s.Multiply(prm.A, s.P);
    s.AZ = s.Multiply(prm.A, s.Z);
    s.Tmp = s.Multiply(s.Pr, s.AZ);
    Lomr pj;
    pj.az_praz = s.Dot(s.AZ, s.Tmp);
    pj.z_az = prm.fixed_pattern ? DotPrR(s, s.AZ) : s.Dot(s.Z, s.AZ);
    if (i > 1) {
      s.Tmp = s.Multiply(s.Pr, s.AP);
      pj.ap_prap = s.Dot(s.AP, s.Tmp);
      pj.az_prap = s.Dot(s.AZ, s.Tmp);
      pj.z_ap = prm.fixed_pattern ? DotPrR(s, s.AP) : s.Dot(s.Z, s.AP);
    }
    double delta = -1, gamma = -1;
    const bool brk = DeltaGamma(pj, i, delta, gamma);
    if (brk || !std::isfinite(delta) || !std::isfinite(gamma)) break;
    if (prm.fixed_pattern) {
      const double d = delta, g = gamma;
      s.M = UnionMerge(
          s.M, s.Z, s.batches, s.threads, false, [d](double l, double r) { return l + d * r; },
          [](double l) { return l; }, [d](double r) { return d * r; });
      s.M = UnionMerge(
          s.M, s.P, s.batches, s.threads, false, [g](double l, double r) { return l + g * r; },
          [](double l) { return l; }, [g](double r) { return g * r; });
      s.R = UnionMerge(
          s.R, s.AZ, s.batches, s.threads, false, [d](double l, double r) { return l + (-d) * r; },
          [](double l) { return l; }, [d](double r) { return (-d) * r; });
      s.R = UnionMerge(
          s.R, s.AP, s.batches, s.threads, false, [g](double l, double r) { return l + (-g) * r; },
          [](double l) { return l; }, [g](double r) { return (-g) * r; });
    } else {
      s.M = AddScaled(s.M, s.Z, delta, s.batches, s.threads);
      s.M = AddScaled(s.M, s.P, gamma, s.batches, s.threads);
      if (prm.enable_dropping) {
        s.ApplyDroppingStrategy();
      } else {
        s.R = AddScaled(s.R, s.AZ, -delta, s.batches, s.threads);
        s.R = AddScaled(s.R, s.AP, -gamma, s.batches, s.threads);
      }
    }
    const double ratio = delta != 0.0 ? gamma / delta : 0.0;
    const double beta = std::isfinite(ratio) ? ratio : 0.0;
    if (prm.fixed_pattern) {
      s.P = UnionMerge(
          s.P, s.Z, s.batches, s.threads, false, [beta](double p, double z) { return p * beta + z; },
          [beta](double p) { return p * beta; }, [](double z) { return z; });
    } else {
      s.P = LinearCombination(s.Z, 1.0, s.P, beta, s.batches, s.threads);
    }
    s.DropForMatrixType(Mat::P);
    if (AppendIteration(out, s, s.P)) break;
    UpdateZ(s);
  }
  out.M = s.M;
  return out;
}

// inner_outer::Run (src/inner_outer/inner_outer_common.hpp:14-107).
Output RunInnerOuter(const Params& prm, i64 threads, bool lomr, i64 inner_iterations) {
  if (inner_iterations <= 0) throw std::invalid_argument("inner_iterations must be positive");
  State s(prm);
  Prepare(s, prm, threads, false);
  Output out;
  out.initial_residual_norm = s.ResidualNorm();
  if (AppendIteration(out, s, s.P)) {
    out.M = s.M;
    return out;
  }
  const i64 n = s.n;
  const bool fixed = prm.fixed_pattern;
  Csc q = fixed ? s.OnS(Csc(n, n)) : Csc(n, n);
  std::vector<double> normalizations(n, 1.0);
  for (i64 iteration = 0; iteration < prm.max_iterations;) {
    const Csc outer_m = s.M;
    s.Residual();
    q = fixed ? s.OnS(Csc(n, n)) : Csc(n, n);
    const i64 inner_count = std::min(inner_iterations, prm.max_iterations - iteration);
    for (i64 inner = 0; inner < inner_count; ++inner, ++iteration) {
      if (fixed) {
        s.Z = s.Sample(Product(outer_m, s.R, s.batches, s.threads, false));
      } else {
        s.Z = s.Multiply(outer_m, s.R);
      }
      s.AZ = s.Multiply(prm.A, s.Z);
      std::vector<double> delta(n), gamma(n, 0.0);
      Csc AQ;
      if (lomr) AQ = Product(prm.A, q, s.batches, s.threads, false);
#pragma omp parallel for schedule(dynamic, 1) num_threads(s.threads) if (s.threads > 1)
      for (i64 b = 0; b < (i64)s.batches.size(); ++b) {
        const Batch& bt = s.batches[b];
        for (i64 j = bt.first; j < bt.first + bt.width; ++j) {
          if (!lomr || inner == 0) {
            delta[j] = SafeQuotient(ColDot(s.R, s.AZ, j), ColSqNorm(s.AZ, j));
            gamma[j] = lomr ? 1.0 : 0.0;
          } else {
            const double coupling = ColDot(s.AZ, AQ, j);
            const double g[2][2] = {{ColSqNorm(s.AZ, j), coupling}, {coupling, ColSqNorm(AQ, j)}};
            const double x[2] = {ColDot(s.R, s.AZ, j), ColDot(s.R, AQ, j)};
            double c[2];
            SolveLS2x2(g, x, c, prm.rank_eps);
            delta[j] = c[0];
            gamma[j] = c[1];
          }
          const double inv = delta[j] != 0.0 ? 1.0 / delta[j] : 1.0;
          normalizations[j] = std::isfinite(inv) ? inv : 1.0;
        }
      }
      if (fixed) {
        // P = Z diag(delta) + q diag(gamma) on S.
        s.P = s.Z;
        for (i64 j = 0; j < n; ++j)
          for (i64 p = s.P.ptr[j]; p < s.P.ptr[j + 1]; ++p) {
            s.P.val[p] = s.Z.val[p] * delta[j];
            if (lomr) s.P.val[p] = s.P.val[p] + q.val[p] * gamma[j];
          }
      } else {
        s.P = s.Z;
        MulScalarColumnwise(s.P, delta, 0);
        if (lomr) {
          MulScalarColumnwise(q, gamma, 0);
          s.P = AddScaled(s.P, q, 1.0, s.batches, s.threads);
        }
      }
      const bool dropped = s.DropForMatrixType(Mat::P);
      s.AP = s.Multiply(prm.A, s.P);
      if (dropped) {
        std::vector<double> corr(n);
        for (i64 j = 0; j < n; ++j) corr[j] = SafeQuotient(ColDot(s.R, s.AP, j), ColSqNorm(s.AP, j));
        MulScalarColumnwise(s.P, corr, 0);
        MulScalarColumnwise(s.AP, corr, 0);
      }
      if (fixed) {
        for (i64 p = 0; p < s.M.nnz(); ++p) s.M.val[p] = s.M.val[p] + s.P.val[p];
        if (lomr) {
          q = s.P;
          for (i64 j = 0; j < n; ++j)
            for (i64 p = q.ptr[j]; p < q.ptr[j + 1]; ++p) q.val[p] = q.val[p] * normalizations[j];
        }
      } else {
        s.M = AddScaled(s.M, s.P, 1.0, s.batches, s.threads);
        if (lomr) {
          q = s.P;
          MulScalarColumnwise(q, normalizations, 0);
        }
      }
      if (prm.enable_dropping) {
        s.ApplyDroppingStrategy();
      } else {
        if (fixed) {
          s.R = UnionMerge(
              s.R, s.AP, s.batches, s.threads, false, [](double l, double r) { return l + (-1.0) * r; },
              [](double l) { return l; }, [](double r) { return (-1.0) * r; });
        } else {
          s.R = AddScaled(s.R, s.AP, -1.0, s.batches, s.threads);
        }
        s.StabilizeResiduals();
      }
      if (AppendIteration(out, s, s.P)) {
        out.M = s.M;
        return out;
      }
    }
  }
  out.M = s.M;
  return out;
}

// ---------------------------------------------------------------------------
// C ABI conversions (src/common/methods_c_api.cpp:16-126)
// ---------------------------------------------------------------------------

Csc FromData(const MatrixData& d) {
  if (d.rows <= 0 || d.cols <= 0 || d.nnz <= 0)
    throw std::invalid_argument("Sparse matrix dimensions and nnz must be positive");
  for (i64 i = 0; i < d.cols; ++i)
    if (d.col_offsets[i] > d.col_offsets[i + 1])
      throw std::invalid_argument("Sparse matrix column offsets must be non-decreasing");
  for (i64 i = 0; i < d.nnz; ++i)
    if (d.row_indices[i] < 0 || d.row_indices[i] >= d.rows)
      throw std::invalid_argument("Sparse matrix row idx is out of range");
  Csc m(d.rows, d.cols);
  const i64 base = d.col_offsets[0];
  for (i64 j = 0; j <= d.cols; ++j) m.ptr[j] = d.col_offsets[j] - base;
  const i64 nz = m.ptr[d.cols];
  m.idx.assign(d.row_indices + base, d.row_indices + base + nz);
  m.val.assign(d.values + base, d.values + base + nz);
  return m;
}

Result ErrorResult(const std::string& msg) {
  Result r{};
  r.error_message = new char[msg.size() + 1];
  std::memcpy(r.error_message, msg.c_str(), msg.size() + 1);
  return r;
}

Result SuccessResult(const Output& o) {
  Result r{};
  r.rows = o.M.rows;
  r.cols = o.M.cols;
  r.nnz = o.M.nnz();
  r.iterations = o.iterations;
  r.converged = o.converged ? 1 : 0;
  r.history_size = (i64)o.history.size();
  r.col_offsets = new int64_t[r.cols + 1];
  r.row_indices = new int64_t[r.nnz > 0 ? r.nnz : 1];
  r.values = new double[r.nnz > 0 ? r.nnz : 1];
  r.history = new SingleIterationData[r.history_size > 0 ? r.history_size : 1];
  std::copy(o.M.ptr.begin(), o.M.ptr.end(), r.col_offsets);
  std::copy(o.M.idx.begin(), o.M.idx.end(), r.row_indices);
  std::copy(o.M.val.begin(), o.M.val.end(), r.values);
  for (i64 i = 0; i < r.history_size; ++i)
    r.history[i] = SingleIterationData{o.history[i].iteration, o.history[i].residual_norm, o.history[i].density_m,
                                       o.history[i].density_direction};
  return r;
}

Result Run(int64_t method, MatrixData A, MatrixData M0, MatrixData Pr, MethodParams cp, SpaiExtParams ext,
           double rank_eps = std::numeric_limits<double>::epsilon()) {
  try {
    Params p;
    p.rank_eps = rank_eps;
    p.A = FromData(A);
    if (cp.use_initial_m) p.M0 = FromData(M0);
    if (cp.use_preconditioner) p.Pr = FromData(Pr);
    p.max_iterations = cp.max_iterations;
    p.tolerance = cp.tolerance;
    p.max_density = cp.max_density;
    p.enable_dropping = cp.enable_dropping != 0;
    p.fixed_pattern = ext.fixed_pattern != 0;
    p.top_k = ext.column_top_k;
    Output o;
    switch (method) {
      case SPAI_GLOBAL_CG: o = RunCG(p, cp.num_threads); break;
      case SPAI_GLOBAL_MR: o = RunMR(p, cp.num_threads); break;
      case SPAI_GLOBAL_LOMR: o = RunLOMR(p, cp.num_threads); break;
      case SPAI_INNER_OUTER_MR: o = RunInnerOuter(p, cp.num_threads, false, cp.inner_iterations); break;
      case SPAI_INNER_OUTER_LOMR: o = RunInnerOuter(p, cp.num_threads, true, cp.inner_iterations); break;
      default: throw std::invalid_argument("unknown method");
    }
    return SuccessResult(o);
  } catch (const std::exception& e) {
    return ErrorResult(e.what());
  } catch (...) {
    return ErrorResult("Unknown failure.");
  }
}

}  // namespace

extern "C" {

Result oracle_run(int64_t method, MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params, SpaiExtParams ext,
                  int64_t omp_threads) {
  if (omp_threads > 0) omp_set_num_threads((int)omp_threads);
  return Run(method, A, M0, Pr, params, ext);
}

// oracle_run with the rank threshold of the inner-outer 2x2 solves
// (SolveLS2x2's rank_eps) set to rank_eps.
Result oracle_run_rank_eps(int64_t method, MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params,
                           SpaiExtParams ext, int64_t omp_threads, double rank_eps) {
  if (omp_threads > 0) omp_set_num_threads((int)omp_threads);
  return Run(method, A, M0, Pr, params, ext, rank_eps);
}

void oracle_free(Result* r) {
  if (!r) return;
  delete[] r->col_offsets;
  delete[] r->row_indices;
  delete[] r->values;
  delete[] r->history;
  delete[] r->error_message;
  r->col_offsets = nullptr;
  r->row_indices = nullptr;
  r->values = nullptr;
  r->history = nullptr;
  r->error_message = nullptr;
}

// Conservative product C = X * Y (optionally the residual I - X * Y), written
// into caller buffers sized by a first call with out arrays NULL: returns nnz.
int64_t oracle_product(MatrixData X, MatrixData Y, int64_t residual, int64_t prune_zero, int64_t* out_ptr,
                       int64_t* out_idx, double* out_val) {
  try {
    const Csc x = FromData(X), y = FromData(Y);
    std::vector<Batch> one{{0, y.cols}};
    const Csc c = Product(x, y, one, 1, prune_zero != 0, residual != 0);
    if (out_ptr) {
      std::copy(c.ptr.begin(), c.ptr.end(), out_ptr);
      std::copy(c.idx.begin(), c.idx.end(), out_idx);
      std::copy(c.val.begin(), c.val.end(), out_val);
    }
    return c.nnz();
  } catch (...) {
    return -1;
  }
}

// Per-column top-k of X (diagonal kept), exposed for pinning.
int64_t oracle_column_topk(MatrixData X, int64_t k, int64_t* out_ptr, int64_t* out_idx, double* out_val) {
  try {
    Csc x = FromData(X);
    State::KeepColumnTopK(x, k, true);
    if (out_ptr) {
      std::copy(x.ptr.begin(), x.ptr.end(), out_ptr);
      std::copy(x.idx.begin(), x.idx.end(), out_idx);
      std::copy(x.val.begin(), x.val.end(), out_val);
    }
    return x.nnz();
  } catch (...) {
    return -1;
  }
}

// SolveLeastSquares2x2 restatement, exposed for pinning.
void oracle_solve_ls2x2(const double* m4, const double* x2, double* out2) {
  const double m[2][2] = {{m4[0], m4[1]}, {m4[2], m4[3]}};
  SolveLS2x2(m, x2, out2);
}

void oracle_solve_ls2x2_rank_eps(const double* m4, const double* x2, double* out2, double rank_eps) {
  const double m[2][2] = {{m4[0], m4[1]}, {m4[2], m4[3]}};
  SolveLS2x2(m, x2, out2, rank_eps);
}

// backward_error.pyx:16-34 (_row_dot/_multiply): y = A x for CSR A.
static void CsrMul(const double* v, const int64_t* ci, const int64_t* ro, const double* x, double* y, int64_t n,
                   int threads) {
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
  for (int64_t r = 0; r < n; ++r) {
    double t = 0.0;
    for (int64_t k = ro[r]; k < ro[r + 1]; ++k) t += v[k] * x[ci[k]];
    y[r] = t;
  }
}

// backward_error.pyx:37-42 (_dot).  Static partition, partial sums combined in
// thread order (sequential for threads == 1).
static double Dot(const double* x, const double* y, int64_t n, int threads) {
  if (threads <= 1) {
    double t = 0.0;
    for (int64_t i = 0; i < n; ++i) t += x[i] * y[i];
    return t;
  }
  std::vector<double> part(threads, 0.0);
#pragma omp parallel num_threads(threads)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s += x[i] * y[i];
    part[t] = s;
  }
  double s = 0.0;
  for (double v : part) s += v;
  return s;
}

// PCG (methods_cython/backward_error.pyx:69-144).  Returns the status; history
// must hold max_iterations entries, *count receives the number written.
int64_t oracle_pcg(CsrData A, CsrData M, const double* b, int64_t max_iterations, double tolerance, int64_t threads_in,
                   double* history, int64_t* count_out, double* x_out) {
  const int64_t n = A.rows;
  if (n <= 0 || max_iterations <= 0 || threads_in <= 0) return -1;
  const int threads = (int)std::min<int64_t>(threads_in, n);
  const int64_t ab = A.row_offsets[0], mb = M.row_offsets[0];
  std::vector<int64_t> aro(n + 1), mro(n + 1);
  for (int64_t i = 0; i <= n; ++i) {
    aro[i] = A.row_offsets[i] - ab;
    mro[i] = M.row_offsets[i] - mb;
  }
  const double* av = A.values + ab;
  const int64_t* ai = A.col_indices + ab;
  const double* mv = M.values + mb;
  const int64_t* mi = M.col_indices + mb;
  std::vector<double> r(b, b + n), z(n), p(n), Ap(n), x(n, 0.0);
  int64_t count = 1;
  int64_t status = 0;
  CsrMul(mv, mi, mro.data(), r.data(), z.data(), n, threads);
  p = z;
  double r_r = Dot(r.data(), r.data(), n, threads);
  const double b_norm = std::sqrt(r_r);
  double r_z = Dot(r.data(), z.data(), n, threads);
  history[0] = 1.0;
  while (count < max_iterations && history[count - 1] > tolerance) {
    CsrMul(av, ai, aro.data(), p.data(), Ap.data(), n, threads);
    const double den = Dot(p.data(), Ap.data(), n, threads);
    if (!std::isfinite(r_z) || !std::isfinite(den) || r_z == 0.0 || den == 0.0) {
      status = 2;
      break;
    }
    const double alpha = r_z / den;
    if (!std::isfinite(alpha)) {
      status = 2;
      break;
    }
    double beta = r_z;
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    r_r = Dot(r.data(), r.data(), n, threads);
    double err = std::sqrt(r_r) / b_norm;
    const bool refreshed = err <= tolerance;
    if (refreshed) {
      CsrMul(av, ai, aro.data(), x.data(), Ap.data(), n, threads);
      for (int64_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
      r_r = Dot(r.data(), r.data(), n, threads);
      err = std::sqrt(r_r) / b_norm;
    }
    CsrMul(mv, mi, mro.data(), r.data(), z.data(), n, threads);
    r_z = Dot(r.data(), z.data(), n, threads);
    beta = refreshed ? 0.0 : r_z / beta;
    history[count] = err;
    count += 1;
    if (!std::isfinite(err)) {
      status = 3;
      break;
    }
    if (!std::isfinite(beta)) {
      status = 2;
      break;
    }
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  CsrMul(av, ai, aro.data(), x.data(), Ap.data(), n, threads);
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
  history[count - 1] = std::sqrt(Dot(r.data(), r.data(), n, threads)) / b_norm;
  if (!std::isfinite(history[count - 1])) {
    status = 3;
  } else if (history[count - 1] <= tolerance) {
    status = 1;
  }
  *count_out = count;
  if (x_out) std::copy(x.begin(), x.end(), x_out);
  return status;
}

}  // extern "C"
