// Reference-semantics builders (pattern of M grows, optional dropping) on the
// device.  Control flow restates the reference method by method; every sparse
// operation is a device kernel from ops.hpp and only scalars and nnz counts
// cross to the host.

#include "engine.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <memory>
#include <stdexcept>

#include "fixed.hpp"
#include "ls2x2.hpp"
#include "ops.hpp"

namespace spai {

namespace {

constexpr double kPruneEps = std::numeric_limits<double>::epsilon();  // methods_common.hpp:32
constexpr double kDensityEps = 1e-9;                                   // methods.hpp:15

bool IsFiniteNonZero(double v, double thr) { return std::isfinite(v) && std::fabs(v) > thr; }

// GetSafeQuotient (src/common/sparse_operations.cpp:14-20).
double SafeQuotient(double num, double den) {
  if (!IsFiniteNonZero(den, 0.0)) return 0.0;
  const double q = num / den;
  return std::isfinite(q) ? q : 0.0;
}

struct SafeMatrix {
  const MatrixData* d = nullptr;
};

class State {
 public:
  State(Ctx& ctx, const EngineParams& p) : ctx_(ctx), prm_(p) {}

  Ctx& ctx_;
  const EngineParams& prm_;
  int64_t n = 0, nb = 1;
  DMat A, Pr;
  bool pr_diag = false;
  DBuf<double> pr_d;
  DBuf<double> a_col_sq;
  DBuf<int> batch_first, batch_of_col;
  DMat M, R, Z, P, AP, AZ, Tmp;
  bool preserve_symmetry = true;
  double initial_residual = 0.0;

  bool Prune() const { return !prm_.enable_dropping; }

  // MultiplyBatches (src/common/state_manager.cpp:47-66).
  DMat MulA(const DMat& Y) { return SpGemm(ctx_, A, Y, Prune(), false); }
  DMat MulPr(const DMat& Y) {
    if (pr_diag) return DiagScale(ctx_, pr_d.get(), Y, Prune());
    return SpGemm(ctx_, Pr, Y, Prune(), false);
  }
  DMat Mul(const DMat& X, const DMat& Y) { return SpGemm(ctx_, X, Y, Prune(), false); }

  // CalculateResidualForBatch (src/common/dropping.cpp:97-104).
  void Residual() { R = SpGemm(ctx_, A, M, true, true); }

  int64_t MaxNonZeros() const {
    const double m = prm_.max_density * (double)n * (double)n;
    return std::max<int64_t>(n, (int64_t)std::floor(m));
  }

  // Trace inner products: tree reductions, or the reference's batch order.
  double Dot(const DMat& X, const DMat& Y) {
    if (prm_.batch_ordered) return BatchOrderedDot(ctx_, X, Y, batch_first.get(), (int)nb, false);
    return FrobDot(ctx_, X, Y);
  }
  double Sq(const DMat& X) {
    if (prm_.batch_ordered) return BatchOrderedDot(ctx_, X, X, batch_first.get(), (int)nb, true);
    return SqNorm(ctx_, X);
  }

  double ResidualNorm() { return std::sqrt(Sq(R)); }

  // StabilizeResiduals (state_manager.cpp:81-95).
  void StabilizeResiduals() {
    const double norm = ResidualNorm();
    const double floor_ = 64.0 * std::numeric_limits<double>::epsilon() * std::sqrt((double)n);
    const double thr = std::max(10.0 * prm_.tolerance, floor_);
    if (std::isfinite(norm) && norm > thr) return;
    Residual();
  }

  // EnsureSymmetry (dropping.cpp:106-120).
  void EnsureSymmetry() {
    if (!preserve_symmetry) return;
    DMat T = Transpose(ctx_, M);
    M = Merge(ctx_, M, T, MergeOp::kSymmetric, 0.0, 0.0, false);
  }

  // ApplyDroppingStrategy (dropping.cpp:122-175).
  void ApplyDroppingStrategy() {
    EnsureSymmetry();
    M = ThresholdColumns(ctx_, M, kPruneEps);
    if (prm_.top_k > 0) {
      M = ColumnTopK(ctx_, M, prm_.top_k);
      Residual();
      return;
    }
    Residual();
    const int64_t maxnz = MaxNonZeros();
    const int64_t nz = M.nnz();
    if (nz <= maxnz) return;
    Tmp = SpGemm(ctx_, A, R, false, false);
    M = DropSmallestByScore(ctx_, M, Tmp, a_col_sq.get(), nz - maxnz);
    Residual();
  }

  // DropForMatrixType (dropping.cpp:177-224).
  bool DropForMatrixType(DMat& X) {
    if (!prm_.enable_dropping) return false;
    const int64_t original = X.nnz();
    auto bsq = ctx_.Alloc<double>(nb);
    BatchSqNorms(ctx_, X, batch_first.get(), (int)nb, bsq.get(), prm_.batch_ordered);
    X = ThresholdBatches(ctx_, X, bsq.get(), batch_of_col.get(), kPruneEps);
    if (prm_.top_k > 0) {
      X = ColumnTopK(ctx_, X, prm_.top_k);
      return X.nnz() != original;
    }
    const int64_t maxnz = MaxNonZeros();
    const int64_t nz = X.nnz();
    if (nz <= maxnz) return nz != original;
    X = DropSmallestByAbs(ctx_, X, nz - maxnz);
    return true;
  }

  // AddScaled (sparse_operations.cpp:41-49).
  DMat AddScaled(const DMat& t, const DMat& d, double c) {
    if (c == 0.0) return Merge(ctx_, t, EmptyMatrix(ctx_, t.rows(), t.cols()), MergeOp::kCopyLeft, 0, 0, true);
    return Merge(ctx_, t, d, MergeOp::kAddScaled, c, 0.0, true);
  }
  // LinearCombination(x, a, y, b) (sparse_operations.cpp:22-26): y*b + a*x.
  DMat LinearCombination(const DMat& x, double a, const DMat& y, double b) {
    return Merge(ctx_, y, x, MergeOp::kLinComb, a, b, true);
  }
};

void Validate(const EngineParams& p, int64_t n, const MatrixData* M0, const MatrixData* Pr) {
  // ValidateProblem (methods_common.cpp:12-36)
  if (M0 && (M0->rows != n || M0->cols != n))
    throw std::invalid_argument("Initial preconditioner M0 must have the same shape as A");
  if (Pr && (Pr->rows != n || Pr->cols != n)) throw std::invalid_argument("Pr must have the same shape as A");
  if (p.max_iterations <= 0) throw std::invalid_argument("max_iterations must be positive");
  if (!std::isfinite(p.tolerance) || p.tolerance < 0.0)
    throw std::invalid_argument("tolerance must be finite and non-negative");
  if (!std::isfinite(p.max_density) || p.max_density <= kDensityEps || p.max_density > 1.0)
    throw std::invalid_argument("max_density must be finite and in the interval (0, 1]");
  const double max_entries = p.max_density * (double)n * (double)n;
  if (p.enable_dropping && max_entries < (double)n)
    throw std::invalid_argument("max_density is too small to preserve the diagonal of M");
  // GetThreadCount (methods_common.cpp:38-48)
  if (p.num_threads <= 0) throw std::invalid_argument("Invalid threads count value");
}

// PrepareStateManager (methods_common.cpp:126-157).
void Prepare(State& s, const MatrixData& Ah, const MatrixData* M0, const MatrixData* Prh, bool preserve_symmetry) {
  Ctx& ctx = s.ctx_;
  const EngineParams& p = s.prm_;
  if (Ah.rows != Ah.cols || Ah.rows == 0) throw std::invalid_argument("A must be a non-empty square matrix");
  const int64_t n = Ah.cols;
  Validate(p, n, M0, Prh);
  s.n = n;
  s.nb = std::min<int64_t>(p.num_threads, n);
  s.preserve_symmetry = preserve_symmetry;
  s.A = Upload(ctx, Ah.rows, Ah.cols, Ah.col_offsets, Ah.row_indices, Ah.values);
  {
    std::vector<int> first(s.nb + 1), of_col(n);
    for (int64_t i = 0; i <= s.nb; ++i) first[i] = (int)(n * i / s.nb);
    for (int64_t b = 0; b < s.nb; ++b)
      for (int64_t j = first[b]; j < first[b + 1]; ++j) of_col[j] = (int)b;
    s.batch_first = ctx.Alloc<int>(s.nb + 1);
    s.batch_of_col = ctx.Alloc<int>(n);
    SPAI_HIP_CHECK(hipMemcpyAsync(s.batch_first.get(), first.data(), sizeof(int) * (s.nb + 1), hipMemcpyHostToDevice,
                                  ctx.stream));
    SPAI_HIP_CHECK(
        hipMemcpyAsync(s.batch_of_col.get(), of_col.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx.stream));
    ctx.Sync();
  }
  // InitializeBatches: M0 or identity.
  if (M0)
    s.M = Upload(ctx, M0->rows, M0->cols, M0->col_offsets, M0->row_indices, M0->values);
  else
    s.M = Identity(ctx, n);
  // GetColumnNormsSquared (only when dropping).
  s.a_col_sq = ctx.Alloc<double>(n);
  if (p.enable_dropping)
    ColumnSqNorms(ctx, s.A, s.a_col_sq.get());
  // GetPreconditioner / GetDiagonalInversePreconditioner.
  s.pr_d = ctx.Alloc<double>(n);
  if (Prh) {
    s.Pr = Upload(ctx, Prh->rows, Prh->cols, Prh->col_offsets, Prh->row_indices, Prh->values);
    s.pr_diag = IsFullDiagonal(ctx, s.Pr);
    if (s.pr_diag) ExtractDiagonal(ctx, s.Pr, s.pr_d.get());
  } else {
    std::vector<double> d(n);
    ExtractDiagonal(ctx, s.A, s.pr_d.get());
    ctx.ReadBack(d.data(), s.pr_d.get(), sizeof(double) * n);
    for (double v : d)
      if (!std::isfinite(v) || v <= 0.0)
        throw std::invalid_argument("A has an invalid diagonal entry; provide Pr explicitly");
    for (double& v : d) {
      v = 1.0 / v;
      if (!std::isfinite(v)) throw std::invalid_argument("A diagonal cannot be inverted safely; provide Pr explicitly");
    }
    SPAI_HIP_CHECK(hipMemcpyAsync(s.pr_d.get(), d.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx.stream));
    ctx.Sync();
    s.pr_diag = true;
  }
  s.R = EmptyMatrix(ctx, n, n);
  s.Z = EmptyMatrix(ctx, n, n);
  s.P = EmptyMatrix(ctx, n, n);
  s.AP = EmptyMatrix(ctx, n, n);
  s.AZ = EmptyMatrix(ctx, n, n);
  s.Tmp = EmptyMatrix(ctx, n, n);
  if (p.enable_dropping)
    s.ApplyDroppingStrategy();
  else
    s.Residual();
  s.initial_residual = s.ResidualNorm();
}

double Dot_(State& s, const DMat& X, const DMat& Y) { return s.Dot(X, Y); }
double Sq_(State& s, const DMat& X) { return s.Sq(X); }

// OutputResult::AppendIteration (src/common/result_operations.cpp:7-67).
bool Append(EngineOutput& out, State& s, const DMat& direction) {
  const int64_t it = (int64_t)out.history.size();
  out.iterations = it;
  const double entries = (double)s.n * (double)s.n;
  const double rn = out.history.empty() ? s.initial_residual : s.ResidualNorm();
  out.history.push_back(SingleIterationData{it, rn, (double)s.M.nnz() / entries, (double)direction.nnz() / entries});
  if (rn < s.prm_.tolerance) return out.converged = true;
  return !s.prm_.enable_dropping && s.prm_.max_density < 1.0 && out.history.back().density_m >= s.prm_.max_density;
}

// RunConjugateGradient (src/global_spai/conjugate_gradient.cpp:15-83).
void RunCG(State& s, EngineOutput& out) {
  s.Z = s.MulPr(s.R);
  s.P = ShareClone(s.ctx_, s.Z);
  double r_z = Dot_(s, s.R, s.Z);
  s.DropForMatrixType(s.P);
  if (Append(out, s, s.P)) return;
  s.AP = s.MulA(s.P);
  for (int64_t i = 1; i <= s.prm_.max_iterations; ++i) {
    const double den = Dot_(s, s.P, s.AP);
    if (!IsFiniteNonZero(r_z, 0.0) || !IsFiniteNonZero(den, 0.0)) break;
    const double alpha = r_z / den;
    if (!std::isfinite(alpha)) break;
    const double prev = r_z;
    s.M = s.AddScaled(s.M, s.P, alpha);
    if (s.prm_.enable_dropping)
      s.ApplyDroppingStrategy();
    else
      s.R = s.AddScaled(s.R, s.AP, -alpha);
    if (Append(out, s, s.P)) break;
    s.Z = s.MulPr(s.R);
    r_z = Dot_(s, s.R, s.Z);
    const double beta = r_z / prev;
    if (!std::isfinite(beta)) break;
    s.P = s.LinearCombination(s.Z, 1.0, s.P, beta);
    s.DropForMatrixType(s.P);
    s.AP = s.MulA(s.P);
  }
}

// RunMinimalResidual (src/global_spai/minimal_residual.cpp:31-93).
void RunMR(State& s, EngineOutput& out) {
  s.Z = s.MulPr(s.R);
  if (Append(out, s, s.Z)) return;
  double num = 0.0, den = 0.0;
  auto project = [&]() {
    s.AZ = s.MulA(s.Z);
    s.Tmp = s.MulPr(s.AZ);
    num = Dot_(s, s.Z, s.Tmp);
    den = Sq_(s, s.Tmp);
  };
  project();
  for (int64_t i = 1; i <= s.prm_.max_iterations; ++i) {
    if (!IsFiniteNonZero(den, 0.0)) break;
    const double alpha = num / den;
    if (!std::isfinite(alpha)) break;
    s.M = s.AddScaled(s.M, s.Z, alpha);
    if (s.prm_.enable_dropping)
      s.ApplyDroppingStrategy();
    else
      s.Residual();
    if (Append(out, s, s.Z)) break;
    s.Z = s.MulPr(s.R);
    project();
  }
}

// RunLocallyOptimalMinimalResidual (src/global_spai/locally_minimal_residual.cpp:54-118).
void RunLOMR(State& s, EngineOutput& out) {
  s.Z = s.MulPr(s.R);
  if (Append(out, s, s.Z)) return;
  for (int64_t i = 1; i <= s.prm_.max_iterations; ++i) {
    s.AP = s.MulA(s.P);
    s.AZ = s.MulA(s.Z);
    s.Tmp = s.MulPr(s.AZ);
    double az_praz = Dot_(s, s.AZ, s.Tmp);
    double z_az = Dot_(s, s.Z, s.AZ);
    double ap_prap = 0.0, az_prap = 0.0, z_ap = 0.0;
    if (i > 1) {
      s.Tmp = s.MulPr(s.AP);
      ap_prap = Dot_(s, s.AP, s.Tmp);
      az_prap = Dot_(s, s.AZ, s.Tmp);
      z_ap = Dot_(s, s.Z, s.AP);
    }
    // GetDeltaGamma (locally_minimal_residual.cpp:36-50)
    double delta = 0.0, gamma = 0.0;
    if (i == 1) {
      if (!IsFiniteNonZero(az_praz, 0.0)) break;
      delta = z_az / az_praz;
      gamma = 1.0;
    } else {
      SolveLeastSquares2x2(az_praz, az_prap, az_prap, ap_prap, z_az, z_ap, delta, gamma);
    }
    if (!std::isfinite(delta) || !std::isfinite(gamma)) break;
    s.M = s.AddScaled(s.M, s.Z, delta);
    s.M = s.AddScaled(s.M, s.P, gamma);
    if (s.prm_.enable_dropping) {
      s.ApplyDroppingStrategy();
    } else {
      s.R = s.AddScaled(s.R, s.AZ, -delta);
      s.R = s.AddScaled(s.R, s.AP, -gamma);
    }
    const double ratio = delta != 0.0 ? gamma / delta : 0.0;
    s.P = s.LinearCombination(s.Z, 1.0, s.P, std::isfinite(ratio) ? ratio : 0.0);
    s.DropForMatrixType(s.P);
    if (Append(out, s, s.P)) break;
    s.Z = s.MulPr(s.R);
  }
}

// inner_outer::Run (src/inner_outer/inner_outer_common.hpp:14-107).
void RunInnerOuter(State& s, EngineOutput& out, bool lomr) {
  Ctx& ctx = s.ctx_;
  const int64_t n = s.n;
  if (Append(out, s, s.P)) return;
  DMat q = EmptyMatrix(ctx, n, n);
  auto cols = ctx.Alloc<double>(8 * n);
  double* r_az = cols.get();
  double* az_sq = r_az + n;
  double* az_aq = az_sq + n;
  double* aq_sq = az_aq + n;
  double* r_aq = aq_sq + n;
  double* delta = r_aq + n;
  double* gamma = delta + n;
  double* normz = gamma + n;
  FillValue(ctx, normz, n, 1.0);
  for (int64_t iteration = 0; iteration < s.prm_.max_iterations;) {
    DMat outer_m = ShareClone(ctx, s.M);
    s.Residual();
    q = EmptyMatrix(ctx, n, n);
    const int64_t inner_count = std::min(s.prm_.inner_iterations, s.prm_.max_iterations - iteration);
    for (int64_t inner = 0; inner < inner_count; ++inner, ++iteration) {
      s.Z = s.Mul(outer_m, s.R);
      s.AZ = s.MulA(s.Z);
      DMat AQ;
      if (lomr) AQ = SpGemm(ctx, s.A, q, false, false);
      ColumnDots(ctx, s.R, s.AZ, r_az);
      ColumnSqNorms(ctx, s.AZ, az_sq);
      const bool two_term = lomr && inner > 0;
      if (two_term) {
        ColumnDots(ctx, s.AZ, AQ, az_aq);
        ColumnSqNorms(ctx, AQ, aq_sq);
        ColumnDots(ctx, s.R, AQ, r_aq);
      }
      InnerOuterSteps(ctx, n, lomr, two_term, r_az, az_sq, az_aq, aq_sq, r_aq, delta, gamma, normz);
      s.P = ColScale(ctx, s.Z, delta, true);
      if (lomr) {
        q = ColScale(ctx, q, gamma, true);
        s.P = s.AddScaled(s.P, q, 1.0);
      }
      const bool dropped = s.DropForMatrixType(s.P);
      s.AP = s.MulA(s.P);
      if (dropped) {
        ColumnDots(ctx, s.R, s.AP, r_az);
        ColumnSqNorms(ctx, s.AP, az_sq);
        SafeQuotients(ctx, n, r_az, az_sq, delta);
        s.P = ColScale(ctx, s.P, delta, true);
        s.AP = ColScale(ctx, s.AP, delta, true);
      }
      s.M = s.AddScaled(s.M, s.P, 1.0);
      if (lomr) q = ColScale(ctx, s.P, normz, true);
      if (s.prm_.enable_dropping) {
        s.ApplyDroppingStrategy();
      } else {
        s.R = s.AddScaled(s.R, s.AP, -1.0);
        s.StabilizeResiduals();
      }
      if (Append(out, s, s.P)) return;
    }
  }
}

// The checks run over slices on the host worker threads; the first failing
// check in the reference's order (offsets, then row indices) is reported.
void ValidateCsc(const MatrixData& d) {
  if (d.rows <= 0 || d.cols <= 0 || d.nnz <= 0)
    throw std::invalid_argument("Sparse matrix dimensions and nnz must be positive");
  std::atomic<bool> bad_offsets{false}, bad_rows{false};
  HostParallelFor(d.cols, [&](int64_t b, int64_t e) {
    bool bad = false;
    for (int64_t i = b; i < e; ++i) bad |= d.col_offsets[i] > d.col_offsets[i + 1];
    if (bad) bad_offsets.store(true);
  });
  if (bad_offsets.load()) throw std::invalid_argument("Sparse matrix column offsets must be non-decreasing");
  HostParallelFor(d.nnz, [&](int64_t b, int64_t e) {
    bool bad = false;
    for (int64_t i = b; i < e; ++i) bad |= d.row_indices[i] < 0 || d.row_indices[i] >= d.rows;
    if (bad) bad_rows.store(true);
  });
  if (bad_rows.load()) throw std::invalid_argument("Sparse matrix row idx is out of range");
}

}  // namespace

void ValidateMatrixData(const MatrixData& d) { ValidateCsc(d); }

EngineParams ToEngineParams(const MethodParams& p, const SpaiExtParams* ext) {
  EngineParams e;
  e.max_iterations = p.max_iterations;
  e.tolerance = p.tolerance;
  e.max_density = p.max_density;
  e.enable_dropping = p.enable_dropping != 0;
  e.num_threads = p.num_threads;
  e.inner_iterations = p.inner_iterations;
  if (ext) {
    e.fixed_pattern = ext->fixed_pattern != 0;
    e.fp32 = ext->precision != 0;
    e.top_k = ext->column_top_k;
    e.batch_ordered = ext->batch_ordered_reductions != 0;
  }
  return e;
}

EngineOutput RunEngine(int64_t method, const MatrixData& A, const MatrixData* M0, const MatrixData* Pr,
                       const EngineParams& params, bool check_csc) {
  if (method < 0 || method > 4) throw std::invalid_argument("unknown method");
  // GetInputParams (methods_c_api.cpp:42-57): A, then M0, then Pr.
  if (check_csc) {
    ValidateCsc(A);
    if (M0) ValidateCsc(*M0);
    if (Pr) ValidateCsc(*Pr);
  }
  const bool inner_outer = method >= SPAI_INNER_OUTER_MR;
  if (inner_outer && params.inner_iterations <= 0) throw std::invalid_argument("inner_iterations must be positive");
  if (params.fp32 && !params.fixed_pattern) throw std::invalid_argument("precision = 1 requires fixed_pattern = 1");
  if (params.fixed_pattern) return RunFixedEngine(method, A, M0, Pr, params);

  Ctx ctx;
  hipEvent_t e0, e1;
  SPAI_HIP_CHECK(hipEventCreate(&e0));
  SPAI_HIP_CHECK(hipEventCreate(&e1));
  SPAI_HIP_CHECK(hipEventRecord(e0, ctx.stream));
  EngineOutput out;
  {
    State s(ctx, params);
    Prepare(s, A, M0, Pr, !inner_outer);
    switch (method) {
      case SPAI_GLOBAL_CG: RunCG(s, out); break;
      case SPAI_GLOBAL_MR: RunMR(s, out); break;
      case SPAI_GLOBAL_LOMR: RunLOMR(s, out); break;
      case SPAI_INNER_OUTER_MR: RunInnerOuter(s, out, false); break;
      default: RunInnerOuter(s, out, true); break;
    }
    out.M.rows = s.M.rows();
    out.M.cols = s.M.cols();
    out.M.ptr.resize(out.M.cols + 1);
    out.M.idx.resize(s.M.nnz());
    out.M.val.resize(s.M.nnz());
    Download(ctx, s.M, out.M.ptr.data(), out.M.idx.data(), out.M.val.data());
  }
  SPAI_HIP_CHECK(hipEventRecord(e1, ctx.stream));
  SPAI_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  SPAI_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  out.device_seconds = ms * 1e-3;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return out;
}

}  // namespace spai
