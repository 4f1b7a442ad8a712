// The reference's C++ test program (tests/test.cpp:1-179) ported to this
// library's include/methods.hpp: the same fixture (tridiag(-1, 2, -1) of
// order 12, M0 = Pr = I), the same five builders and the same checks --
// batch-count independence, the absolute stopping rule, the nnz budget of the
// dropping mode, the density stop without dropping, and the reported residual
// of inner-outer LOMR against ||I - A M||_F recomputed from M.  Only the
// matrix construction differs: methods::common::Triplet stands in for
// Eigen::Triplet.  Exits non-zero on the first failed check.
#include "methods.hpp"

#include <cmath>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

namespace mc = methods::common;
namespace mgs = methods::global_spai;
namespace mio = methods::inner_outer;

using Builder = std::function<mc::OutputResult(const mc::InputParams&, int64_t)>;

int g_failures = 0;

void Expect(bool ok, const std::string& what) {
  if (!ok) {
    std::cout << "FAILED: " << what << "\n";
    ++g_failures;
  }
}

mc::SparseMatrix Tridiagonal(int64_t n) {
  std::vector<mc::Triplet<double, int64_t>> t;
  for (int64_t i = 0; i < n; ++i) {
    t.emplace_back(i, i, 2.0);
    if (i > 0) t.emplace_back(i, i - 1, -1.0);
    if (i + 1 < n) t.emplace_back(i, i + 1, -1.0);
  }
  mc::SparseMatrix A(n, n);
  A.setFromTriplets(t.begin(), t.end());
  return A;
}

mc::SparseMatrix Identity(int64_t n) {
  mc::SparseMatrix I(n, n);
  I.setIdentity();
  return I;
}

double ResidualNorm(const mc::SparseMatrix& A, const mc::SparseMatrix& M) {
  if (A.rows() != A.cols() || M.rows() != A.cols() || M.cols() != A.rows())
    throw std::invalid_argument("A and M must be square matrices of the same shape");
  return (Identity(A.rows()) - A * M).norm();
}

// history rows 0..iterations, residual below the start
void CheckHistory(const std::string& name, const mc::OutputResult& r, double initial) {
  Expect(!r.history.empty(), name + ": history present");
  if (r.history.empty()) return;
  Expect((int64_t)r.history.size() == r.iterations + 1, name + ": one row per iteration plus the start");
  Expect(r.history.front().iteration == 0, name + ": first row is iteration 0");
  Expect(r.history.back().iteration == r.iterations, name + ": last row is the last iteration");
  Expect(r.history.back().residual_norm < initial, name + ": residual decreased");
  std::cout << name << " start " << initial << " end " << r.history.back().residual_norm << " iterations "
            << r.iterations << "\n";
}

void CheckSame(const std::string& name, const mc::OutputResult& a, const mc::OutputResult& b) {
  const double dm = (a.M - b.M).norm();
  const double dr = std::fabs(a.history.back().residual_norm - b.history.back().residual_norm);
  Expect(a.iterations == b.iterations, name + ": same iteration count");
  Expect(dm < 1e-10, name + ": same M (" + std::to_string(dm) + ")");
  Expect(dr < 1e-10, name + ": same final residual (" + std::to_string(dr) + ")");
}

void CheckBudget(const std::string& name, const mc::OutputResult& r, double max_density) {
  const int64_t n = r.M.rows();
  const int64_t budget = std::max(n, static_cast<int64_t>(std::floor(max_density * n * n)));
  Expect(r.M.nonZeros() <= budget, name + ": nnz(M) within max(n, floor(rho n^2))");
  Expect(r.iterations == 3, name + ": three iterations");
  std::cout << name << " nnz(M) " << r.M.nonZeros() << " budget " << budget << "\n";
}

void CheckDensityStop(const std::string& name, const mc::OutputResult& r, double max_density) {
  Expect(!r.history.empty(), name + ": history present");
  if (r.history.empty()) return;
  const double density = r.history.back().density_m;
  Expect(!r.converged, name + ": not converged");
  Expect(r.iterations > 0 && r.iterations < 8, name + ": stopped by density before the limit");
  Expect(density >= max_density, name + ": density reached");
  std::cout << name << " density " << density << " iterations " << r.iterations << "\n";
}

}  // namespace

int main() {
  try {
    const int64_t n = 12;
    const mc::SparseMatrix A = Tridiagonal(n);
    Expect(A.outerSize() == A.cols() && A.isCompressed() && A.nonZeros() == 3 * n - 2, "fixture shape");
    const mc::InputParams params(A, Identity(n), Identity(n), 8, 1e-14, 1.0);
    const double initial = ResidualNorm(params.A, params.M0);

    const std::vector<std::pair<std::string, Builder>> builders = {
        {"global_cg", [](const mc::InputParams& p, int64_t t) { return mgs::RunConjugateGradient(p, t); }},
        {"global_mr", [](const mc::InputParams& p, int64_t t) { return mgs::RunMinimalResidual(p, t); }},
        {"global_lomr",
         [](const mc::InputParams& p, int64_t t) { return mgs::RunLocallyOptimalMinimalResidual(p, t); }},
        {"inner_outer_mr", [](const mc::InputParams& p, int64_t t) { return mio::RunMinimalResidual(p, t); }},
        {"inner_outer_lomr",
         [](const mc::InputParams& p, int64_t t) { return mio::RunLocallyOptimalMinimalResidual(p, t); }},
    };

    // batch-count independence (reference test.cpp:74-81, 145-149)
    for (const auto& [name, run] : builders) {
      const auto one = run(params, 1), two = run(params, 2);
      CheckHistory(name + "_p1", one, initial);
      CheckHistory(name + "_p2", two, initial);
      CheckSame(name, one, two);
    }

    // absolute stopping rule (test.cpp:151-158)
    mc::InputParams absolute = params;
    absolute.tolerance = 0.9 * initial;
    const auto stopped = mgs::RunMinimalResidual(absolute, 2);
    Expect(stopped.converged && stopped.iterations == 1 && stopped.history.size() == 2, "absolute stop after 1");
    Expect(stopped.history.back().residual_norm < absolute.tolerance, "absolute stop below tolerance");

    // dropping: nnz budget, 1 and 4 batches agree (test.cpp:84-90, 160-170)
    mc::InputParams sparse = params;
    sparse.max_iterations = 3;
    sparse.max_density = 0.12;
    sparse.tolerance = 0.0;
    for (const auto& [name, run] : builders) {
      const auto one = run(sparse, 1), four = run(sparse, 4);
      CheckBudget(name + "_drop_p1", one, sparse.max_density);
      CheckBudget(name + "_drop_p4", four, sparse.max_density);
      CheckSame(name + "_drop", one, four);
    }

    // density stop without dropping (test.cpp:172-183)
    mc::InputParams dense = params;
    dense.max_iterations = 8;
    dense.tolerance = 0.0;
    dense.max_density = 0.2;
    dense.enable_dropping = false;
    for (const auto& [name, run] : builders) CheckDensityStop(name + "_no_drop", run(dense, 1), dense.max_density);

    // reported vs recomputed residual of inner-outer LOMR (test.cpp:185-196)
    mc::InputParams reliable = params;
    reliable.max_iterations = 20;
    reliable.tolerance = 1e-20;
    const auto r = mio::RunLocallyOptimalMinimalResidual(reliable, 2);
    const double actual = ResidualNorm(A, r.M);
    Expect(r.iterations == reliable.max_iterations && !r.converged, "inner-outer LOMR runs to the limit");
    Expect(std::fabs(r.history.back().residual_norm - actual) < 1e-9, "reported residual equals ||I - A M||_F");
    std::cout << "reliable residual " << r.history.back().residual_norm << " actual " << actual << "\n";

    // the reference's validation messages (methods_common.cpp:12-48,
    // inner_outer_common.hpp:16) reach C++ callers as std::invalid_argument
    auto message = [](const std::function<void()>& f) -> std::string {
      try {
        f();
      } catch (const std::invalid_argument& e) {
        return e.what();
      }
      return "";
    };
    mc::InputParams bad = params;
    bad.max_iterations = 0;
    Expect(message([&] { mgs::RunMinimalResidual(bad, 1); }) == "max_iterations must be positive", "iterations msg");
    Expect(message([&] { mgs::RunMinimalResidual(params, 0); }) == "Invalid threads count value", "threads msg");
    Expect(message([&] { mio::RunMinimalResidual(params, 1, 0); }) == "inner_iterations must be positive",
           "inner iterations msg");
    mc::InputParams wrong_shape(A, Identity(n + 1));
    Expect(message([&] { mgs::RunConjugateGradient(wrong_shape, 1); }) ==
               "Initial preconditioner M0 must have the same shape as A",
           "M0 shape msg");
  } catch (const std::exception& e) {
    std::cout << "FAILED with exception: " << e.what() << "\n";
    return 2;
  }
  std::cout << (g_failures ? "FAILED " + std::to_string(g_failures) + " checks" : std::string("all checks passed"))
            << "\n";
  return g_failures ? 1 : 0;
}
