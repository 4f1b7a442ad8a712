// Host-side checks of methods::common::SparseMatrix (include/methods.hpp)
// against dense brute force: triplet assembly with duplicates summed in input
// order, CSC invariants (offsets, sorted rows), identity, coeff, transpose,
// sums, differences, scaling, the sparse product and the Frobenius norm, on
// seeded random matrices of several shapes.  Runs without a device.
#include "methods.hpp"

#include <cmath>
#include <cstdint>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

namespace mc = methods::common;
using Dense = std::vector<std::vector<double>>;

int g_failures = 0;

void Expect(bool ok, const std::string& what) {
  if (!ok) {
    std::cout << "FAILED: " << what << "\n";
    ++g_failures;
  }
}

struct Pair {
  mc::SparseMatrix S;
  Dense D;
};

// random matrix with repeated entries (duplicates must add up)
Pair Random(int64_t rows, int64_t cols, int64_t entries, uint32_t seed) {
  std::mt19937_64 g(seed);
  std::uniform_int_distribution<int64_t> ri(0, rows - 1), ci(0, cols - 1);
  std::uniform_int_distribution<int> vi(-4, 4);
  std::vector<mc::Triplet<double, int64_t>> t;
  Dense D(rows, std::vector<double>(cols, 0.0));
  for (int64_t e = 0; e < entries; ++e) {
    const int64_t r = ri(g), c = ci(g);
    const double v = 0.5 * vi(g);  // exact in binary: sums do not depend on order
    t.emplace_back(r, c, v);
    D[r][c] += v;
  }
  mc::SparseMatrix S(rows, cols);
  S.setFromTriplets(t.begin(), t.end());
  return {S, D};
}

bool WellFormed(const mc::SparseMatrix& S) {
  const int64_t* p = S.outerIndexPtr();
  const int64_t* r = S.innerIndexPtr();
  if (p[0] != 0 || p[S.cols()] != S.nonZeros()) return false;
  for (int64_t j = 0; j < S.cols(); ++j) {
    if (p[j] > p[j + 1]) return false;
    for (int64_t q = p[j]; q < p[j + 1]; ++q) {
      if (r[q] < 0 || r[q] >= S.rows()) return false;
      if (q > p[j] && r[q - 1] >= r[q]) return false;
    }
  }
  return true;
}

bool Equal(const mc::SparseMatrix& S, const Dense& D, double tol = 0.0) {
  if (S.rows() != (int64_t)D.size()) return false;
  for (int64_t i = 0; i < S.rows(); ++i)
    for (int64_t j = 0; j < S.cols(); ++j)
      if (std::fabs(S.coeff(i, j) - D[i][j]) > tol) return false;
  // no entry outside the dense support unless it is a summed zero
  int64_t nz = 0;
  for (const auto& row : D)
    for (double v : row) nz += v != 0.0;
  return S.nonZeros() >= nz;
}

Dense Mul(const Dense& X, const Dense& Y) {
  const size_t n = X.size(), k = Y.size(), m = Y.empty() ? 0 : Y[0].size();
  Dense Z(n, std::vector<double>(m, 0.0));
  for (size_t i = 0; i < n; ++i)
    for (size_t l = 0; l < k; ++l)
      for (size_t j = 0; j < m; ++j) Z[i][j] += X[i][l] * Y[l][j];
  return Z;
}

}  // namespace

int main() {
  const int64_t shapes[][3] = {{7, 5, 20}, {30, 30, 200}, {64, 17, 300}, {1, 9, 4}, {40, 40, 0}};
  uint32_t seed = 1;
  for (const auto& s : shapes) {
    const std::string tag = std::to_string(s[0]) + "x" + std::to_string(s[1]);
    Pair x = Random(s[0], s[1], s[2], seed++);
    Pair y = Random(s[0], s[1], s[2], seed++);
    Pair z = Random(s[1], s[0], s[2] / 2 + 1, seed++);
    Expect(WellFormed(x.S) && WellFormed(y.S) && WellFormed(z.S), tag + ": well formed");
    Expect(Equal(x.S, x.D), tag + ": triplets with duplicates summed");
    Dense sum = x.D, diff = x.D, scaled = x.D;
    for (int64_t i = 0; i < s[0]; ++i)
      for (int64_t j = 0; j < s[1]; ++j) {
        sum[i][j] += y.D[i][j];
        diff[i][j] -= y.D[i][j];
        scaled[i][j] *= -2.5;
      }
    const mc::SparseMatrix ps = x.S + y.S, pd = x.S - y.S, pc = -2.5 * x.S;
    Expect(WellFormed(ps) && Equal(ps, sum), tag + ": sum");
    Expect(WellFormed(pd) && Equal(pd, diff), tag + ": difference");
    Expect(WellFormed(pc) && Equal(pc, scaled), tag + ": scaling");
    const mc::SparseMatrix t = x.S.transpose();
    Dense td(s[1], std::vector<double>(s[0], 0.0));
    for (int64_t i = 0; i < s[0]; ++i)
      for (int64_t j = 0; j < s[1]; ++j) td[j][i] = x.D[i][j];
    Expect(WellFormed(t) && Equal(t, td), tag + ": transpose");
    const mc::SparseMatrix prod = x.S * z.S;
    Expect(WellFormed(prod) && Equal(prod, Mul(x.D, z.D)), tag + ": product");
    double sq = 0.0;
    for (const auto& row : x.D)
      for (double v : row) sq += v * v;
    Expect(std::fabs(x.S.squaredNorm() - sq) <= 1e-12 * (1.0 + sq), tag + ": squared norm");
  }
  mc::SparseMatrix I(5, 3);
  I.setIdentity();
  Expect(WellFormed(I) && I.nonZeros() == 3 && I.coeff(2, 2) == 1.0 && I.coeff(3, 2) == 0.0, "rectangular identity");
  bool threw = false;
  try {
    std::vector<mc::Triplet<double, int64_t>> bad = {{0, 0, 1.0}, {5, 0, 1.0}};
    mc::SparseMatrix B(5, 5);
    B.setFromTriplets(bad.begin(), bad.end());
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  Expect(threw, "out-of-range triplet rejected");
  const mc::InputParams p(I);
  Expect(p.max_non_zeros == 5, "max_non_zeros budget of the given A");
  std::cout << (g_failures ? "FAILED " + std::to_string(g_failures) + " checks" : std::string("all checks passed"))
            << "\n";
  return g_failures ? 1 : 0;
}
