// Preconditioned CG (methods_cython/backward_error.pyx:69-144) on the device.
#pragma once

#include <cstdint>
#include <vector>

#include "ops.hpp"

namespace spai {

// CSR matrix on the device (int32 offsets / columns); values FP64, plus an
// optional FP32 copy for the FP32-bulk preconditioner apply.
struct DCsr {
  int64_t n = 0, nnz = 0;
  DBuf<int> ptr, col;
  DBuf<double> val;
  DBuf<float> val32;
};

// Sliced ELLPACK copy of a CSR matrix for one-thread-per-row SpMV: slices of
// 64 consecutive rows (one wavefront), each slice width = its longest row,
// entries stored column-major inside a slice (entry u of row r at
// off[s] + 64 u + (r mod 64)), so the u-th loads of a wavefront are
// contiguous.  Padding has col = -1.  src[k] is the CSR position of slot k
// (-1 for padding), used to refresh the values.
struct DSell {
  int64_t n = 0, nslices = 0, len = 0;
  DBuf<int> off, col, src;
  DBuf<double> val;
  DBuf<float> val32;
};

// Diagonal layout of a CSR matrix whose entries lie on few distinct diagonals
// (grid stencils): diagonal d has offset off[d] = col - row (ascending), and
// its value in row r sits at d * n + r (0 where the row has no entry there);
// bit d of mask[r] says whether row r has one.  One thread per row reads it
// with every address known before any load returns (values, mask and the
// vector at r + off[d]), and the loads of a wavefront are contiguous.  src[k]
// is the CSR position of slot k (-1 for none), used to refresh the values.
constexpr int kDiaMaxDiagonals = 32;
struct DDia {
  int64_t n = 0, nd = 0;  // nd = 0: the matrix does not take this layout
  std::vector<int> off;   // host copy (the kernels take it as an argument)
  DBuf<int> src;
  DBuf<unsigned> mask;
  DBuf<double> val;
  DBuf<float> val32;
};

DCsr CsrFromHost(Ctx& ctx, int64_t n, int64_t nnz, const int64_t* ptr, const int64_t* col, const double* val);
// CSR of the matrix stored as CSC in m (i.e. the transpose of its CSC arrays).
DCsr CsrFromCsc(Ctx& ctx, const DMat& m);
void MakeFp32Copy(Ctx& ctx, DCsr& m);
// dst's values (and FP32 copy, if it has one) from src of the same pattern;
// dst's buffers stay where they are, so graphs recorded on them stay valid.
void CopyValues(Ctx& ctx, DCsr& dst, const DCsr& src);
// SELL layout of m; values (and an FP32 copy when fp32) gathered from m.
DSell SellFromCsr(Ctx& ctx, const DCsr& m, bool fp32);
// Re-gathers the values of s from m (same pattern).
void SellValues(Ctx& ctx, DSell& s, const DCsr& m);
// Diagonal layout of m when its entries lie on at most kDiaMaxDiagonals
// diagonals and the layout has at most 1.5 nnz slots (else nd = 0); values
// (and an FP32 copy when fp32) gathered from m.
DDia DiaFromCsr(Ctx& ctx, const DCsr& m, bool fp32);
// Re-gathers the values of d from m (same pattern).
void DiaValues(Ctx& ctx, DDia& d, const DCsr& m);
// Row layout of the PCG iteration kernels: CSR (SG lanes per row), sliced
// ELLPACK or diagonals (one thread per row).  The environment variable
// SPAI_PCG_LAYOUT caps the choice: "csr", "sell", or "auto" (the default:
// diagonals when the matrix takes them, else SELL for short rows, else CSR).
enum class PcgLayout : int { kCsr = 0, kSell = 1, kDia = 2 };
// y = A x (methods_cython/backward_error.pyx:27-34 _multiply).
void SpMV(Ctx& ctx, const DCsr& A, const double* x, double* y);

class PcgSolver {
 public:
  PcgSolver(Ctx& ctx, const DCsr& A, const DCsr& M, bool m_fp32);
  ~PcgSolver();
  // Runs from x0 = 0 for b (device pointer, length n).  Returns the status.
  int Solve(const double* b_dev, int64_t max_iterations, double tolerance, int64_t* count);
  void History(double* host, int64_t count);
  const double* X() const { return x_.get(); }
  // After M's values changed in place (same pattern): refresh the solver's
  // copies of them.  Recorded graphs stay valid.
  void RefreshPreconditioner();
  // Average device time of the A-SpMV kernel in the solve's layout
  // (microseconds; one event pair around `reps` back-to-back launches).
  double TimeSpmvA(int reps);
  PcgLayout LayoutA() const { return a_layout_; }
  PcgLayout LayoutM() const { return m_layout_; }

 private:
  void Record(int first_parity, int64_t iterations);
  void DropGraphs();
  Ctx& ctx_;
  const DCsr& A_;
  const DCsr& M_;
  bool m_fp32_;
  int64_t n_;
  unsigned grid_ = 1;
  PcgLayout a_layout_ = PcgLayout::kCsr, m_layout_ = PcgLayout::kCsr;
  DSell a_sell_, m_sell_;
  DDia a_dia_, m_dia_;
  DBuf<double> x_, r_, z_, p_, ap_, part_, hist_;
  DBuf<double> scal_;
  DBuf<int> ist_;
  const double* b_ = nullptr;
  hipGraph_t graph_[2] = {nullptr, nullptr};
  hipGraphExec_t exec_[2] = {nullptr, nullptr};
  // kTailChunk-iteration graphs, launched after the tolerance-crossing
  // refresh, when only a few iterations remain
  hipGraph_t tail_graph_[2] = {nullptr, nullptr};
  hipGraphExec_t tail_exec_[2] = {nullptr, nullptr};
  const double* graph_key_b_ = nullptr;
  int64_t last_iterations_ = 0;  // iterations of the previous solve, for the launch-ahead
};

}  // namespace spai
