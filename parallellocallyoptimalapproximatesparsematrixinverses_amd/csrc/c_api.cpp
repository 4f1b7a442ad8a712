// C ABI of libspai_hip.so (include/methods_c_api.h, include/spai_ext.h).
//
// The five builders keep the reference's by-value structs and its error
// contract (src/common/methods_c_api.cpp:59-126): nothing throws across the ABI;
// a failure returns a zeroed Result whose error_message holds the reason.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

#include "engine.hpp"
#include "fixed.hpp"
#include "ops.hpp"
#include "methods_c_api.h"
#include "pcg.hpp"
#include "spai_ext.h"

namespace {

thread_local std::string g_last_error;

Result ErrorResult(const std::string& msg) {
  Result out{};
  out.error_message = new char[msg.size() + 1];
  std::memcpy(out.error_message, msg.c_str(), msg.size() + 1);
  return out;
}

// Hands M's arrays over to the Result (free_mem releases them with delete[])
// and copies the history.
Result SuccessResult(spai::EngineOutput& o) {
  Result r{};
  r.rows = o.M.rows;
  r.cols = o.M.cols;
  r.nnz = (int64_t)o.M.idx.size();
  r.iterations = o.iterations;
  r.converged = o.converged ? 1 : 0;
  r.history_size = (int64_t)o.history.size();
  if (o.M.ptr.size() != size_t(r.cols + 1)) throw std::logic_error("result offsets have the wrong length");
  if (r.nnz == 0) {
    o.M.idx.resize(1);
    o.M.val.resize(1);
  }
  r.history = new SingleIterationData[r.history_size > 0 ? r.history_size : 1];
  for (int64_t i = 0; i < r.history_size; ++i) r.history[i] = o.history[i];
  r.col_offsets = o.M.ptr.release();
  r.row_indices = o.M.idx.release();
  r.values = o.M.val.release();
  return r;
}

Result RunMethod(int64_t method, const MatrixData& A, const MatrixData& M0, const MatrixData& Pr,
                 const MethodParams& p, const SpaiExtParams* ext, double* seconds) {
  try {
    const spai::EngineParams ep = spai::ToEngineParams(p, ext);
    auto out = spai::RunEngine(method, A, p.use_initial_m ? &M0 : nullptr, p.use_preconditioner ? &Pr : nullptr,
                               ep);
    if (seconds) *seconds = out.device_seconds;
    return SuccessResult(out);
  } catch (const std::exception& ex) {
    return ErrorResult(ex.what());
  } catch (...) {
    return ErrorResult("Unknown failure.");
  }
}

int64_t Fail(const std::string& msg) {
  g_last_error = msg;
  return -1;
}

void ValidateCsr(const CsrData& m) {
  if (m.rows <= 0 || m.cols != m.rows) throw std::invalid_argument("A must be a nonempty square matrix");
  for (int64_t i = 0; i < m.rows; ++i)
    if (m.row_offsets[i] > m.row_offsets[i + 1]) throw std::invalid_argument("row offsets must be non-decreasing");
  const int64_t nnz = m.row_offsets[m.rows] - m.row_offsets[0];
  for (int64_t k = 0; k < nnz; ++k) {
    const int64_t c = m.col_indices[m.row_offsets[0] + k];
    if (c < 0 || c >= m.cols) throw std::invalid_argument("column index out of range");
  }
}

}  // namespace

struct SpaiSession {
  std::unique_ptr<spai::FixedProblem> prob;
  std::unique_ptr<spai::DCsr> a_csr, m_csr;
  std::unique_ptr<spai::PcgSolver> pcg;
  spai::DBuf<double> rhs;
  spai::DBuf<int> m_perm;  // position on S of each entry of m_csr
  bool fp32 = false;
  bool m_stale = true;
  int64_t last_count = 0;
};

struct SpaiPcgSystem {
  spai::Ctx ctx;
  std::unique_ptr<spai::DCsr> a, m;
  std::unique_ptr<spai::PcgSolver> solver;
  spai::DBuf<double> b, x, y;
  int64_t n = 0;
};

extern "C" {

Result spai_global_cg(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) {
  return RunMethod(SPAI_GLOBAL_CG, A, M0, Pr, params, nullptr, nullptr);
}

Result spai_global_mr(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) {
  return RunMethod(SPAI_GLOBAL_MR, A, M0, Pr, params, nullptr, nullptr);
}

Result spai_global_lomr(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) {
  return RunMethod(SPAI_GLOBAL_LOMR, A, M0, Pr, params, nullptr, nullptr);
}

Result inner_outer_mr(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) {
  return RunMethod(SPAI_INNER_OUTER_MR, A, M0, Pr, params, nullptr, nullptr);
}

Result inner_outer_lomr(MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params) {
  return RunMethod(SPAI_INNER_OUTER_LOMR, A, M0, Pr, params, nullptr, nullptr);
}

void free_mem(Result* result) {
  if (result == nullptr) return;
  delete[] result->col_offsets;
  delete[] result->row_indices;
  delete[] result->values;
  delete[] result->history;
  delete[] result->error_message;
  result->col_offsets = nullptr;
  result->row_indices = nullptr;
  result->values = nullptr;
  result->history = nullptr;
  result->error_message = nullptr;
}

Result spai_run_ex(int64_t method, MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params,
                   SpaiExtParams ext, double* seconds_out) {
  return RunMethod(method, A, M0, Pr, params, &ext, seconds_out);
}

Result spai_spgemm(MatrixData X, MatrixData Y, int64_t residual, int64_t prune_zero, int64_t repeats,
                   double* seconds_out) {
  try {
    spai::ValidateMatrixData(X);
    spai::ValidateMatrixData(Y);
    if (X.cols != Y.rows) throw std::invalid_argument("Sparse matrix dimensions do not match");
    if (residual && X.rows != Y.cols) throw std::invalid_argument("residual product needs a square result");
    if (repeats < 1) throw std::invalid_argument("repeats must be positive");
    spai::Ctx ctx;
    const spai::DMat x = spai::Upload(ctx, X.rows, X.cols, X.col_offsets, X.row_indices, X.values);
    const spai::DMat y = spai::Upload(ctx, Y.rows, Y.cols, Y.col_offsets, Y.row_indices, Y.values);
    spai::DMat c = spai::SpGemm(ctx, x, y, prune_zero != 0, residual != 0);
    ctx.Sync();
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t r = 0; r < repeats; ++r) c = spai::SpGemm(ctx, x, y, prune_zero != 0, residual != 0);
    ctx.Sync();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (seconds_out) *seconds_out = secs / double(repeats);
    spai::EngineOutput out;
    out.M.rows = c.rows();
    out.M.cols = c.cols();
    out.M.ptr.resize(c.cols() + 1);
    out.M.idx.resize(c.nnz());
    out.M.val.resize(c.nnz());
    spai::Download(ctx, c, out.M.ptr.data(), out.M.idx.data(), out.M.val.data());
    return SuccessResult(out);
  } catch (const std::exception& ex) {
    return ErrorResult(ex.what());
  } catch (...) {
    return ErrorResult("Unknown failure.");
  }
}

Result spai_column_topk(MatrixData X, int64_t k) {
  try {
    spai::ValidateMatrixData(X);
    if (k < 0) throw std::invalid_argument("k must be non-negative");
    spai::Ctx ctx;
    const spai::DMat x = spai::Upload(ctx, X.rows, X.cols, X.col_offsets, X.row_indices, X.values);
    const spai::DMat c = spai::ColumnTopK(ctx, x, k);
    spai::EngineOutput out;
    out.M.rows = c.rows();
    out.M.cols = c.cols();
    out.M.ptr.resize(c.cols() + 1);
    out.M.idx.resize(c.nnz());
    out.M.val.resize(c.nnz());
    spai::Download(ctx, c, out.M.ptr.data(), out.M.idx.data(), out.M.val.data());
    return SuccessResult(out);
  } catch (const std::exception& ex) {
    return ErrorResult(ex.what());
  } catch (...) {
    return ErrorResult("Unknown failure.");
  }
}

int64_t spai_column_topk_noalloc(int64_t rows, int64_t cols, int64_t nnz, const int32_t* col_offsets_dev,
                                 const int32_t* row_indices_dev, const double* values_dev, int64_t k,
                                 int32_t* out_col_offsets_dev, int32_t* out_row_indices_dev, double* out_values_dev,
                                 uint64_t stream) {
  try {
    spai::CheckTopKNoAllocSizes(rows, cols, nnz, k);
    if (!col_offsets_dev || !out_col_offsets_dev) throw std::invalid_argument("null column offsets");
    if (nnz > 0 && (!row_indices_dev || !values_dev || !out_row_indices_dev || !out_values_dev))
      throw std::invalid_argument("null entry arrays");
    spai::ColumnTopKNoAlloc(reinterpret_cast<hipStream_t>(stream), rows, cols, nnz, col_offsets_dev,
                            row_indices_dev, values_dev, k, out_col_offsets_dev, out_row_indices_dev,
                            out_values_dev);
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  } catch (...) {
    return Fail("Unknown failure.");
  }
}

int64_t spai_pcg(CsrData A, CsrData M, const double* b, int64_t max_iterations, double tolerance, double* history,
                 int64_t* count, double* x_out, int64_t m_fp32) {
  try {
    if (A.rows <= 0 || max_iterations <= 0) throw std::invalid_argument("n, max_iterations and threads must be positive");
    ValidateCsr(A);
    ValidateCsr(M);
    if (M.rows != A.rows) throw std::invalid_argument("A and M must have the same shape");
    spai::Ctx ctx;
    const int64_t n = A.rows;
    spai::DCsr a = spai::CsrFromHost(ctx, n, A.nnz, A.row_offsets, A.col_indices, A.values);
    spai::DCsr m = spai::CsrFromHost(ctx, n, M.nnz, M.row_offsets, M.col_indices, M.values);
    if (m_fp32) spai::MakeFp32Copy(ctx, m);
    auto bd = ctx.Alloc<double>(n);
    spai::CopyToDevice(ctx, bd.get(), b, sizeof(double) * n);
    spai::PcgSolver solver(ctx, a, m, m_fp32 != 0);
    int64_t cnt = 0;
    const int status = solver.Solve(bd.get(), max_iterations, tolerance, &cnt);
    solver.History(history, cnt);
    if (x_out) {
      spai::CopyToHost(ctx, x_out, solver.X(), sizeof(double) * n);
    }
    *count = cnt;
    return status;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  } catch (...) {
    return Fail("Unknown failure.");
  }
}

int64_t spai_matvec(CsrData A, const double* x, double* y) {
  try {
    ValidateCsr(A);
    spai::Ctx ctx;
    spai::DCsr a = spai::CsrFromHost(ctx, A.rows, A.nnz, A.row_offsets, A.col_indices, A.values);
    auto xd = ctx.Alloc<double>(A.rows);
    auto yd = ctx.Alloc<double>(A.rows);
    SPAI_HIP_CHECK(hipMemcpyAsync(xd.get(), x, sizeof(double) * A.rows, hipMemcpyHostToDevice, ctx.stream));
    spai::SpMV(ctx, a, xd.get(), yd.get());
    SPAI_HIP_CHECK(hipMemcpyAsync(y, yd.get(), sizeof(double) * A.rows, hipMemcpyDeviceToHost, ctx.stream));
    ctx.Sync();
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  } catch (...) {
    return Fail("Unknown failure.");
  }
}

SpaiPcgSystem* spai_pcg_prepare(CsrData A, CsrData M, int64_t m_fp32) {
  try {
    ValidateCsr(A);
    ValidateCsr(M);
    if (M.rows != A.rows) throw std::invalid_argument("A and M must have the same shape");
    auto s = std::make_unique<SpaiPcgSystem>();
    s->n = A.rows;
    s->a = std::make_unique<spai::DCsr>(
        spai::CsrFromHost(s->ctx, A.rows, A.nnz, A.row_offsets, A.col_indices, A.values));
    s->m = std::make_unique<spai::DCsr>(
        spai::CsrFromHost(s->ctx, M.rows, M.nnz, M.row_offsets, M.col_indices, M.values));
    if (m_fp32) spai::MakeFp32Copy(s->ctx, *s->m);
    s->b = s->ctx.Alloc<double>(s->n);
    s->x = s->ctx.Alloc<double>(s->n);
    s->y = s->ctx.Alloc<double>(s->n);
    s->solver = std::make_unique<spai::PcgSolver>(s->ctx, *s->a, *s->m, m_fp32 != 0);
    s->ctx.Sync();
    return s.release();
  } catch (const std::exception& ex) {
    g_last_error = ex.what();
    return nullptr;
  }
}

void spai_pcg_release(SpaiPcgSystem* system) { delete system; }

int64_t spai_pcg_solve(SpaiPcgSystem* s, const double* b, int64_t max_iterations, double tolerance, double* history,
                       int64_t* count, double* x_out) {
  try {
    if (!s) throw std::invalid_argument("null PCG system");
    if (max_iterations <= 0) throw std::invalid_argument("n, max_iterations and threads must be positive");
    spai::CopyToDevice(s->ctx, s->b.get(), b, sizeof(double) * s->n);
    int64_t cnt = 0;
    const int status = s->solver->Solve(s->b.get(), max_iterations, tolerance, &cnt);
    if (history) s->solver->History(history, cnt);
    if (x_out) {
      spai::CopyToHost(s->ctx, x_out, s->solver->X(), sizeof(double) * s->n);
    }
    s->ctx.Sync();
    if (count) *count = cnt;
    return status;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  } catch (...) {
    return Fail("Unknown failure.");
  }
}

int64_t spai_pcg_matvec(SpaiPcgSystem* s, const double* x, double* y) {
  try {
    if (!s) throw std::invalid_argument("null PCG system");
    spai::CopyToDevice(s->ctx, s->x.get(), x, sizeof(double) * s->n);
    spai::SpMV(s->ctx, *s->a, s->x.get(), s->y.get());
    spai::CopyToHost(s->ctx, y, s->y.get(), sizeof(double) * s->n);
    s->ctx.Sync();
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  } catch (...) {
    return Fail("Unknown failure.");
  }
}

int64_t spai_session_set_kernel_timing(SpaiSession* s, int64_t on) {
  try {
    if (!s) throw std::invalid_argument("null session");
    s->prob->SetKernelTiming(on != 0);
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

double spai_session_build_kernel_us(SpaiSession* s, int64_t* launches) {
  if (!s) {
    g_last_error = "null session";
    return -1.0;
  }
  return s->prob->KernelMicros(launches);
}

int64_t spai_session_info(SpaiSession* s, int64_t* out, int64_t len) {
  try {
    if (!s) throw std::invalid_argument("null session");
    if (!out || len < 0) throw std::invalid_argument("null output");
    int64_t info[spai::kFixedInfoLen];
    s->prob->FusedInfo(info);
    for (int64_t i = 0; i < len && i < spai::kFixedInfoLen; ++i) out[i] = info[i];
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

int64_t spai_stream_bandwidth(int64_t bytes_per_array, int64_t repeats, double* gbs_out) {
  try {
    if (!gbs_out) throw std::invalid_argument("null output");
    spai::StreamBandwidth(bytes_per_array, repeats, gbs_out);
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

const char* spai_last_error(void) { return g_last_error.c_str(); }

int64_t spai_device_name(char* buffer, int64_t size) {
  try {
    int dev = 0;
    SPAI_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SPAI_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    // the marketing name can be empty under restricted device access
    const std::string name = prop.name[0] != '\0' ? std::string(prop.name) : std::string("AMD GPU");
    const std::string s = name + " (" + prop.gcnArchName + ", " + std::to_string(prop.multiProcessorCount) +
                          " CUs, " + std::to_string(prop.totalGlobalMem >> 30) + " GiB)";
    if (buffer && size > 0) {
      std::strncpy(buffer, s.c_str(), (size_t)size - 1);
      buffer[size - 1] = '\0';
    }
    return (int64_t)s.size();
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

SpaiSession* spai_session_create(int64_t method, MatrixData A, MatrixData M0, MatrixData Pr, MethodParams params,
                                 SpaiExtParams ext) {
  try {
    const spai::EngineParams ep = spai::ToEngineParams(params, &ext);
    if (!ep.fixed_pattern) throw std::invalid_argument("device-resident sessions need fixed_pattern = 1");
    spai::ValidateMatrixData(A);
    if (params.use_initial_m) spai::ValidateMatrixData(M0);
    if (params.use_preconditioner) spai::ValidateMatrixData(Pr);
    auto s = std::make_unique<SpaiSession>();
    s->prob = spai::MakeFixedProblem(method, A, params.use_initial_m ? &M0 : nullptr,
                                     params.use_preconditioner ? &Pr : nullptr, ep);
    s->fp32 = ep.fp32;
    spai::Ctx& c = s->prob->context();
    s->a_csr = std::make_unique<spai::DCsr>(spai::CsrFromCsc(c, s->prob->MatrixA()));
    s->rhs = c.Alloc<double>(A.rows);
    c.Sync();
    return s.release();
  } catch (const std::exception& ex) {
    g_last_error = ex.what();
    return nullptr;
  }
}

void spai_session_destroy(SpaiSession* s) { delete s; }

int64_t spai_session_build(SpaiSession* s) {
  try {
    if (!s) throw std::invalid_argument("null session");
    spai::Ctx& c = s->prob->context();
    // PCG operand: CSR of M.  The pattern is fixed, so after the first build
    // only the values move, through the transposition kept from the first
    // build (one gather instead of a transpose), queued before the build's
    // final read-back; the solver (and its recorded graphs) is kept.
    if (!s->m_csr) {
      const int64_t it = s->prob->Build();
      const spai::DMat Md = s->prob->MatrixM();
      s->m_csr = std::make_unique<spai::DCsr>(spai::CsrFromCsc(c, Md));
      if (s->fp32) spai::MakeFp32Copy(c, *s->m_csr);
      s->m_perm = spai::TransposePositions(c, Md);
      s->pcg = std::make_unique<spai::PcgSolver>(c, *s->a_csr, *s->m_csr, s->fp32);
      return it;
    }
    const int64_t it = s->prob->BuildThen([&] {
      s->prob->GatherM(s->m_perm.get(), s->m_csr->nnz, s->m_csr->val.get(),
                       s->fp32 ? s->m_csr->val32.get() : nullptr);
      s->pcg->RefreshPreconditioner();
    });
    c.Sync();
    return it;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

int64_t spai_session_set_rhs(SpaiSession* s, const double* b) {
  try {
    if (!s) throw std::invalid_argument("null session");
    if (!b) throw std::invalid_argument("null right-hand side");
    spai::Ctx& c = s->prob->context();
    spai::CopyToDevice(c, s->rhs.get(), b, sizeof(double) * s->a_csr->n);
    c.Sync();
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

int64_t spai_session_pcg(SpaiSession* s, int64_t max_iterations, double tolerance, int64_t* count) {
  try {
    if (!s || !s->pcg) throw std::invalid_argument("session has no built preconditioner");
    int64_t cnt = 0;
    const int st = s->pcg->Solve(s->rhs.get(), max_iterations, tolerance, &cnt);
    s->last_count = cnt;
    if (count) *count = cnt;
    return st;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

int64_t spai_session_pcg_output(SpaiSession* s, double* history, double* x) {
  try {
    if (!s || !s->pcg) throw std::invalid_argument("session has no built preconditioner");
    spai::Ctx& c = s->prob->context();
    if (history) s->pcg->History(history, s->last_count);
    if (x) {
      spai::CopyToHost(c, x, s->pcg->X(), sizeof(double) * s->a_csr->n);
      c.Sync();
    }
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

Result spai_session_result(SpaiSession* s) {
  try {
    spai::EngineOutput out;
    s->prob->Output(out);
    return SuccessResult(out);
  } catch (const std::exception& ex) {
    return ErrorResult(ex.what());
  }
}

double spai_session_pcg_iteration_us(SpaiSession* s) {
  try {
    if (!s || !s->pcg) throw std::invalid_argument("session has no built preconditioner");
    return s->pcg->TimeSpmvA(20);
  } catch (const std::exception& ex) {
    g_last_error = ex.what();
    return -1.0;
  }
}

int64_t spai_session_pcg_layouts(SpaiSession* s, int64_t* out) {
  try {
    if (!s || !s->pcg) throw std::invalid_argument("session has no built preconditioner");
    if (!out) throw std::invalid_argument("null output array");
    out[0] = int64_t(s->pcg->LayoutA());
    out[1] = int64_t(s->pcg->LayoutM());
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

int64_t spai_pcg_layouts(SpaiPcgSystem* s, int64_t* out) {
  try {
    if (!s) throw std::invalid_argument("null PCG system");
    if (!out) throw std::invalid_argument("null output array");
    out[0] = int64_t(s->solver->LayoutA());
    out[1] = int64_t(s->solver->LayoutM());
    return 0;
  } catch (const std::exception& ex) {
    return Fail(ex.what());
  }
}

}  // extern "C"
