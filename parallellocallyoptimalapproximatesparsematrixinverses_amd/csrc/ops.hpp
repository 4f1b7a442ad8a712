// Host-side entry points of the device kernels (one call = one or a few
// launches on ctx.stream).  Every matrix is CSC with sorted rows.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "device.hpp"

namespace spai {

struct Ctx {
  hipStream_t stream = nullptr;
  Pool pool;
  int device = 0;
  int num_cus = 256;
  // pinned host staging for small device-to-host reads
  void* pinned = nullptr;
  void* pinned_dev = nullptr;  // its device address (mapped)
  size_t pinned_bytes = 0;
  // Device error code in mapped host memory: kernels store a nonzero code on
  // a failed consistency check (never otherwise); ReadBack and Sync raise it
  // after their synchronisation, so such checks cost no read-back of their
  // own.  It sits after the staging area; err_dev is its device address.
  int* err_host = nullptr;
  int* err_dev = nullptr;

  Ctx();
  // Runs on a caller-owned stream (not destroyed with the context).
  explicit Ctx(hipStream_t external);
  ~Ctx();
  bool owns_stream = true;
  Ctx(const Ctx&) = delete;
  Ctx& operator=(const Ctx&) = delete;

  template <class T>
  DBuf<T> Alloc(size_t n) {
    return DBuf<T>(&pool, n);
  }
  // Synchronous small device-to-host copy through pinned memory.
  void ReadBack(void* host, const void* dev, size_t bytes);
  // Two small reads (at most 256 bytes each) with one synchronisation.
  void ReadBack2(void* h1, const void* d1, size_t b1, void* h2, const void* d2, size_t b2);
  void Sync();

 private:
  void InitHostBuffers();
  void CheckErrorCode();
};

// ---------------------------------------------------------------- patterns
PatternPtr MakePattern(Ctx& ctx, int64_t rows, int64_t cols, int64_t nnz);
// Upload a CSC matrix with int64 host indices (validated by the caller).
DMat Upload(Ctx& ctx, int64_t rows, int64_t cols, const int64_t* ptr, const int64_t* idx, const double* val);
// Download into host arrays (ptr: cols + 1, idx/val: nnz).
void Download(Ctx& ctx, const DMat& m, int64_t* ptr, int64_t* idx, double* val);

// Copies between pageable host arrays and device memory through pinned
// chunks, the host side split over worker threads (transfer.cpp).  Each
// returns when the copy is complete; the ones to the host are queued after
// the work already on ctx.stream.
void CopyToDevice(Ctx& ctx, void* dst, const void* src, size_t bytes);
// dst[i] = int(src[i] - sub); throws when a value leaves [0, 2^31 - 1].
void NarrowToDevice(Ctx& ctx, int* dst, const int64_t* src, int64_t n, int64_t sub);
void CopyToHost(Ctx& ctx, void* dst, const void* src, size_t bytes);
// dst[i] = src[i] (int32 on the device, int64 on the host).
void WidenToHost(Ctx& ctx, int64_t* dst, const int* src, int64_t n);
// f(begin, end) over slices of [0, n) on the library's host worker threads
// (and the calling thread); returns when every slice is done.
void HostParallelFor(int64_t n, const std::function<void(int64_t, int64_t)>& f);
DMat EmptyMatrix(Ctx& ctx, int64_t rows, int64_t cols);
DMat Identity(Ctx& ctx, int64_t n);
DMat Clone(Ctx& ctx, const DMat& m);
// Values-only copy sharing the pattern.
DMat ShareClone(Ctx& ctx, const DMat& m);

// ptr[0..n] = exclusive scan of counts[0..n) (ptr[n] = the sum); returns the
// sum read back to the host, or -1 without that read-back when read_total is
// false (for callers that know it, do not need it, or read it later from
// total_dev, which then receives the sum as int64).
int64_t ScanCounts(Ctx& ctx, const int* counts, int64_t n, int* ptr, bool read_total = true,
                   int64_t* total_dev = nullptr);

// ---------------------------------------------------------------- products
// Conservative product C = X * Y (Eigen's column-wise Gustavson order:
// sum over Y(:, j)'s rows k ascending, first product then +=).
// residual: C = I - X * Y with the diagonal always present, then exact zeros
// pruned.  prune_zero: drop exact zeros from C.
DMat SpGemm(Ctx& ctx, const DMat& X, const DMat& Y, bool prune_zero, bool residual);
// C = diag(d) * Y, d on the device (length Y.rows).
DMat DiagScale(Ctx& ctx, const double* d, const DMat& Y, bool prune_zero);

// ---------------------------------------------------------------- unions
enum class MergeOp : int {
  kAddScaled = 0,  // l + a * r        (AddScaled; l + 0.0 / 0.0 + a * r)
  kLinComb = 1,    // l * b + a * r    (LinearCombination(y=l, x=r))
  kSymmetric = 2,  // 0.5 * (l + r)    (EnsureSymmetry)
  kCopyLeft = 3,   // l               (prune only)
};
DMat Merge(Ctx& ctx, const DMat& L, const DMat& R, MergeOp op, double a, double b, bool prune_zero);
// Drop exact zeros (returns the input when it has none).
DMat PruneZeros(Ctx& ctx, DMat&& X);
// X(:, j) *= f[j] (f on the device), then drop exact zeros.
DMat ColScale(Ctx& ctx, const DMat& X, const double* f, bool prune_zero);
DMat Transpose(Ctx& ctx, const DMat& X);
// For every entry of Transpose(X), in its order, the position of the same
// entry in X (a fixed pattern's transposition, to move values without
// transposing again).
DBuf<int> TransposePositions(Ctx& ctx, const DMat& X);

// ---------------------------------------------------------------- reductions
double FrobDot(Ctx& ctx, const DMat& X, const DMat& Y);
double SqNorm(Ctx& ctx, const DMat& X);
// Per-column intersection dots / squared norms into out[0..cols) (device).
void ColumnDots(Ctx& ctx, const DMat& X, const DMat& Y, double* out);
void ColumnSqNorms(Ctx& ctx, const DMat& X, double* out);
// Frobenius dot (or squared norm when same) summed as the reference does:
// sequentially within each batch [first[b], first[b+1]) in CSC order, then over
// the batches in order (methods_common.cpp:148-154).
double BatchOrderedDot(Ctx& ctx, const DMat& X, const DMat& Y, const int* batch_first, int nbatches, bool same);
// Per-batch sequential squared norms for the batch column ranges (device out).
// ordered: sequential sum per batch in CSC order (the reference's order);
// otherwise a deterministic block tree per batch.
void BatchSqNorms(Ctx& ctx, const DMat& X, const int* batch_first, int nbatches, double* out, bool ordered);

// ---------------------------------------------------------------- dropping
// ApplyDroppingStrategy stage 1: keep v != 0 and (diagonal or |v| >= eps*||col||).
DMat ThresholdColumns(Ctx& ctx, const DMat& M, double eps);
// DropForMatrixType stage 1: keep |v| > sqrt(batch_sq[b]) * eps for the batch
// holding the column (batch_of_col on device).
DMat ThresholdBatches(Ctx& ctx, const DMat& X, const double* batch_sq, const int* batch_of_col, double eps);
// Per-column top-k: keep the diagonal and the k largest |v| off the diagonal.
DMat ColumnTopK(Ctx& ctx, const DMat& X, int64_t k);
// The same on caller-owned device arrays, allocating nothing: the keep flags
// and scan tile sums live in a file-scope device array of
// kTopKNoAllocScratchBytes (one per GPU).  Enqueued on `stream`, no
// synchronisation; out_ptr[cols] holds the result's nnz when it completes.
// Inputs must be CSC with int32 offsets/rows sorted per column; the outputs
// need cols + 1 offsets and room for nnz entries.  Calls that overlap on one
// GPU race on the shared scratch.
constexpr int64_t kTopKNoAllocMaxNnz = int64_t(1) << 24;
constexpr int64_t kTopKNoAllocMaxCols = int64_t(1) << 22;
constexpr size_t kTopKNoAllocScratchBytes = size_t(kTopKNoAllocMaxNnz) + 4 * 4096 + 256;
void CheckTopKNoAllocSizes(int64_t rows, int64_t cols, int64_t nnz, int64_t k);
void ColumnTopKNoAlloc(hipStream_t stream, int64_t rows, int64_t cols, int64_t nnz, const int* ptr, const int* idx,
                       const double* v, int64_t k, int* out_ptr, int* out_idx, double* out_v);
// Global select: among candidates (drop keys in CSC order), drops the
// `drop` smallest under (key, position); candidates are the off-diagonal
// entries (score mode, scores m^2 c_k + 2 m (A R)_kl) or every entry (|v| mode).
DMat DropSmallestByScore(Ctx& ctx, const DMat& M, const DMat& AR, const double* a_col_sq, int64_t drop);
DMat DropSmallestByAbs(Ctx& ctx, const DMat& X, int64_t drop);

// ---------------------------------------------------------------- misc
// Achievable streaming bandwidth (GB/s, median over `repeats` launches):
// out[0] read, out[1] copy, out[2] triad, on arrays of bytes_per_array.
void StreamBandwidth(int64_t bytes_per_array, int64_t repeats, double out[3]);
void FillValue(Ctx& ctx, double* p, int64_t n, double v);
// Inner-outer per-column step sizes (inner_outer_common.hpp:50-69): delta,
// gamma and the normalisations 1/delta from the column dots.  two_term selects
// the LOMR 2x2 solve.
void InnerOuterSteps(Ctx& ctx, int64_t n, bool lomr, bool two_term, const double* r_az, const double* az_sq,
                     const double* az_aq, const double* aq_sq, const double* r_aq, double* delta, double* gamma,
                     double* normalization);
// out[j] = GetSafeQuotient(num[j], den[j]).
void SafeQuotients(Ctx& ctx, int64_t n, const double* num, const double* den, double* out);
// Device vector of the diagonal entries of A (0 where absent).
void ExtractDiagonal(Ctx& ctx, const DMat& A, double* d);
// True when X has exactly one entry per column, on the diagonal.
bool IsFullDiagonal(Ctx& ctx, const DMat& X);

}  // namespace spai
