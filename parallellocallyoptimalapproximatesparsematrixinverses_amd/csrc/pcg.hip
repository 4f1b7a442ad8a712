// Preconditioned conjugate gradient of methods_cython/backward_error.pyx:69-144
// on the device.
//
// One iteration is five kernels: SpMV(A, p) + <p, Ap>; the x / r update +
// <r, r>; the true-residual refresh (a no-op unless the recursive residual
// crossed the tolerance, backward_error.pyx:116-122); SpMV(M, r) + <r, z>; and
// the direction update.  Each reduction is finished by the last block of its
// kernel (fixed-order sum of the block partials, so results are deterministic),
// which also applies the reference's breakdown and stop rules to a device-side
// state.  Iterations are replayed in chunks from a captured hipGraph; once the
// state says done, the remaining kernels of a chunk return at once.

#include "pcg.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "dev_common.hpp"

namespace spai {

namespace {

constexpr int kBlk = 256;
constexpr int kChunk = 32;  // iterations per captured graph

enum DSlot : int { kRz = 0, kBnorm, kAlpha, kBeta, kErr, kTol, kNumD };
enum ISlot : int { kCount = 0, kStatus, kDone, kRefreshed, kMax, kNumI };
enum CSlot : int { kCntSpmvA = 0, kCntUpdate, kCntRefresh, kCntSpmvM, kCntInit, kCntFinal, kNumC };

template <int SG, class MV>
__device__ __forceinline__ double RowDot(const int* __restrict__ ptr, const int* __restrict__ col,
                                         const MV* __restrict__ val, const double* __restrict__ x, int row, int sl) {
  double s = 0.0;
  const int e = ptr[row + 1];
  for (int k = ptr[row] + sl; k < e; k += SG) s += double(val[k]) * x[col[k]];
#pragma unroll
  for (int off = SG / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, SG);
  return s;
}

// Block sum of v, partial to part[block]; returns true in the grid's last block
// (after which part[] holds every block's partial).
__device__ __forceinline__ bool LastBlock(double v, double* __restrict__ part, unsigned* __restrict__ counter) {
  __shared__ double red[16];
  __shared__ int last;
  v = BlockSum<double>(v, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v;
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (last) __threadfence();
  return last != 0;
}

// Fixed-order sum of the partials (last block only); resets the counter.
__device__ __forceinline__ double FinishSum(const double* __restrict__ part, unsigned* __restrict__ counter) {
  __shared__ double red[16];
  double s = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
    s += __hip_atomic_load(&part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s = BlockSum<double>(s, red);
  if (threadIdx.x == 0) *counter = 0u;
  return s;
}

__device__ __forceinline__ bool Finite(double v) { return isfinite(v); }

// r = b, x = 0, z = M b, p = z; <r, r>, <r, z>; history[0] = 1.
template <int SG, class MV>
__global__ __launch_bounds__(kBlk) void PcgInit(int n, const int* __restrict__ mp, const int* __restrict__ mc,
                                                const MV* __restrict__ mv, const double* __restrict__ b,
                                                double* __restrict__ x, double* __restrict__ r,
                                                double* __restrict__ z, double* __restrict__ p,
                                                double* __restrict__ part, unsigned* __restrict__ cnt,
                                                double* __restrict__ ds, int* __restrict__ is,
                                                double* __restrict__ hist) {
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double rr = 0.0, rz = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double zr = RowDot<SG, MV>(mp, mc, mv, b, row, sl);
    if (sl == 0) {
      const double br = b[row];
      r[row] = br;
      x[row] = 0.0;
      z[row] = zr;
      p[row] = zr;
      rr += br * br;
      rz += br * zr;
    }
  }
  // two reductions: rr into part[0..G), rz into part[G..2G)
  const bool last1 = LastBlock(rr, part, cnt + kCntInit);
  __shared__ double rrs;
  if (last1) {
    const double s = FinishSum(part, cnt + kCntInit);
    if (threadIdx.x == 0) rrs = s;
  }
  const bool last2 = LastBlock(rz, part + gridDim.x, cnt + kCntFinal);
  if (last2) {
    const double s = FinishSum(part + gridDim.x, cnt + kCntFinal);
    if (threadIdx.x == 0) {
      // both sums are final in the block that finishes the second one
      double rr_total = 0.0;
      for (unsigned bb = 0; bb < gridDim.x; ++bb)
        rr_total += __hip_atomic_load(&part[bb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      (void)rrs;
      ds[kBnorm] = sqrt(rr_total);
      ds[kRz] = s;
      hist[0] = 1.0;
      is[kCount] = 1;
      is[kStatus] = 0;
      is[kRefreshed] = 0;
      is[kDone] = (1 < is[kMax] && 1.0 > ds[kTol]) ? 0 : 1;
    }
  }
}

// Ap = A p and <p, Ap>; last block: alpha (backward_error.pyx:101-109).
template <int SG>
__global__ __launch_bounds__(kBlk) void PcgSpmvA(int n, const int* __restrict__ ap_, const int* __restrict__ ac,
                                                 const double* __restrict__ av, const double* __restrict__ p,
                                                 double* __restrict__ Ap, double* __restrict__ part,
                                                 unsigned* __restrict__ cnt, double* __restrict__ ds,
                                                 int* __restrict__ is, int force) {
  if (!force && is[kDone]) return;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, double>(ap_, ac, av, p, row, sl);
    if (sl == 0) {
      Ap[row] = s;
      acc += p[row] * s;
    }
  }
  if (!LastBlock(acc, part, cnt + kCntSpmvA)) return;
  const double den = FinishSum(part, cnt + kCntSpmvA);
  if (threadIdx.x != 0 || force) return;
  const double r_z = ds[kRz];
  if (!Finite(r_z) || !Finite(den) || r_z == 0.0 || den == 0.0) {
    is[kStatus] = 2;
    is[kDone] = 1;
    return;
  }
  const double alpha = r_z / den;
  if (!Finite(alpha)) {
    is[kStatus] = 2;
    is[kDone] = 1;
    return;
  }
  ds[kAlpha] = alpha;
}

// x += alpha p; r -= alpha Ap; error = ||r|| / ||b|| (backward_error.pyx:111-116).
__global__ __launch_bounds__(kBlk) void PcgUpdate(int n, double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ p, const double* __restrict__ Ap,
                                                  double* __restrict__ part, unsigned* __restrict__ cnt,
                                                  double* __restrict__ ds, int* __restrict__ is) {
  if (is[kDone]) return;
  const double alpha = ds[kAlpha];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    acc += ri * ri;
  }
  if (!LastBlock(acc, part, cnt + kCntUpdate)) return;
  const double rr = FinishSum(part, cnt + kCntUpdate);
  if (threadIdx.x != 0) return;
  const double err = sqrt(rr) / ds[kBnorm];
  ds[kErr] = err;
  is[kRefreshed] = err <= ds[kTol] ? 1 : 0;
}

// Refresh at the tolerance crossing: r = b - A x (backward_error.pyx:117-122).
template <int SG>
__global__ __launch_bounds__(kBlk) void PcgRefresh(int n, const int* __restrict__ ap_, const int* __restrict__ ac,
                                                   const double* __restrict__ av, const double* __restrict__ x,
                                                   const double* __restrict__ b, double* __restrict__ r,
                                                   double* __restrict__ Ap, double* __restrict__ part,
                                                   unsigned* __restrict__ cnt, double* __restrict__ ds,
                                                   int* __restrict__ is) {
  if (is[kDone] || !is[kRefreshed]) return;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, double>(ap_, ac, av, x, row, sl);
    if (sl == 0) {
      Ap[row] = s;
      const double ri = b[row] - s;
      r[row] = ri;
      acc += ri * ri;
    }
  }
  if (!LastBlock(acc, part, cnt + kCntRefresh)) return;
  const double rr = FinishSum(part, cnt + kCntRefresh);
  if (threadIdx.x != 0) return;
  ds[kErr] = sqrt(rr) / ds[kBnorm];
}

// z = M r, <r, z>; last block: beta, history, stop rules (backward_error.pyx:123-133,100).
template <int SG, class MV>
__global__ __launch_bounds__(kBlk) void PcgSpmvM(int n, const int* __restrict__ mp, const int* __restrict__ mc,
                                                 const MV* __restrict__ mv, const double* __restrict__ r,
                                                 double* __restrict__ z, double* __restrict__ part,
                                                 unsigned* __restrict__ cnt, double* __restrict__ ds,
                                                 int* __restrict__ is, double* __restrict__ hist) {
  if (is[kDone]) return;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, MV>(mp, mc, mv, r, row, sl);
    if (sl == 0) {
      z[row] = s;
      acc += r[row] * s;
    }
  }
  if (!LastBlock(acc, part, cnt + kCntSpmvM)) return;
  const double r_z = FinishSum(part, cnt + kCntSpmvM);
  if (threadIdx.x != 0) return;
  const bool refreshed = is[kRefreshed] != 0;
  const double beta = refreshed ? 0.0 : r_z / ds[kRz];
  ds[kRz] = r_z;
  const double err = ds[kErr];
  const int count = is[kCount];
  hist[count] = err;
  is[kCount] = count + 1;
  ds[kBeta] = beta;
  if (!Finite(err)) {
    is[kStatus] = 3;
    is[kDone] = 1;
    return;
  }
  if (!Finite(beta)) {
    is[kStatus] = 2;
    is[kDone] = 1;
    return;
  }
  if (!(count + 1 < is[kMax] && err > ds[kTol])) is[kDone] = 1;
}

// p = z + beta p (backward_error.pyx:134-135).
__global__ __launch_bounds__(kBlk) void PcgDirection(int n, double* __restrict__ p, const double* __restrict__ z,
                                                     const double* __restrict__ ds, const int* __restrict__ is) {
  if (is[kDone]) return;
  const double beta = ds[kBeta];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
}

// Final true residual overwrites the last history entry (backward_error.pyx:136-143).
template <int SG>
__global__ __launch_bounds__(kBlk) void PcgFinal(int n, const int* __restrict__ ap_, const int* __restrict__ ac,
                                                 const double* __restrict__ av, const double* __restrict__ x,
                                                 const double* __restrict__ b, double* __restrict__ r,
                                                 double* __restrict__ Ap, double* __restrict__ part,
                                                 unsigned* __restrict__ cnt, double* __restrict__ ds,
                                                 int* __restrict__ is, double* __restrict__ hist) {
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, double>(ap_, ac, av, x, row, sl);
    if (sl == 0) {
      Ap[row] = s;
      const double ri = b[row] - s;
      r[row] = ri;
      acc += ri * ri;
    }
  }
  if (!LastBlock(acc, part, cnt + kCntFinal)) return;
  const double rr = FinishSum(part, cnt + kCntFinal);
  if (threadIdx.x != 0) return;
  const int count = is[kCount];
  const double h = sqrt(rr) / ds[kBnorm];
  hist[count - 1] = h;
  if (!Finite(h))
    is[kStatus] = 3;
  else if (h <= ds[kTol])
    is[kStatus] = 1;
}

__global__ void CsrFromInt64(int64_t n, const int64_t* __restrict__ in, int64_t base, int* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = int(in[i] - base);
}

__global__ void ToFloat(int64_t n, const double* __restrict__ in, float* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = float(in[i]);
}

unsigned Grid(int64_t rows, int rows_per_block, int cus) {
  int64_t g = (rows + rows_per_block - 1) / rows_per_block;
  const int64_t cap = int64_t(cus) * 8;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// Lanes per row chosen from the mean row length.
int SubgroupFor(int64_t n, int64_t nnz) {
  const double mean = n > 0 ? double(nnz) / double(n) : 0.0;
  if (mean <= 6.0) return 4;
  if (mean <= 16.0) return 8;
  if (mean <= 48.0) return 16;
  return 32;
}

template <int SG>
__global__ __launch_bounds__(kBlk) void SpmvKernel(int n, const int* __restrict__ ap_, const int* __restrict__ ac,
                                                   const double* __restrict__ av, const double* __restrict__ x,
                                                   double* __restrict__ y) {
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, double>(ap_, ac, av, x, row, sl);
    if (sl == 0) y[row] = s;
  }
}

}  // namespace

DCsr CsrFromHost(Ctx& ctx, int64_t n, int64_t nnz, const int64_t* ptr, const int64_t* col, const double* val) {
  if (nnz > kMaxDeviceNnz || n > kMaxDeviceNnz) throw std::runtime_error("matrix exceeds the int32 device range");
  DCsr m;
  m.n = n;
  const int64_t base = ptr[0];
  m.nnz = ptr[n] - base;
  m.ptr = ctx.Alloc<int>(n + 1);
  m.col = ctx.Alloc<int>(std::max<int64_t>(m.nnz, 1));
  m.val = ctx.Alloc<double>(std::max<int64_t>(m.nnz, 1));
  auto tmp = ctx.Alloc<int64_t>(std::max<int64_t>(n + 1, m.nnz));
  SPAI_HIP_CHECK(hipMemcpyAsync(tmp.get(), ptr, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, ctx.stream));
  CsrFromInt64<<<(unsigned)((n + 1 + 255) / 256), 256, 0, ctx.stream>>>(n + 1, tmp.get(), base, m.ptr.get());
  ctx.Sync();
  if (m.nnz > 0) {
    SPAI_HIP_CHECK(hipMemcpyAsync(tmp.get(), col + base, sizeof(int64_t) * m.nnz, hipMemcpyHostToDevice, ctx.stream));
    CsrFromInt64<<<(unsigned)((m.nnz + 255) / 256), 256, 0, ctx.stream>>>(m.nnz, tmp.get(), 0, m.col.get());
    SPAI_HIP_CHECK(hipMemcpyAsync(m.val.get(), val + base, sizeof(double) * m.nnz, hipMemcpyHostToDevice, ctx.stream));
  }
  ctx.Sync();
  return m;
}

DCsr CsrFromCsc(Ctx& ctx, const DMat& m) {
  DMat t = Transpose(ctx, m);
  DCsr c;
  c.n = m.rows();
  c.nnz = m.nnz();
  c.ptr = std::move(t.pat->ptr);
  c.col = std::move(t.pat->idx);
  c.val = std::move(t.val);
  return c;
}

void MakeFp32Copy(Ctx& ctx, DCsr& m) {
  m.val32 = ctx.Alloc<float>(std::max<int64_t>(m.nnz, 1));
  if (m.nnz > 0)
    ToFloat<<<(unsigned)((m.nnz + 255) / 256), 256, 0, ctx.stream>>>(m.nnz, m.val.get(), m.val32.get());
}

PcgSolver::PcgSolver(Ctx& ctx, const DCsr& A, const DCsr& M, bool m_fp32)
    : ctx_(ctx), A_(A), M_(M), m_fp32_(m_fp32), n_(A.n) {
  if (M.n != A.n) throw std::invalid_argument("A and M must have the same shape");
  if (m_fp32 && !M.val32.get()) throw std::invalid_argument("FP32 copy of M missing");
  grid_ = Grid(n_, 64, ctx.num_cus);
  x_ = ctx.Alloc<double>(n_);
  r_ = ctx.Alloc<double>(n_);
  z_ = ctx.Alloc<double>(n_);
  p_ = ctx.Alloc<double>(n_);
  ap_ = ctx.Alloc<double>(n_);
  part_ = ctx.Alloc<double>(2 * grid_ + 2);
  scal_ = ctx.Alloc<double>(kNumD);
  ist_ = ctx.Alloc<int>(kNumI);
  counters_ = ctx.Alloc<unsigned>(kNumC);
  SPAI_HIP_CHECK(hipMemsetAsync(counters_.get(), 0, sizeof(unsigned) * kNumC, ctx.stream));
}

PcgSolver::~PcgSolver() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
}

#define SPAI_SG_DISPATCH(sg, ...) \
  switch (sg) {                   \
    case 4: {                     \
      constexpr int SG = 4;       \
      __VA_ARGS__;                \
    } break;                      \
    case 8: {                     \
      constexpr int SG = 8;       \
      __VA_ARGS__;                \
    } break;                      \
    case 16: {                    \
      constexpr int SG = 16;      \
      __VA_ARGS__;                \
    } break;                      \
    default: {                    \
      constexpr int SG = 32;      \
      __VA_ARGS__;                \
    } break;                      \
  }

void PcgSolver::Record(bool fused, int64_t iterations) {
  (void)fused;
  const int sga = SubgroupFor(A_.n, A_.nnz), sgm = SubgroupFor(M_.n, M_.nnz);
  hipStream_t st = ctx_.stream;
  const int n = (int)n_;
  for (int64_t it = 0; it < iterations; ++it) {
    SPAI_SG_DISPATCH(sga, (PcgSpmvA<SG><<<grid_, kBlk, 0, st>>>(n, A_.ptr.get(), A_.col.get(), A_.val.get(),
                                                                p_.get(), ap_.get(), part_.get(), counters_.get(),
                                                                scal_.get(), ist_.get(), 0)));
    PcgUpdate<<<grid_, kBlk, 0, st>>>(n, x_.get(), r_.get(), p_.get(), ap_.get(), part_.get(), counters_.get(),
                                      scal_.get(), ist_.get());
    SPAI_SG_DISPATCH(sga, (PcgRefresh<SG><<<grid_, kBlk, 0, st>>>(n, A_.ptr.get(), A_.col.get(), A_.val.get(),
                                                                  x_.get(), b_, r_.get(), ap_.get(), part_.get(),
                                                                  counters_.get(), scal_.get(), ist_.get())));
    if (m_fp32_) {
      SPAI_SG_DISPATCH(sgm, (PcgSpmvM<SG, float><<<grid_, kBlk, 0, st>>>(
                                n, M_.ptr.get(), M_.col.get(), M_.val32.get(), r_.get(), z_.get(), part_.get(),
                                counters_.get(), scal_.get(), ist_.get(), hist_.get())));
    } else {
      SPAI_SG_DISPATCH(sgm, (PcgSpmvM<SG, double><<<grid_, kBlk, 0, st>>>(
                                n, M_.ptr.get(), M_.col.get(), M_.val.get(), r_.get(), z_.get(), part_.get(),
                                counters_.get(), scal_.get(), ist_.get(), hist_.get())));
    }
    PcgDirection<<<grid_, kBlk, 0, st>>>(n, p_.get(), z_.get(), scal_.get(), ist_.get());
  }
}

int PcgSolver::Solve(const double* b_dev, int64_t max_iterations, double tolerance, int64_t* count) {
  if (max_iterations <= 0) throw std::invalid_argument("n, max_iterations and threads must be positive");
  if (max_iterations > (int64_t(1) << 30)) throw std::invalid_argument("max_iterations too large");
  hipStream_t st = ctx_.stream;
  const int n = (int)n_;
  if ((int64_t)hist_.size() < max_iterations) {
    hist_ = ctx_.Alloc<double>(max_iterations);
    if (exec_) {
      (void)hipGraphExecDestroy(exec_);
      exec_ = nullptr;
    }
    if (graph_) {
      (void)hipGraphDestroy(graph_);
      graph_ = nullptr;
    }
  }
  b_ = b_dev;
  if (graph_key_b_ != b_dev && exec_) {
    (void)hipGraphExecDestroy(exec_);
    exec_ = nullptr;
    (void)hipGraphDestroy(graph_);
    graph_ = nullptr;
  }
  const double hd[kNumD] = {0, 0, 0, 0, 0, tolerance};
  const int hi[kNumI] = {0, 0, 0, 0, (int)max_iterations};
  SPAI_HIP_CHECK(hipMemcpyAsync(scal_.get(), hd, sizeof(hd), hipMemcpyHostToDevice, st));
  SPAI_HIP_CHECK(hipMemcpyAsync(ist_.get(), hi, sizeof(hi), hipMemcpyHostToDevice, st));
  const int sgm = SubgroupFor(M_.n, M_.nnz), sga = SubgroupFor(A_.n, A_.nnz);
  if (m_fp32_) {
    SPAI_SG_DISPATCH(sgm, (PcgInit<SG, float><<<grid_, kBlk, 0, st>>>(
                              n, M_.ptr.get(), M_.col.get(), M_.val32.get(), b_dev, x_.get(), r_.get(), z_.get(),
                              p_.get(), part_.get(), counters_.get(), scal_.get(), ist_.get(), hist_.get())));
  } else {
    SPAI_SG_DISPATCH(sgm, (PcgInit<SG, double><<<grid_, kBlk, 0, st>>>(
                              n, M_.ptr.get(), M_.col.get(), M_.val.get(), b_dev, x_.get(), r_.get(), z_.get(),
                              p_.get(), part_.get(), counters_.get(), scal_.get(), ist_.get(), hist_.get())));
  }
  if (!exec_) {
    SPAI_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    Record(true, kChunk);
    SPAI_HIP_CHECK(hipStreamEndCapture(st, &graph_));
    SPAI_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    graph_key_b_ = b_dev;
  }
  int hst[kNumI];
  int64_t launched = 0;
  while (true) {
    SPAI_HIP_CHECK(hipGraphLaunch(exec_, st));
    launched += kChunk;
    ctx_.ReadBack(hst, ist_.get(), sizeof(hst));
    if (hst[kDone] || launched >= max_iterations) break;
  }
  SPAI_SG_DISPATCH(sga, (PcgFinal<SG><<<grid_, kBlk, 0, st>>>(n, A_.ptr.get(), A_.col.get(), A_.val.get(), x_.get(),
                                                              b_dev, r_.get(), ap_.get(), part_.get(),
                                                              counters_.get(), scal_.get(), ist_.get(),
                                                              hist_.get())));
  SPAI_HIP_CHECK(hipGetLastError());
  ctx_.ReadBack(hst, ist_.get(), sizeof(hst));
  *count = hst[kCount];
  return hst[kStatus];
}

void PcgSolver::History(double* host, int64_t count) {
  ctx_.ReadBack(host, hist_.get(), sizeof(double) * count);
}

double PcgSolver::TimeSpmvA(int reps) {
  const int sga = SubgroupFor(A_.n, A_.nnz);
  hipStream_t st = ctx_.stream;
  const int n = (int)n_;
  std::vector<hipEvent_t> ev(2 * reps);
  for (auto& e : ev) SPAI_HIP_CHECK(hipEventCreate(&e));
  for (int i = 0; i < reps; ++i) {
    SPAI_HIP_CHECK(hipEventRecord(ev[2 * i], st));
    SPAI_SG_DISPATCH(sga, (PcgSpmvA<SG><<<grid_, kBlk, 0, st>>>(n, A_.ptr.get(), A_.col.get(), A_.val.get(),
                                                                p_.get(), ap_.get(), part_.get(), counters_.get(),
                                                                scal_.get(), ist_.get(), 1)));
    SPAI_HIP_CHECK(hipEventRecord(ev[2 * i + 1], st));
  }
  SPAI_HIP_CHECK(hipStreamSynchronize(st));
  double total = 0.0;
  for (int i = 0; i < reps; ++i) {
    float ms = 0.f;
    SPAI_HIP_CHECK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
    total += ms;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return total * 1e3 / reps;
}

void SpMV(Ctx& ctx, const DCsr& A, const double* x, double* y) {
  const int sga = SubgroupFor(A.n, A.nnz);
  const unsigned g = Grid(A.n, 64, ctx.num_cus);
  SPAI_SG_DISPATCH(sga, (SpmvKernel<SG><<<g, kBlk, 0, ctx.stream>>>((int)A.n, A.ptr.get(), A.col.get(), A.val.get(),
                                                                    x, y)));
}

double PcgSolver::TimeIterationUs(int reps) {
  (void)reps;
  return 0.0;
}

}  // namespace spai
