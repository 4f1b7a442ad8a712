// Preconditioned conjugate gradient of methods_cython/backward_error.pyx:69-144
// on the device.
//
// One iteration is three kernels:
//   K1  finish the previous iteration (beta, history, stop rules), then
//       p = z + beta p fused into Ap = A p, and <p, Ap>;
//   K2  alpha, x += alpha p, r -= alpha Ap, <r, r>;
//   K3  error = ||r|| / ||b||, z = M r, <r, z>.
// Each kernel writes one partial sum per block; every block of the next kernel
// finishes that reduction itself (the same fixed-order sum in every block, so
// all blocks take the same decision and results are deterministic), and block
// 0 records outcomes in a device-side state that only later kernels read;
// values that one kernel reads and the next rewrites are double-buffered by
// iteration parity.  When the recursive residual crosses the tolerance, K3
// pauses the device loop and the host runs the true-residual refresh of
// backward_error.pyx:116-122 (three kernels, once per solve) before resuming.
// Iterations are replayed in chunks from captured hipGraphs (one per starting
// parity); once the state says done, the remaining kernels return at once.

#include "pcg.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "dev_common.hpp"

namespace spai {

namespace {

constexpr int kBlk = 256;
constexpr int kChunk = 16;  // iterations per captured graph (even)

enum DSlot : int { kBnorm = 0, kTol, kRz0, kRz1, kErrPre, kNumD };
enum ISlot : int {
  kDone = 0,
  kStatus,
  kMax,
  kCount0,
  kCount1,
  kCountLatest,
  kPend0,
  kPend1,
  kNeedRefresh,
  kIter,
  kNumI
};
// partial-sum arrays, G doubles each
enum PSlot : int { kP1 = 0, kP2, kP4, kPR, kPA, kPB, kPF, kNumP };

// Row dot of a CSR matrix with x.  SG == 1 (short rows): one thread per row,
// entries taken 8 at a time with all loads and gathers of a group issued
// before the sums, which are then added in column order -- the sequential
// order of backward_error.pyx:16-24.  SG > 1: SG lanes per row, strided
// partial sums combined by shuffles.
template <int SG, class MV>
__device__ __forceinline__ double RowDot(const int* __restrict__ ptr, const int* __restrict__ col,
                                         const MV* __restrict__ val, const double* __restrict__ x, int row, int sl) {
  double s = 0.0;
  if constexpr (SG == 1) {
    const int b = ptr[row], e = ptr[row + 1];
    for (int k0 = b; k0 < e; k0 += 8) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u;
        t[u] = k < e ? double(val[k]) * x[col[k]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u < e) s += t[u];
    }
  } else {
    const int e = ptr[row + 1];
    for (int k = ptr[row] + sl; k < e; k += SG) s += double(val[k]) * x[col[k]];
#pragma unroll
    for (int off = SG / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, SG);
  }
  return s;
}

// Row of A times (z + beta p): the direction update fused into the product.
template <int SG>
__device__ __forceinline__ double RowDotDir(const int* __restrict__ ptr, const int* __restrict__ col,
                                            const double* __restrict__ val, const double* __restrict__ z,
                                            const double* __restrict__ p, double beta, int row, int sl) {
  double s = 0.0;
  if constexpr (SG == 1) {
    const int b = ptr[row], e = ptr[row + 1];
    for (int k0 = b; k0 < e; k0 += 8) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u;
        if (k < e) {
          const int c = col[k];
          t[u] = val[k] * (z[c] + beta * p[c]);
        } else {
          t[u] = 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u < e) s += t[u];
    }
  } else {
    const int e = ptr[row + 1];
    for (int k = ptr[row] + sl; k < e; k += SG) {
      const int c = col[k];
      s += val[k] * (z[c] + beta * p[c]);
    }
#pragma unroll
    for (int off = SG / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, SG);
  }
  return s;
}

// Block sum of v written to part[blockIdx.x].
__device__ __forceinline__ void StorePartial(double v, double* __restrict__ part) {
  __shared__ double red[16];
  v = BlockSum<double>(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Every block: fixed-order sum of the previous kernel's G partials.
__device__ __forceinline__ double SumPartials(const double* __restrict__ part, int G) {
  __shared__ double red[16];
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) s += part[b];
  return BlockSum<double>(s, red);
}

__device__ __forceinline__ bool Finite(double v) { return isfinite(v); }

// End of an iteration (backward_error.pyx:123-133,100): beta, history entry,
// counters, status.  Returns beta; sets stop when the loop must end.  Writes
// happen in block 0 only; slots written are never read by the same kernel.
__device__ __forceinline__ double EndIteration(int prev, double r_z_new, double err, bool refreshed,
                                               double* __restrict__ ds, int* __restrict__ is,
                                               double* __restrict__ hist, bool& stop, bool writer) {
  const double beta = refreshed ? 0.0 : r_z_new / ds[prev ? kRz1 : kRz0];
  const int count = is[prev ? kCount1 : kCount0];
  int status = 0;
  if (!Finite(err))
    status = 3;
  else if (!Finite(beta))
    status = 2;
  stop = status != 0 || !(count + 1 < is[kMax] && err > ds[kTol]);
  if (writer) {
    hist[count] = err;
    is[prev ? kCount0 : kCount1] = count + 1;
    is[kCountLatest] = count + 1;
    ds[prev ? kRz0 : kRz1] = r_z_new;
    if (status) is[kStatus] = status;
    if (stop) is[kDone] = 1;
  }
  return beta;
}

// r = b, x = 0, z = M b, p = 0; partials of <r, r> and <r, z>.
template <int SG, class MV>
__global__ __launch_bounds__(kBlk) void PcgInit(int n, const int* __restrict__ mp, const int* __restrict__ mc,
                                                const MV* __restrict__ mv, const double* __restrict__ b,
                                                double* __restrict__ x, double* __restrict__ r,
                                                double* __restrict__ z, double* __restrict__ p0,
                                                double* __restrict__ part) {
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double rr = 0.0, rz = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double zr = RowDot<SG, MV>(mp, mc, mv, b, row, sl);
    if (sl == 0) {
      const double br = b[row];
      r[row] = br;
      x[row] = 0.0;
      z[row] = zr;
      p0[row] = 0.0;
      rr += br * br;
      rz += br * zr;
    }
  }
  StorePartial(rr, part + kPA * gridDim.x);
  StorePartial(rz, part + kPB * gridDim.x);
}

// ||b||, r.z, history[0] = 1 (backward_error.pyx:91-100).
__global__ __launch_bounds__(kBlk) void PcgInitFinish(const double* __restrict__ part, int G, double* __restrict__ ds,
                                                      int* __restrict__ is, double* __restrict__ hist) {
  const double rr = SumPartials(part + kPA * G, G);
  const double rz = SumPartials(part + kPB * G, G);
  if (threadIdx.x != 0) return;
  ds[kBnorm] = sqrt(rr);
  ds[kRz0] = rz;
  hist[0] = 1.0;
  is[kCount0] = 1;
  is[kCountLatest] = 1;
  is[kStatus] = 0;
  is[kPend0] = 0;
  is[kPend1] = 0;
  is[kNeedRefresh] = 0;
  is[kIter] = 0;
  is[kDone] = (1 < is[kMax] && 1.0 > ds[kTol]) ? 0 : 1;
}

// K1 of iteration with parity par: end of the previous iteration, then
// p_out = z + beta p_in and Ap = A p_out; partials of <p_out, Ap>.
template <int SG>
__global__ __launch_bounds__(kBlk) void PcgDirSpmvA(int n, int par, const int* __restrict__ ap_,
                                                    const int* __restrict__ ac, const double* __restrict__ av,
                                                    const double* __restrict__ z, const double* __restrict__ p_in,
                                                    double* __restrict__ p_out, double* __restrict__ Ap,
                                                    double* __restrict__ part, double* __restrict__ ds,
                                                    int* __restrict__ is, double* __restrict__ hist) {
  if (is[kDone]) return;
  const int G = gridDim.x;
  const int prev = par ^ 1;
  double beta = 0.0;
  if (is[prev ? kPend1 : kPend0]) {
    const double r_z = SumPartials(part + kP4 * G, G);
    bool stop = false;
    beta = EndIteration(prev, r_z, ds[kErrPre], false, ds, is, hist, stop, blockIdx.x == 0 && threadIdx.x == 0);
    if (stop) return;
  }
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDotDir<SG>(ap_, ac, av, z, p_in, beta, row, sl);
    if (sl == 0) {
      const double pr = z[row] + beta * p_in[row];
      p_out[row] = pr;
      Ap[row] = s;
      acc += pr * s;
    }
  }
  StorePartial(acc, part + kP1 * G);
}

// K2: alpha (backward_error.pyx:102-110), x += alpha p, r -= alpha Ap; <r, r>.
__global__ __launch_bounds__(kBlk) void PcgUpdate(int n, int par, double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ p, const double* __restrict__ Ap,
                                                  double* __restrict__ part, double* __restrict__ ds,
                                                  int* __restrict__ is) {
  if (is[kDone]) return;
  const int G = gridDim.x;
  const double den = SumPartials(part + kP1 * G, G);
  const double r_z = ds[par ? kRz1 : kRz0];
  bool brk = !Finite(r_z) || !Finite(den) || r_z == 0.0 || den == 0.0;
  const double alpha = brk ? 0.0 : r_z / den;
  brk = brk || !Finite(alpha);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    is[par ? kPend0 : kPend1] = 0;  // consumed by K1 of this iteration
    if (brk) {
      is[kStatus] = 2;
      is[kDone] = 1;
    } else {
      is[kIter] = is[kIter] + 1;
    }
  }
  if (brk) return;
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    acc += ri * ri;
  }
  StorePartial(acc, part + kP2 * G);
}

// K3: error; at the tolerance crossing pause for the host-side refresh;
// otherwise z = M r and <r, z>.
template <int SG, class MV>
__global__ __launch_bounds__(kBlk) void PcgSpmvM(int n, int par, const int* __restrict__ mp,
                                                 const int* __restrict__ mc, const MV* __restrict__ mv,
                                                 const double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ part, double* __restrict__ ds,
                                                 int* __restrict__ is) {
  if (is[kDone]) return;
  const int G = gridDim.x;
  const double err = sqrt(SumPartials(part + kP2 * G, G)) / ds[kBnorm];
  const bool refresh = err <= ds[kTol];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ds[kErrPre] = err;
    if (refresh) {
      is[kNeedRefresh] = 1;
      is[kDone] = 1;
    } else {
      is[par ? kPend1 : kPend0] = 1;
    }
  }
  if (refresh) return;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, MV>(mp, mc, mv, r, row, sl);
    if (sl == 0) {
      z[row] = s;
      acc += r[row] * s;
    }
  }
  StorePartial(acc, part + kP4 * G);
}

// Host-launched refresh (backward_error.pyx:117-122): r = b - A x; <r, r>.
template <int SG>
__global__ __launch_bounds__(kBlk) void PcgRefresh(int n, const int* __restrict__ ap_, const int* __restrict__ ac,
                                                   const double* __restrict__ av, const double* __restrict__ x,
                                                   const double* __restrict__ b, double* __restrict__ r,
                                                   double* __restrict__ part) {
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, double>(ap_, ac, av, x, row, sl);
    if (sl == 0) {
      const double ri = b[row] - s;
      r[row] = ri;
      acc += ri * ri;
    }
  }
  StorePartial(acc, part + kPR * gridDim.x);
}

// After the refresh: error from the true residual, z = M r, <r, z>.
template <int SG, class MV>
__global__ __launch_bounds__(kBlk) void PcgRefreshSpmvM(int n, const int* __restrict__ mp, const int* __restrict__ mc,
                                                        const MV* __restrict__ mv, const double* __restrict__ r,
                                                        double* __restrict__ z, double* __restrict__ part,
                                                        double* __restrict__ ds) {
  const int G = gridDim.x;
  const double err = sqrt(SumPartials(part + kPR * G, G)) / ds[kBnorm];
  if (blockIdx.x == 0 && threadIdx.x == 0) ds[kErrPre] = err;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, MV>(mp, mc, mv, r, row, sl);
    if (sl == 0) {
      z[row] = s;
      acc += r[row] * s;
    }
  }
  StorePartial(acc, part + kP4 * G);
}

// End of the refreshed iteration (beta = 0); resumes the loop unless it stops.
__global__ __launch_bounds__(kBlk) void PcgRefreshFinish(int par, const double* __restrict__ part, int G,
                                                         double* __restrict__ ds, int* __restrict__ is,
                                                         double* __restrict__ hist) {
  const double r_z = SumPartials(part + kP4 * G, G);
  if (threadIdx.x != 0) return;
  is[kNeedRefresh] = 0;
  is[kDone] = 0;
  bool stop = false;
  EndIteration(par, r_z, ds[kErrPre], true, ds, is, hist, stop, true);
}

// Final true residual (backward_error.pyx:136-139): partials of ||b - A x||^2.
template <int SG>
__global__ __launch_bounds__(kBlk) void PcgFinal(int n, const int* __restrict__ ap_, const int* __restrict__ ac,
                                                 const double* __restrict__ av, const double* __restrict__ x,
                                                 const double* __restrict__ b, double* __restrict__ part) {
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, double>(ap_, ac, av, x, row, sl);
    if (sl == 0) {
      const double ri = b[row] - s;
      acc += ri * ri;
    }
  }
  StorePartial(acc, part + kPF * gridDim.x);
}

// Overwrites the last history entry with the true residual; status 1 / 3
// (backward_error.pyx:139-143).
__global__ __launch_bounds__(kBlk) void PcgFinalFinish(const double* __restrict__ part, int G,
                                                       const double* __restrict__ ds, int* __restrict__ is,
                                                       double* __restrict__ hist) {
  const double rr = SumPartials(part + kPF * G, G);
  if (threadIdx.x != 0) return;
  const int count = is[kCountLatest];
  const double h = sqrt(rr) / ds[kBnorm];
  hist[count - 1] = h;
  if (!Finite(h))
    is[kStatus] = 3;
  else if (h <= ds[kTol])
    is[kStatus] = 1;
}

// Plain Ap = A p with <p, Ap> partials (timing of the SpMV itself).
template <int SG>
__global__ __launch_bounds__(kBlk) void PcgSpmvA(int n, const int* __restrict__ ap_, const int* __restrict__ ac,
                                                 const double* __restrict__ av, const double* __restrict__ p,
                                                 double* __restrict__ Ap, double* __restrict__ part) {
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, double>(ap_, ac, av, p, row, sl);
    if (sl == 0) {
      Ap[row] = s;
      acc += p[row] * s;
    }
  }
  StorePartial(acc, part + kP1 * gridDim.x);
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
  if (mean <= 16.0) return 1;
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
  grid_ = Grid(n_, kBlk, ctx.num_cus);
  x_ = ctx.Alloc<double>(n_);
  r_ = ctx.Alloc<double>(n_);
  z_ = ctx.Alloc<double>(n_);
  p_ = ctx.Alloc<double>(2 * n_);
  ap_ = ctx.Alloc<double>(n_);
  part_ = ctx.Alloc<double>(size_t(kNumP) * grid_);
  scal_ = ctx.Alloc<double>(kNumD);
  ist_ = ctx.Alloc<int>(kNumI);
}

PcgSolver::~PcgSolver() { DropGraphs(); }

void PcgSolver::DropGraphs() {
  for (int i = 0; i < 2; ++i) {
    if (exec_[i]) (void)hipGraphExecDestroy(exec_[i]);
    if (graph_[i]) (void)hipGraphDestroy(graph_[i]);
    exec_[i] = nullptr;
    graph_[i] = nullptr;
  }
}

#define SPAI_SG_DISPATCH(sg, ...) \
  switch (sg) {                   \
    case 1: {                     \
      constexpr int SG = 1;       \
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

void PcgSolver::Record(int first_parity, int64_t iterations) {
  const int sga = SubgroupFor(A_.n, A_.nnz), sgm = SubgroupFor(M_.n, M_.nnz);
  hipStream_t st = ctx_.stream;
  const int n = (int)n_;
  for (int64_t it = 0; it < iterations; ++it) {
    const int par = int((first_parity + it) & 1);
    double* p_in = p_.get() + size_t(par) * n_;
    double* p_out = p_.get() + size_t(par ^ 1) * n_;
    SPAI_SG_DISPATCH(sga, (PcgDirSpmvA<SG><<<grid_, kBlk, 0, st>>>(n, par, A_.ptr.get(), A_.col.get(),
                                                                   A_.val.get(), z_.get(), p_in, p_out, ap_.get(),
                                                                   part_.get(), scal_.get(), ist_.get(),
                                                                   hist_.get())));
    PcgUpdate<<<grid_, kBlk, 0, st>>>(n, par, x_.get(), r_.get(), p_out, ap_.get(), part_.get(), scal_.get(),
                                      ist_.get());
    if (m_fp32_) {
      SPAI_SG_DISPATCH(sgm, (PcgSpmvM<SG, float><<<grid_, kBlk, 0, st>>>(n, par, M_.ptr.get(), M_.col.get(),
                                                                         M_.val32.get(), r_.get(), z_.get(),
                                                                         part_.get(), scal_.get(), ist_.get())));
    } else {
      SPAI_SG_DISPATCH(sgm, (PcgSpmvM<SG, double><<<grid_, kBlk, 0, st>>>(n, par, M_.ptr.get(), M_.col.get(),
                                                                          M_.val.get(), r_.get(), z_.get(),
                                                                          part_.get(), scal_.get(), ist_.get())));
    }
  }
}

int PcgSolver::Solve(const double* b_dev, int64_t max_iterations, double tolerance, int64_t* count) {
  if (max_iterations <= 0) throw std::invalid_argument("n, max_iterations and threads must be positive");
  if (max_iterations > (int64_t(1) << 30)) throw std::invalid_argument("max_iterations too large");
  hipStream_t st = ctx_.stream;
  const int n = (int)n_;
  if ((int64_t)hist_.size() < max_iterations || b_dev != graph_key_b_) {
    DropGraphs();
    if ((int64_t)hist_.size() < max_iterations) hist_ = ctx_.Alloc<double>(max_iterations);
  }
  b_ = b_dev;
  const double hd[kNumD] = {0.0, tolerance, 0.0, 0.0, 0.0};
  int hi[kNumI] = {0};
  hi[kMax] = (int)max_iterations;
  SPAI_HIP_CHECK(hipMemcpyAsync(scal_.get(), hd, sizeof(hd), hipMemcpyHostToDevice, st));
  SPAI_HIP_CHECK(hipMemcpyAsync(ist_.get(), hi, sizeof(hi), hipMemcpyHostToDevice, st));
  const int sgm = SubgroupFor(M_.n, M_.nnz), sga = SubgroupFor(A_.n, A_.nnz);
  if (m_fp32_) {
    SPAI_SG_DISPATCH(sgm, (PcgInit<SG, float><<<grid_, kBlk, 0, st>>>(n, M_.ptr.get(), M_.col.get(), M_.val32.get(),
                                                                      b_dev, x_.get(), r_.get(), z_.get(), p_.get(),
                                                                      part_.get())));
  } else {
    SPAI_SG_DISPATCH(sgm, (PcgInit<SG, double><<<grid_, kBlk, 0, st>>>(n, M_.ptr.get(), M_.col.get(), M_.val.get(),
                                                                       b_dev, x_.get(), r_.get(), z_.get(),
                                                                       p_.get(), part_.get())));
  }
  PcgInitFinish<<<1, kBlk, 0, st>>>(part_.get(), (int)grid_, scal_.get(), ist_.get(), hist_.get());
  if (!exec_[0]) {
    for (int par = 0; par < 2; ++par) {
      SPAI_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      Record(par, kChunk);
      SPAI_HIP_CHECK(hipStreamEndCapture(st, &graph_[par]));
      SPAI_HIP_CHECK(hipGraphInstantiate(&exec_[par], graph_[par], nullptr, nullptr, 0));
    }
    graph_key_b_ = b_dev;
  }
  int hst[kNumI];
  ctx_.ReadBack(hst, ist_.get(), sizeof(hst));
  int64_t chunks = 0;
  const int64_t max_chunks = max_iterations / kChunk + 4;
  while (!hst[kDone] && chunks < max_chunks) {
    SPAI_HIP_CHECK(hipGraphLaunch(exec_[hst[kIter] & 1], st));
    ++chunks;
    ctx_.ReadBack(hst, ist_.get(), sizeof(hst));
    if (hst[kNeedRefresh]) {
      const int par = (hst[kIter] - 1) & 1;
      SPAI_SG_DISPATCH(sga, (PcgRefresh<SG><<<grid_, kBlk, 0, st>>>(n, A_.ptr.get(), A_.col.get(), A_.val.get(),
                                                                    x_.get(), b_dev, r_.get(), part_.get())));
      if (m_fp32_) {
        SPAI_SG_DISPATCH(sgm, (PcgRefreshSpmvM<SG, float><<<grid_, kBlk, 0, st>>>(
                                  n, M_.ptr.get(), M_.col.get(), M_.val32.get(), r_.get(), z_.get(), part_.get(),
                                  scal_.get())));
      } else {
        SPAI_SG_DISPATCH(sgm, (PcgRefreshSpmvM<SG, double><<<grid_, kBlk, 0, st>>>(
                                  n, M_.ptr.get(), M_.col.get(), M_.val.get(), r_.get(), z_.get(), part_.get(),
                                  scal_.get())));
      }
      PcgRefreshFinish<<<1, kBlk, 0, st>>>(par, part_.get(), (int)grid_, scal_.get(), ist_.get(), hist_.get());
      ctx_.ReadBack(hst, ist_.get(), sizeof(hst));
    }
  }
  SPAI_SG_DISPATCH(sga, (PcgFinal<SG><<<grid_, kBlk, 0, st>>>(n, A_.ptr.get(), A_.col.get(), A_.val.get(), x_.get(),
                                                              b_dev, part_.get())));
  PcgFinalFinish<<<1, kBlk, 0, st>>>(part_.get(), (int)grid_, scal_.get(), ist_.get(), hist_.get());
  SPAI_HIP_CHECK(hipGetLastError());
  ctx_.ReadBack(hst, ist_.get(), sizeof(hst));
  *count = hst[kCountLatest];
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
                                                                x_.get(), ap_.get(), part_.get())));
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
  const unsigned g = Grid(A.n, kBlk, ctx.num_cus);
  SPAI_SG_DISPATCH(sga, (SpmvKernel<SG><<<g, kBlk, 0, ctx.stream>>>((int)A.n, A.ptr.get(), A.col.get(), A.val.get(),
                                                                    x, y)));
}


}  // namespace spai
