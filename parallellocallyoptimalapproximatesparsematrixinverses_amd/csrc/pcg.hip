// Preconditioned conjugate gradient of methods_cython/backward_error.pyx:69-144
// on the device.
//
// One iteration is two kernels:
//   K1  finish the previous iteration (error, refresh decision, beta, history,
//       stop rules), then p = z + beta p fused into Ap = A p, and <p, Ap>;
//   K2  alpha, x += alpha p, r' = r - alpha Ap, and z = M r' with the entries
//       of r' that a row of M needs formed on the fly from r and Ap (the same
//       expression, so the same bits), then <r', r'> and <r', z>.
// r is double-buffered by iteration parity so that K2 reads the old r while
// it writes the new one.  Each kernel writes one partial sum per block; every
// block of the next kernel finishes that reduction itself (the same
// fixed-order sum in every block, so all blocks take the same decision and
// results are deterministic), and block 0 records outcomes in a device-side
// state that only later kernels read.  When the recursive residual crosses
// the tolerance, K1 pauses the device loop and the host runs the
// true-residual refresh of backward_error.pyx:116-122 (three kernels, once per
// solve) before resuming.  Iterations are replayed in chunks from captured
// hipGraphs (one per starting parity); once the state says done, the
// remaining kernels return at once.  Rows are mapped to blocks so that each
// XCD (blocks b = x mod 8) works on one contiguous eighth of the rows, which
// keeps a stencil's plane neighbours in the same L2.

#include "pcg.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "dev_common.hpp"

namespace spai {

namespace {

constexpr int kBlk = 256;
constexpr int kChunk = 16;     // iterations per captured graph (even)
constexpr int kTailChunk = 2;  // iterations per graph after the refresh (even)

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
  kFinalDone,
  kRefreshes,
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
  // BlockSum's order (wavefront sums, then the wavefronts in order) with one
  // barrier: called once, at the end of a kernel
  __shared__ double red[16];
  v = WaveSum(v);
  if (LaneId() == 0) red[WaveInBlock()] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int(blockDim.x) + kWave - 1) / kWave;
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += red[i];
    part[blockIdx.x] = t;
  }
}

// Block sums of a and b written to pa[blockIdx.x] and pb[blockIdx.x]
// (BlockSum's order for each, one barrier; at the end of a kernel).
__device__ __forceinline__ void StorePartial2(double a, double b, double* __restrict__ pa, double* __restrict__ pb) {
  __shared__ double red[2][16];
  a = WaveSum(a);
  b = WaveSum(b);
  if (LaneId() == 0) {
    red[0][WaveInBlock()] = a;
    red[1][WaveInBlock()] = b;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int nw = (int(blockDim.x) + kWave - 1) / kWave;
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += red[threadIdx.x][i];
    (threadIdx.x == 0 ? pa : pb)[blockIdx.x] = t;
  }
}

// Every block: the fixed-order sums of two arrays of G partials at once
// (each equal to SumPartials of it).
__device__ __forceinline__ void SumPartials2(const double* __restrict__ pa, const double* __restrict__ pb, int G,
                                             double& a, double& b) {
  __shared__ double red[32];
  a = 0.0;
  b = 0.0;
  for (int i = threadIdx.x; i < G; i += blockDim.x) {
    a += pa[i];
    b += pb[i];
  }
  BlockSum2<double>(a, b, red);
}

// Every block: fixed-order sum of the previous kernel's G partials.
__device__ __forceinline__ double SumPartials(const double* __restrict__ part, int G) {
  __shared__ double red[16];
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) s += part[b];
  return BlockSum<double>(s, red);
}

__device__ __forceinline__ bool Finite(double v) { return isfinite(v); }

// Logical block of this block: consecutive logical blocks share an XCD
// (workgroups are dealt to the 8 XCDs round robin).
__device__ __forceinline__ int XcdBlock() {
  const int G = gridDim.x, b = blockIdx.x;
  return (G & 7) ? b : (b & 7) * (G >> 3) + (b >> 3);
}

// End of an iteration (backward_error.pyx:123-133,100): beta, history entry,
// counters, status.  Returns beta; sets stop when the loop must end.  Writes
// happen in block 0 only; slots written are never read by the same kernel.
// rz_prev, count, max_it and tol are the slots kRz/kCount of parity prev,
// kMax and kTol as the kernel read them.
__device__ __forceinline__ double EndIteration(int prev, double r_z_new, double err, bool refreshed, double rz_prev,
                                               int count, int max_it, double tol, double* __restrict__ ds,
                                               int* __restrict__ is, double* __restrict__ hist, bool& stop,
                                               bool writer) {
  const double beta = refreshed ? 0.0 : r_z_new / rz_prev;
  int status = 0;
  if (!Finite(err))
    status = 3;
  else if (!Finite(beta))
    status = 2;
  stop = status != 0 || !(count + 1 < max_it && err > tol);
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

// A thread's share of G <= kMaxPartials * kBlk block partials (entries
// threadIdx.x + u kBlk), loaded without waiting on anything: the loads take
// no per-lane branch (lanes past G read entry G - 1, which AddPartials
// skips), so none of them is waited on before the others are issued.
constexpr int kMaxPartials = 2;
__device__ __forceinline__ void LoadPartials(const double* __restrict__ part, int G, double (&v)[kMaxPartials]) {
#pragma unroll
  for (int u = 0; u < kMaxPartials; ++u) v[u] = part[min(int(threadIdx.x) + u * kBlk, G - 1)];
}

// This thread's entries of a partial array in SumPartials' order.
__device__ __forceinline__ double AddPartials(const double (&v)[kMaxPartials], int G) {
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < kMaxPartials; ++u)
    if (int(threadIdx.x) + u * kBlk < G) s += v[u];
  return s;
}

// What K1's prologue reads, loaded at kernel start, none waiting on another
// (the state slots of the previous iteration and the previous K2's partials
// of <r', r'> and <r', z>), so the prologue costs one memory round trip.
struct DirLoads {
  int done, pend, count, max_it;
  double bnorm, tol, rz_prev;
  double rr[kMaxPartials], rz[kMaxPartials];
  __device__ __forceinline__ void Load(int par, const double* __restrict__ part, int G, const double* __restrict__ ds,
                                       const int* __restrict__ is) {
    const int prev = par ^ 1;
    done = is[kDone];
    pend = is[prev ? kPend1 : kPend0];
    count = is[prev ? kCount1 : kCount0];
    max_it = is[kMax];
    bnorm = ds[kBnorm];
    tol = ds[kTol];
    rz_prev = ds[prev ? kRz1 : kRz0];
    LoadPartials(part + kP2 * G, G, rr);
    LoadPartials(part + kP4 * G, G, rz);
  }
};

// Prologue of K1 (iteration with parity par): when the previous iteration is
// pending, its error from the previous K2's <r', r'> partials, the tolerance
// crossing (pause for the host refresh) or its end (beta, history, stop
// rules).  Returns false when the kernel must return without work.
__device__ __forceinline__ bool DirPrologue(int par, const DirLoads& L, int G, double* __restrict__ ds,
                                            int* __restrict__ is, double* __restrict__ hist, double& beta) {
  const int prev = par ^ 1;
  const bool writer = blockIdx.x == 0 && threadIdx.x == 0;
  beta = 0.0;
  if (L.done) return false;
  if (!L.pend) return true;
  __shared__ double red[32];
  double rr = AddPartials(L.rr, G), r_z = AddPartials(L.rz, G);
  BlockSum2<double>(rr, r_z, red);
  const double err = sqrt(rr) / L.bnorm;
  if (err <= L.tol) {
    if (writer) {
      ds[kErrPre] = err;
      is[kNeedRefresh] = 1;
      is[kDone] = 1;
    }
    return false;
  }
  bool stop = false;
  beta = EndIteration(prev, r_z, err, false, L.rz_prev, L.count, L.max_it, L.tol, ds, is, hist, stop, writer);
  return !stop;
}

// What K2's prologue reads, loaded at kernel start (K1's <p, Ap> partials and
// the state slots).
struct UpdateLoads {
  int done, iter;
  double rz;
  double pap[kMaxPartials];
  __device__ __forceinline__ void Load(int par, const double* __restrict__ part, int G, const double* __restrict__ ds,
                                       const int* __restrict__ is) {
    done = is[kDone];
    iter = is[kIter];
    rz = ds[par ? kRz1 : kRz0];
    LoadPartials(part + kP1 * G, G, pap);
  }
};

// Prologue of K2: alpha = <r, z> / <p, Ap> from K1's partials
// (backward_error.pyx:102-110); breakdown ends the solve with status 2.
// Returns false when the kernel must return without work.
__device__ __forceinline__ bool UpdatePrologue(int par, const UpdateLoads& L, int G, int* __restrict__ is,
                                               double& alpha) {
  alpha = 0.0;
  if (L.done) return false;
  __shared__ double red[16];
  const double den = BlockSum<double>(AddPartials(L.pap, G), red);
  const double r_z = L.rz;
  bool brk = !Finite(r_z) || !Finite(den) || r_z == 0.0 || den == 0.0;
  alpha = brk ? 0.0 : r_z / den;
  brk = brk || !Finite(alpha);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    is[par ? kPend0 : kPend1] = 0;  // consumed by K1 of this iteration
    if (brk) {
      is[kStatus] = 2;
      is[kDone] = 1;
    } else {
      is[kIter] = L.iter + 1;
      is[par ? kPend1 : kPend0] = 1;
    }
  }
  return !brk;
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

// Entries of one row, as seen by one thread: entry u sits at base + u * stride
// for u < count; a column index < 0 marks padding.
struct RowView {
  int base, stride, count;
};

// Entries past the end of a row: kPadded rows can be read there (the arrays
// hold kSellPad entries of padding, so reads of 8 consecutive entries never
// leave them); others are only read inside the row.
constexpr int kSellPad = 8 * 64;

// CSR rows.
struct CsrRows {
  static constexpr bool kPadded = false;
  const int* __restrict__ ptr;
  __device__ __forceinline__ RowView Row(int r) const { return RowView{ptr[r], 1, ptr[r + 1] - ptr[r]}; }
};

// Sliced ELLPACK rows (DSell): 64-row slices, column-major inside a slice.
struct SellRows {
  static constexpr bool kPadded = true;
  const int* __restrict__ off;
  __device__ __forceinline__ RowView Row(int r) const {
    const int s = r >> 6, b = off[s];
    return RowView{b + (r & 63), 64, (off[s + 1] - b) >> 6};
  }
};

// sum_u val[k_u] * f(col[k_u]) over a row's entries in order, 8 loads in
// flight at a time; u < 8 come from `first` when given (prefetched).
// kEager: the 8 column indices and values of a group are loaded first, then
// the gathers at max(c, 0) (in bounds for padding too), all without branches
// on loaded data, so the loads of a group are in flight together; this holds
// more registers.  Otherwise each entry's loads wait for its column index
// (fewer registers, more waves).
template <class MV, bool kPadded, bool kEager, class F>
__device__ __forceinline__ double RowSum(const RowView rv, const int* __restrict__ col, const MV* __restrict__ val,
                                         int u_begin, F f, double s = 0.0) {
  for (int u0 = u_begin; u0 < rv.count; u0 += 8) {
    if constexpr (!kEager) {
      double t[8];
      bool in[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = rv.base + (u0 + u) * rv.stride;
        const int c = u0 + u < rv.count ? col[k] : -1;
        in[u] = c >= 0;
        t[u] = in[u] ? double(val[k]) * f(c) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (in[u]) s += t[u];
      continue;
    }
    int c[8];
    MV v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = rv.base + (u0 + u) * rv.stride;
      const bool have = u0 + u < rv.count;
      if constexpr (kPadded) {
        c[u] = col[k];
        v[u] = val[k];
        c[u] = have ? c[u] : -1;
      } else {
        c[u] = have ? col[k] : -1;
        v[u] = have ? val[k] : MV(0);
      }
    }
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = double(v[u]) * f(max(c[u], 0));
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (c[u] >= 0) s += t[u];
  }
  return s;
}

// Operands of a thread's first row (one thread per row): the row's first 8
// values and the two gathered vectors at their columns, loaded before the
// kernel's prologue reduction so that their latency overlaps it.
template <class MV>
struct RowPrefetch {
  RowView rv;
  MV v[8];
  double x0[8], x1[8];
  bool in[8];
  template <bool kEager, class Rows>
  __device__ __forceinline__ void Load(const Rows& rows, const int* __restrict__ col, const MV* __restrict__ val,
                                       const double* __restrict__ y0, const double* __restrict__ y1, int row) {
    rv = row >= 0 ? rows.Row(row) : RowView{0, 1, 0};
    if constexpr (!kEager) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = rv.base + u * rv.stride;
        const int c = u < rv.count ? col[k] : -1;
        in[u] = c >= 0;
        v[u] = in[u] ? val[k] : MV(0);
        x0[u] = in[u] ? y0[c] : 0.0;
        x1[u] = in[u] ? y1[c] : 0.0;
      }
      return;
    }
    int c[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = rv.base + u * rv.stride;
      const bool have = u < rv.count;
      if constexpr (Rows::kPadded) {
        c[u] = col[k];
        v[u] = val[k];
        c[u] = have ? c[u] : -1;
      } else {
        c[u] = have ? col[k] : -1;
        v[u] = have ? val[k] : MV(0);
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      in[u] = c[u] >= 0;
      const int cc = max(c[u], 0);
      x0[u] = y0[cc];
      x1[u] = y1[cc];
    }
  }
};

// K1 of iteration with parity par: end of the previous iteration (error from
// the previous K2's <r', r'>; at the tolerance crossing, pause for the host
// refresh), then p_out = z + beta p_in and Ap = A p_out; partials of <p_out, Ap>.
// SG == 1: one thread per row through `rows` (CSR or SELL); SG > 1: SG lanes
// per CSR row (rows.ptr).
template <int SG, class Rows>
__global__ __launch_bounds__(kBlk) void PcgDirSpmvA(int n, int par, Rows rows, const int* __restrict__ ac,
                                                    const double* __restrict__ av, const double* __restrict__ z,
                                                    const double* __restrict__ p_in, double* __restrict__ p_out,
                                                    double* __restrict__ Ap, double* __restrict__ part,
                                                    double* __restrict__ ds, int* __restrict__ is,
                                                    double* __restrict__ hist) {
  const int G = gridDim.x;
  DirLoads L;
  L.Load(par, part, G, ds, is);
  const int row0 = XcdBlock() * (kBlk / SG) + threadIdx.x / SG;
  RowPrefetch<double> pf;
  double z0 = 0.0, p0 = 0.0;
  if constexpr (SG == 1) {
    const bool have = row0 < n;
    pf.template Load<false>(rows, ac, av, z, p_in, have ? row0 : -1);
    z0 = have ? z[row0] : 0.0;
    p0 = have ? p_in[row0] : 0.0;
  }
  double beta = 0.0;
  if (!DirPrologue(par, L, G, ds, is, hist, beta)) return;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = row0; row < n; row += G * groups) {
    double s, pr;
    if constexpr (SG == 1) {
      auto dir = [&](int c) { return z[c] + beta * p_in[c]; };
      if (row == row0) {
        // the first 8 entries from the prefetched operands, then the rest
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = pf.in[u] ? pf.v[u] * (pf.x0[u] + beta * pf.x1[u]) : 0.0;
        s = 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (pf.in[u]) s += t[u];
        // entries past the first 8 continue the same sequential sum
        if (pf.rv.count > 8) s = RowSum<double, Rows::kPadded, false>(pf.rv, ac, av, 8, dir, s);
        pr = z0 + beta * p0;
      } else {
        s = RowSum<double, Rows::kPadded, false>(rows.Row(row), ac, av, 0, dir);
        pr = z[row] + beta * p_in[row];
      }
    } else {
      s = RowDotDir<SG>(rows.ptr, ac, av, z, p_in, beta, row, sl);
      pr = z[row] + beta * p_in[row];
    }
    if (sl == 0) {
      p_out[row] = pr;
      Ap[row] = s;
      acc += pr * s;
    }
  }
  StorePartial(acc, part + kP1 * G);
}

// Row of M times r' = r - alpha Ap, r' formed per entry (SG > 1, CSR).
template <int SG, class MV>
__device__ __forceinline__ double RowDotUpdated(const int* __restrict__ ptr, const int* __restrict__ col,
                                                const MV* __restrict__ val, const double* __restrict__ r,
                                                const double* __restrict__ Ap, double alpha, int row, int sl) {
  double s = 0.0;
  const int e = ptr[row + 1];
  for (int k = ptr[row] + sl; k < e; k += SG) {
    const int c = col[k];
    s += double(val[k]) * (r[c] - alpha * Ap[c]);
  }
#pragma unroll
  for (int off = SG / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, SG);
  return s;
}

// K2: alpha (backward_error.pyx:102-110), x += alpha p, r' = r - alpha Ap,
// z = M r'; partials of <r', r'> and <r', z> (backward_error.pyx:111-124).
template <int SG, class MV, class Rows>
__global__ __launch_bounds__(kBlk) void PcgUpdateSpmvM(int n, int par, Rows rows, const int* __restrict__ mc,
                                                       const MV* __restrict__ mv, double* __restrict__ x,
                                                       const double* __restrict__ r_in, double* __restrict__ r_out,
                                                       const double* __restrict__ p, const double* __restrict__ Ap,
                                                       double* __restrict__ z, double* __restrict__ part,
                                                       double* __restrict__ ds, int* __restrict__ is) {
  // FP32 values leave room for the registers of eager loads; FP64 ones do not
  constexpr bool kEagerM = sizeof(MV) == 4 && Rows::kPadded;
  const int G = gridDim.x;
  UpdateLoads L;
  L.Load(par, part, G, ds, is);
  const int row0 = XcdBlock() * (kBlk / SG) + threadIdx.x / SG;
  RowPrefetch<MV> pf;
  double x0 = 0.0, p0 = 0.0, r0 = 0.0, a0 = 0.0;
  if constexpr (SG == 1) {
    const bool have = row0 < n;
    pf.template Load<kEagerM>(rows, mc, mv, r_in, Ap, have ? row0 : -1);
    x0 = have ? x[row0] : 0.0;
    p0 = have ? p[row0] : 0.0;
    r0 = have ? r_in[row0] : 0.0;
    a0 = have ? Ap[row0] : 0.0;
  }
  double alpha;
  if (!UpdatePrologue(par, L, G, is, alpha)) return;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double rr = 0.0, rz = 0.0;
  for (int row = row0; row < n; row += G * groups) {
    double zs, xv, ri;
    if constexpr (SG == 1) {
      auto upd = [&](int c) { return r_in[c] - alpha * Ap[c]; };
      if (row == row0) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = pf.in[u] ? double(pf.v[u]) * (pf.x0[u] - alpha * pf.x1[u]) : 0.0;
        zs = 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (pf.in[u]) zs += t[u];
        if (pf.rv.count > 8) zs = RowSum<MV, Rows::kPadded, kEagerM>(pf.rv, mc, mv, 8, upd, zs);
        xv = x0 + alpha * p0;
        ri = r0 - alpha * a0;
      } else {
        zs = RowSum<MV, Rows::kPadded, kEagerM>(rows.Row(row), mc, mv, 0, upd);
        xv = x[row] + alpha * p[row];
        ri = r_in[row] - alpha * Ap[row];
      }
    } else {
      zs = RowDotUpdated<SG, MV>(rows.ptr, mc, mv, r_in, Ap, alpha, row, sl);
      xv = x[row] + alpha * p[row];
      ri = r_in[row] - alpha * Ap[row];
    }
    if (sl == 0) {
      x[row] = xv;
      r_out[row] = ri;
      z[row] = zs;
      rr += ri * ri;
      rz += ri * zs;
    }
  }
  StorePartial2(rr, rz, part + kP2 * G, part + kP4 * G);
}

// Diagonal layout as the kernels take it: the offsets by value (kernel
// arguments, loaded with the others at wave start), the row masks and the
// diagonal count.
struct DiaView {
  int off[kDiaMaxDiagonals];
  const unsigned* __restrict__ mask;
  int nd;
};

// Operands of diagonals [D0, D0 + 8) of one row: the values and NV vectors at
// row + off[d], the column clamped into [0, n) (an entry the row does not
// have is loaded but never enters a sum).  No address depends on a load, so
// all of them are in flight together.
template <class MV, int NV>
struct DiaChunk {
  MV v[8];
  double x[NV][8];
  template <int D0>
  __device__ __forceinline__ void Load(const DiaView& dv, const MV* __restrict__ val, const double* __restrict__ y0,
                                       const double* __restrict__ y1, int row, int n) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (D0 + u < dv.nd) {
        const int c = min(max(row + dv.off[D0 + u], 0), n - 1);
        v[u] = val[size_t(D0 + u) * size_t(n) + row];
        x[0][u] = y0[c];
        if constexpr (NV > 1) x[1][u] = y1[c];
      } else {
        v[u] = MV(0);
#pragma unroll
        for (int w = 0; w < NV; ++w) x[w][u] = 0.0;
      }
    }
  }
  // s + sum of v[u] * f(x[0][u], x[NV-1][u]) over the diagonals the row has,
  // in diagonal order (= column order, the sequential order of the CSR row)
  template <int D0, class F>
  __device__ __forceinline__ double Sum(unsigned m, int nd, F f, double s) const {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = double(v[u]) * f(x[0][u], x[NV - 1][u]);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (D0 + u < nd && ((m >> (D0 + u)) & 1u)) s += t[u];
    return s;
  }
};

// Row sum over diagonals 8.. of a row (the first chunk is summed by the
// caller): chunk by chunk, each chunk's loads issued together.
template <class MV, int NV, class F>
__device__ __forceinline__ double DiaRest(const DiaView& dv, const MV* __restrict__ val,
                                          const double* __restrict__ y0, const double* __restrict__ y1, int row,
                                          int n, unsigned m, F f, double s) {
  static_assert(kDiaMaxDiagonals == 32, "chunks below cover 32 diagonals");
  if (dv.nd > 8) {
    DiaChunk<MV, NV> c;
    c.template Load<8>(dv, val, y0, y1, row, n);
    s = c.template Sum<8>(m, dv.nd, f, s);
  }
  if (dv.nd > 16) {
    DiaChunk<MV, NV> c;
    c.template Load<16>(dv, val, y0, y1, row, n);
    s = c.template Sum<16>(m, dv.nd, f, s);
  }
  if (dv.nd > 24) {
    DiaChunk<MV, NV> c;
    c.template Load<24>(dv, val, y0, y1, row, n);
    s = c.template Sum<24>(m, dv.nd, f, s);
  }
  return s;
}

// K1 on the diagonal layout of A (one thread per row): the same work as
// PcgDirSpmvA, with the first row's operands loaded before the prologue.
// kPipe (more than one row a thread): each row's operands are loaded while
// the row before is summed and stored.
template <int kChunks, bool kPipe>
__global__ __launch_bounds__(kBlk) void PcgDirSpmvADia(int n, int par, DiaView dv, const double* __restrict__ av,
                                                       const double* __restrict__ z, const double* __restrict__ p_in,
                                                       double* __restrict__ p_out, double* __restrict__ Ap,
                                                       double* __restrict__ part, double* __restrict__ ds,
                                                       int* __restrict__ is, double* __restrict__ hist) {
  const int G = gridDim.x;
  DirLoads L;
  L.Load(par, part, G, ds, is);
  const int row0 = XcdBlock() * kBlk + threadIdx.x;
  DiaChunk<double, 2> c0;
  unsigned m0 = 0;
  double z0 = 0.0, p0 = 0.0;
  if (row0 < n) {
    c0.Load<0>(dv, av, z, p_in, row0, n);
    m0 = dv.mask[row0];
    z0 = z[row0];
    p0 = p_in[row0];
  }
  double beta;
  if (!DirPrologue(par, L, G, ds, is, hist, beta)) return;
  auto dir = [&](double zc, double pc) { return zc + beta * pc; };
  double acc = 0.0;
  for (int row = row0; row < n; row += G * kBlk) {
    if constexpr (kPipe) {
      // this row's operands are in c0; the next row's go into c1 now
      const int nxt = row + G * kBlk;
      DiaChunk<double, 2> c1;
      unsigned m1 = 0;
      double z1 = 0.0, p1 = 0.0;
      if (nxt < n) {
        c1.Load<0>(dv, av, z, p_in, nxt, n);
        m1 = dv.mask[nxt];
        z1 = z[nxt];
        p1 = p_in[nxt];
      }
      double s = c0.Sum<0>(m0, dv.nd, dir, 0.0);
      if constexpr (kChunks > 1) s = DiaRest<double, 2>(dv, av, z, p_in, row, n, m0, dir, s);
      const double pr = z0 + beta * p0;
      p_out[row] = pr;
      Ap[row] = s;
      acc += pr * s;
      c0 = c1;
      m0 = m1;
      z0 = z1;
      p0 = p1;
      continue;
    }
    if (row != row0) {
      c0.Load<0>(dv, av, z, p_in, row, n);
      m0 = dv.mask[row];
      z0 = z[row];
      p0 = p_in[row];
    }
    double s = c0.Sum<0>(m0, dv.nd, dir, 0.0);
    if constexpr (kChunks > 1) s = DiaRest<double, 2>(dv, av, z, p_in, row, n, m0, dir, s);
    const double pr = z0 + beta * p0;
    p_out[row] = pr;
    Ap[row] = s;
    acc += pr * s;
  }
  StorePartial(acc, part + kP1 * G);
}

// K2 on the diagonal layout of M (one thread per row): the same work as
// PcgUpdateSpmvM, with the first row's operands loaded before the prologue.
// kPipe as in PcgDirSpmvADia.
template <class MV, int kChunks, bool kPipe>
__global__ __launch_bounds__(kBlk) void PcgUpdateSpmvMDia(int n, int par, DiaView dv, const MV* __restrict__ mv,
                                                          double* __restrict__ x, const double* __restrict__ r_in,
                                                          double* __restrict__ r_out, const double* __restrict__ p,
                                                          const double* __restrict__ Ap, double* __restrict__ z,
                                                          double* __restrict__ part, double* __restrict__ ds,
                                                          int* __restrict__ is) {
  const int G = gridDim.x;
  UpdateLoads L;
  L.Load(par, part, G, ds, is);
  const int row0 = XcdBlock() * kBlk + threadIdx.x;
  DiaChunk<MV, 2> c0;
  unsigned m0 = 0;
  double x0 = 0.0, p0 = 0.0, r0 = 0.0, a0 = 0.0;
  if (row0 < n) {
    c0.template Load<0>(dv, mv, r_in, Ap, row0, n);
    m0 = dv.mask[row0];
    x0 = x[row0];
    p0 = p[row0];
    r0 = r_in[row0];
    a0 = Ap[row0];
  }
  double alpha;
  if (!UpdatePrologue(par, L, G, is, alpha)) return;
  auto upd = [&](double rc, double ac) { return rc - alpha * ac; };
  double rr = 0.0, rz = 0.0;
  for (int row = row0; row < n; row += G * kBlk) {
    if constexpr (kPipe) {
      // this row's operands are in c0; the next row's go into c1 now
      const int nxt = row + G * kBlk;
      DiaChunk<MV, 2> c1;
      unsigned m1 = 0;
      double x1 = 0.0, p1 = 0.0, r1 = 0.0, a1 = 0.0;
      if (nxt < n) {
        c1.template Load<0>(dv, mv, r_in, Ap, nxt, n);
        m1 = dv.mask[nxt];
        x1 = x[nxt];
        p1 = p[nxt];
        r1 = r_in[nxt];
        a1 = Ap[nxt];
      }
      double zs = c0.template Sum<0>(m0, dv.nd, upd, 0.0);
      if constexpr (kChunks > 1) zs = DiaRest<MV, 2>(dv, mv, r_in, Ap, row, n, m0, upd, zs);
      const double ri = r0 - alpha * a0;
      x[row] = x0 + alpha * p0;
      r_out[row] = ri;
      z[row] = zs;
      rr += ri * ri;
      rz += ri * zs;
      c0 = c1;
      m0 = m1;
      x0 = x1;
      p0 = p1;
      r0 = r1;
      a0 = a1;
      continue;
    }
    if (row != row0) {
      c0.template Load<0>(dv, mv, r_in, Ap, row, n);
      m0 = dv.mask[row];
      x0 = x[row];
      p0 = p[row];
      r0 = r_in[row];
      a0 = Ap[row];
    }
    double zs = c0.template Sum<0>(m0, dv.nd, upd, 0.0);
    if constexpr (kChunks > 1) zs = DiaRest<MV, 2>(dv, mv, r_in, Ap, row, n, m0, upd, zs);
    const double ri = r0 - alpha * a0;
    x[row] = x0 + alpha * p0;
    r_out[row] = ri;
    z[row] = zs;
    rr += ri * ri;
    rz += ri * zs;
  }
  StorePartial2(rr, rz, part + kP2 * G, part + kP4 * G);
}

// Plain Ap = A p with <p, Ap> partials on the diagonal layout (the solve's
// A-SpMV replayed on its own for timing).
template <int kChunks>
__global__ __launch_bounds__(kBlk) void PcgSpmvADia(int n, DiaView dv, const double* __restrict__ av,
                                                    const double* __restrict__ p, double* __restrict__ Ap,
                                                    double* __restrict__ part) {
  auto id = [](double v, double) { return v; };
  double acc = 0.0;
  for (int row = XcdBlock() * kBlk + threadIdx.x; row < n; row += gridDim.x * kBlk) {
    DiaChunk<double, 1> c;
    c.Load<0>(dv, av, p, p, row, n);
    const unsigned m = dv.mask[row];
    double s = c.Sum<0>(m, dv.nd, id, 0.0);
    if constexpr (kChunks > 1) s = DiaRest<double, 1>(dv, av, p, p, row, n, m, id, s);
    Ap[row] = s;
    acc += p[row] * s;
  }
  StorePartial(acc, part + kP1 * gridDim.x);
}

// Diagonal layout construction: flag[col - row + n - 1] = 1 for every entry.
__global__ void DiaMark(int n, const int* __restrict__ ptr, const int* __restrict__ col, int* __restrict__ flag) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  for (int k = ptr[row]; k < ptr[row + 1]; ++k) flag[int64_t(col[k]) - row + n - 1] = 1;
}

// off[pos[i]] = i - (n - 1) for every flagged i (pos: exclusive scan of flag).
__global__ void DiaOffsets(int64_t len, int n, const int* __restrict__ flag, const int* __restrict__ pos,
                           int* __restrict__ off) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len && flag[i]) off[pos[i]] = int(i - (n - 1));
}

// Per row: the CSR position of each diagonal slot (-1 for none) and the mask.
__global__ void DiaFill(int n, int nd, const int* __restrict__ ptr, const int* __restrict__ col,
                        const int* __restrict__ pos, int* __restrict__ src, unsigned* __restrict__ mask) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  for (int d = 0; d < nd; ++d) src[size_t(d) * n + row] = -1;
  unsigned m = 0;
  for (int k = ptr[row]; k < ptr[row + 1]; ++k) {
    const int d = pos[int64_t(col[k]) - row + n - 1];
    src[size_t(d) * n + row] = k;
    m |= 1u << d;
  }
  mask[row] = m;
}

// Row dot products of the once-per-solve kernels (start, refresh, final
// residual): CSR with SG lanes per row, or the diagonal layout with one
// thread per row.  Both add a row's entries in column order.
template <int SG, class MV>
struct CsrDot {
  static constexpr int kSG = SG;
  const int* __restrict__ ptr;
  const int* __restrict__ col;
  const MV* __restrict__ val;
  __device__ __forceinline__ double operator()(const double* __restrict__ x, int row, int sl, int) const {
    return RowDot<SG, MV>(ptr, col, val, x, row, sl);
  }
};

template <class MV, int kChunks>
struct DiaDot {
  static constexpr int kSG = 1;
  DiaView dv;
  const MV* __restrict__ val;
  __device__ __forceinline__ double operator()(const double* __restrict__ x, int row, int, int n) const {
    auto id = [](double v, double) { return v; };
    DiaChunk<MV, 1> c;
    c.template Load<0>(dv, val, x, x, row, n);
    const unsigned m = dv.mask[row];
    double s = c.template Sum<0>(m, dv.nd, id, 0.0);
    if constexpr (kChunks > 1) s = DiaRest<MV, 1>(dv, val, x, x, row, n, m, id, s);
    return s;
  }
};

// r = b, x = 0, z = M b, p = 0; partials of <r, r> and <r, z>.
template <class Dot>
__global__ __launch_bounds__(kBlk) void PcgInit(int n, Dot m, const double* __restrict__ b, double* __restrict__ x,
                                                double* __restrict__ r, double* __restrict__ z,
                                                double* __restrict__ p0, double* __restrict__ part) {
  constexpr int SG = Dot::kSG;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double rr = 0.0, rz = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double zr = m(b, row, sl, n);
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
  StorePartial2(rr, rz, part + kPA * gridDim.x, part + kPB * gridDim.x);
}

// The refresh of backward_error.pyx:116-122, enqueued after every batch of
// iterations and run only when K1 found the tolerance crossing (kNeedRefresh):
// r = b - A x, <r, r>.  The last K2 had parity (kIter - 1) & 1 and left r' in
// buffer ((kIter - 1) & 1) ^ 1, which is refreshed in place.
template <class Dot>
__global__ __launch_bounds__(kBlk) void PcgRefresh(int n, Dot a, const double* __restrict__ x,
                                                   const double* __restrict__ b, double* __restrict__ r2,
                                                   double* __restrict__ part, const int* __restrict__ is) {
  if (!is[kNeedRefresh]) return;
  constexpr int SG = Dot::kSG;
  double* __restrict__ r = r2 + size_t(((is[kIter] - 1) & 1) ^ 1) * size_t(n);
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = a(x, row, sl, n);
    if (sl == 0) {
      const double ri = b[row] - s;
      r[row] = ri;
      acc += ri * ri;
    }
  }
  StorePartial(acc, part + kPR * gridDim.x);
}

// After the refresh: error from the true residual, z = M r, <r, z>.
template <class Dot>
__global__ __launch_bounds__(kBlk) void PcgRefreshSpmvM(int n, Dot m, const double* __restrict__ r2,
                                                        double* __restrict__ z, double* __restrict__ part,
                                                        double* __restrict__ ds, const int* __restrict__ is) {
  if (!is[kNeedRefresh]) return;
  constexpr int SG = Dot::kSG;
  const double* __restrict__ r = r2 + size_t(((is[kIter] - 1) & 1) ^ 1) * size_t(n);
  const int G = gridDim.x;
  const double err = sqrt(SumPartials(part + kPR * G, G)) / ds[kBnorm];
  if (blockIdx.x == 0 && threadIdx.x == 0) ds[kErrPre] = err;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = m(r, row, sl, n);
    if (sl == 0) {
      z[row] = s;
      acc += r[row] * s;
    }
  }
  StorePartial(acc, part + kP4 * G);
}

// End of the refreshed iteration (beta = 0); resumes the loop unless it stops.
__global__ __launch_bounds__(kBlk) void PcgRefreshFinish(const double* __restrict__ part, int G,
                                                         double* __restrict__ ds, int* __restrict__ is,
                                                         double* __restrict__ hist) {
  if (!is[kNeedRefresh]) return;
  const int par = (is[kIter] - 1) & 1;
  const double r_z = SumPartials(part + kP4 * G, G);
  if (threadIdx.x != 0) return;
  is[kNeedRefresh] = 0;
  is[kDone] = 0;
  is[kRefreshes] = is[kRefreshes] + 1;
  is[par ? kPend1 : kPend0] = 0;  // finished here; K1 of the next iteration starts with beta = 0
  bool stop = false;
  EndIteration(par, r_z, ds[kErrPre], true, 0.0, is[par ? kCount1 : kCount0], is[kMax], ds[kTol], ds, is, hist,
               stop, true);
}

// Final true residual (backward_error.pyx:136-139): partials of ||b - A x||^2,
// enqueued after every batch and run once the solve has ended (done, no
// refresh pending, not yet finished).
__device__ __forceinline__ bool FinalDue(const int* __restrict__ is) {
  return is[kDone] && !is[kNeedRefresh] && !is[kFinalDone];
}

template <class Dot>
__global__ __launch_bounds__(kBlk) void PcgFinal(int n, Dot a, const double* __restrict__ x,
                                                 const double* __restrict__ b, double* __restrict__ part,
                                                 const int* __restrict__ is) {
  if (!FinalDue(is)) return;
  constexpr int SG = Dot::kSG;
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = blockIdx.x * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = a(x, row, sl, n);
    if (sl == 0) {
      const double ri = b[row] - s;
      acc += ri * ri;
    }
  }
  StorePartial(acc, part + kPF * gridDim.x);
}

// Overwrites the last history entry with the true residual; status 1 / 3
// (backward_error.pyx:139-143); marks the solve finished.
__global__ __launch_bounds__(kBlk) void PcgFinalFinish(const double* __restrict__ part, int G,
                                                       const double* __restrict__ ds, int* __restrict__ is,
                                                       double* __restrict__ hist) {
  if (!FinalDue(is)) return;
  const double rr = SumPartials(part + kPF * G, G);
  if (threadIdx.x != 0) return;
  const int count = is[kCountLatest];
  const double h = sqrt(rr) / ds[kBnorm];
  hist[count - 1] = h;
  if (!Finite(h))
    is[kStatus] = 3;
  else if (h <= ds[kTol])
    is[kStatus] = 1;
  is[kFinalDone] = 1;
}

// Plain Ap = A p with <p, Ap> partials (timing of the SpMV itself).
template <int SG>
__global__ __launch_bounds__(kBlk) void PcgSpmvA(int n, const int* __restrict__ ap_, const int* __restrict__ ac,
                                                 const double* __restrict__ av, const double* __restrict__ p,
                                                 double* __restrict__ Ap, double* __restrict__ part) {
  const int groups = kBlk / SG, sl = threadIdx.x % SG;
  double acc = 0.0;
  for (int row = XcdBlock() * groups + threadIdx.x / SG; row < n; row += gridDim.x * groups) {
    const double s = RowDot<SG, double>(ap_, ac, av, p, row, sl);
    if (sl == 0) {
      Ap[row] = s;
      acc += p[row] * s;
    }
  }
  StorePartial(acc, part + kP1 * gridDim.x);
}

// The same SpMV + <p, Ap> on the sliced-ELLPACK copy of A, one thread per row
// as the iteration kernels read it (8 column indices and values in flight,
// then the gathers).
__global__ __launch_bounds__(kBlk) void PcgSpmvASell(int n, SellRows rows, const int* __restrict__ ac,
                                                     const double* __restrict__ av, const double* __restrict__ p,
                                                     double* __restrict__ Ap, double* __restrict__ part) {
  double acc = 0.0;
  auto gather = [&](int c) { return p[c]; };
  for (int row = XcdBlock() * kBlk + threadIdx.x; row < n; row += gridDim.x * kBlk) {
    const double s = RowSum<double, SellRows::kPadded, true>(rows.Row(row), ac, av, 0, gather);
    Ap[row] = s;
    acc += p[row] * s;
  }
  StorePartial(acc, part + kP1 * gridDim.x);
}

// Start of a solve: every scalar slot 0 except the tolerance and the
// iteration limit.
__global__ void PcgSetScalars(double* __restrict__ ds, int* __restrict__ is, double tolerance, int max_iterations) {
  const int t = threadIdx.x;
  if (t < kNumD) ds[t] = t == kTol ? tolerance : 0.0;
  if (t < kNumI) is[t] = t == kMax ? max_iterations : 0;
}

// SELL construction: per-slice width (x 64 entries), then the slots.
__global__ void SellWidths(int64_t n, int64_t nslices, const int* __restrict__ ptr, int* __restrict__ slots) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  int w = 0;
  const int64_t r1 = min(n, (s + 1) * 64);
  for (int64_t r = s * 64; r < r1; ++r) w = max(w, ptr[r + 1] - ptr[r]);
  slots[s] = w * 64;
}

__global__ void SellFill(int64_t n, int64_t nslices, const int* __restrict__ off, const int* __restrict__ ptr,
                         const int* __restrict__ col, int* __restrict__ scol, int* __restrict__ ssrc) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= nslices * 64) return;
  const int s = int(r >> 6), l = int(r & 63);
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  const int rb = r < n ? ptr[r] : 0, len = r < n ? ptr[r + 1] - rb : 0;
  for (int u = 0; u < w; ++u) {
    const int k = b + u * 64 + l;
    scol[k] = u < len ? col[rb + u] : -1;
    ssrc[k] = u < len ? rb + u : -1;
  }
}

template <class T>
__global__ void SellGather(int64_t len, const int* __restrict__ src, const T* __restrict__ v, T* __restrict__ out) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < len) out[k] = src[k] >= 0 ? v[src[k]] : T(0);
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
  NarrowToDevice(ctx, m.ptr.get(), ptr, n + 1, base);
  if (m.nnz > 0) {
    NarrowToDevice(ctx, m.col.get(), col + base, m.nnz, 0);
    CopyToDevice(ctx, m.val.get(), val + base, sizeof(double) * size_t(m.nnz));
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

void CopyValues(Ctx& ctx, DCsr& dst, const DCsr& src) {
  if (dst.n != src.n || dst.nnz != src.nnz) throw std::invalid_argument("CopyValues: patterns differ");
  if (dst.nnz == 0) return;
  SPAI_HIP_CHECK(hipMemcpyAsync(dst.val.get(), src.val.get(), sizeof(double) * dst.nnz, hipMemcpyDeviceToDevice,
                                ctx.stream));
  if (dst.val32.get())
    ToFloat<<<(unsigned)((dst.nnz + 255) / 256), 256, 0, ctx.stream>>>(dst.nnz, dst.val.get(), dst.val32.get());
}

DSell SellFromCsr(Ctx& ctx, const DCsr& m, bool fp32) {
  DSell s;
  s.n = m.n;
  s.nslices = (m.n + 63) / 64;
  auto slots = ctx.Alloc<int>(std::max<int64_t>(s.nslices, 1));
  SellWidths<<<(unsigned)((s.nslices + 255) / 256), 256, 0, ctx.stream>>>(m.n, s.nslices, m.ptr.get(), slots.get());
  s.off = ctx.Alloc<int>(s.nslices + 1);
  s.len = ScanCounts(ctx, slots.get(), s.nslices, s.off.get());
  if (s.len >= (int64_t(1) << 31)) throw std::runtime_error("SELL copy exceeds the int32 range");
  // kSellPad entries of padding (col = src = -1, values 0): a thread reads 8
  // entries of its row at a time, also past the row's end
  s.col = ctx.Alloc<int>(s.len + kSellPad);
  s.src = ctx.Alloc<int>(s.len + kSellPad);
  SPAI_HIP_CHECK(hipMemsetAsync(s.col.get() + s.len, 0xFF, sizeof(int) * kSellPad, ctx.stream));
  SPAI_HIP_CHECK(hipMemsetAsync(s.src.get() + s.len, 0xFF, sizeof(int) * kSellPad, ctx.stream));
  SellFill<<<(unsigned)((s.nslices * 64 + 255) / 256), 256, 0, ctx.stream>>>(m.n, s.nslices, s.off.get(),
                                                                            m.ptr.get(), m.col.get(), s.col.get(),
                                                                            s.src.get());
  s.val = ctx.Alloc<double>(s.len + kSellPad);
  SPAI_HIP_CHECK(hipMemsetAsync(s.val.get() + s.len, 0, sizeof(double) * kSellPad, ctx.stream));
  if (fp32) {
    s.val32 = ctx.Alloc<float>(s.len + kSellPad);
    SPAI_HIP_CHECK(hipMemsetAsync(s.val32.get() + s.len, 0, sizeof(float) * kSellPad, ctx.stream));
  }
  SellValues(ctx, s, m);
  return s;
}

void SellValues(Ctx& ctx, DSell& s, const DCsr& m) {
  if (s.len == 0) return;
  const unsigned g = (unsigned)((s.len + 255) / 256);
  SellGather<double><<<g, 256, 0, ctx.stream>>>(s.len, s.src.get(), m.val.get(), s.val.get());
  if (s.val32.get()) SellGather<float><<<g, 256, 0, ctx.stream>>>(s.len, s.src.get(), m.val32.get(), s.val32.get());
}

DDia DiaFromCsr(Ctx& ctx, const DCsr& m, bool fp32) {
  DDia d;
  d.n = m.n;
  if (m.n <= 0 || m.nnz <= 0 || m.n > (int64_t(1) << 29)) return d;
  const int n = (int)m.n;
  const int64_t len = 2 * m.n - 1;  // offsets -(n-1) .. n-1
  auto flag = ctx.Alloc<int>(len);
  auto pos = ctx.Alloc<int>(len + 1);
  SPAI_HIP_CHECK(hipMemsetAsync(flag.get(), 0, sizeof(int) * len, ctx.stream));
  const unsigned g = (unsigned)((m.n + 255) / 256);
  DiaMark<<<g, 256, 0, ctx.stream>>>(n, m.ptr.get(), m.col.get(), flag.get());
  const int64_t nd = ScanCounts(ctx, flag.get(), len, pos.get());
  if (nd > kDiaMaxDiagonals || double(nd) * double(m.n) > 1.5 * double(m.nnz)) return d;
  auto off = ctx.Alloc<int>(nd);
  DiaOffsets<<<(unsigned)((len + 255) / 256), 256, 0, ctx.stream>>>(len, n, flag.get(), pos.get(), off.get());
  d.off.resize(nd);
  ctx.ReadBack(d.off.data(), off.get(), sizeof(int) * nd);
  d.src = ctx.Alloc<int>(nd * m.n);
  d.mask = ctx.Alloc<unsigned>(m.n);
  DiaFill<<<g, 256, 0, ctx.stream>>>(n, (int)nd, m.ptr.get(), m.col.get(), pos.get(), d.src.get(), d.mask.get());
  d.val = ctx.Alloc<double>(nd * m.n);
  if (fp32) d.val32 = ctx.Alloc<float>(nd * m.n);
  d.nd = nd;
  DiaValues(ctx, d, m);
  return d;
}

void DiaValues(Ctx& ctx, DDia& d, const DCsr& m) {
  const int64_t len = d.nd * d.n;
  if (len == 0) return;
  const unsigned g = (unsigned)((len + 255) / 256);
  SellGather<double><<<g, 256, 0, ctx.stream>>>(len, d.src.get(), m.val.get(), d.val.get());
  if (d.val32.get()) SellGather<float><<<g, 256, 0, ctx.stream>>>(len, d.src.get(), m.val32.get(), d.val32.get());
}

namespace {

DiaView ViewOf(const DDia& d) {
  DiaView v{};
  for (int64_t i = 0; i < d.nd; ++i) v.off[i] = d.off[i];
  v.mask = d.mask.get();
  v.nd = (int)d.nd;
  return v;
}

// The cap on the layout choice from SPAI_PCG_LAYOUT ("csr", "sell", "auto").
PcgLayout LayoutCap() {
  const char* e = std::getenv("SPAI_PCG_LAYOUT");
  if (!e || !*e || std::strcmp(e, "auto") == 0) return PcgLayout::kDia;
  if (std::strcmp(e, "sell") == 0) return PcgLayout::kSell;
  if (std::strcmp(e, "csr") == 0) return PcgLayout::kCsr;
  throw std::invalid_argument("SPAI_PCG_LAYOUT must be csr, sell or auto");
}

// f(dot): the row-dot policy of matrix m for the once-per-solve kernels
// (diagonals when the solver reads m in that layout, else CSR with the lanes
// per row of its mean row length); val / dval are its CSR / diagonal values.
template <class MV, class F>
void WithDot(const DCsr& m, const MV* val, const DDia& d, const MV* dval, PcgLayout layout, F&& f) {
  if (layout == PcgLayout::kDia) {
    if (d.nd <= 8)
      f(DiaDot<MV, 1>{ViewOf(d), dval});
    else
      f(DiaDot<MV, 4>{ViewOf(d), dval});
    return;
  }
  switch (SubgroupFor(m.n, m.nnz)) {
    case 1: f(CsrDot<1, MV>{m.ptr.get(), m.col.get(), val}); break;
    case 16: f(CsrDot<16, MV>{m.ptr.get(), m.col.get(), val}); break;
    default: f(CsrDot<32, MV>{m.ptr.get(), m.col.get(), val}); break;
  }
}

}  // namespace

PcgSolver::PcgSolver(Ctx& ctx, const DCsr& A, const DCsr& M, bool m_fp32)
    : ctx_(ctx), A_(A), M_(M), m_fp32_(m_fp32), n_(A.n) {
  if (M.n != A.n) throw std::invalid_argument("A and M must have the same shape");
  if (m_fp32 && !M.val32.get()) throw std::invalid_argument("FP32 copy of M missing");
  grid_ = Grid(n_, kBlk, ctx.num_cus);
  // At most kMaxPartials * kBlk = 512 blocks: the iteration kernels'
  // prologues sum every block's partial, and fewer, fuller blocks (two rows a
  // thread once n > 131,072) measured faster on the 64^3 solve than one row
  // a thread: 1.76 against 1.89 ms with 1,024 blocks of 256 threads and
  // 2.03 ms with 512 blocks of 512.
  if (grid_ > unsigned(kMaxPartials * kBlk)) grid_ = unsigned(kMaxPartials * kBlk);
  x_ = ctx.Alloc<double>(n_);
  r_ = ctx.Alloc<double>(2 * n_);
  z_ = ctx.Alloc<double>(n_);
  p_ = ctx.Alloc<double>(2 * n_);
  ap_ = ctx.Alloc<double>(n_);
  part_ = ctx.Alloc<double>(size_t(kNumP) * grid_);
  scal_ = ctx.Alloc<double>(kNumD);
  ist_ = ctx.Alloc<int>(kNumI);
  // the iteration kernels' row layouts: diagonals (no index loads) when the
  // matrix lies on few diagonals, else SELL copies for short rows
  // (coalesced, one thread per row), else CSR with lanes per row
  const PcgLayout cap = LayoutCap();
  auto choose = [&](const DCsr& m, bool fp32, DSell& sell, DDia& dia) {
    if (cap == PcgLayout::kDia) {
      dia = DiaFromCsr(ctx, m, fp32);
      if (dia.nd > 0) return PcgLayout::kDia;
    }
    if (cap != PcgLayout::kCsr && SubgroupFor(m.n, m.nnz) == 1) {
      sell = SellFromCsr(ctx, m, fp32);
      return PcgLayout::kSell;
    }
    return PcgLayout::kCsr;
  };
  a_layout_ = choose(A, false, a_sell_, a_dia_);
  m_layout_ = choose(M, m_fp32, m_sell_, m_dia_);
  ctx.Sync();
}

void PcgSolver::RefreshPreconditioner() {
  if (m_layout_ == PcgLayout::kSell) SellValues(ctx_, m_sell_, M_);
  if (m_layout_ == PcgLayout::kDia) DiaValues(ctx_, m_dia_, M_);
}

PcgSolver::~PcgSolver() { DropGraphs(); }

void PcgSolver::DropGraphs() {
  for (int i = 0; i < 2; ++i) {
    if (exec_[i]) (void)hipGraphExecDestroy(exec_[i]);
    if (graph_[i]) (void)hipGraphDestroy(graph_[i]);
    if (tail_exec_[i]) (void)hipGraphExecDestroy(tail_exec_[i]);
    if (tail_graph_[i]) (void)hipGraphDestroy(tail_graph_[i]);
    exec_[i] = nullptr;
    graph_[i] = nullptr;
    tail_exec_[i] = nullptr;
    tail_graph_[i] = nullptr;
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

// Diagonal kernels: one chunk of 8 diagonals (nd <= 8) or up to four.
#define SPAI_DIA_DISPATCH(nd, ...)   \
  if ((nd) <= 8) {                   \
    constexpr int kChunks = 1;       \
    __VA_ARGS__;                     \
  } else {                           \
    constexpr int kChunks = 4;       \
    __VA_ARGS__;                     \
  }

// The iteration kernels on diagonals, also by whether a thread has more
// than one row (kPipe: each row's loads overlap the row before).
#define SPAI_DIA_PIPE_DISPATCH(nd, pipe, ...) \
  if (pipe) {                                 \
    constexpr bool kPipe = true;              \
    SPAI_DIA_DISPATCH(nd, __VA_ARGS__);       \
  } else {                                    \
    constexpr bool kPipe = false;             \
    SPAI_DIA_DISPATCH(nd, __VA_ARGS__);       \
  }

void PcgSolver::Record(int first_parity, int64_t iterations) {
  const int sga = SubgroupFor(A_.n, A_.nnz), sgm = SubgroupFor(M_.n, M_.nnz);
  hipStream_t st = ctx_.stream;
  const int n = (int)n_;
  const bool pipe = n_ > int64_t(grid_) * kBlk;
  for (int64_t it = 0; it < iterations; ++it) {
    const int par = int((first_parity + it) & 1);
    double* p_in = p_.get() + size_t(par) * n_;
    double* p_out = p_.get() + size_t(par ^ 1) * n_;
    const double* r_in = r_.get() + size_t(par) * n_;
    double* r_out = r_.get() + size_t(par ^ 1) * n_;
    if (a_layout_ == PcgLayout::kDia) {
      SPAI_DIA_PIPE_DISPATCH(a_dia_.nd, pipe, (PcgDirSpmvADia<kChunks, kPipe><<<grid_, kBlk, 0, st>>>(
                                       n, par, ViewOf(a_dia_), a_dia_.val.get(), z_.get(), p_in, p_out, ap_.get(),
                                       part_.get(), scal_.get(), ist_.get(), hist_.get())));
    } else if (a_layout_ == PcgLayout::kSell) {
      PcgDirSpmvA<1, SellRows><<<grid_, kBlk, 0, st>>>(n, par, SellRows{a_sell_.off.get()}, a_sell_.col.get(),
                                                       a_sell_.val.get(), z_.get(), p_in, p_out, ap_.get(),
                                                       part_.get(), scal_.get(), ist_.get(), hist_.get());
    } else {
      SPAI_SG_DISPATCH(sga, (PcgDirSpmvA<SG, CsrRows><<<grid_, kBlk, 0, st>>>(
                                n, par, CsrRows{A_.ptr.get()}, A_.col.get(), A_.val.get(), z_.get(), p_in, p_out,
                                ap_.get(), part_.get(), scal_.get(), ist_.get(), hist_.get())));
    }
    if (m_layout_ == PcgLayout::kDia && m_fp32_) {
      SPAI_DIA_PIPE_DISPATCH(m_dia_.nd, pipe, (PcgUpdateSpmvMDia<float, kChunks, kPipe><<<grid_, kBlk, 0, st>>>(
                                       n, par, ViewOf(m_dia_), m_dia_.val32.get(), x_.get(), r_in, r_out, p_out,
                                       ap_.get(), z_.get(), part_.get(), scal_.get(), ist_.get())));
    } else if (m_layout_ == PcgLayout::kDia) {
      SPAI_DIA_PIPE_DISPATCH(m_dia_.nd, pipe, (PcgUpdateSpmvMDia<double, kChunks, kPipe><<<grid_, kBlk, 0, st>>>(
                                       n, par, ViewOf(m_dia_), m_dia_.val.get(), x_.get(), r_in, r_out, p_out,
                                       ap_.get(), z_.get(), part_.get(), scal_.get(), ist_.get())));
    } else if (m_layout_ == PcgLayout::kSell && m_fp32_) {
      PcgUpdateSpmvM<1, float, SellRows><<<grid_, kBlk, 0, st>>>(
          n, par, SellRows{m_sell_.off.get()}, m_sell_.col.get(), m_sell_.val32.get(), x_.get(), r_in, r_out, p_out,
          ap_.get(), z_.get(), part_.get(), scal_.get(), ist_.get());
    } else if (m_layout_ == PcgLayout::kSell) {
      PcgUpdateSpmvM<1, double, SellRows><<<grid_, kBlk, 0, st>>>(
          n, par, SellRows{m_sell_.off.get()}, m_sell_.col.get(), m_sell_.val.get(), x_.get(), r_in, r_out, p_out,
          ap_.get(), z_.get(), part_.get(), scal_.get(), ist_.get());
    } else if (m_fp32_) {
      SPAI_SG_DISPATCH(sgm, (PcgUpdateSpmvM<SG, float, CsrRows><<<grid_, kBlk, 0, st>>>(
                                n, par, CsrRows{M_.ptr.get()}, M_.col.get(), M_.val32.get(), x_.get(), r_in, r_out,
                                p_out, ap_.get(), z_.get(), part_.get(), scal_.get(), ist_.get())));
    } else {
      SPAI_SG_DISPATCH(sgm, (PcgUpdateSpmvM<SG, double, CsrRows><<<grid_, kBlk, 0, st>>>(
                                n, par, CsrRows{M_.ptr.get()}, M_.col.get(), M_.val.get(), x_.get(), r_in, r_out,
                                p_out, ap_.get(), z_.get(), part_.get(), scal_.get(), ist_.get())));
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
  // the solve's scalars, written by a kernel (no copy from pageable memory)
  PcgSetScalars<<<1, kWave, 0, st>>>(scal_.get(), ist_.get(), tolerance, (int)max_iterations);
  auto with_a = [&](auto&& f) { WithDot<double>(A_, A_.val.get(), a_dia_, a_dia_.val.get(), a_layout_, f); };
  auto with_m = [&](auto&& f) {
    if (m_fp32_)
      WithDot<float>(M_, M_.val32.get(), m_dia_, m_dia_.val32.get(), m_layout_, f);
    else
      WithDot<double>(M_, M_.val.get(), m_dia_, m_dia_.val.get(), m_layout_, f);
  };
  with_m([&](auto dot) {
    PcgInit<decltype(dot)><<<grid_, kBlk, 0, st>>>(n, dot, b_dev, x_.get(), r_.get(), z_.get(), p_.get(), part_.get());
  });
  PcgInitFinish<<<1, kBlk, 0, st>>>(part_.get(), (int)grid_, scal_.get(), ist_.get(), hist_.get());
  if (!exec_[0]) {
    for (int par = 0; par < 2; ++par) {
      SPAI_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      Record(par, kChunk);
      SPAI_HIP_CHECK(hipStreamEndCapture(st, &graph_[par]));
      SPAI_HIP_CHECK(hipGraphInstantiate(&exec_[par], graph_[par], nullptr, nullptr, 0));
      SPAI_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      Record(par, kTailChunk);
      SPAI_HIP_CHECK(hipStreamEndCapture(st, &tail_graph_[par]));
      SPAI_HIP_CHECK(hipGraphInstantiate(&tail_exec_[par], tail_graph_[par], nullptr, nullptr, 0));
    }
    graph_key_b_ = b_dev;
  }
  // PcgInitFinish leaves iteration 0 pending unless the limit or the
  // tolerance ends the solve at once: the host knows which without reading
  int hst[kNumI] = {0};
  hst[kDone] = (1 < max_iterations && 1.0 > tolerance) ? 0 : 1;
  hst[kIter] = 0;
  // The device ends the solve by itself: every iteration it runs either
  // appends a history entry or stops (limit, tolerance crossing pending a
  // refresh, breakdown).  Each round launches a batch of iterations, then the
  // refresh (runs only when K1 paused at the tolerance crossing) and the final
  // true residual (runs only once the solve has ended), then reads the state
  // back once: a solve whose crossing falls in its first batch and ends at
  // the refresh takes one read-back.  Launches past the stop return at once.
  bool tail = false, first = true;
  int64_t ahead = 1;
  for (;;) {
    const int count_before = hst[kCountLatest], refreshes_before = hst[kRefreshes];
    if (!hst[kDone]) {
      int64_t full = 0, rest = 0;
      if (first) {
        // as many iterations as the previous solve ran, plus the one whose K1
        // finds the tolerance crossing (at least one chunk, at most this
        // solve's limit), whole chunks and then tail chunks
        const int64_t want =
            std::min<int64_t>(max_iterations + 1, std::max<int64_t>(kChunk, last_iterations_ + 1));
        full = want / kChunk;
        rest = (want % kChunk + kTailChunk - 1) / kTailChunk;
      } else if (tail) {
        // past the refresh: tail chunks, doubling each round in case the
        // solve goes on
        rest = ahead;
        ahead = std::min<int64_t>(2 * ahead, 64);
      } else {
        full = 1;  // then one chunk at a time
      }
      // chunks hold an even number of iterations: each starts at the same parity
      const int par0 = hst[kIter] & 1;
      for (int64_t c = 0; c < full; ++c) SPAI_HIP_CHECK(hipGraphLaunch(exec_[par0], st));
      for (int64_t c = 0; c < rest; ++c) SPAI_HIP_CHECK(hipGraphLaunch(tail_exec_[par0], st));
    }
    first = false;
    with_a([&](auto dot) {
      PcgRefresh<decltype(dot)><<<grid_, kBlk, 0, st>>>(n, dot, x_.get(), b_dev, r_.get(), part_.get(), ist_.get());
    });
    with_m([&](auto dot) {
      PcgRefreshSpmvM<decltype(dot)><<<grid_, kBlk, 0, st>>>(n, dot, r_.get(), z_.get(), part_.get(), scal_.get(),
                                                            ist_.get());
    });
    PcgRefreshFinish<<<1, kBlk, 0, st>>>(part_.get(), (int)grid_, scal_.get(), ist_.get(), hist_.get());
    with_a([&](auto dot) {
      PcgFinal<decltype(dot)><<<grid_, kBlk, 0, st>>>(n, dot, x_.get(), b_dev, part_.get(), ist_.get());
    });
    PcgFinalFinish<<<1, kBlk, 0, st>>>(part_.get(), (int)grid_, scal_.get(), ist_.get(), hist_.get());
    SPAI_HIP_CHECK(hipGetLastError());
    ctx_.ReadBack(hst, ist_.get(), sizeof(hst));
    if (hst[kFinalDone]) break;
    if (hst[kRefreshes] != refreshes_before) {
      tail = true;  // past the tolerance crossing only a few iterations remain
      ahead = 1;
    }
    if (hst[kCountLatest] <= count_before)
      throw std::runtime_error("PCG: launched iterations appended no history entry");
  }
  *count = hst[kCountLatest];
  last_iterations_ = hst[kIter];
  return hst[kStatus];
}

void PcgSolver::History(double* host, int64_t count) {
  ctx_.ReadBack(host, hist_.get(), sizeof(double) * count);
}

double PcgSolver::TimeSpmvA(int reps) {
  const int sga = SubgroupFor(A_.n, A_.nnz);
  hipStream_t st = ctx_.stream;
  const int n = (int)n_;
  auto launch = [&] {
    if (a_layout_ == PcgLayout::kDia) {
      SPAI_DIA_DISPATCH(a_dia_.nd, (PcgSpmvADia<kChunks><<<grid_, kBlk, 0, st>>>(
                                       n, ViewOf(a_dia_), a_dia_.val.get(), x_.get(), ap_.get(), part_.get())));
    } else if (a_layout_ == PcgLayout::kSell) {
      PcgSpmvASell<<<grid_, kBlk, 0, st>>>(n, SellRows{a_sell_.off.get()}, a_sell_.col.get(), a_sell_.val.get(),
                                           x_.get(), ap_.get(), part_.get());
    } else {
      SPAI_SG_DISPATCH(sga, (PcgSpmvA<SG><<<grid_, kBlk, 0, st>>>(n, A_.ptr.get(), A_.col.get(), A_.val.get(),
                                                                  x_.get(), ap_.get(), part_.get())));
    }
  };
  // one launch to warm the caches' translations, then `reps` back-to-back
  // launches between one pair of events (the gaps between launches are
  // included, the host's launch latency of a lone kernel is not)
  launch();
  hipEvent_t ev[2];
  for (auto& e : ev) SPAI_HIP_CHECK(hipEventCreate(&e));
  SPAI_HIP_CHECK(hipEventRecord(ev[0], st));
  for (int i = 0; i < reps; ++i) launch();
  SPAI_HIP_CHECK(hipEventRecord(ev[1], st));
  SPAI_HIP_CHECK(hipEventSynchronize(ev[1]));
  float ms = 0.f;
  SPAI_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  return double(ms) * 1e3 / reps;
}

void SpMV(Ctx& ctx, const DCsr& A, const double* x, double* y) {
  const int sga = SubgroupFor(A.n, A.nnz);
  const unsigned g = Grid(A.n, kBlk, ctx.num_cus);
  SPAI_SG_DISPATCH(sga, (SpmvKernel<SG><<<g, kBlk, 0, ctx.stream>>>((int)A.n, A.ptr.get(), A.col.get(), A.val.get(),
                                                                    x, y)));
}


}  // namespace spai
