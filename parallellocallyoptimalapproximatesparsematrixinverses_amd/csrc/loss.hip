// f(M, A) = ||I - A M||_F^2 on fixed patterns and its gradients:
//   df/dM = -2 (A^T R) sampled on pattern(M),   df/dA = -2 (R M^T) sampled on pattern(A),
// with R = I - A M.  The first is the sampled product that the reference's drop
// score uses (src/common/dropping.cpp:19-29,145: (A R)_kl for symmetric A).
//
// R is formed per block of consecutive columns and never stored whole: the
// forward pass sums ||R_block||^2; the backward pass re-forms each block, writes
// the gradient entries of M in the block's columns and adds the block's share
// of R M^T into an FP64 accumulator per entry of A (blocks in column order, so
// the result is deterministic).  Values are FP64 or FP32 ("FP32 bulk": FP32
// products and storage, FP64 reductions and gradient accumulation).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dev_common.hpp"
#include "ops.hpp"
#include "spai_ext.h"

namespace spai {
namespace {

constexpr int kBlk = 256;

// Sorted-intersection dot of x[xb, xe) (rows xi) and y[yb, ye) (rows yi).
template <class TX, class TY>
__device__ __forceinline__ double Intersect(const int* __restrict__ xi, const TX* __restrict__ xv, int xb, int xe,
                                            const int* __restrict__ yi, const TY* __restrict__ yv, int yb, int ye) {
  double s = 0.0;
  while (xb < xe && yb < ye) {
    const int a = xi[xb], b = yi[yb];
    if (a == b) {
      s += double(xv[xb]) * double(yv[yb]);
      ++xb;
      ++yb;
    } else if (a < b) {
      ++xb;
    } else {
      ++yb;
    }
  }
  return s;
}

// R(i, j) = [i == j] - sum_k A(i, k) M(k, j) for the S2 entries q in
// [q0, q1): A^T column i (rows k, values gathered through atperm) against
// column j of M (rows k), k ascending.
template <class V>
__global__ __launch_bounds__(kBlk) void ResidualBlock(int64_t q0, int64_t q1, const int* __restrict__ s2i,
                                                      const int* __restrict__ s2c, const int* __restrict__ atp,
                                                      const int* __restrict__ ati, const int* __restrict__ atperm,
                                                      const V* __restrict__ av, const int* __restrict__ sp,
                                                      const int* __restrict__ si, const V* __restrict__ mv,
                                                      V* __restrict__ rblk, double* __restrict__ part) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t q = q0 + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < q1; q += int64_t(gridDim.x) * blockDim.x) {
    const int i = s2i[q], j = s2c[q];
    V s = V(0);
    int a = atp[i], ae = atp[i + 1], b = sp[j], be = sp[j + 1];
    while (a < ae && b < be) {
      const int ka = ati[a], kb = si[b];
      if (ka == kb) {
        s += av[atperm[a]] * mv[b];
        ++a;
        ++b;
      } else if (ka < kb) {
        ++a;
      } else {
        ++b;
      }
    }
    V r = -s;
    if (i == j) r += V(1);
    rblk[q - q0] = r;
    acc += double(r) * double(r);
  }
  acc = BlockSum<double>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// dM(k, l) = -2 scale * sum_i A(i, k) R(i, l) for the S entries p in [p0, p1).
template <class V>
__global__ __launch_bounds__(kBlk) void GradMBlock(int64_t p0, int64_t p1, int64_t q0, const int* __restrict__ si,
                                                   const int* __restrict__ scol, const int* __restrict__ ap,
                                                   const int* __restrict__ ai, const V* __restrict__ av,
                                                   const int* __restrict__ s2p, const int* __restrict__ s2i,
                                                   const V* __restrict__ rblk, double scale, V* __restrict__ gm) {
  for (int64_t p = p0 + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < p1; p += int64_t(gridDim.x) * blockDim.x) {
    const int k = si[p], l = scol[p];
    const int rb = s2p[l], re = s2p[l + 1];
    const double s = Intersect(ai, av, ap[k], ap[k + 1], s2i, rblk - q0, rb, re);
    gm[p] = V(-2.0 * scale * s);
  }
}

// acc(i, k) += sum_{j in [j0, j1)} R(i, j) M(k, j) for every entry (i, k) of A:
// row i of R (column i of S2^T, columns j) against row k of M (column k of S^T).
template <class V>
__global__ __launch_bounds__(kBlk) void GradABlock(int64_t nnz_a, int j0, int j1, int64_t q0, const int* __restrict__ ai,
                                                   const int* __restrict__ acol, const int* __restrict__ s2tp,
                                                   const int* __restrict__ s2ti, const int* __restrict__ s2tperm,
                                                   const V* __restrict__ rblk, const int* __restrict__ stp,
                                                   const int* __restrict__ sti, const int* __restrict__ stperm,
                                                   const V* __restrict__ mv, double* __restrict__ acc) {
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nnz_a; e += int64_t(gridDim.x) * blockDim.x) {
    const int i = ai[e], k = acol[e];
    int x = LowerBound(s2ti, s2tp[i], s2tp[i + 1], j0), xe = LowerBound(s2ti, x, s2tp[i + 1], j1);
    int y = LowerBound(sti, stp[k], stp[k + 1], j0), ye = LowerBound(sti, y, stp[k + 1], j1);
    double s = 0.0;
    while (x < xe && y < ye) {
      const int a = s2ti[x], b = sti[y];
      if (a == b) {
        s += double(rblk[s2tperm[x] - q0]) * double(mv[stperm[y]]);
        ++x;
        ++y;
      } else if (a < b) {
        ++x;
      } else {
        ++y;
      }
    }
    acc[e] += s;
  }
}

template <class V>
__global__ void ScaleOut(int64_t n, const double* __restrict__ acc, double scale, V* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = V(-2.0 * scale * acc[i]);
}

__global__ void ColOf(int n, const int* __restrict__ ptr, int* __restrict__ col) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int p = ptr[j]; p < ptr[j + 1]; ++p) col[p] = j;
}

__global__ void Iota(int64_t n, double* __restrict__ v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = double(i);
}

__global__ void ToInt(int64_t n, const double* __restrict__ v, int* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = int(v[i]);
}

// Grid of the grid-stride kernels (capped; partial sums fit `part`).
unsigned GridOf(int64_t n) {
  int64_t g = (n + kBlk - 1) / kBlk;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 4096));
}

// Grid of the one-thread-per-element kernels: every element needs a thread.
unsigned GridAll(int64_t n) {
  const int64_t g = (n + kBlk - 1) / kBlk;
  if (g > (int64_t(1) << 31) - 1) throw std::runtime_error("grid too large");
  return (unsigned)std::max<int64_t>(1, g);
}

}  // namespace
}  // namespace spai

struct SpaiLossPlan {
  std::unique_ptr<spai::Ctx> ctx;
  int64_t n = 0, block_cols = 0;
  bool fp32 = false;
  spai::DMat A, S, S2;                 // patterns (values unused)
  spai::PatternPtr AT, S2T, ST;        // transposed patterns
  spai::DBuf<int> atperm, s2tperm, stperm, s2col, scol, acol;
  spai::DBuf<char> rblk;               // one column block of R
  spai::DBuf<double> part, acc;
  int64_t max_block_entries = 0;
  std::vector<int> blocks;             // column boundaries
};

namespace {

thread_local std::string g_loss_error;

int64_t LossFail(const std::string& m) {
  g_loss_error = m;
  return -1;
}

// Transposed pattern of X plus, for each transposed entry, its position in X.
spai::PatternPtr TransposeWithPerm(spai::Ctx& c, const spai::DMat& X, spai::DBuf<int>& perm) {
  spai::DMat P;
  P.pat = X.pat;
  P.val = c.Alloc<double>(std::max<int64_t>(X.nnz(), 1));
  spai::Iota<<<spai::GridAll(X.nnz()), spai::kBlk, 0, c.stream>>>(X.nnz(), P.val.get());
  spai::DMat T = spai::Transpose(c, P);
  perm = c.Alloc<int>(std::max<int64_t>(X.nnz(), 1));
  spai::ToInt<<<spai::GridAll(X.nnz()), spai::kBlk, 0, c.stream>>>(X.nnz(), T.val.get(), perm.get());
  return T.pat;
}

template <class V>
int64_t Forward(SpaiLossPlan* p, const V* m, const V* a, double* loss, bool backward, double scale, V* gm, V* ga) {
  using namespace spai;
  Ctx& c = *p->ctx;
  hipStream_t st = c.stream;
  V* rblk = reinterpret_cast<V*>(p->rblk.get());
  const int* s2p = p->S2.ptr();
  if (backward) SPAI_HIP_CHECK(hipMemsetAsync(p->acc.get(), 0, sizeof(double) * p->A.nnz(), st));
  double total = 0.0;
  for (size_t b = 0; b + 1 < p->blocks.size(); ++b) {
    const int j0 = p->blocks[b], j1 = p->blocks[b + 1];
    int qq[2];
    SPAI_HIP_CHECK(hipMemcpyAsync(qq, s2p + j0, sizeof(int), hipMemcpyDeviceToHost, st));
    SPAI_HIP_CHECK(hipMemcpyAsync(qq + 1, s2p + j1, sizeof(int), hipMemcpyDeviceToHost, st));
    SPAI_HIP_CHECK(hipStreamSynchronize(st));
    const int64_t q0 = qq[0], q1 = qq[1];
    const unsigned g = GridOf(q1 - q0);
    ResidualBlock<V><<<g, kBlk, 0, st>>>(q0, q1, p->S2.idx(), p->s2col.get(), p->AT->ptr.get(), p->AT->idx.get(),
                                         p->atperm.get(), a, p->S.ptr(), p->S.idx(), m, rblk, p->part.get());
    std::vector<double> h(g);
    SPAI_HIP_CHECK(hipMemcpyAsync(h.data(), p->part.get(), sizeof(double) * g, hipMemcpyDeviceToHost, st));
    SPAI_HIP_CHECK(hipStreamSynchronize(st));
    double bs = 0.0;
    for (double v : h) bs += v;
    total += bs;
    if (backward) {
      int pp[2];
      SPAI_HIP_CHECK(hipMemcpyAsync(pp, p->S.ptr() + j0, sizeof(int), hipMemcpyDeviceToHost, st));
      SPAI_HIP_CHECK(hipMemcpyAsync(pp + 1, p->S.ptr() + j1, sizeof(int), hipMemcpyDeviceToHost, st));
      SPAI_HIP_CHECK(hipStreamSynchronize(st));
      GradMBlock<V><<<GridOf(pp[1] - pp[0]), kBlk, 0, st>>>(pp[0], pp[1], q0, p->S.idx(), p->scol.get(), p->A.ptr(),
                                                            p->A.idx(), a, s2p, p->S2.idx(), rblk, scale, gm);
      GradABlock<V><<<GridOf(p->A.nnz()), kBlk, 0, st>>>(p->A.nnz(), j0, j1, q0, p->A.idx(), p->acol.get(),
                                                         p->S2T->ptr.get(), p->S2T->idx.get(), p->s2tperm.get(), rblk,
                                                         p->ST->ptr.get(), p->ST->idx.get(), p->stperm.get(), m,
                                                         p->acc.get());
    }
  }
  if (backward)
    ScaleOut<V><<<GridAll(p->A.nnz()), kBlk, 0, st>>>(p->A.nnz(), p->acc.get(), scale, ga);
  SPAI_HIP_CHECK(hipGetLastError());
  SPAI_HIP_CHECK(hipStreamSynchronize(st));
  if (loss) *loss = total;
  return 0;
}

}  // namespace

extern "C" {

SpaiLossPlan* spai_loss_plan_create(MatrixData A, MatrixData M, int64_t precision, int64_t block_cols,
                                    uint64_t stream) {
  try {
    using namespace spai;
    if (A.rows != A.cols || M.rows != A.rows || M.cols != A.cols || A.rows <= 0)
      throw std::invalid_argument("A and M must be square matrices of the same shape");
    auto p = std::make_unique<SpaiLossPlan>();
    p->ctx = std::make_unique<Ctx>(reinterpret_cast<hipStream_t>(stream));
    Ctx& c = *p->ctx;
    p->n = A.rows;
    p->fp32 = precision != 0;
    p->block_cols = block_cols > 0 ? std::min<int64_t>(block_cols, p->n) : p->n;
    p->A = Upload(c, A.rows, A.cols, A.col_offsets, A.row_indices, A.values);
    p->S = Upload(c, M.rows, M.cols, M.col_offsets, M.row_indices, M.values);
    DMat S2 = SpGemm(c, p->A, p->S, false, true);
    p->S2 = std::move(S2);
    p->AT = TransposeWithPerm(c, p->A, p->atperm);
    p->S2T = TransposeWithPerm(c, p->S2, p->s2tperm);
    p->ST = TransposeWithPerm(c, p->S, p->stperm);
    p->s2col = c.Alloc<int>(std::max<int64_t>(p->S2.nnz(), 1));
    p->scol = c.Alloc<int>(std::max<int64_t>(p->S.nnz(), 1));
    p->acol = c.Alloc<int>(std::max<int64_t>(p->A.nnz(), 1));
    ColOf<<<GridAll(p->n), kBlk, 0, c.stream>>>((int)p->n, p->S2.ptr(), p->s2col.get());
    ColOf<<<GridAll(p->n), kBlk, 0, c.stream>>>((int)p->n, p->S.ptr(), p->scol.get());
    ColOf<<<GridAll(p->n), kBlk, 0, c.stream>>>((int)p->n, p->A.ptr(), p->acol.get());
    std::vector<int> s2p(p->n + 1);
    SPAI_HIP_CHECK(hipMemcpyAsync(s2p.data(), p->S2.ptr(), sizeof(int) * (p->n + 1), hipMemcpyDeviceToHost, c.stream));
    c.Sync();
    for (int64_t j = 0; j < p->n; j += p->block_cols) p->blocks.push_back((int)j);
    p->blocks.push_back((int)p->n);
    for (size_t b = 0; b + 1 < p->blocks.size(); ++b)
      p->max_block_entries =
          std::max<int64_t>(p->max_block_entries, s2p[p->blocks[b + 1]] - s2p[p->blocks[b]]);
    p->rblk = c.Alloc<char>(std::max<int64_t>(p->max_block_entries, 1) * (p->fp32 ? 4 : 8));
    p->part = c.Alloc<double>(4096);
    p->acc = c.Alloc<double>(std::max<int64_t>(p->A.nnz(), 1));
    c.Sync();
    return p.release();
  } catch (const std::exception& ex) {
    g_loss_error = ex.what();
    return nullptr;
  }
}

void spai_loss_plan_destroy(SpaiLossPlan* p) { delete p; }

int64_t spai_loss_plan_info(SpaiLossPlan* p, int64_t* out4) {
  if (!p) return LossFail("null plan");
  out4[0] = p->S2.nnz();
  out4[1] = p->max_block_entries;
  out4[2] = (int64_t)p->blocks.size() - 1;
  out4[3] = p->S.nnz();
  return 0;
}

int64_t spai_loss_forward(SpaiLossPlan* p, const void* m_values_dev, const void* a_values_dev, double* loss_out) {
  try {
    if (!p) throw std::invalid_argument("null plan");
    if (p->fp32)
      return Forward<float>(p, (const float*)m_values_dev, (const float*)a_values_dev, loss_out, false, 0.0, nullptr,
                            nullptr);
    return Forward<double>(p, (const double*)m_values_dev, (const double*)a_values_dev, loss_out, false, 0.0, nullptr,
                           nullptr);
  } catch (const std::exception& ex) {
    return LossFail(ex.what());
  }
}

int64_t spai_loss_backward(SpaiLossPlan* p, const void* m_values_dev, const void* a_values_dev, double grad_scale,
                           void* grad_m_dev, void* grad_a_dev) {
  try {
    if (!p) throw std::invalid_argument("null plan");
    if (p->fp32)
      return Forward<float>(p, (const float*)m_values_dev, (const float*)a_values_dev, nullptr, true, grad_scale,
                            (float*)grad_m_dev, (float*)grad_a_dev);
    return Forward<double>(p, (const double*)m_values_dev, (const double*)a_values_dev, nullptr, true, grad_scale,
                           (double*)grad_m_dev, (double*)grad_a_dev);
  } catch (const std::exception& ex) {
    return LossFail(ex.what());
  }
}

const char* spai_loss_last_error(void) { return g_loss_error.c_str(); }

}  // extern "C"
