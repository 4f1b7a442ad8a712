// f(M, A) = ||I - A M||_F^2 on fixed patterns and its gradients:
//   df/dM = -2 (A^T R) sampled on pattern(M),   df/dA = -2 (R M^T) sampled on pattern(A),
// with R = I - A M.  The first is the sampled product that the reference's drop
// score uses (src/common/dropping.cpp:19-29,145: (A R)_kl for symmetric A).
//
// Every product walks the terms A(i, k) M(k, j) (k over pattern(M)(:, j),
// i over A(:, k)) with the slot plan of the fixed patterns: the position of
// row i in S2(:, j) = pattern(A M) | I for each term, built once per plan.
// R is formed per block of consecutive columns, column by column in LDS, and
// never stored whole: the forward pass sums ||R||^2; the backward pass also
// computes df/dM on the column's M entries from the LDS copy (the terms of
// the product again, now a dot per entry of M), stores the block of R, and a
// second kernel adds the block's share of R M^T into an FP64 accumulator per
// entry of A (blocks in column order, so the result is deterministic).
// Values are FP64 or FP32 ("FP32 bulk": FP32 products and storage of R, FP64
// reductions and gradient accumulation).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dev_common.hpp"
#include "ops.hpp"
#include "slot_plan.hpp"
#include "spai_ext.h"

namespace spai {
namespace {

constexpr int kBlk = 256;

// Offsets of the product terms inside each column's plan: for the S entry e
// = (k, l), poff[e] = the number of terms of column l before those of A(:, k).
// One thread per column, the A column bounds of eight S entries loaded
// together.
__global__ void TermOffsets(int n, const int* __restrict__ sp, const int* __restrict__ si,
                            const int* __restrict__ ap, int* __restrict__ poff) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  const int e0 = sp[l], e1 = sp[l + 1];
  int off = 0;
  for (int e = e0; e < e1; e += 8) {
    int k[8], len[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) k[t] = e + t < e1 ? si[e + t] : -1;
#pragma unroll
    for (int t = 0; t < 8; ++t) len[t] = k[t] >= 0 ? ap[k[t] + 1] - ap[k[t]] : 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (e + t < e1) poff[e + t] = off;
      off += len[t];
    }
  }
}

// Position of row j inside S2(:, j) for every column (-1 when absent).
__global__ void DiagSlots(int n, const int* __restrict__ s2p, const int* __restrict__ s2i, int* __restrict__ s2d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int q0 = s2p[j], q1 = s2p[j + 1];
  const int p = LowerBound(s2i, q0, q1, j);
  s2d[j] = p < q1 && s2i[p] == j ? p - q0 : -1;
}

// Sum over the GL lanes of a group (a fixed tree, so deterministic).
template <int GL>
__device__ __forceinline__ double GroupSum(double v) {
#pragma unroll
  for (int off = GL / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, GL);
  return v;
}

// Columns j in [j0, j1) of R = I - A M, GL lanes per column (a power of two
// at least the longest column of A): R(:, j) is accumulated on S2(:, j) in
// LDS by the slot plan -- for each S entry (k, m) in order, the lanes add
// A(u, k) m at the slot of row u -- then ||R(:, j)||^2 goes to the block's
// FP64 partial.  kBack: R(:, j) is also stored in the block buffer rblk (at
// the column's S2 offset minus q_base), and df/dM(k, j) = -2 scale
// sum_u A(u, k) R(u, j) for every S entry of the column: the same terms
// walked again, each a product with R read back from LDS at its slot, summed
// over the group's lanes in FP64.  kNarrow: every column of A has at most GL
// entries (one term per lane and step).  The S entries' A column bounds and M
// values are staged in LDS per chunk of GL entries, so that a step reads them
// with one broadcast load.
constexpr int kLossBatch = 8;  // steps whose loads are in flight together

// One step of a column's product: the S entry's A column (start, length) and
// its M value, staged in LDS so that a step reads it with one broadcast load.
template <class V>
struct alignas(16) LossStep {
  int a0, c;
  V m;
};
template <class V, class SlotT, int GL, bool kBack, bool kNarrow>
__global__ __launch_bounds__(kBlk) void LossColumns(int j0, int j1, const int* __restrict__ sp,
                                                    const int* __restrict__ si, const V* __restrict__ mv,
                                                    const int* __restrict__ ap, const V* __restrict__ av,
                                                    const int64_t* __restrict__ plan_ptr,
                                                    const SlotT* __restrict__ plan, const int* __restrict__ s2p,
                                                    const int* __restrict__ s2d, int64_t q_base, int lmax,
                                                    V* __restrict__ rblk, double scale, V* __restrict__ gm,
                                                    double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ double red_s[16];
  const int groups = blockDim.x / GL;
  const int g = threadIdx.x / GL, gl = threadIdx.x % GL;
  // LDS: the groups' step tables, then their R columns (lmax entries each)
  LossStep<V>* const steps = reinterpret_cast<LossStep<V>*>(lds_raw) + size_t(g) * GL;
  V* const acc = reinterpret_cast<V*>(lds_raw + size_t(groups) * GL * sizeof(LossStep<V>)) + size_t(g) * lmax;
  double red = 0.0;
  for (int j = j0 + blockIdx.x * groups + g; j < j1; j += gridDim.x * groups) {
    const int q0 = s2p[j], L = s2p[j + 1] - q0;
    for (int t = gl; t < L; t += GL) acc[t] = V(0);
    WaveFence();
    const int e0 = sp[j], e1 = sp[j + 1];
    int64_t pp = plan_ptr[j];
    for (int eb = e0; eb < e1; eb += GL) {
      // the chunk's steps: lane gl stages S entry eb + gl
      if (eb + gl < e1) {
        const int k = si[eb + gl];
        const int a0 = ap[k];
        steps[gl] = LossStep<V>{a0, ap[k + 1] - a0, mv[eb + gl]};
      }
      WaveFence();
      const int ns = min(GL, e1 - eb);
      if (kNarrow) {
        // every column of A within GL: lane gl holds term gl of a step; the
        // slots and values of kLossBatch steps are loaded before their
        // additions, which then follow in step order
        for (int sb = 0; sb < ns; sb += kLossBatch) {
          int slot[kLossBatch];
          V prod[kLossBatch];
#pragma unroll
          for (int b = 0; b < kLossBatch; ++b) {
            const int st = sb + b;
            const LossStep<V> sd = steps[st < ns ? st : 0];
            const bool has = st < ns && gl < sd.c;
            slot[b] = has ? int(plan[pp + gl]) : -1;
            prod[b] = has ? av[sd.a0 + gl] * sd.m : V(0);
            if (st < ns) pp += sd.c;
          }
#pragma unroll
          for (int b = 0; b < kLossBatch; ++b)
            if (slot[b] >= 0) acc[slot[b]] += prod[b];
        }
      } else {
        for (int st = 0; st < ns; ++st) {
          const LossStep<V> sd = steps[st];
          // distinct rows within one column of A: no two lanes share a slot
          for (int u = gl; u < sd.c; u += GL) acc[plan[pp + u]] += av[sd.a0 + u] * sd.m;
          pp += sd.c;
        }
      }
      WaveFence();
    }
    WaveFence();
    const int dj = s2d[j];  // slot of the diagonal entry in S2(:, j)
    for (int t = gl; t < L; t += GL) {
      V r = -acc[t];
      if (t == dj) r += V(1);
      red += double(r) * double(r);
      if (kBack) {
        acc[t] = r;
        rblk[int64_t(q0) - q_base + t] = r;
      }
    }
    if (kBack) {
      WaveFence();
      pp = plan_ptr[j];
      for (int eb = e0; eb < e1; eb += GL) {
        if (eb + gl < e1) {
          const int k = si[eb + gl];
          const int a0 = ap[k];
          steps[gl] = LossStep<V>{a0, ap[k + 1] - a0, V(0)};
        }
        WaveFence();
        const int ns = min(GL, e1 - eb);
        if (kNarrow) {
          // kLossBatch dots at a time: the loads first, then the lane sums
          for (int sb = 0; sb < ns; sb += kLossBatch) {
            double s[kLossBatch];
#pragma unroll
            for (int b = 0; b < kLossBatch; ++b) {
              const int st = sb + b;
              const LossStep<V> sd = steps[st < ns ? st : 0];
              const bool has = st < ns && gl < sd.c;
              const int slot = has ? int(plan[pp + gl]) : 0;
              s[b] = has ? double(av[sd.a0 + gl]) * double(acc[slot]) : 0.0;
              if (st < ns) pp += sd.c;
            }
            double mine = 0.0;
#pragma unroll
            for (int b = 0; b < kLossBatch; ++b) {
              const double t = GroupSum<GL>(s[b]);
              if (gl == b) mine = t;
            }
            if (gl < kLossBatch && sb + gl < ns) gm[eb + sb + gl] = V(-2.0 * scale * mine);
          }
        } else {
          double mine = 0.0;  // lane st keeps the dot of step st
          for (int st = 0; st < ns; ++st) {
            const LossStep<V> sd = steps[st];
            double s = 0.0;
            for (int u = gl; u < sd.c; u += GL) s += double(av[sd.a0 + u]) * double(acc[plan[pp + u]]);
            s = GroupSum<GL>(s);
            if (gl == st) mine = s;
            pp += sd.c;
          }
          if (gl < ns) gm[eb + gl] = V(-2.0 * scale * mine);
        }
        WaveFence();
      }
    }
    WaveFence();
  }
  red = BlockSum<double>(red, red_s);
  if (threadIdx.x == 0) part[blockIdx.x] = red;
}

// df/dA for the A columns k in [k0, k1), contributions of the R block's
// columns [j0, j1): acc(u, k) += sum_l M(k, l) R(u, l) over the entries l of
// row k of S (column k of S^T) inside the block, l ascending.  R(u, l) is
// the block buffer at the slot the plan gives the term (l, k, u), found
// through the term offset of the S entry (k, l).  GL lanes per column of A,
// lane gl holds rows gl + GL r (r < kRegs); FP64 products and sums; blocks
// add in column order, so the result is deterministic.
constexpr int kGradARegs = 4;

// One step of LossGradA: the term offset of the S entry (k, l) in the plan,
// the S2 offset of column l inside the R block, and M(k, l).
template <class V>
struct alignas(8) GradStep {
  int64_t off;
  int rb;
  V m;
};

template <class V, class SlotT, int GL>
__global__ __launch_bounds__(kBlk) void LossGradA(int k0, int k1, int j0, int j1, const int* __restrict__ ap,
                                                  const int* __restrict__ stp, const int* __restrict__ sti,
                                                  const int* __restrict__ stperm, const V* __restrict__ mv,
                                                  const int64_t* __restrict__ plan_ptr,
                                                  const SlotT* __restrict__ plan, const int* __restrict__ poff,
                                                  const int* __restrict__ s2p, int64_t q_base,
                                                  const V* __restrict__ rblk, double* __restrict__ acc_a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int groups = blockDim.x / GL;
  const int g = threadIdx.x / GL, gl = threadIdx.x % GL;
  GradStep<V>* const steps = reinterpret_cast<GradStep<V>*>(lds_raw) + size_t(g) * GL;
  for (int k = k0 + blockIdx.x * groups + g; k < k1; k += gridDim.x * groups) {
    const int a0 = ap[k], c = ap[k + 1] - a0;
    const int x = LowerBound(sti, stp[k], stp[k + 1], j0);
    const int xe = LowerBound(sti, x, stp[k + 1], j1);
    double acc[kGradARegs];
#pragma unroll
    for (int r = 0; r < kGradARegs; ++r) acc[r] = 0.0;
    for (int tb = x; tb < xe; tb += GL) {
      if (tb + gl < xe) {
        const int p = stperm[tb + gl], l = sti[tb + gl];
        steps[gl] = GradStep<V>{plan_ptr[l] + poff[p], int(int64_t(s2p[l]) - q_base), mv[p]};
      }
      WaveFence();
      const int ns = min(GL, xe - tb);
      if (c <= GL) {
        // one row per lane: the slots, then the R values of kLossBatch
        // steps are loaded before their additions (in step order)
        for (int sb = 0; sb < ns; sb += kLossBatch) {
          int at[kLossBatch];
          double m[kLossBatch], rv[kLossBatch];
#pragma unroll
          for (int b = 0; b < kLossBatch; ++b) {
            const int st = sb + b;
            const GradStep<V> sd = steps[st < ns ? st : 0];
            const bool has = st < ns && gl < c;
            m[b] = double(sd.m);
            at[b] = has ? sd.rb + int(plan[sd.off + gl]) : -1;
          }
#pragma unroll
          for (int b = 0; b < kLossBatch; ++b) rv[b] = at[b] >= 0 ? double(rblk[at[b]]) : 0.0;
#pragma unroll
          for (int b = 0; b < kLossBatch; ++b) acc[0] += m[b] * rv[b];
        }
      } else {
        for (int st = 0; st < ns; ++st) {
          const GradStep<V> sd = steps[st];
#pragma unroll
          for (int r = 0; r < kGradARegs; ++r) {
            const int u = gl + GL * r;
            if (u < c) acc[r] += double(sd.m) * double(rblk[sd.rb + int(plan[sd.off + u])]);
          }
        }
      }
      WaveFence();
    }
#pragma unroll
    for (int r = 0; r < kGradARegs; ++r) {
      const int u = gl + GL * r;
      if (u < c) acc_a[a0 + u] += acc[r];
    }
  }
}

// Row range of S over the columns of each block b of block_cols columns:
// the columns k of A that a block's R reaches in LossGradA.  blockIdx.y
// strides over the column blocks, blockIdx.x over chunks of one block's
// columns; one atomic pair per workgroup.
__global__ __launch_bounds__(kBlk) void BlockRowRanges(int n, const int* __restrict__ sp, const int* __restrict__ si,
                                                      int block_cols, int nb, int* __restrict__ kmin,
                                                      int* __restrict__ kmax) {
  __shared__ int s_lo[kBlk / kWave], s_hi[kBlk / kWave];
  for (int b = blockIdx.y; b < nb; b += gridDim.y) {
    const int64_t c0 = int64_t(b) * block_cols, c1 = std::min<int64_t>(n, c0 + block_cols);
    int lo = INT_MAX, hi = -1;
    for (int64_t j = c0 + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < c1;
         j += int64_t(gridDim.x) * blockDim.x) {
      const int e0 = sp[j], e1 = sp[j + 1];
      if (e1 > e0) {  // rows sorted inside a column
        lo = min(lo, si[e0]);
        hi = max(hi, si[e1 - 1]);
      }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      lo = min(lo, __shfl_xor(lo, off, kWave));
      hi = max(hi, __shfl_xor(hi, off, kWave));
    }
    if (LaneId() == 0) {
      s_lo[WaveInBlock()] = lo;
      s_hi[WaveInBlock()] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < int(blockDim.x) / kWave; ++w) {
        lo = min(lo, s_lo[w]);
        hi = max(hi, s_hi[w]);
      }
      if (hi >= 0) {
        atomicMin(&kmin[b], lo);
        atomicMax(&kmax[b], hi);
      }
    }
    __syncthreads();
  }
}

// Deterministic sum of the per-block partials (one block of kSumThreads;
// each thread adds its strided partials in index order, eight loads at a
// time).
constexpr int kSumThreads = 1024;
__global__ __launch_bounds__(kSumThreads) void SumLoss(const double* __restrict__ part, int np,
                                                      double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int i0 = 0; i0 < np; i0 += 8 * kSumThreads) {
    double v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int i = i0 + t * kSumThreads + int(threadIdx.x);
      v[t] = i < np ? part[i] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) s += v[t];
  }
  s = BlockSum<double>(s, red);
  if (threadIdx.x == 0) *out = s;
}

template <class V>
__global__ void ScaleOut(int64_t n, const double* __restrict__ acc, double scale, V* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = V(-2.0 * scale * acc[i]);
}

__global__ void Iota(int64_t n, double* __restrict__ v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = double(i);
}

__global__ void ToInt(int64_t n, const double* __restrict__ v, int* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = int(v[i]);
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
  spai::SlotPlan plan;                 // slots of the terms of A S on S2
  spai::DBuf<int> poff;                // per S entry: its first term inside its column's plan
  spai::DBuf<int> s2d;                 // per column j: the slot of row j in S2(:, j)
  spai::PatternPtr ST;                 // S^T
  spai::DBuf<int> stperm;              // S^T entry -> S entry
  spai::DBuf<char> rblk;               // one column block of R
  spai::DBuf<double> part, acc, loss;
  int64_t max_block_entries = 0;
  int gl = 8, lmax = 1, block_threads = spai::kBlk;
  std::vector<int> blocks;             // column boundaries
  std::vector<int64_t> q_base;         // first S2 entry of each block
  std::vector<int> kmin, kmax;         // rows of S in each block's columns
  std::vector<unsigned> grid;          // column-kernel blocks of each column block
  std::vector<int> part_off;           // their partials' offsets
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

template <class V, class SlotT, int GL>
void LaunchBlock(SpaiLossPlan* p, size_t b, const V* m, const V* a, bool backward, double scale, V* gm) {
  using namespace spai;
  Ctx& c = *p->ctx;
  const int j0 = p->blocks[b], j1 = p->blocks[b + 1];
  const SlotT* plan = std::is_same<SlotT, uint8_t>::value ? reinterpret_cast<const SlotT*>(p->plan.plan8.get())
                                                          : reinterpret_cast<const SlotT*>(p->plan.plan16.get());
  V* rblk = reinterpret_cast<V*>(p->rblk.get());
  const size_t lds = size_t(p->block_threads / GL) * (GL * sizeof(LossStep<V>) + size_t(p->lmax) * sizeof(V));
  double* part = p->part.get() + p->part_off[b];
#define SPAI_LOSS_COLUMNS(BACK, NARROW)                                                                       \
  LossColumns<V, SlotT, GL, BACK, NARROW><<<p->grid[b], p->block_threads, lds, c.stream>>>(                     \
      j0, j1, p->S.ptr(), p->S.idx(), m, p->A.ptr(), a, p->plan.ptr.get(), plan, p->S2.ptr(), p->s2d.get(), \
      p->q_base[b], p->lmax, rblk, scale, gm, part)
  const bool narrow = p->plan.max_a_col <= GL;
  if (backward) {
    if (narrow)
      SPAI_LOSS_COLUMNS(true, true);
    else
      SPAI_LOSS_COLUMNS(true, false);
  } else {
    if (narrow)
      SPAI_LOSS_COLUMNS(false, true);
    else
      SPAI_LOSS_COLUMNS(false, false);
  }
#undef SPAI_LOSS_COLUMNS
  if (backward && p->kmax[b] >= p->kmin[b]) {
    const int k0 = p->kmin[b], k1 = p->kmax[b] + 1;
    const int groups = p->block_threads / GL;
    const unsigned ga = (unsigned)std::max<int64_t>(1, std::min<int64_t>((k1 - k0 + groups - 1) / groups, 8192));
    const size_t lds_a = size_t(p->block_threads) * sizeof(GradStep<V>);
    LossGradA<V, SlotT, GL><<<ga, p->block_threads, lds_a, c.stream>>>(
        k0, k1, j0, j1, p->A.ptr(), p->ST->ptr.get(), p->ST->idx.get(), p->stperm.get(), m, p->plan.ptr.get(), plan,
        p->poff.get(), p->S2.ptr(), p->q_base[b], rblk, p->acc.get());
  }
}

template <class V, class SlotT>
void LaunchBlockGl(SpaiLossPlan* p, size_t b, const V* m, const V* a, bool backward, double scale, V* gm) {
  switch (p->gl) {
    case 8: LaunchBlock<V, SlotT, 8>(p, b, m, a, backward, scale, gm); break;
    case 16: LaunchBlock<V, SlotT, 16>(p, b, m, a, backward, scale, gm); break;
    case 32: LaunchBlock<V, SlotT, 32>(p, b, m, a, backward, scale, gm); break;
    default: LaunchBlock<V, SlotT, 64>(p, b, m, a, backward, scale, gm); break;
  }
}

// Forward (loss to *loss) or forward and backward (gradients to gm, ga),
// blocks of R in column order on the plan's stream.
template <class V>
int64_t Forward(SpaiLossPlan* p, const V* m, const V* a, double* loss, bool backward, double scale, V* gm, V* ga) {
  using namespace spai;
  Ctx& c = *p->ctx;
  hipStream_t st = c.stream;
  if (backward) SPAI_HIP_CHECK(hipMemsetAsync(p->acc.get(), 0, sizeof(double) * p->A.nnz(), st));
  for (size_t b = 0; b + 1 < p->blocks.size(); ++b) {
    if (p->plan.slot16)
      LaunchBlockGl<V, uint16_t>(p, b, m, a, backward, scale, gm);
    else
      LaunchBlockGl<V, uint8_t>(p, b, m, a, backward, scale, gm);
  }
  if (backward) ScaleOut<V><<<GridAll(p->A.nnz()), kBlk, 0, st>>>(p->A.nnz(), p->acc.get(), scale, ga);
  SumLoss<<<1, kSumThreads, 0, st>>>(p->part.get(), p->part_off.back(), p->loss.get());
  SPAI_HIP_CHECK(hipGetLastError());
  double total = 0.0;
  c.ReadBack(&total, p->loss.get(), sizeof(double));
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
    p->plan = BuildSlotPlan(c, p->n, p->S.ptr(), p->S.idx(), p->A.ptr(), p->A.idx(), p->S2.ptr(), p->S2.idx());
    if (!p->plan.built) throw std::invalid_argument("the differentiation form needs columns of A M of at most 65535 entries");
    if (p->plan.max_a_col > 64 * kGradARegs)
      throw std::invalid_argument("the differentiation form needs columns of A of at most 256 entries");
    p->gl = p->plan.max_a_col <= 8 ? 8 : p->plan.max_a_col <= 16 ? 16 : p->plan.max_a_col <= 32 ? 32 : 64;
    p->lmax = std::max(p->plan.max_s2_col, 1);
    // step tables and R columns in LDS: fewer lanes per block when 256
    // would not fit
    const size_t vsize = p->fp32 ? 4 : 8;
    const size_t per_group = size_t(p->gl) * (p->fp32 ? sizeof(LossStep<float>) : sizeof(LossStep<double>)) +
                             size_t(p->lmax) * vsize;
    p->block_threads = kBlk;
    while (p->block_threads > p->gl && size_t(p->block_threads / p->gl) * per_group > 48 * 1024)
      p->block_threads /= 2;
    if (size_t(p->block_threads / p->gl) * per_group > 64 * 1024)
      throw std::invalid_argument("the differentiation form needs columns of A M within 64 KiB of LDS");
    p->poff = c.Alloc<int>(std::max<int64_t>(p->S.nnz(), 1));
    TermOffsets<<<GridAll(p->n), kBlk, 0, c.stream>>>((int)p->n, p->S.ptr(), p->S.idx(), p->A.ptr(), p->poff.get());
    p->s2d = c.Alloc<int>(size_t(p->n));
    DiagSlots<<<GridAll(p->n), kBlk, 0, c.stream>>>((int)p->n, p->S2.ptr(), p->S2.idx(), p->s2d.get());
    p->ST = TransposeWithPerm(c, p->S, p->stperm);
    for (int64_t j = 0; j < p->n; j += p->block_cols) p->blocks.push_back((int)j);
    p->blocks.push_back((int)p->n);
    const size_t nb = p->blocks.size() - 1;
    // per block: S2 offsets, the A columns its R reaches, the column grid
    std::vector<int> s2p(p->n + 1);
    SPAI_HIP_CHECK(hipMemcpyAsync(s2p.data(), p->S2.ptr(), sizeof(int) * (p->n + 1), hipMemcpyDeviceToHost, c.stream));
    auto kr = c.Alloc<int>(2 * nb);
    {
      std::vector<int> init(2 * nb);
      for (size_t b = 0; b < nb; ++b) {
        init[b] = INT_MAX;
        init[nb + b] = -1;
      }
      SPAI_HIP_CHECK(hipMemcpyAsync(kr.get(), init.data(), sizeof(int) * 2 * nb, hipMemcpyHostToDevice, c.stream));
      const dim3 grr((unsigned)std::min<int64_t>((p->block_cols + kBlk - 1) / kBlk, 16),
                     (unsigned)std::min<size_t>(nb, 65535));
      BlockRowRanges<<<grr, kBlk, 0, c.stream>>>((int)p->n, p->S.ptr(), p->S.idx(), (int)p->block_cols, (int)nb,
                                                 kr.get(), kr.get() + nb);
      std::vector<int> h(2 * nb);
      c.ReadBack(h.data(), kr.get(), sizeof(int) * 2 * nb);
      p->kmin.assign(h.begin(), h.begin() + nb);
      p->kmax.assign(h.begin() + nb, h.end());
    }
    const int groups = p->block_threads / p->gl;
    int total_parts = 0;
    for (size_t b = 0; b < nb; ++b) {
      p->q_base.push_back(s2p[p->blocks[b]]);
      p->max_block_entries = std::max<int64_t>(p->max_block_entries, s2p[p->blocks[b + 1]] - s2p[p->blocks[b]]);
      const int64_t cols = p->blocks[b + 1] - p->blocks[b];
      p->grid.push_back((unsigned)std::max<int64_t>(1, std::min<int64_t>((cols + groups - 1) / groups, 2048)));
      p->part_off.push_back(total_parts);
      total_parts += (int)p->grid.back();
    }
    p->part_off.push_back(total_parts);
    p->rblk = c.Alloc<char>(std::max<int64_t>(p->max_block_entries, 1) * vsize);
    p->part = c.Alloc<double>(std::max(total_parts, 1));
    p->loss = c.Alloc<double>(1);
    p->acc = c.Alloc<double>(std::max<int64_t>(p->A.nnz(), 1));
    c.Sync();
    return p.release();
  } catch (const std::exception& ex) {
    g_loss_error = ex.what();
    return nullptr;
  }
}

int64_t spai_loss_set_stream(SpaiLossPlan* p, uint64_t stream) {
  if (!p) return LossFail("null plan");
  p->ctx->stream = reinterpret_cast<hipStream_t>(stream);
  return 0;
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
