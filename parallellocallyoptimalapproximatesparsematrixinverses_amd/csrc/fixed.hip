// Fixed-pattern Global SPAI builders (CG, MR, LOMR) on the device, FP64 or
// FP32-bulk.
//
// M, Z and P live on S = pattern(A) | pattern(M0) | I; the residual R and the
// products A Z, A P live on S2 = pattern(A S) | S | I.  Both patterns are fixed
// for the whole build, so every kernel is a flat pass over S2 (or S) and the
// iteration runs without host synchronisation: step sizes, the stop flag and
// the residual history stay in device memory.
//
// Products are inner products over k ascending -- (A Y)(i, j) =
// sum_k A^T(k, i) Y(k, j) merged from column i of A^T and column j of Y -- which
// is the same summation order as the reference's column-wise Gustavson
// product (state_manager.cpp:60).  Reductions (trace inner products, ||R||_F)
// accumulate in FP64 in both forms.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "dev_common.hpp"
#include "fixed.hpp"
#include "slot_plan.hpp"
#include "ls2x2.hpp"
#include "ops.hpp"

namespace spai {

namespace {

constexpr int kBlk = 256;
constexpr int kMaxGrid = 2048;
constexpr int kMaxFusedGrid = 2048;

// Scalar slots in device memory.
enum Slot : int {
  kDelta = 0,
  kGamma,
  kBeta,
  kAlpha,
  kRz,
  kPrevRz,
  kNumSlots
};

struct Flags {
  int stop;        // 0 running, 1 breakdown, 2 converged
  int iterations;  // history rows appended - 1
};

// (A Y)(i, j) = sum_k A^T(k, i) Y(k, j) over k ascending (the reference
// product's order).  Row i of A (column i of A^T) is read in groups of 8;
// each of its k is located in the sorted column Y(:, j) by an independent
// binary search, so the 8 searches of a group overlap in flight; the products
// are then added in k order.
constexpr int kIlp = 8;

template <class V, int NY>
__device__ __forceinline__ void Gather(const int* __restrict__ ati, const V* __restrict__ atv, int ab, int ae,
                                       const int* __restrict__ si, const V* __restrict__ y0,
                                       const V* __restrict__ y1, int sb, int se, V& s0, V& s1) {
  s0 = V(0);
  s1 = V(0);
  for (int t0 = ab; t0 < ae; t0 += kIlp) {
    int pos[kIlp];
    V a[kIlp];
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      const int t = t0 + u;
      pos[u] = -1;
      if (t < ae) {
        const int k = ati[t];
        a[u] = atv[t];
        int lo = sb, hi = se;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (si[mid] < k)
            lo = mid + 1;
          else
            hi = mid;
        }
        if (lo < se && si[lo] == k) pos[u] = lo;
      }
    }
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      if (pos[u] >= 0) {
        s0 += a[u] * y0[pos[u]];
        if (NY > 1) s1 += a[u] * y1[pos[u]];
      }
    }
  }
}

template <class V>
__device__ __forceinline__ V InnerProduct(const int* __restrict__ ai, const V* __restrict__ av, int ab, int ae,
                                          const int* __restrict__ si, const V* __restrict__ y, int sb, int se) {
  V s0, s1;
  Gather<V, 1>(ai, av, ab, ae, si, y, y, sb, se, s0, s1);
  return s0;
}

template <class V>
__device__ __forceinline__ void InnerProduct2(const int* __restrict__ ai, const V* __restrict__ av, int ab, int ae,
                                              const int* __restrict__ si, const V* __restrict__ y0,
                                              const V* __restrict__ y1, int sb, int se, V& s0, V& s1) {
  Gather<V, 2>(ai, av, ab, ae, si, y0, y1, sb, se, s0, s1);
}

// The block's sums of K quantities into part[k * gridDim.x + blockIdx.x]:
// BlockSum's order for each (wavefront sums, then the wavefronts in order),
// all K reduced with one barrier; thread k stores quantity k.  Called once,
// at the end of a kernel.
template <int K>
__device__ __forceinline__ void StorePartials(double (&acc)[K], double* __restrict__ part) {
  static_assert(K <= kBlk, "one thread per quantity");
  __shared__ double red[K][16];
  const int w = WaveInBlock(), nw = (int(blockDim.x) + kWave - 1) / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double v = WaveSum(acc[k]);
    if (LaneId() == 0) red[k][w] = v;
  }
  __syncthreads();
  if (int(threadIdx.x) < K) {
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += red[threadIdx.x][i];
    part[int(threadIdx.x) * gridDim.x + blockIdx.x] = t;
  }
}

// Deterministic final sum of K quantities over G block partials (1 block).
template <int K>
__device__ __forceinline__ void SumPartialsBlock(const double* __restrict__ part, int G, double (&out)[K]) {
  __shared__ double red[16];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < G; b += blockDim.x) s += part[k * G + b];
    out[k] = BlockSum<double>(s, red);
  }
}

// As SumPartialsBlock, with all K sums read in one pass and reduced together
// (the same per-thread and per-wave order for each k).
template <int K>
__device__ __forceinline__ void SumPartialsBlockFused(const double* __restrict__ part, int G, double (&out)[K]) {
  __shared__ double red[K][16];
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += part[k * G + b];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = WaveSum(v[k]);
  const int w = WaveInBlock(), nw = (blockDim.x + kWave - 1) / kWave;
  if (LaneId() == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[k][w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += red[k][i];
    out[k] = t;
  }
}

// ------------------------------------------------------------ slot product
// (A Y)(:, j) on S2(:, j) by column-wise Gustavson: for each S(:, j) entry
// (k, y) in order -- k ascending, the reference product's order -- the group's
// lanes stream column A(:, k) and add a * y into an LDS accumulator.  The
// accumulator slot of every contribution (its row's position in S2(:, j)) is
// precomputed once for the fixed pattern ("plan", one uint8/uint16 per product
// term, laid out in the order the terms are visited), so the iteration does no
// searching: A's columns, the plan and Y are read contiguously.  Rows inside
// one A(:, k) are distinct, so lanes never collide within a step.  With every
// A column at most GL long (narrow) and GL <= 16, a lane holds at most one
// term per step: the slots and values of all GL steps are loaded first and
// the additions follow in step order.
template <class V, class SlotT, int GL, int NY, class Epi>
__global__ __launch_bounds__(kBlk) void SlotProduct(int n, const int* __restrict__ sp, const int* __restrict__ si,
                                                    const V* __restrict__ y0, const V* __restrict__ y1,
                                                    const int* __restrict__ ap, const V* __restrict__ av,
                                                    const int64_t* __restrict__ plan_ptr,
                                                    const SlotT* __restrict__ plan, const int* __restrict__ s2p,
                                                    const int* __restrict__ s2i, int lmax, Epi epi,
                                                    const double* __restrict__ sc, double* __restrict__ part,
                                                    const Flags* __restrict__ flags, int narrow) {
  if (flags->stop) return;
  epi.Load(sc);  // epilogues that use the step sizes read them here
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int kGroups = kBlk / GL;
  const int g = threadIdx.x / GL, gl = threadIdx.x % GL;
  V* acc = reinterpret_cast<V*>(smem_raw) + size_t(g) * NY * lmax;
  double red[Epi::K];
#pragma unroll
  for (int k = 0; k < Epi::K; ++k) red[k] = 0.0;
  for (int j = blockIdx.x * kGroups + g; j < n; j += gridDim.x * kGroups) {
    const int q0 = s2p[j], L = s2p[j + 1] - q0;
    for (int t = gl; t < L; t += GL) {
      acc[t] = V(0);
      if (NY > 1) acc[lmax + t] = V(0);
    }
    int64_t pp = plan_ptr[j];
    const int e1 = sp[j + 1];
    for (int eb = sp[j]; eb < e1; eb += GL) {
      // each lane fetches one S entry of the column; the steps broadcast them
      int my_a0 = 0, my_c = 0;
      V my_ya = V(0), my_yb = V(0);
      if (eb + gl < e1) {
        const int k = si[eb + gl];
        my_a0 = ap[k];
        my_c = ap[k + 1] - my_a0;
        my_ya = y0[eb + gl];
        if (NY > 1) my_yb = y1[eb + gl];
      }
      const int ns = min(GL, e1 - eb);
      if constexpr (GL <= 16) {
        if (narrow) {
          int slot_r[GL];
          V a_r[GL], ya_r[GL], yb_r[GL];
          bool has_r[GL];
#pragma unroll
          for (int st = 0; st < GL; ++st) {
            const int a0 = __shfl(my_a0, st, GL);
            const int c = __shfl(my_c, st, GL);
            ya_r[st] = __shfl(my_ya, st, GL);
            yb_r[st] = NY > 1 ? __shfl(my_yb, st, GL) : V(0);
            has_r[st] = st < ns && gl < c;
            slot_r[st] = has_r[st] ? int(plan[pp + gl]) : 0;
            a_r[st] = has_r[st] ? av[a0 + gl] : V(0);
            if (st < ns) pp += c;
          }
#pragma unroll
          for (int st = 0; st < GL; ++st) {
            if (has_r[st]) {
              acc[slot_r[st]] += a_r[st] * ya_r[st];
              if (NY > 1) acc[lmax + slot_r[st]] += a_r[st] * yb_r[st];
            }
          }
          continue;
        }
      }
      for (int st = 0; st < ns; ++st) {
        const int a0 = __shfl(my_a0, st, GL);
        const int c = __shfl(my_c, st, GL);
        const V ya = __shfl(my_ya, st, GL);
        const V yb = NY > 1 ? __shfl(my_yb, st, GL) : V(0);
        // LDS operations of one wavefront complete in issue order, so step st+1
        // observes step st's sums without a fence.
        for (int u = gl; u < c; u += GL) {
          const int slot = plan[pp + u];
          const V a = av[a0 + u];
          acc[slot] += a * ya;
          if (NY > 1) acc[lmax + slot] += a * yb;
        }
        pp += c;
      }
    }
    for (int t = gl; t < L; t += GL) {
      const int64_t q = q0 + t;
      epi(q, s2i[q], j, acc[t], NY > 1 ? acc[lmax + t] : V(0), red);
    }
  }
  StorePartials<Epi::K>(red, part);
}

// ------------------------------------------------------- fused LOMR / MR pass
// One kernel per LOMR iteration (locally_minimal_residual.cpp:57-114 on the
// fixed pattern); MR (minimal_residual.cpp:56-92) takes the same pass with
// M' = M + alpha Z, R' = I - A M' recomputed and one direction.  What an iteration does to column j -- A [Z P](:, j), the
// updates of R(:, j), M(:, j), P(:, j), Z(:, j), and the products
// A [Z' P'](:, j) whose traces give the next step sizes -- reads column j
// alone, so a wavefront takes W = 64 / GL consecutive columns (GL lanes each),
// keeps their accumulators and new directions in LDS and touches R, M, P and
// Z once per iteration.  Only the five trace products and ||R||^2 cross
// columns: each block stores its partial sums and a one-block kernel
// (LomrFusedStep) sums them in a fixed order, records the history row and
// the stop decision and turns the traces into the next step sizes.  The
// R/S-sized reads and writes of a tile run over the tile's contiguous range
// with all 64 lanes, so stores cover whole lines.

// Wavefront-local LDS of the fused pass (byte offsets from the wavefront's
// region), every array 16-byte aligned:
//   acc    V2[W * lmax]  the products on the tile's S2 range (two sums side by
//                        side), indexed by S2 position minus the tile's first;
//   rr     V [W * lmax]  R' on the tile's S2 range, same indexing;
//   yzp    V2[W * smax]  (Z', P') on the tile's S range, indexed likewise;
//   stage  step operands of a product, step-major ([GL][W] int2, then [GL][W]
//          V2), so the W columns of one step read W distinct banks;
//   stash  FusedTerm[smax][64]: with the stash on, product 1 records every
//          lane's (A value, slot) of every step and product 2 reads them back
//          instead of loading the plan and A again.  Product 2 then needs no
//          staging, so the staging overlays rr and yzp (unused in product 1).
__host__ __device__ constexpr int Align16(int b) { return (b + 15) / 16 * 16; }

struct FusedLayout {
  int acc, rr, yzp, stage, stage_y, stash, bytes;
};

__host__ __device__ constexpr FusedLayout MakeFusedLayout(int lmax, int smax, int gl, int vsize, bool stash) {
  const int w = kWave / gl;
  // S2 positions count from the tile's first entry rounded down to 16 bytes,
  // and the update loops cover whole 16-byte chunks: 2 chunks of slack
  const int lq = w * lmax + 2 * (16 / vsize);
  const int rr = Align16(2 * lq * vsize);
  const int yzp = rr + Align16(lq * vsize);
  const int free_end = yzp + Align16(2 * w * smax * vsize);
  const int stage = stash ? rr : free_end;
  const int stage_y = stage + Align16(kWave * 8);
  const int stage_end = stage_y + Align16(kWave * 2 * vsize);
  const int stash_at = stage_end > free_end ? stage_end : free_end;
  const int bytes = stash ? stash_at + Align16(smax * kWave * 2 * vsize) : stage_end;
  return FusedLayout{0, rr, yzp, stage, stage_y, stash_at, bytes};
}

template <class V>
struct alignas(2 * sizeof(V)) V2 {
  V a, b;
};

// 16 bytes of consecutive S2 entries (values, or pairs of sums): the unit of
// the update and trace loops, one 16-byte access per array.
template <class V>
struct alignas(16) VecQ {
  V v[16 / sizeof(V)];
};
template <class V>
struct alignas(16) VecQ2 {
  V2<V> v[16 / sizeof(V)];
};

// The lane's terms of the last step batch of product 1 (bit s of act: the
// lane has a term in step s of the batch).  A column of at most kFusedPre
// steps has one batch, so product 2 takes its terms from these registers.
constexpr int kFusedPre = 8;
template <class V, class SlotT>
struct FusedTerms {
  SlotT sl[kFusedPre];
  V a[kFusedPre];
  uint32_t act;
};

// One term of product 1 as product 2 reuses it; slot -1: the lane had no term.
template <class V>
struct alignas(2 * sizeof(V)) FusedTerm {
  V a;
  int slot;
};

// acc[slot] += a * (y.a, y.b) (NY == 1: the first sum only), as one LDS read
// and one write per sum pair.
template <class V, int NY>
__device__ __forceinline__ void AddTerm(V2<V>* acc, int slot, V a, V2<V> y) {
#pragma clang fp contract(fast)
  if constexpr (NY > 1) {
    V2<V> t = acc[slot];
    t.a += a * y.a;
    t.b += a * y.b;
    acc[slot] = t;
  } else {
    acc[slot].a += a * y.a;
  }
}

// Stages S entries [eb, eb + GL) of a column for FusedProduct1: lane gl
// writes entry eb + gl's (A(:, k) start, length) and Y values.
template <class V, int GL, class YGet>
__device__ __forceinline__ void FusedStage(int g, int gl, int eb, int e1, const int2* __restrict__ a0c, int2* st_ac,
                                           V2<V>* st_y, YGet yget) {
  constexpr int W = kWave / GL;
  if (eb + gl < e1) {
    V ya = V(0), yb = V(0);
    yget(eb + gl, ya, yb);
    st_ac[gl * W + g] = a0c[eb + gl];
    st_y[gl * W + g] = V2<V>{ya, yb};
  }
}

// Column-wise Gustavson accumulation of A [y0 y1](:, j) into acc (column j's
// S2 range), in the order of SlotProduct (S entries of the column in order, k
// ascending).  YGet(e, ya, yb) returns S entry e's values.  The group stages
// up to GL S entries -- (A(:, k) start, length) and (ya, yb) -- in LDS (the
// caller stages the first GL), then every step (one S entry, column k of A)
// reads them with one broadcast LDS load; when |A(:, k)| <= GL lane gl takes
// the term u = gl, and the plan slots and A values of up to kPre steps are
// loaded before any of them is added, so a column waits for one load latency
// per kPre steps.  The column is S entries [e0, e1) with plan terms from pp
// on.  kStash and stash_write: every lane's (A value, slot) of every step
// goes to stash[step][lane] (no column of A is longer than GL then).  `last`
// receives the lane's terms of the last batch.
template <class V, class SlotT, int GL, int NY, bool kStash, class YGet>
__device__ __forceinline__ void FusedProduct1(int g, int gl, int e0, int e1, uint32_t pp,
                                              const int2* __restrict__ a0c, const V* __restrict__ av,
                                              const SlotT* __restrict__ plan, V2<V>* acc, int2* st_ac,
                                              V2<V>* st_y, FusedTerm<V>* stash, bool stash_write,
                                              FusedTerms<V, SlotT>& last, YGet yget) {
  constexpr int W = kWave / GL;
  constexpr int kPre = kFusedPre;
  const int lane = g * GL + gl;
  for (int eb = e0; eb < e1; eb += GL) {
    __builtin_amdgcn_wave_barrier();
    const int ns = min(GL, e1 - eb);
    for (int s0 = 0; s0 < ns; s0 += kPre) {
      SlotT sl[kPre];
      V a[kPre];
      uint32_t act = 0;  // bit s: the lane has a term in step s0 + s
      uint32_t off = pp;
      bool wide = false;
#pragma unroll
      for (int s = 0; s < kPre; ++s) {
        sl[s] = SlotT(0);
        a[s] = V(0);
        if (s0 + s < ns) {
          const int2 ac = st_ac[(s0 + s) * W + g];
          if constexpr (!kStash) wide = wide || ac.y > GL;
          if (gl < ac.y) {
            act |= 1u << s;
            sl[s] = plan[off + uint32_t(gl)];
            a[s] = av[uint32_t(ac.x + gl)];
          }
          off += uint32_t(ac.y);
        }
      }
      if (!wide) {
#pragma unroll
        for (int s = 0; s < kPre; ++s) {
          if (s0 + s < ns) {
            const bool act_s = (act >> s) & 1u;
            if constexpr (kStash) {
              if (stash_write) stash[(eb - e0 + s0 + s) * kWave + lane] = FusedTerm<V>{a[s], act_s ? int(sl[s]) : -1};
            }
            if (act_s) AddTerm<V, NY>(acc, int(sl[s]), a[s], st_y[(s0 + s) * W + g]);
          }
        }
#pragma unroll
        for (int s = 0; s < kPre; ++s) {
          last.sl[s] = sl[s];
          last.a[s] = a[s];
        }
        last.act = act;
      } else {
        // a column of A longer than GL: lanes stride through each step
        uint32_t o2 = pp;
        for (int s = 0; s < min(kPre, ns - s0); ++s) {
          const int2 ac = st_ac[(s0 + s) * W + g];
          const V2<V> y = st_y[(s0 + s) * W + g];
          for (int u = gl; u < ac.y; u += GL) AddTerm<V, NY>(acc, int(plan[o2 + uint32_t(u)]), av[uint32_t(ac.x + u)], y);
          o2 += uint32_t(ac.y);
        }
      }
      pp = off;
    }
    __builtin_amdgcn_wave_barrier();
    if (eb + GL < e1) FusedStage<V, GL>(g, gl, eb + GL, e1, a0c, st_ac, st_y, yget);
  }
}

// Product 2 from the stash: steps in the order of product 1, the step's Y
// values (Z', P') of S entry s of the column read from LDS.
template <class V, int NY>
__device__ __forceinline__ void FusedProduct2Stash(int ns, int lane, const FusedTerm<V>* stash, const V2<V>* y,
                                                   V2<V>* acc) {
  constexpr int kPre = 4;
  for (int s0 = 0; s0 < ns; s0 += kPre) {
    FusedTerm<V> t[kPre];
    V2<V> yv[kPre];
#pragma unroll
    for (int s = 0; s < kPre; ++s) {
      t[s] = FusedTerm<V>{V(0), -1};
      yv[s] = V2<V>{V(0), V(0)};
      if (s0 + s < ns) {
        t[s] = stash[(s0 + s) * kWave + lane];
        yv[s] = y[s0 + s];
      }
    }
#pragma unroll
    for (int s = 0; s < kPre; ++s)
      if (t[s].slot >= 0) AddTerm<V, NY>(acc, t[s].slot, t[s].a, yv[s]);
  }
}

// kInit: M = M0 already on S; R = I - A M, Z = Pr R on S, P = 0, and the
// traces of A Z (the pass before iteration 1).  Otherwise one iteration with
// the step sizes in sc: update, then the traces of the next iteration.
// stq[p] (S entry p of column j): position of p's row in S2(:, j), counted
// from the first S2 entry of j's tile rounded down to a multiple of 16 bytes
// (kVec entries).  s2d[j]: position of the diagonal in S2(:, j).  R and prS2
// hold kVec - 1 entries of padding past the end, so whole chunks load.
// kTerms: how product 2 gets product 1's terms.  kTermsReload: from
// registers for columns of at most kFusedPre steps (narrow), else re-staged
// from a0c with the plan and A values loaded again; kTermsStash: from
// registers or the LDS stash; kTermsRegs: every column has at most kFusedPre
// steps and every column of A at most GL entries, so always from registers
// (no stash in LDS, no re-staging code).
constexpr int kTermsReload = 0, kTermsStash = 1, kTermsRegs = 2;
template <class V, class SlotT, int GL, bool kInit, bool kMr, int kTerms>
__global__ __launch_bounds__(kBlk, sizeof(V) == 4 ? 5 : 4) void FusedPass(
    int n, const int* __restrict__ sp, const int2* __restrict__ a0c, const uint16_t* __restrict__ stq,
    const int* __restrict__ s2d, const int* __restrict__ s2p, const V* __restrict__ av,
    const int64_t* __restrict__ plan_ptr, const SlotT* __restrict__ plan, int lmax, int smax,
    const V* __restrict__ prS, const V* __restrict__ prS2, V* __restrict__ R, V* __restrict__ M, V* __restrict__ Z,
    V* __restrict__ P, const double* __restrict__ sc, double* __restrict__ part_out,
    const Flags* __restrict__ flags, int narrow) {
  // Fused multiply-adds here (the file is built without contraction): the
  // fixed-pattern builders are held to the oracle within a tolerance, not
  // bit for bit, and the FP64 trace sums and the updates take half the
  // instructions.
#pragma clang fp contract(fast)
  constexpr int W = kWave / GL;
  constexpr int kVec = 16 / int(sizeof(V));  // S2 entries per chunk
  constexpr int RQ = sizeof(V) == 4 ? 1 : 2;  // S2 chunks of a tile held in registers per lane
  constexpr int RS = 2;                        // S entries of a tile held in registers per lane
  constexpr bool kStash = kTerms == kTermsStash;
  constexpr bool kRegsOnly = kTerms == kTermsRegs;
  if (flags->stop) return;
  // LOMR: R -= delta AZ + gamma AP (R read); MR: R = I - A M recomputed
  constexpr bool kLomrIter = !kInit && !kMr;
  constexpr int NY = kLomrIter ? 2 : 1;
  V d = V(0), gm = V(0), be = V(0), nd = V(0), ng = V(0), alpha = V(0);
  if constexpr (kLomrIter) {
    d = V(sc[kDelta]);
    gm = V(sc[kGamma]);
    be = V(sc[kBeta]);
    nd = V(-sc[kDelta]);
    ng = V(-sc[kGamma]);
  }
  if constexpr (kMr && !kInit) alpha = V(sc[kAlpha]);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = LaneId(), wib = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int g = lane / GL, gl = lane % GL;
  const FusedLayout lay = MakeFusedLayout(lmax, smax, GL, int(sizeof(V)), kStash);
  unsigned char* const wave_lds = smem_raw + size_t(wib) * size_t(lay.bytes);
  V2<V>* const acc = reinterpret_cast<V2<V>*>(wave_lds + lay.acc);
  V* const rr = reinterpret_cast<V*>(wave_lds + lay.rr);
  V2<V>* const yzp = reinterpret_cast<V2<V>*>(wave_lds + lay.yzp);
  int2* const st_ac = reinterpret_cast<int2*>(wave_lds + lay.stage);
  V2<V>* const st_y = reinterpret_cast<V2<V>*>(wave_lds + lay.stage_y);
  FusedTerm<V>* const stash = reinterpret_cast<FusedTerm<V>*>(wave_lds + lay.stash);
  double red[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // Tiles: when the grid is a multiple of 8, the blocks of each XCD (blocks
  // b with b mod 8 = x, as workgroups are dealt to the XCDs round robin)
  // walk one contiguous eighth of the tiles, so that the columns of A and
  // the operand entries that neighbouring tiles share stay in one L2.
  // ntiles is the end of this block's tile range.
  constexpr int kWpb = kBlk / kWave;
  const int ntiles_all = (n + W - 1) / W;
  int ntiles = ntiles_all, wstride = int(gridDim.x) * kWpb, tile_first = int(blockIdx.x) * kWpb + wib;
  if ((gridDim.x & 7u) == 0) {
    const int x = int(blockIdx.x & 7u);
    const int tile_begin = int(int64_t(ntiles_all) * x / 8);
    ntiles = int(int64_t(ntiles_all) * (x + 1) / 8);
    wstride = int(gridDim.x >> 3) * kWpb;
    tile_first = tile_begin + int(blockIdx.x >> 3) * kWpb + wib;
  }
  // Product 1's Y values of S entry e
  auto yget1 = [&](int e, V& ya, V& yb) {
    if constexpr (kInit) {
      ya = M[e];
    } else if constexpr (kMr) {
      ya = M[e] + alpha * Z[e];
    } else {
      ya = Z[e];
      yb = P[e];
    }
  };
  // What a tile needs before its loads can start -- the tile's S2 and S
  // bounds and, per group, the column's S range, first plan term, first S2
  // entry and diagonal -- and the lane's operands of product 1's first step
  // chunk are loaded one tile ahead, so they arrive during the previous
  // tile's updates and product 2.
  int nqa = 0, nqe = 0, npa = 0, npe = 0, ns2j = 0, ne0 = 0, ne1 = 0, ndj = 0;
  uint32_t npp = 0;
  auto load_tile = [&](int t) {
    const int j0 = t * W, jj = j0 + g, je = min(j0 + W, n);
    nqa = s2p[j0];
    nqe = s2p[je];
    npa = sp[j0];
    npe = sp[je];
    const bool ok = jj < n;
    ns2j = ok ? s2p[jj] : 0;
    ne0 = ok ? sp[jj] : 0;
    ne1 = ok ? sp[jj + 1] : 0;
    npp = ok ? uint32_t(plan_ptr[jj]) : 0u;
    ndj = (!kLomrIter && ok && gl == 0) ? s2d[jj] : 0;
  };
  int2 nac = make_int2(0, 0);
  V2<V> ny = V2<V>{V(0), V(0)};
  auto load_stage = [&]() {
    nac = make_int2(0, 0);
    ny = V2<V>{V(0), V(0)};
    if (ne0 + gl < ne1) {
      nac = a0c[ne0 + gl];
      yget1(ne0 + gl, ny.a, ny.b);
    }
  };
  int tile = tile_first;
  if (tile < ntiles) {
    load_tile(tile);
    load_stage();
  }
  for (; tile < ntiles; tile += wstride) {
    const int j0 = tile * W;
    const int j = j0 + g;
    const bool valid = j < n;
    const int qa = __builtin_amdgcn_readfirstlane(nqa), qe = __builtin_amdgcn_readfirstlane(nqe);
    const int pa = __builtin_amdgcn_readfirstlane(npa), pe = __builtin_amdgcn_readfirstlane(npe);
    const int qal = qa & ~(kVec - 1);
    const int cb2 = ns2j - qal;  // column j's S2 range in acc / rr
    const int cb = ne0 - pa;     // column j's S range in yzp
    const int e0 = ne0, e1 = ne1, dj = ndj;
    const uint32_t pp = npp;
    const int2 pre_ac = nac;
    const V2<V> pre_y = ny;
    // operands of the updates, issued before the product so their latency
    // overlaps it: R and Pr(i) on the lane's S2 chunks, M, Z, P, Pr(i) and
    // the tile's S2 position of the row on its S entries
    VecQ<V> rq[RQ], pq[RQ];
#pragma unroll
    for (int r = 0; r < RQ; ++r) {
      const int qb = qal + (lane + r * kWave) * kVec;
      rq[r] = VecQ<V>{};
      pq[r] = VecQ<V>{};
      if (qb < qe) {
        if constexpr (kLomrIter) rq[r] = *reinterpret_cast<const VecQ<V>*>(R + qb);
        pq[r] = *reinterpret_cast<const VecQ<V>*>(prS2 + qb);
      }
    }
    V mp[RS], zp[RS], ppv[RS], ps[RS];
    int tp[RS];
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int p = pa + lane + r * kWave;
      const bool in = p < pe;
      mp[r] = in && !kInit ? M[p] : V(0);
      zp[r] = in && !kInit ? Z[p] : V(0);
      ppv[r] = in && kLomrIter ? P[p] : V(0);
      ps[r] = in ? prS[p] : V(0);
      tp[r] = in ? int(stq[p]) : 0;
    }
    const int ncs = e1 - e0;
    // product 2 takes a one-batch column's terms from product 1's registers
    // (FP64 LOMR keeps 16 doubles of direction state already; holding the
    // terms as well spills, so it reloads or stashes them)
    constexpr bool kRegTerms = !(sizeof(V) == 8 && kLomrIter);
    static_assert(!kRegsOnly || kRegTerms, "kTermsRegs needs product 1's terms kept in registers");
    const bool reg2 = kRegTerms && (kRegsOnly || (narrow && ncs <= kFusedPre));
    FusedTerms<V, SlotT> terms;
    // product 1: A [Z P](:, j)  (kInit: A M(:, j); MR: A (M + alpha Z)(:, j))
    for (int t = lane; t < qe - qal; t += kWave) {
      if constexpr (NY > 1)
        acc[t] = V2<V>{V(0), V(0)};
      else
        acc[t].a = V(0);
    }
    __builtin_amdgcn_wave_barrier();
    if (valid) {
      if (e0 + gl < e1) {
        st_ac[gl * W + g] = pre_ac;
        st_y[gl * W + g] = pre_y;
      }
      // (with kTermsRegs no column of A is longer than GL, as with the stash)
      FusedProduct1<V, SlotT, GL, NY, kStash || kRegsOnly>(g, gl, e0, e1, pp, a0c, av, plan, acc + cb2, st_ac, st_y,
                                                           stash, kStash && !reg2, terms, yget1);
      // R = I - A M: the diagonal's sum becomes (A M)(j, j) - 1, so that
      // R = -sum below rounds as 1 - (A M)(j, j)
      if (!kLomrIter && gl == 0) acc[cb2 + dj].a -= V(1);
    }
    __builtin_amdgcn_wave_barrier();
    // R' = R - delta A Z - gamma A P over the tile's S2 range (kInit, MR:
    // I - A M), one chunk of kVec entries per lane; entries outside [qa, qe)
    // are computed in LDS slack and neither stored nor summed
    auto update_r = [&](int qb, const VecQ<V>& rold) {
      const int t0 = qb - qal;
      const VecQ2<V> s = *reinterpret_cast<const VecQ2<V>*>(acc + t0);
      VecQ<V> out;
#pragma unroll
      for (int v = 0; v < kVec; ++v) {
        V rv;
        if constexpr (!kLomrIter) {
          rv = -s.v[v].a;
        } else {
          rv = rold.v[v];
          rv = rv + nd * s.v[v].a;
          rv = rv + ng * s.v[v].b;
        }
        out.v[v] = rv;
        if (qb + v >= qa && qb + v < qe) red[5] += double(rv) * double(rv);
      }
      *reinterpret_cast<VecQ<V>*>(rr + t0) = out;
      if (qb >= qa && qb + kVec <= qe) {
        *reinterpret_cast<VecQ<V>*>(R + qb) = out;
      } else {
#pragma unroll
        for (int v = 0; v < kVec; ++v)
          if (qb + v >= qa && qb + v < qe) R[qb + v] = out.v[v];
      }
    };
#pragma unroll
    for (int r = 0; r < RQ; ++r) {
      const int qb = qal + (lane + r * kWave) * kVec;
      if (qb < qe) update_r(qb, rq[r]);
    }
    for (int r = RQ; qal + r * kWave * kVec < qe; ++r) {
      const int qb = qal + (lane + r * kWave) * kVec;
      if (qb < qe) update_r(qb, kLomrIter ? *reinterpret_cast<const VecQ<V>*>(R + qb) : VecQ<V>{});
    }
    __builtin_amdgcn_wave_barrier();
    const int next = tile + wstride;
    if (next < ntiles) load_tile(next);
    // M' = M + delta Z + gamma P, P' = beta P + Z, Z' = Pr R' over the S range
    auto update_s = [&](int p, V mo, V zo, V po, V pri, int tq) {
      const V zn = pri * rr[tq];
      V pn = V(0);
      if constexpr (kLomrIter) {
        V m = mo;
        m = m + d * zo;
        m = m + gm * po;
        M[p] = m;
        pn = po * be + zo;
      } else if constexpr (kMr && !kInit) {
        M[p] = mo + alpha * zo;
      }
      Z[p] = zn;
      if constexpr (!kMr) P[p] = pn;
      yzp[p - pa] = V2<V>{zn, pn};
    };
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int p = pa + lane + r * kWave;
      if (p < pe) update_s(p, mp[r], zp[r], ppv[r], ps[r], tp[r]);
    }
    for (int r = RS; pa + r * kWave < pe; ++r) {
      const int p = pa + lane + r * kWave;
      if (p < pe) {
        if constexpr (kInit)
          update_s(p, V(0), V(0), V(0), prS[p], int(stq[p]));
        else if constexpr (kMr)
          update_s(p, M[p], Z[p], V(0), prS[p], int(stq[p]));
        else
          update_s(p, M[p], Z[p], P[p], prS[p], int(stq[p]));
      }
    }
    __builtin_amdgcn_wave_barrier();
    // product 2: A [Z' P'](:, j) from LDS (MR: A Z'(:, j))
    constexpr int NY2 = kMr ? 1 : 2;
    for (int t = lane; t < qe - qal; t += kWave) {
      if constexpr (NY2 > 1)
        acc[t] = V2<V>{V(0), V(0)};
      else
        acc[t].a = V(0);
    }
    __builtin_amdgcn_wave_barrier();
    if (valid) {
      if (reg2) {
        // one batch: product 1's terms are still in registers
#pragma unroll
        for (int s = 0; s < kFusedPre; ++s)
          if ((terms.act >> s) & 1u) AddTerm<V, NY2>(acc + cb2, int(terms.sl[s]), terms.a[s], yzp[cb + s]);
      } else if constexpr (kStash) {
        FusedProduct2Stash<V, NY2>(ncs, lane, stash, yzp + cb, acc + cb2);
      } else if constexpr (!kRegsOnly) {
        const V2<V>* const ycol = yzp + cb - e0;
        auto yget2 = [&](int e, V& ya, V& yb) {
          const V2<V> y = ycol[e];
          ya = y.a;
          yb = y.b;
        };
        FusedStage<V, GL>(g, gl, e0, e1, a0c, st_ac, st_y, yget2);
        FusedProduct1<V, SlotT, GL, NY2, false>(g, gl, e0, e1, pp, a0c, av, plan, acc + cb2, st_ac, st_y, stash,
                                                false, terms, yget2);
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (next < ntiles) load_stage();
    // traces of the next step (LomrProjectEpi order), by chunks as above
    auto traces = [&](int qb, const VecQ<V>& pri) {
      const int t0 = qb - qal;
      const VecQ2<V> s = *reinterpret_cast<const VecQ2<V>*>(acc + t0);
      const VecQ<V> rv = *reinterpret_cast<const VecQ<V>*>(rr + t0);
#pragma unroll
      for (int v = 0; v < kVec; ++v) {
        if (qb + v < qa || qb + v >= qe) continue;
        if constexpr (kMr) {
          // MrProjectEpi order: <Pr R, Pr A Z>, ||Pr A Z||^2
          const double tmp = double(pri.v[v] * s.v[v].a);
          red[0] += double(pri.v[v] * rv.v[v]) * tmp;
          red[1] += tmp * tmp;
        } else {
          const double daz = s.v[v].a, dap = s.v[v].b, dpr = pri.v[v], dprr = double(pri.v[v] * rv.v[v]);
          red[0] += daz * (dpr * daz);
          red[1] += dap * (dpr * dap);
          red[2] += daz * (dpr * dap);
          red[3] += dprr * daz;
          red[4] += dprr * dap;
        }
      }
    };
#pragma unroll
    for (int r = 0; r < RQ; ++r) {
      const int qb = qal + (lane + r * kWave) * kVec;
      if (qb < qe) traces(qb, pq[r]);
    }
    for (int r = RQ; qal + r * kWave * kVec < qe; ++r) {
      const int qb = qal + (lane + r * kWave) * kVec;
      if (qb < qe) traces(qb, *reinterpret_cast<const VecQ<V>*>(prS2 + qb));
    }
    __builtin_amdgcn_wave_barrier();
  }
  StorePartials<6>(red, part_out);
}

// Threads of the one-block kernels between the fused passes.
constexpr int kStepBlk = 256;

// Between passes it - 1 and it, from the sums s of the pass's partials:
// ||R|| of row it - 1 (result_operations.cpp:7-67, converged iff ||R|| <
// tol: stop 2), then GetDeltaGamma (locally_minimal_residual.cpp:36-50) and
// the direction ratio gamma / delta for iteration it (breakdown: stop 1).
struct LomrStep {
  int stop;
  double rn, delta, gamma, beta;
};
__device__ LomrStep LomrStepValues(const double (&s)[6], int it, double tol) {
  LomrStep r{0, sqrt(s[5]), 0.0, 0.0, 0.0};
  if (r.rn < tol) {
    r.stop = 2;
    return r;
  }
  const bool first = it == 1;
  const double az_praz = s[0], ap_prap = first ? 0.0 : s[1], az_prap = first ? 0.0 : s[2], z_az = s[3],
               z_ap = first ? 0.0 : s[4];
  double delta = 0.0, gamma = 0.0;
  if (first) {
    if (!(isfinite(az_praz) && fabs(az_praz) > 0.0)) {
      r.stop = 1;
      return r;
    }
    delta = z_az / az_praz;
    gamma = 1.0;
  } else {
    SolveLeastSquares2x2(az_praz, az_prap, az_prap, ap_prap, z_az, z_ap, delta, gamma);
  }
  if (!isfinite(delta) || !isfinite(gamma)) {
    r.stop = 1;
    return r;
  }
  const double ratio = delta != 0.0 ? gamma / delta : 0.0;
  r.delta = delta;
  r.gamma = gamma;
  r.beta = isfinite(ratio) ? ratio : 0.0;
  return r;
}

// LomrStepValues with its results stored: history row it - 1, the iteration
// count and stop flag, and the step values (thread 0 of one block).
__device__ void LomrStepFromSums(const double (&s)[6], int it, double tol, double* __restrict__ hist,
                                 double* __restrict__ sc, Flags* __restrict__ flags) {
  const LomrStep r = LomrStepValues(s, it, tol);
  hist[it - 1] = r.rn;
  flags->iterations = it - 1;
  if (r.stop) {
    flags->stop = r.stop;
    return;
  }
  sc[kDelta] = r.delta;
  sc[kGamma] = r.gamma;
  sc[kBeta] = r.beta;
}

// History row it of the last pass (AppendResidual), from its summed partials.
__device__ void LomrFinishFromSums(const double (&s)[6], int it, double tol, double* __restrict__ hist,
                                   Flags* __restrict__ flags) {
  const double rn = sqrt(s[5]);
  hist[it] = rn;
  flags->iterations = it;
  if (rn < tol) flags->stop = 2;
}

// LomrStepFromSums as a kernel of its own, between the fused passes.
__global__ __launch_bounds__(kStepBlk) void LomrFusedStep(const double* __restrict__ part, int G, int it, double tol,
                                                      double* __restrict__ hist, double* __restrict__ sc,
                                                      Flags* __restrict__ flags) {
  if (flags->stop) return;
  double s[6];
  SumPartialsBlockFused<6>(part, G, s);
  if (threadIdx.x != 0) return;
  LomrStepFromSums(s, it, tol, hist, sc, flags);
}

// MR between passes it - 1 and it: history row it - 1, then the step
// alpha = <Pr R, Pr A Z> / ||Pr A Z||^2 (minimal_residual.cpp:58-65).
__global__ __launch_bounds__(kStepBlk) void MrFusedStep(const double* __restrict__ part, int G, int it, double tol,
                                                    double* __restrict__ hist, double* __restrict__ sc,
                                                    Flags* __restrict__ flags) {
  if (flags->stop) return;
  double s[6];
  SumPartialsBlockFused<6>(part, G, s);
  if (threadIdx.x != 0) return;
  const double rn = sqrt(s[5]);
  hist[it - 1] = rn;
  flags->iterations = it - 1;
  if (rn < tol) {
    flags->stop = 2;
    return;
  }
  const double num = s[0], den = s[1];
  if (!(isfinite(den) && fabs(den) > 0.0)) {
    flags->stop = 1;
    return;
  }
  const double alpha = num / den;
  if (!isfinite(alpha)) {
    flags->stop = 1;
    return;
  }
  sc[kAlpha] = alpha;
}

// Clears the stop flags (a kernel, so that recorded graphs hold no memset).
__global__ void ResetFlags(Flags* __restrict__ flags) {
  flags->stop = 0;
  flags->iterations = 0;
}

// History row of the last pass (AppendResidual).
__global__ __launch_bounds__(kStepBlk) void LomrFusedFinish(const double* __restrict__ part, int G, int it, double tol,
                                                        double* __restrict__ hist, Flags* __restrict__ flags) {
  if (flags->stop) return;
  double s[6];
  SumPartialsBlock<6>(part, G, s);
  if (threadIdx.x != 0) return;
  LomrFinishFromSums(s, it, tol, hist, flags);
}

// R = I - A M; Z = Pr R on S (mode 1: P = Z and <R, Z>).
template <class V>
struct ResidualEpi {
  static constexpr int K = 2;
  const int* s2_to_s;
  const V* pr;
  V* R;
  V* Z;
  V* P;
  int mode;
  __device__ void Load(const double*) {}
  __device__ void operator()(int64_t q, int i, int j, V a0, V, double* acc) const {
    V r = -a0;
    if (i == j) r += V(1);
    R[q] = r;
    acc[0] += double(r) * double(r);
    const int p = s2_to_s[q];
    if (p >= 0) {
      const V z = pr[i] * r;
      Z[p] = z;
      if (mode == 1) {
        P[p] = z;
        acc[1] += double(r) * double(z);
      }
    }
  }
};

// AZ, AP and the five trace products of the LOMR step.
template <class V>
struct LomrProjectEpi {
  static constexpr int K = 5;
  const V* R;
  const V* pr;
  __device__ void Load(const double*) {}
  __device__ void operator()(int64_t q, int i, int, V az, V ap, double* acc) const {
    const double daz = az, dap = ap, dpr = pr[i], dprr = double(pr[i] * R[q]);
    acc[0] += daz * (dpr * daz);
    acc[1] += dap * (dpr * dap);
    acc[2] += daz * (dpr * dap);
    acc[3] += dprr * daz;
    acc[4] += dprr * dap;
  }
};

// LOMR update from recomputed AZ, AP (locally_minimal_residual.cpp:92-107):
// R -= delta AZ + gamma AP; on S: M += delta Z + gamma P, P = beta P + Z, and
// Z = Pr R for the next step.  Every entry touched belongs to the column whose
// product was just formed, so no other group reads it in this kernel.
template <class V>
struct LomrUpdateEpi {
  static constexpr int K = 1;
  const int* s2_to_s;
  const V* pr;
  V* R;
  V* M;
  V* Z;
  V* P;
  V d, g, b, nd, ng;
  __device__ void Load(const double* sc) {
    d = V(sc[kDelta]);
    g = V(sc[kGamma]);
    b = V(sc[kBeta]);
    nd = V(-sc[kDelta]);
    ng = V(-sc[kGamma]);
  }
  __device__ void operator()(int64_t q, int i, int, V az, V ap, double* acc) const {
    V r = R[q];
    r = r + nd * az;
    r = r + ng * ap;
    R[q] = r;
    acc[0] += double(r) * double(r);
    const int p = s2_to_s[q];
    if (p >= 0) {
      const V z = Z[p], pp = P[p];
      V m = M[p];
      m = m + d * z;
      m = m + g * pp;
      M[p] = m;
      P[p] = pp * b + z;
      Z[p] = pr[i] * r;
    }
  }
};

// num = <Pr R, Pr A Z>, den = ||Pr A Z||^2.
template <class V>
struct MrProjectEpi {
  static constexpr int K = 2;
  const V* R;
  const V* pr;
  __device__ void Load(const double*) {}
  __device__ void operator()(int64_t q, int i, int, V az, V, double* acc) const {
    const double tmp = double(pr[i] * az);
    acc[0] += double(pr[i] * R[q]) * tmp;
    acc[1] += tmp * tmp;
  }
};

// AP and <P, AP>.
template <class V>
struct CgProjectEpi {
  static constexpr int K = 1;
  const int* s2_to_s;
  const V* P;
  V* AP;
  __device__ void Load(const double*) {}
  __device__ void operator()(int64_t q, int, int, V ap, V, double* acc) const {
    AP[q] = ap;
    const int p = s2_to_s[q];
    if (p >= 0) acc[0] += double(P[p]) * double(ap);
  }
};

// ------------------------------------------------------------ inner-outer
// Z = M_outer R sampled on S: (i, j) -> sum_k M_outer(i, k) R(k, j) over k
// ascending (row i of M_outer = column i of S^T, gathered through stperm,
// against column j of R on S2).
constexpr int kOuterRow = 8, kOuterCol = 16;  // list lengths of the register path
template <class V>
__global__ __launch_bounds__(kBlk) void SampledOuterProduct(int64_t nnz_s, const int* __restrict__ si,
                                                            const int* __restrict__ scol, const int* __restrict__ stp,
                                                            const int* __restrict__ sti,
                                                            const int* __restrict__ stperm,
                                                            const V* __restrict__ mout, const int* __restrict__ s2p,
                                                            const int* __restrict__ s2i, const V* __restrict__ R,
                                                            V* __restrict__ Z) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < nnz_s; p += int64_t(gridDim.x) * blockDim.x) {
    const int i = si[p], j = scol[p];
    int a = stp[i], ae = stp[i + 1], b = s2p[j], be = s2p[j + 1];
    V acc = V(0);
    if (ae - a <= kOuterRow && be - b <= kOuterCol) {
      // short lists: both loaded at once, matched by comparisons, the matches
      // added in k order (the row's order), as the merge below adds them
      int ka[kOuterRow], mp[kOuterRow], kb[kOuterCol];
#pragma unroll
      for (int u = 0; u < kOuterRow; ++u) {
        ka[u] = a + u < ae ? sti[a + u] : -1;
        mp[u] = a + u < ae ? stperm[a + u] : 0;
      }
#pragma unroll
      for (int v = 0; v < kOuterCol; ++v) kb[v] = b + v < be ? s2i[b + v] : -2;
#pragma unroll
      for (int u = 0; u < kOuterRow; ++u) {
        int m = -1;
#pragma unroll
        for (int v = 0; v < kOuterCol; ++v) m = kb[v] == ka[u] ? v : m;
        if (m >= 0) acc += mout[mp[u]] * R[b + m];
      }
      Z[p] = acc;
      continue;
    }
    while (a < ae && b < be) {
      const int ka = sti[a], kb = s2i[b];
      if (ka == kb) {
        acc += mout[stperm[a]] * R[b];
        ++a;
        ++b;
      } else if (ka < kb) {
        ++a;
      } else {
        ++b;
      }
    }
    Z[p] = acc;
  }
}

// Stores A Z and A q on S2.
template <class V>
struct StoreTwoEpi {
  static constexpr int K = 1;
  V* AZ;
  V* AQ;
  __device__ void Load(const double*) {}
  __device__ void operator()(int64_t q, int, int, V az, V aq, double*) const {
    AZ[q] = az;
    if (AQ) AQ[q] = aq;
  }
};

// Per-column step sizes of inner_outer_common.hpp:50-69: sequential column dots
// over S2 (row order), delta = <r, az> / ||az||^2 or the 2x2 solve, and the
// normalisation 1 / delta.  FP64 throughout.  A wavefront takes 64
// consecutive columns: their S2 entries are contiguous, so the wavefront
// stages them through LDS kIoChunk at a time with coalesced loads, and each
// lane adds its own column's entries from there in row order.
constexpr int kIoChunk = 512;
template <class V>
__global__ __launch_bounds__(kBlk) void IoColumnSteps(int n, const int* __restrict__ s2p, const V* __restrict__ R,
                                                      const V* __restrict__ AZ, const V* __restrict__ AQ, int lomr,
                                                      int two_term, double rank_eps, double* __restrict__ delta,
                                                      double* __restrict__ gamma, double* __restrict__ norm) {
  constexpr int W = kBlk / kWave;
  __shared__ V sR[W][kIoChunk];
  __shared__ V sZ[W][kIoChunk];
  __shared__ V sQ[W][kIoChunk];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int j0 = (blockIdx.x * W + w) * kWave;
  if (j0 >= n) return;  // the whole wavefront leaves
  const int j = j0 + lane;
  const int span0 = s2p[j0], span1 = s2p[min(n, j0 + kWave)];
  const int b = j < n ? s2p[j] : span1, e = j < n ? s2p[j + 1] : span1;
  double r_az = 0.0, az_sq = 0.0, az_aq = 0.0, aq_sq = 0.0, r_aq = 0.0;
  for (int c0 = span0; c0 < span1; c0 += kIoChunk) {
    const int c1 = min(span1, c0 + kIoChunk);
    for (int q = c0 + lane; q < c1; q += kWave) {
      sR[w][q - c0] = R[q];
      sZ[w][q - c0] = AZ[q];
      if (two_term) sQ[w][q - c0] = AQ[q];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const int qe = min(e, c1);
    for (int q = max(b, c0); q < qe; ++q) {
      const double r = sR[w][q - c0], az = sZ[w][q - c0];
      r_az += r * az;
      az_sq += az * az;
      if (two_term) {
        const double aq = sQ[w][q - c0];
        az_aq += az * aq;
        aq_sq += aq * aq;
        r_aq += r * aq;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
  if (j >= n) return;
  double d, g;
  if (!two_term) {
    d = (isfinite(az_sq) && fabs(az_sq) > 0.0) ? r_az / az_sq : 0.0;
    if (!isfinite(d)) d = 0.0;
    g = lomr ? 1.0 : 0.0;
  } else {
    SolveLeastSquares2x2(az_sq, az_aq, az_aq, aq_sq, r_az, r_aq, d, g, rank_eps);
  }
  delta[j] = d;
  gamma[j] = g;
  const double inv = d != 0.0 ? 1.0 / d : 1.0;
  norm[j] = isfinite(inv) ? inv : 1.0;
}

// P = Z diag(delta) + q diag(gamma) (inner_outer_common.hpp:70-75), formed in
// FP64 and rounded once: near-collinear columns give large delta and gamma of
// opposite sign whose products cancel.
template <class V>
__global__ void IoDirection(int64_t nnz_s, const int* __restrict__ scol, const V* __restrict__ Z,
                            const V* __restrict__ Q, const double* __restrict__ delta,
                            const double* __restrict__ gamma, int lomr, V* __restrict__ P) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < nnz_s; p += int64_t(gridDim.x) * blockDim.x) {
    const int j = scol[p];
    double v = double(Z[p]) * delta[j];
    if (lomr) v = v + 1.0 * (double(Q[p]) * gamma[j]);
    P[p] = V(v);
  }
}

// R -= A P; M += P; q = P diag(1/delta) (inner_outer_common.hpp:90-98).
template <class V>
struct IoUpdateEpi {
  static constexpr int K = 1;
  const int* s2_to_s;
  const int* ncol_unused;
  const V* P;
  const double* norm;
  V* R;
  V* M;
  V* Q;
  int lomr;
  __device__ void Load(const double*) {}
  __device__ void operator()(int64_t q, int, int j, V ap, V, double* acc) const {
    const V r = R[q] + V(-1) * ap;
    R[q] = r;
    acc[0] += double(r) * double(r);
    const int p = s2_to_s[q];
    if (p >= 0) {
      const V pv = P[p];
      M[p] = M[p] + V(1) * pv;
      if (lomr) Q[p] = pv * V(norm[j]);
    }
  }
};

// --------------------------------------------------------------- residual
// R = I - A M on S2 (value -(sum) + [i == j]); Z = Pr R sampled on S.
// mode 1 (CG setup) also sets P = Z and accumulates <R, Z>.
template <class V>
__global__ __launch_bounds__(kBlk) void ResidualKernel(
    int64_t nnz2, const int* __restrict__ s2row, const int* __restrict__ s2col, const int* __restrict__ s2_to_s,
    const int* __restrict__ atp, const int* __restrict__ ati, const V* __restrict__ atv, const int* __restrict__ sp,
    const int* __restrict__ si, const V* __restrict__ M, const V* __restrict__ pr, V* __restrict__ R,
    V* __restrict__ Z, V* __restrict__ P, int mode, double* __restrict__ part, const Flags* __restrict__ flags) {
  if (flags->stop) return;
  double acc[2] = {0.0, 0.0};
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz2; q += int64_t(gridDim.x) * blockDim.x) {
    const int i = s2row[q], j = s2col[q];
    V r = -InnerProduct<V>(ati, atv, atp[i], atp[i + 1], si, M, sp[j], sp[j + 1]);
    if (i == j) r += V(1);
    R[q] = r;
    acc[0] += double(r) * double(r);
    const int p = s2_to_s[q];
    if (p >= 0) {
      const V z = pr[i] * r;
      Z[p] = z;
      if (mode == 1) {
        P[p] = z;
        acc[1] += double(r) * double(z);
      }
    }
  }
  StorePartials<2>(acc, part);
}

// --------------------------------------------------------------- LOMR
// AZ, AP on S2 and the five Pr-weighted trace products of
// locally_minimal_residual.cpp:66-87 (numerators projected: <Pr R, A Z>, <Pr R, A P>).
template <class V>
__global__ __launch_bounds__(kBlk) void LomrProject(
    int64_t nnz2, const int* __restrict__ s2row, const int* __restrict__ s2col, const int* __restrict__ atp,
    const int* __restrict__ ati, const V* __restrict__ atv, const int* __restrict__ sp, const int* __restrict__ si,
    const V* __restrict__ Z, const V* __restrict__ P, const V* __restrict__ R, const V* __restrict__ pr,
    V* __restrict__ AZ, V* __restrict__ AP, double* __restrict__ part, const Flags* __restrict__ flags) {
  if (flags->stop) return;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz2; q += int64_t(gridDim.x) * blockDim.x) {
    const int i = s2row[q], j = s2col[q];
    V az, ap;
    InnerProduct2<V>(ati, atv, atp[i], atp[i + 1], si, Z, P, sp[j], sp[j + 1], az, ap);
    AZ[q] = az;
    AP[q] = ap;
    const double daz = az, dap = ap, dpr = pr[i], dprr = double(pr[i] * R[q]);
    acc[0] += daz * (dpr * daz);   // <AZ, Pr AZ>
    acc[1] += dap * (dpr * dap);   // <AP, Pr AP>
    acc[2] += daz * (dpr * dap);   // <AZ, Pr AP>
    acc[3] += dprr * daz;          // <Pr R, AZ>
    acc[4] += dprr * dap;          // <Pr R, AP>
  }
  StorePartials<5>(acc, part);
}

// GetDeltaGamma (locally_minimal_residual.cpp:36-50) on the device.
__global__ __launch_bounds__(kBlk) void LomrSteps(const double* __restrict__ part, int G, int first,
                                                  double* __restrict__ sc, Flags* __restrict__ flags) {
  if (flags->stop) return;
  double s[5];
  SumPartialsBlock<5>(part, G, s);
  if (threadIdx.x != 0) return;
  const double az_praz = s[0], ap_prap = first ? 0.0 : s[1], az_prap = first ? 0.0 : s[2], z_az = s[3],
               z_ap = first ? 0.0 : s[4];
  double delta, gamma;
  if (first) {
    if (!(isfinite(az_praz) && fabs(az_praz) > 0.0)) {
      flags->stop = 1;
      return;
    }
    delta = z_az / az_praz;
    gamma = 1.0;
  } else {
    SolveLeastSquares2x2(az_praz, az_prap, az_prap, ap_prap, z_az, z_ap, delta, gamma);
  }
  if (!isfinite(delta) || !isfinite(gamma)) {
    flags->stop = 1;
    return;
  }
  const double ratio = delta != 0.0 ? gamma / delta : 0.0;
  sc[kDelta] = delta;
  sc[kGamma] = gamma;
  sc[kBeta] = isfinite(ratio) ? ratio : 0.0;
}

// R -= delta AZ + gamma AP; M += delta Z + gamma P; P = beta P + Z; Z = Pr R.
template <class V>
__global__ __launch_bounds__(kBlk) void LomrUpdate(int64_t nnz2, const int* __restrict__ s2row,
                                                   const int* __restrict__ s2_to_s, const V* __restrict__ AZ,
                                                   const V* __restrict__ AP, const V* __restrict__ pr,
                                                   V* __restrict__ R, V* __restrict__ M, V* __restrict__ Z,
                                                   V* __restrict__ P, const double* __restrict__ sc,
                                                   double* __restrict__ part, const Flags* __restrict__ flags) {
  if (flags->stop) return;
  const V d = V(sc[kDelta]), g = V(sc[kGamma]), b = V(sc[kBeta]);
  const V nd = V(-sc[kDelta]), ng = V(-sc[kGamma]);
  double acc[1] = {0.0};
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz2; q += int64_t(gridDim.x) * blockDim.x) {
    V r = R[q];
    r = r + nd * AZ[q];
    r = r + ng * AP[q];
    R[q] = r;
    acc[0] += double(r) * double(r);
    const int p = s2_to_s[q];
    if (p >= 0) {
      const V z = Z[p], pp = P[p];
      V m = M[p];
      m = m + d * z;
      m = m + g * pp;
      M[p] = m;
      P[p] = pp * b + z;
      Z[p] = pr[s2row[q]] * r;
    }
  }
  StorePartials<1>(acc, part);
}

// AppendIteration on the device: history[it] = ||R||_F, converged if < tol.
__global__ __launch_bounds__(kBlk) void AppendResidual(const double* __restrict__ part, int G, double tol,
                                                       double* __restrict__ hist, Flags* __restrict__ flags) {
  if (flags->stop) return;
  double s[1];
  SumPartialsBlock<1>(part, G, s);
  if (threadIdx.x != 0) return;
  const double rn = sqrt(s[0]);
  const int it = flags->iterations + 1;
  hist[it] = rn;
  flags->iterations = it;
  if (rn < tol) flags->stop = 2;
}

__global__ __launch_bounds__(kBlk) void AppendInitial(const double* __restrict__ part, int G, double tol,
                                                      double* __restrict__ hist, double* __restrict__ sc,
                                                      Flags* __restrict__ flags) {
  double s[2];
  SumPartialsBlock<2>(part, G, s);
  if (threadIdx.x != 0) return;
  const double rn = sqrt(s[0]);
  hist[0] = rn;
  flags->iterations = 0;
  sc[kRz] = s[1];
  flags->stop = rn < tol ? 2 : 0;
}

// --------------------------------------------------------------- MR
// num = <Pr R, Pr A Z>, den = ||Pr A Z||^2 (minimal_residual.cpp:44-54, projected numerator).
template <class V>
__global__ __launch_bounds__(kBlk) void MrProject(int64_t nnz2, const int* __restrict__ s2row,
                                                  const int* __restrict__ s2col, const int* __restrict__ atp,
                                                  const int* __restrict__ ati, const V* __restrict__ atv,
                                                  const int* __restrict__ sp, const int* __restrict__ si,
                                                  const V* __restrict__ Z, const V* __restrict__ R,
                                                  const V* __restrict__ pr, double* __restrict__ part,
                                                  const Flags* __restrict__ flags) {
  if (flags->stop) return;
  double acc[2] = {0.0, 0.0};
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz2; q += int64_t(gridDim.x) * blockDim.x) {
    const int i = s2row[q], j = s2col[q];
    const V az = InnerProduct<V>(ati, atv, atp[i], atp[i + 1], si, Z, sp[j], sp[j + 1]);
    const double tmp = double(pr[i] * az);
    acc[0] += double(pr[i] * R[q]) * tmp;
    acc[1] += tmp * tmp;
  }
  StorePartials<2>(acc, part);
}

__global__ __launch_bounds__(kBlk) void MrStep(const double* __restrict__ part, int G, double* __restrict__ sc,
                                               Flags* __restrict__ flags) {
  if (flags->stop) return;
  double s[2];
  SumPartialsBlock<2>(part, G, s);
  if (threadIdx.x != 0) return;
  const double num = s[0], den = s[1];
  if (!(isfinite(den) && fabs(den) > 0.0)) {
    flags->stop = 1;
    return;
  }
  const double alpha = num / den;
  if (!isfinite(alpha)) {
    flags->stop = 1;
    return;
  }
  sc[kAlpha] = alpha;
}

// M += alpha * D over S.
template <class V>
__global__ __launch_bounds__(kBlk) void AxpyS(int64_t nnz, V* __restrict__ M, const V* __restrict__ D,
                                              const double* __restrict__ sc, const Flags* __restrict__ flags) {
  if (flags->stop) return;
  const V a = V(sc[kAlpha]);
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < nnz; p += int64_t(gridDim.x) * blockDim.x)
    M[p] = M[p] + a * D[p];
}

// --------------------------------------------------------------- CG
// AP on S2 and <P, AP> (conjugate_gradient.cpp:36-39).
template <class V>
__global__ __launch_bounds__(kBlk) void CgProject(int64_t nnz2, const int* __restrict__ s2row,
                                                  const int* __restrict__ s2col, const int* __restrict__ s2_to_s,
                                                  const int* __restrict__ atp, const int* __restrict__ ati,
                                                  const V* __restrict__ atv, const int* __restrict__ sp,
                                                  const int* __restrict__ si, const V* __restrict__ P,
                                                  V* __restrict__ AP, double* __restrict__ part,
                                                  const Flags* __restrict__ flags) {
  if (flags->stop) return;
  double acc[1] = {0.0};
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz2; q += int64_t(gridDim.x) * blockDim.x) {
    const int i = s2row[q], j = s2col[q];
    const V ap = InnerProduct<V>(ati, atv, atp[i], atp[i + 1], si, P, sp[j], sp[j + 1]);
    AP[q] = ap;
    const int p = s2_to_s[q];
    if (p >= 0) acc[0] += double(P[p]) * double(ap);
  }
  StorePartials<1>(acc, part);
}

__global__ __launch_bounds__(kBlk) void CgAlpha(const double* __restrict__ part, int G, double* __restrict__ sc,
                                                Flags* __restrict__ flags) {
  if (flags->stop) return;
  double s[1];
  SumPartialsBlock<1>(part, G, s);
  if (threadIdx.x != 0) return;
  const double r_z = sc[kRz], den = s[0];
  if (!(isfinite(r_z) && fabs(r_z) > 0.0) || !(isfinite(den) && fabs(den) > 0.0)) {
    flags->stop = 1;
    return;
  }
  const double alpha = r_z / den;
  if (!isfinite(alpha)) {
    flags->stop = 1;
    return;
  }
  sc[kAlpha] = alpha;
  sc[kPrevRz] = r_z;
}

// R -= alpha AP; M += alpha P; Z = Pr R on S; partials ||R||^2 and <R, Z>.
template <class V>
__global__ __launch_bounds__(kBlk) void CgUpdate(int64_t nnz2, const int* __restrict__ s2row,
                                                 const int* __restrict__ s2_to_s, const V* __restrict__ AP,
                                                 const V* __restrict__ pr, V* __restrict__ R, V* __restrict__ M,
                                                 const V* __restrict__ P, V* __restrict__ Z,
                                                 const double* __restrict__ sc, double* __restrict__ part,
                                                 const Flags* __restrict__ flags) {
  if (flags->stop) return;
  const V a = V(sc[kAlpha]), na = V(-sc[kAlpha]);
  double acc[2] = {0.0, 0.0};
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz2; q += int64_t(gridDim.x) * blockDim.x) {
    const V r = R[q] + na * AP[q];
    R[q] = r;
    acc[0] += double(r) * double(r);
    const int p = s2_to_s[q];
    if (p >= 0) {
      M[p] = M[p] + a * P[p];
      const V z = pr[s2row[q]] * r;
      Z[p] = z;
      acc[1] += double(r) * double(z);
    }
  }
  StorePartials<2>(acc, part);
}

// History row, then beta = <R, Z>_new / <R, Z>_old (conjugate_gradient.cpp:57-71).
__global__ __launch_bounds__(kBlk) void CgAppendBeta(const double* __restrict__ part, int G, double tol,
                                                     double* __restrict__ hist, double* __restrict__ sc,
                                                     Flags* __restrict__ flags) {
  if (flags->stop) return;
  double s[2];
  SumPartialsBlock<2>(part, G, s);
  if (threadIdx.x != 0) return;
  const double rn = sqrt(s[0]);
  const int it = flags->iterations + 1;
  hist[it] = rn;
  flags->iterations = it;
  if (rn < tol) {
    flags->stop = 2;
    return;
  }
  const double r_z = s[1];
  sc[kRz] = r_z;
  const double beta = r_z / sc[kPrevRz];
  if (!isfinite(beta)) {
    flags->stop = 1;
    return;
  }
  sc[kBeta] = beta;
}

// P = P * beta + Z over S (LinearCombination(Z, 1, P, beta)).
template <class V>
__global__ __launch_bounds__(kBlk) void CgDirection(int64_t nnz, V* __restrict__ P, const V* __restrict__ Z,
                                                    const double* __restrict__ sc, const Flags* __restrict__ flags) {
  if (flags->stop) return;
  const V b = V(sc[kBeta]);
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < nnz; p += int64_t(gridDim.x) * blockDim.x)
    P[p] = P[p] * b + Z[p];
}

// --------------------------------------------------------------- setup
__global__ void ColumnOfEntry(int ncols, const int* __restrict__ ptr, int* __restrict__ col) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  for (int p = ptr[j]; p < ptr[j + 1]; ++p) col[p] = j;
}

// s_to_s2 / s2_to_s maps (S is a subset of S2 column by column).
__global__ void MapSubPattern(int ncols, const int* __restrict__ sp, const int* __restrict__ si,
                              const int* __restrict__ s2p, const int* __restrict__ s2i, int* __restrict__ s2_to_s,
                              int* __restrict__ error_flag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  for (int p = sp[j]; p < sp[j + 1]; ++p) {
    const int q = FindSorted(s2i, s2p[j], s2p[j + 1], si[p]);
    if (q < 0) {
      *error_flag = 5;  // the context's error code (mapped host memory)
      continue;
    }
    s2_to_s[q] = p;
  }
}

// Facts that decide how S and S2 are formed, and the checks of Pr =
// diag(A)^-1 (methods_common.cpp:50-63) on d = diag(A), inverted in place:
// facts[0] = 1 when a column of A lacks its diagonal entry, facts[1] = 1 when
// M0 has an entry off the diagonal, facts[2] = 1 when a diagonal entry is
// not finite and positive, facts[3] = 1 when its inverse is not finite.
__global__ void PatternFacts(int n, const int* __restrict__ ap, const int* __restrict__ ai,
                             const int* __restrict__ mp, const int* __restrict__ mi, double* __restrict__ d,
                             int* __restrict__ facts) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (FindSorted(ai, ap[j], ap[j + 1], j) < 0) facts[0] = 1;
  if (mp)
    for (int e = mp[j]; e < mp[j + 1]; ++e)
      if (mi[e] != j) facts[1] = 1;
  if (d) {
    const double v = d[j];
    if (!isfinite(v) || v <= 0.0) {
      facts[2] = 1;
    } else {
      const double r = 1.0 / v;
      if (!isfinite(r)) facts[3] = 1;
      d[j] = r;
    }
  }
}

// Values of X sampled on the pattern S (0 where X has no entry).
// kSampleLanes lanes per column: a column's entries go to consecutive lanes,
// so a wavefront's stores cover contiguous ranges of out.
constexpr int kSampleLanes = 8;
template <class V>
__global__ void SampleOnPattern(int ncols, const int* __restrict__ sp, const int* __restrict__ si,
                                const int* __restrict__ xp, const int* __restrict__ xi, const double* __restrict__ xv,
                                V* __restrict__ out) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int j = int(t / kSampleLanes), gl = int(t % kSampleLanes);
  if (j >= ncols) return;
  const int x0 = xp[j], x1 = xp[j + 1];
  for (int p = sp[j] + gl; p < sp[j + 1]; p += kSampleLanes) {
    const int q = FindSorted(xi, x0, x1, si[p]);
    out[p] = q >= 0 ? V(xv[q]) : V(0);
  }
}

__global__ void IotaD(int64_t n, double* __restrict__ v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = double(i);
}

__global__ void ToIntD(int64_t n, const double* __restrict__ v, int* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = int(v[i]);
}

template <class V>
__global__ void ConvertValues(int64_t n, const double* __restrict__ in, V* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = V(in[i]);
}

// out_d[i] = m[perm[i]] in FP64, out_f (nullable) the same rounded to FP32
// (as converting to FP64 and then to FP32 rounds).
template <class V>
__global__ void GatherValues(int64_t len, const int* __restrict__ perm, const V* __restrict__ m,
                             double* __restrict__ out_d, float* __restrict__ out_f) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const double v = double(m[perm[i]]);
  out_d[i] = v;
  if (out_f) out_f[i] = float(v);
}

template <class V>
__global__ void ToDouble(int64_t n, const V* __restrict__ in, double* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = double(in[i]);
}

template <class V>
__global__ void FillZero(int64_t n, V* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = V(0);
}

// Fused-pass operand tables: (A(:, k) start, length) of each S entry,
// Pr(i) of each S and S2 entry, and each S entry's position in its S2 column.
__global__ void SEntryA(int64_t nnz_s, const int* __restrict__ si, const int* __restrict__ ap,
                        int2* __restrict__ a0c) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nnz_s) return;
  const int k = si[e];
  a0c[e] = make_int2(ap[k], ap[k + 1] - ap[k]);
}

template <class V>
__global__ void GatherRows(int64_t nnz, const int* __restrict__ rows, const V* __restrict__ pr, V* __restrict__ out) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < nnz) out[e] = pr[rows[e]];
}

// Tile tables of the fused pass (W = 64 / GL columns per tile): stq[p], for
// every S entry p of column j, is the position of p's row in S2(:, j) counted
// from the first S2 entry of j's tile rounded down to a multiple of vec; s2d[j]
// is the position of the diagonal in S2(:, j) (S2 holds I).
__global__ void TileInfo(int64_t nnz2, int W, int vec, const int* __restrict__ s2_to_s, const int* __restrict__ s2col,
                         const int* __restrict__ s2i, const int* __restrict__ s2p, uint16_t* __restrict__ stq,
                         int* __restrict__ s2d) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= nnz2) return;
  const int j = s2col[q];
  const int p = s2_to_s[q];
  if (p >= 0) stq[p] = uint16_t(q - (s2p[j - j % W] & ~(vec - 1)));
  if (s2i[q] == j) s2d[j] = int(q - s2p[j]);
}

unsigned GridFor(int64_t n) {
  int64_t g = (n + kBlk - 1) / kBlk;
  if (g < 1) g = 1;
  if (g > kMaxGrid) g = kMaxGrid;
  return (unsigned)g;
}

unsigned Blocks1(int64_t n) {
  int64_t g = (n + kBlk - 1) / kBlk;
  return (unsigned)std::max<int64_t>(g, 1);
}

// ------------------------------------------------------------ diagonal pass
// LOMR and MR on a fixed pattern whose matrices are a grid stencil's: A and
// S on the diagonals of a Laplacian (0, +-1, +-m, +-m^2, up to the grid's
// dimension) or of the 2-dimensional box stencil (every offset a + b m with
// a, b in {-1, 0, 1}), o = col - row ascending, and S2 on all their pairwise
// sums.  Every matrix is held by columns and diagonals: entry (j - o_d, j)
// sits at d ld + j, 0 where the column has no entry there (a bit per
// diagonal of S's column mask says which are entries).  Entry (i, j) of A Y
// (Y on S) is the sum over the S entries (k, j) of A(i, k) Y(k, j); with the
// diagonals in this order the S2 diagonal that pair (S diagonal es, A
// diagonal ea) lands on is a constant of the shape, so one thread per column
// forms its products with every index known at compile time, every load
// address known before any load returns, and the loads of a wavefront
// contiguous.  Absent entries are exact zeros: a product term with one adds
// nothing, and no mask is needed for A or S2.
constexpr int kDpS = 9;    // at most this many diagonals in A and in S
constexpr int kDpS2 = 25;  // at most this many in S2
constexpr int kDpPairs = kDpS * kDpS;

struct DiaShape {
  int nd = 0, nd2 = 0, np = 0;
  long long off[kDpS] = {};        // diagonals of A and S, with a stand-in grid side
  long long off2[kDpS2] = {};      // S2: their pairwise sums, ascending
  signed char s2OfS[kDpS] = {};    // S2 diagonal of off[es]
  signed char sOfS2[kDpS2] = {};   // S diagonal of off2[t] (-1: none)
  // the pairs (S diagonal es, A diagonal ea): S2 diagonal t ascending, and
  // for one t, es descending (S rows k = j - off[es] ascending, the order of
  // the reference's products)
  unsigned char pt[kDpPairs] = {}, pes[kDpPairs] = {}, pea[kDpPairs] = {};
};

// Shape kinds: 1, 2, 3 = Laplacian in 1, 2, 3 dimensions (3, 5, 7
// diagonals); 5 = box stencil in 2 dimensions (9).  (The 27-point box, 729
// terms a column, held 230+ VGPRs and ran 2.5 times slower than the fused
// pass on a 128^3 MR build; it takes the fused pass.)  The side 2^12
// stands in for the real one: the order of the offsets and of their sums is
// the same for any side of at least 5, and the host checks the real offsets
// against the shape.
constexpr DiaShape MakeDiaShape(int kind) {
  DiaShape s{};
  long long o[kDpS] = {};
  int n = 0;
  const bool box = kind > 3;
  const int dims = box ? kind - 3 : kind;
  if (box) {
    int cnt = 1;
    for (int d = 0; d < dims; ++d) cnt *= 3;
    for (int c = 0; c < cnt; ++c) {
      long long v = 0, unit = 1;
      int x = c;
      for (int d = 0; d < dims; ++d) {
        v += (x % 3 - 1) * unit;
        x /= 3;
        unit *= 4096;
      }
      o[n++] = v;
    }
  } else {
    o[n++] = 0;
    long long unit = 1;
    for (int d = 0; d < dims; ++d) {
      o[n++] = unit;
      o[n++] = -unit;
      unit *= 4096;
    }
  }
  for (int i = 1; i < n; ++i)
    for (int k = i; k > 0 && o[k - 1] > o[k]; --k) {
      const long long x = o[k];
      o[k] = o[k - 1];
      o[k - 1] = x;
    }
  // the distinct sums, kept sorted (binary search, then insertion)
  int n2 = 0;
  for (int i = 0; i < n; ++i)
    for (int e = 0; e < n; ++e) {
      const long long v = o[i] + o[e];
      int lo = 0, hi = n2;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (s.off2[mid] < v)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < n2 && s.off2[lo] == v) continue;
      for (int k = n2; k > lo; --k) s.off2[k] = s.off2[k - 1];
      s.off2[lo] = v;
      ++n2;
    }
  s.nd = n;
  s.nd2 = n2;
  for (int k = 0; k < n2; ++k) s.sOfS2[k] = -1;
  // S2 diagonal of every pair, then the pairs placed by t (counting sort;
  // es descending within one t)
  int tof[kDpS][kDpS] = {};
  int count[kDpS2 + 1] = {};
  for (int i = 0; i < n; ++i)
    for (int e = 0; e < n; ++e) {
      const long long v = o[i] + o[e];
      int lo = 0, hi = n2;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (s.off2[mid] < v)
          lo = mid + 1;
        else
          hi = mid;
      }
      tof[i][e] = lo;
      ++count[lo + 1];
    }
  for (int k = 0; k < n2; ++k) count[k + 1] += count[k];
  for (int i = n - 1; i >= 0; --i)
    for (int e = 0; e < n; ++e) {
      const int q = count[tof[i][e]]++;
      s.pt[q] = (unsigned char)tof[i][e];
      s.pes[q] = (unsigned char)i;
      s.pea[q] = (unsigned char)e;
    }
  s.np = n * n;
  for (int i = 0; i < n; ++i) {
    s.off[i] = o[i];
    s.s2OfS[i] = (signed char)tof[i][n / 2];  // o[n / 2] = 0
    s.sOfS2[tof[i][n / 2]] = (signed char)i;
  }
  return s;
}

// The real offsets of S (= A's) and S2 (kernel arguments).
struct DiaPassArgs {
  int offS[kDpS];
  int offS2[kDpS2];
};

// The diagonal pass's grid is at most kDiaMaxGrid blocks, so that a block of
// kBlk threads sums the previous pass's partials with every load in flight
// at once (kDiaMaxGrid / kBlk per thread and quantity, branch-free: the
// clamped extra loads are not added).
constexpr int kDiaMaxGrid = 1024;
constexpr int kDiaPartials = kDiaMaxGrid / kBlk;

// Sums of the 6 quantities over G <= kDiaMaxGrid block partials, the same in
// every block that calls it (one fixed order).
__device__ __forceinline__ void SumDiaPartials(const double* __restrict__ part, int G, double (&out)[6]) {
  __shared__ double red[6][kBlk / kWave];
  double v[6][kDiaPartials];
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int u = 0; u < kDiaPartials; ++u) v[k][u] = part[k * G + min(int(threadIdx.x) + u * kBlk, G - 1)];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < kDiaPartials; ++u)
      if (int(threadIdx.x) + u * kBlk < G) t += v[k][u];
    t = WaveSum(t);
    if (LaneId() == 0) red[k][WaveInBlock()] = t;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kBlk / kWave; ++w) t += red[k][w];
    out[k] = t;
  }
}

// MR's step between passes it - 1 and it from the summed partials (thread
// 0): ||R|| of row it - 1 (converged iff ||R|| < tol: stop 2) and alpha =
// <Pr R, Pr A Z> / ||Pr A Z||^2 (minimal_residual.cpp:58-65; breakdown: stop
// 1), as MrFusedStep.
struct MrStepOut {
  int stop;
  double rn, alpha;
};
__device__ MrStepOut MrStepValues(const double (&s)[6], double tol) {
  MrStepOut r{0, sqrt(s[5]), 0.0};
  if (r.rn < tol) {
    r.stop = 2;
    return r;
  }
  const double num = s[0], den = s[1];
  if (!(isfinite(den) && fabs(den) > 0.0)) {
    r.stop = 1;
    return r;
  }
  r.alpha = num / den;
  if (!isfinite(r.alpha)) r.stop = 1;
  return r;
}

// One LOMR iteration (kMr false; kInit: R = I - A M, Z = Pr R on S, P = 0,
// and the first traces) on the diagonal layout, the work of FusedPass: M' =
// M + delta Z + gamma P, P' = beta P + Z on S; R' = I - A M' on S2 (||R'||^2
// into partial 5); Z' = Pr R' on S; then A [Z' P'] and the traces of the
// next step (LomrProjectEpi order).  With kMr, one MR iteration: M' = M +
// alpha Z, R' = I - A M', Z' = Pr R', A Z' and MrProjectEpi's two traces.
// R' is formed from M' rather than updated as R - delta A Z - gamma A P
// (locally_minimal_residual.cpp:92-107; MR recomputes it as the reference
// does, minimal_residual.cpp:31-93): the two are equal in exact arithmetic,
// and R (25 diagonals for a 3-D Laplacian or a 9-point box) is never
// stored -- the pass reads and writes only M, Z (and P).  A pass after the
// first also does the step kernel's work (history row, stop decision, the
// step of its iteration) from the previous pass's partials, so no kernel runs
// between passes.
template <class V, bool kInit, int kShape, bool kOneCol, bool kMr>
__global__ __launch_bounds__(kBlk) void DiaPass(int n, int64_t ld, DiaPassArgs g, const unsigned* __restrict__ maskS,
                                                const V* __restrict__ av, const V* __restrict__ pr,
                                                V* __restrict__ M, V* __restrict__ Z, V* __restrict__ P,
                                                const double* __restrict__ part_in, int g_in, int it, double tol,
                                                double* __restrict__ hist, double* __restrict__ part_out,
                                                Flags* __restrict__ flags) {
#pragma clang fp contract(fast)
  constexpr DiaShape sh = MakeDiaShape(kShape);
  constexpr int ND = sh.nd, NP = sh.np;
  constexpr bool kDir = !kMr;  // LOMR's direction P
  const size_t ns = size_t(ld);
  const int j0 = blockIdx.x * kBlk + threadIdx.x;
  // a column's mask and M, Z, P; with kOneCol (the grid covers n, one column
  // a thread) loaded before the stop flag and the step values arrive
  unsigned mS = 0;
  V mo[ND], zo[ND], po[ND];
  auto load = [&](int j) {
    mS = maskS[j];
#pragma unroll
    for (int e = 0; e < ND; ++e) {
      const size_t at = size_t(e) * ns + size_t(j);
      mo[e] = M[at];
      zo[e] = kInit ? V(0) : Z[at];
      po[e] = (kInit || !kDir) ? V(0) : P[at];
    }
  };
  if constexpr (kOneCol) load(min(j0, n - 1));
  if (flags->stop) return;
  V d = V(0), gm = V(0), be = V(0);
  if constexpr (!kInit) {
    // The step of iteration it (LomrFusedStep, MrFusedStep) from the
    // previous pass's partials, in every block: the same sums in the same
    // order, so the same values everywhere.  Block 0 stores the history row,
    // the iteration count and the stop flag (a block of this pass that reads
    // the flag once set stops, as it would from its own values).
    __shared__ double step[3];
    __shared__ int stop;
    double s[6];
    SumDiaPartials(part_in, g_in, s);
    if (threadIdx.x == 0) {
      int st = 0;
      double rn = 0.0;
      if constexpr (kMr) {
        const MrStepOut r = MrStepValues(s, tol);
        st = r.stop;
        rn = r.rn;
        step[0] = r.alpha;
        step[1] = step[2] = 0.0;
      } else {
        const LomrStep r = LomrStepValues(s, it, tol);
        st = r.stop;
        rn = r.rn;
        step[0] = r.delta;
        step[1] = r.gamma;
        step[2] = r.beta;
      }
      if (blockIdx.x == 0) {
        hist[it - 1] = rn;
        flags->iterations = it - 1;
        if (st) flags->stop = st;
      }
      stop = st;
    }
    __syncthreads();
    if (stop) return;
    d = V(step[0]);  // MR: alpha
    gm = V(step[1]);
    be = V(step[2]);
  }
  double red[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = j0; j < n; j += gridDim.x * kBlk) {
    if constexpr (!kOneCol) load(j);
    int kk[ND];  // S rows j - off (clamped: out of range only where S has no entry)
    V mv[ND], pn[ND], zn[ND];
#pragma unroll
    for (int e = 0; e < ND; ++e) {
      kk[e] = min(max(j - g.offS[e], 0), n - 1);
      const size_t at = size_t(e) * ns + size_t(j);
      mv[e] = mo[e];
      pn[e] = V(0);
      if constexpr (!kInit) {
        V m = mv[e];
        m = m + d * zo[e];
        if constexpr (kDir) {
          m = m + gm * po[e];
          pn[e] = po[e] * be + zo[e];
        }
        mv[e] = m;
        M[at] = m;
      }
    }
    // A's values of the column's pairs (es, ea): FP32 with one column a
    // thread and at most 49 pairs loads all of them once and holds them for
    // both products; otherwise they are loaded where they are used (holding
    // them would cost waves per SIMD)
    constexpr bool kHold = sizeof(V) == 4 && kOneCol && NP <= 49;
    constexpr int NH = kHold ? ND : 1;
    V aval[NH][NH];
    if constexpr (kHold) {
#pragma unroll
      for (int es = 0; es < ND; ++es)
#pragma unroll
        for (int ea = 0; ea < ND; ++ea) aval[es][ea] = av[size_t(ea) * ns + size_t(kk[es])];
    }
    auto a_of = [&](int es, int ea) -> V {
      if constexpr (kHold)
        return aval[es][ea];
      else
        return av[size_t(ea) * ns + size_t(kk[es])];
    };
    // Z' = Pr R' on S: R' = I - A M' on the S2 diagonals that are S's
    {
      V am = V(0);
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const int t = sh.pt[q], es = sh.pes[q], ea = sh.pea[q];
        if (sh.sOfS2[t] < 0) continue;
        am += a_of(es, ea) * mv[es];
        if (q + 1 < NP && sh.pt[q + 1] == t) continue;
        V r = -am;
        if (sh.off2[t] == 0) r += V(1);
#pragma unroll
        for (int e = 0; e < ND; ++e)
          if (sh.sOfS2[t] == e) zn[e] = ((mS >> e) & 1u) ? pr[kk[e]] * r : V(0);
        am = V(0);
      }
    }
#pragma unroll
    for (int e = 0; e < ND; ++e) {
      const size_t at = size_t(e) * ns + size_t(j);
      Z[at] = zn[e];
      if constexpr (kDir) P[at] = pn[e];
    }
    // R' = I - A M' on all of S2 (the same terms in the same order, so the
    // same values on S's diagonals), ||R'||^2, A [Z' P'] and the traces
    // (LomrProjectEpi / MrProjectEpi order)
    {
      V am = V(0), az = V(0), ap = V(0);
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const int t = sh.pt[q], es = sh.pes[q], ea = sh.pea[q];
        const V a = a_of(es, ea);
        am += a * mv[es];
        az += a * zn[es];
        if constexpr (kDir) ap += a * pn[es];
        if (q + 1 < NP && sh.pt[q + 1] == t) continue;
        V r = -am;
        if (sh.off2[t] == 0) r += V(1);
        red[5] += double(r) * double(r);
        const V p2 = pr[min(max(j - g.offS2[t], 0), n - 1)];
        if constexpr (kMr) {
          const double tmp = double(p2 * az);
          red[0] += double(p2 * r) * tmp;
          red[1] += tmp * tmp;
        } else {
          const double daz = az, dap = ap, dpr = p2, dprr = double(p2 * r);
          red[0] += daz * (dpr * daz);
          red[1] += dap * (dpr * dap);
          red[2] += daz * (dpr * dap);
          red[3] += dprr * daz;
          red[4] += dprr * dap;
        }
        am = az = ap = V(0);
      }
    }
  }
  StorePartials<6>(red, part_out);
}

// Diagonal layout construction (columns): flag[j - i + n - 1] = 1 for every
// entry (i, j) of a CSC pattern.
__global__ void ColDiaMark(int n, const int* __restrict__ ptr, const int* __restrict__ idx, int* __restrict__ flag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int p = ptr[j]; p < ptr[j + 1]; ++p) flag[int64_t(j) - idx[p] + n - 1] = 1;
}

// off[pos[i]] = i - (n - 1) for every flagged i.
__global__ void ColDiaOffsets(int64_t len, int n, const int* __restrict__ flag, const int* __restrict__ pos,
                              int* __restrict__ off) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len && flag[i]) off[pos[i]] = int(i - (n - 1));
}

// Per column: the CSC position of each diagonal slot (-1 for none) and the mask.
__global__ void ColDiaFill(int n, int nd, const int* __restrict__ ptr, const int* __restrict__ idx,
                           const int* __restrict__ pos, int* __restrict__ src, unsigned* __restrict__ mask) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int e = 0; e < nd; ++e) src[size_t(e) * n + j] = -1;
  unsigned m = 0;
  for (int p = ptr[j]; p < ptr[j + 1]; ++p) {
    const int e = pos[int64_t(j) - idx[p] + n - 1];
    src[size_t(e) * n + j] = p;
    m |= 1u << e;
  }
  mask[j] = m;
}

// Diagonal slot k = d n + j of a layout with leading dimension ld (d ld + j).
__device__ __forceinline__ int64_t DiaSlot(int64_t k, int n, int64_t ld) { return (k / n) * ld + k % n; }

// out[slot k] = src[k] >= 0 ? v[src[k]] : 0 (CSC values into diagonal slots).
template <class VI, class VO>
__global__ void ToDiagonals(int64_t len, int n, int64_t ld, const int* __restrict__ src, const VI* __restrict__ v,
                            VO* __restrict__ out) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < len) out[DiaSlot(k, n, ld)] = src[k] >= 0 ? VO(v[src[k]]) : VO(0);
}

// out[src[k]] = v[k] for the slots that hold an entry (diagonals back to CSC).
template <class V>
__global__ void FromDiagonals(int64_t len, int n, int64_t ld, const int* __restrict__ src, const V* __restrict__ v,
                              V* __restrict__ out) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < len && src[k] >= 0) out[src[k]] = v[DiaSlot(k, n, ld)];
}

// Leading dimension of the diagonal pass's arrays: n rounded up to 64 values
// plus kDiaPad, so that the diagonals do not start a power of two apart.
constexpr int64_t kDiaPad = 0;
int64_t DiaStride(int64_t n) { return (n + 63) / 64 * 64 + kDiaPad; }

// Column-diagonal layout of a CSC pattern: offsets (host), masks and slot
// sources (device; only the offsets when slots is false); nd = 0 when it has
// more than max_nd diagonals or more than 1.5 nnz slots.
struct ColDia {
  int nd = 0;
  std::vector<int> off;
  DBuf<int> src;
  DBuf<unsigned> mask;
};

ColDia MakeColDia(Ctx& c, int64_t n, const Pattern& pat, int max_nd, bool slots = true) {
  ColDia out;
  const int64_t len = 2 * n - 1;
  auto flag = c.Alloc<int>(len);
  auto pos = c.Alloc<int>(len + 1);
  SPAI_HIP_CHECK(hipMemsetAsync(flag.get(), 0, sizeof(int) * len, c.stream));
  ColDiaMark<<<Blocks1(n), kBlk, 0, c.stream>>>((int)n, pat.ptr.get(), pat.idx.get(), flag.get());
  const int64_t nd = ScanCounts(c, flag.get(), len, pos.get());
  if (nd > max_nd || double(nd) * double(n) > 1.5 * double(pat.nnz)) return out;
  auto off = c.Alloc<int>(nd);
  ColDiaOffsets<<<Blocks1(len), kBlk, 0, c.stream>>>(len, (int)n, flag.get(), pos.get(), off.get());
  out.off.resize(nd);
  c.ReadBack(out.off.data(), off.get(), sizeof(int) * nd);
  out.nd = (int)nd;
  if (!slots) return out;
  out.src = c.Alloc<int>(nd * n);
  out.mask = c.Alloc<unsigned>(n);
  ColDiaFill<<<Blocks1(n), kBlk, 0, c.stream>>>((int)n, (int)nd, pat.ptr.get(), pat.idx.get(), pos.get(),
                                                out.src.get(), out.mask.get());
  out.nd = (int)nd;
  return out;
}

template <class V>
class FixedImpl final : public FixedProblem {
 public:
  FixedImpl(int64_t method, const MatrixData& Ah, const MatrixData* M0h, const MatrixData* Prh,
            const EngineParams& p)
      : method_(method), prm_(p) {
    if (method < SPAI_GLOBAL_CG || method > SPAI_INNER_OUTER_LOMR) throw std::invalid_argument("unknown method");
    if (method >= SPAI_INNER_OUTER_MR && p.inner_iterations <= 0)
      throw std::invalid_argument("inner_iterations must be positive");
    Ctx& c = ctx_;
    if (Ah.rows != Ah.cols || Ah.rows == 0) throw std::invalid_argument("A must be a non-empty square matrix");
    n_ = Ah.cols;
    if (M0h && (M0h->rows != n_ || M0h->cols != n_))
      throw std::invalid_argument("Initial preconditioner M0 must have the same shape as A");
    if (Prh && (Prh->rows != n_ || Prh->cols != n_)) throw std::invalid_argument("Pr must have the same shape as A");
    if (p.max_iterations <= 0) throw std::invalid_argument("max_iterations must be positive");
    if (!std::isfinite(p.tolerance) || p.tolerance < 0.0)
      throw std::invalid_argument("tolerance must be finite and non-negative");
    if (!std::isfinite(p.max_density) || p.max_density <= 1e-9 || p.max_density > 1.0)
      throw std::invalid_argument("max_density must be finite and in the interval (0, 1]");
    if (p.enable_dropping) throw std::invalid_argument("fixed_pattern requires enable_dropping = 0");
    if (p.num_threads <= 0) throw std::invalid_argument("Invalid threads count value");

    A_ = Upload(c, Ah.rows, Ah.cols, Ah.col_offsets, Ah.row_indices, Ah.values);
    DMat AT = Transpose(c, A_);
    at_pat_ = AT.pat;
    at_d_ = std::move(AT.val);
    // Pr: diagonal (given, or diag(A)^-1 as methods_common.cpp:50-63).
    pr_d_ = c.Alloc<double>(n_);
    if (Prh) {
      DMat Pr = Upload(c, Prh->rows, Prh->cols, Prh->col_offsets, Prh->row_indices, Prh->values);
      if (!IsFullDiagonal(c, Pr)) throw std::invalid_argument("fixed_pattern requires a diagonal Pr");
      ExtractDiagonal(c, Pr, pr_d_.get());
    } else {
      ExtractDiagonal(c, A_, pr_d_.get());
    }
    if (M0h) m0_ = Upload(c, M0h->rows, M0h->cols, M0h->col_offsets, M0h->row_indices, M0h->values);
    // one read-back decides the patterns and checks the diagonal
    int facts[4] = {0, 0, 0, 0};
    {
      auto fd = c.Alloc<int>(4);
      SPAI_HIP_CHECK(hipMemsetAsync(fd.get(), 0, 4 * sizeof(int), c.stream));
      PatternFacts<<<Blocks1(n_), kBlk, 0, c.stream>>>((int)n_, A_.ptr(), A_.idx(), M0h ? m0_.ptr() : nullptr,
                                                       M0h ? m0_.idx() : nullptr, Prh ? nullptr : pr_d_.get(),
                                                       fd.get());
      c.ReadBack(facts, fd.get(), sizeof(facts));
    }
    if (facts[2]) throw std::invalid_argument("A has an invalid diagonal entry; provide Pr explicitly");
    if (facts[3]) throw std::invalid_argument("A diagonal cannot be inverted safely; provide Pr explicitly");
    // S = pattern(A) | pattern(M0) | I and S2 = pattern(A S) | S | I.  When A
    // has its whole diagonal and M0 (if given) none off it, S is pattern(A)
    // and pattern(A S) already holds S (each (i, j) of S meets A(i, i)), so
    // the unions are skipped.
    if (!M0h) m0_ = Identity(c, n_);
    if (!facts[0] && !facts[1]) {
      s_pat_ = A_.pat;
      DMat AS = SpGemm(c, A_, A_, false, true);  // only the pattern is kept
      s2_pat_ = AS.pat;
    } else {
      DMat I = Identity(c, n_);
      DMat S = Merge(c, A_, I, MergeOp::kCopyLeft, 0, 0, false);
      if (M0h) S = Merge(c, S, m0_, MergeOp::kCopyLeft, 0, 0, false);
      s_pat_ = S.pat;
      DMat AS = SpGemm(c, A_, S, false, true);
      DMat S2 = Merge(c, AS, S, MergeOp::kCopyLeft, 0, 0, false);
      s2_pat_ = S2.pat;
    }
    nnz_s_ = s_pat_->nnz;
    nnz_s2_ = s2_pat_->nnz;
    s2col_ = c.Alloc<int>(nnz_s2_);
    ColumnOfEntry<<<Blocks1(n_), kBlk, 0, c.stream>>>((int)n_, s2_pat_->ptr.get(), s2col_.get());
    s2_to_s_ = c.Alloc<int>(nnz_s2_);
    SPAI_HIP_CHECK(hipMemsetAsync(s2_to_s_.get(), 0xFF, sizeof(int) * nnz_s2_, c.stream));
    // a miss is reported through the context's error code at its next
    // synchronisation
    MapSubPattern<<<Blocks1(n_), kBlk, 0, c.stream>>>((int)n_, s_pat_->ptr.get(), s_pat_->idx.get(),
                                                      s2_pat_->ptr.get(), s2_pat_->idx.get(), s2_to_s_.get(),
                                                      c.err_dev);
    // slot plan of the products (fixed pattern)
    {
      SlotPlan plan = BuildSlotPlan(c, n_, s_pat_->ptr.get(), s_pat_->idx.get(), A_.ptr(), A_.idx(),
                                    s2_pat_->ptr.get(), s2_pat_->idx.get());
      const int64_t total = plan.terms;
      const int h[3] = {plan.max_a_col, plan.max_s2_col, plan.max_s_col};
      plan_ptr_ = std::move(plan.ptr);
      maxcol_ = std::max(h[0], 1);
      lmax_ = std::max(h[1], 1);
      gl_ = maxcol_ <= 8 ? 8 : maxcol_ <= 16 ? 16 : maxcol_ <= 32 ? 32 : 64;
      smem_ = size_t(kBlk / gl_) * 2 * size_t(lmax_) * sizeof(V);
      slot16_ = plan.slot16;
      use_plan_ = plan.built && smem_ <= 64 * 1024;
      if (use_plan_) {
        plan8_ = std::move(plan.plan8);
        plan16_ = std::move(plan.plan16);
        const int64_t groups = kBlk / gl_;
        gridp_ = (unsigned)std::min<int64_t>(std::max<int64_t>((n_ + groups - 1) / groups, 1), kMaxGrid);
        // fused LOMR pass: W = 64 / GL columns per wavefront
        smax_ = std::max(h[2], 1);
        // LDS under 64 KiB with the kernel's static part; the tile tables
        // need tiles of at most 65535 S2 entries.  The stash (product 2
        // reuses product 1's terms) needs every column of A within GL and
        // is taken while it leaves room for 4 blocks per CU (160 KiB).
        const int vsize = int(sizeof(V));
        const size_t waves = kBlk / kWave;
        const size_t with = waves * size_t(MakeFusedLayout(lmax_, smax_, gl_, vsize, true).bytes);
        const size_t without = waves * size_t(MakeFusedLayout(lmax_, smax_, gl_, vsize, false).bytes);
        // every column within kFusedPre steps and every column of A within
        // GL (FP32, byte slots, GL <= 32): product 2 always from registers
        const bool regs_only =
            sizeof(V) == 4 && maxcol_ <= gl_ && smax_ <= kFusedPre && !plan.slot16 && gl_ <= 32;
        terms_ = regs_only ? kTermsRegs : (maxcol_ <= gl_ && with <= 40 * 1024) ? kTermsStash : kTermsReload;
        smemf_ = terms_ == kTermsStash ? with : without;
        fused_ok_ = smemf_ <= 60 * 1024 && int64_t(kWave / gl_) * lmax_ + 16 <= 65535 &&
                    total < (int64_t(1) << 32);
        // the pass walks its tiles with a grid that the device holds
        // resident at once (blocks per CU of the iteration kernel x CUs)
        const int64_t ntiles = (n_ + (kWave / gl_) - 1) / (kWave / gl_);
        int64_t resident = kMaxFusedGrid;
        if (fused_ok_) {
          int dev = 0, cus = 0;
          SPAI_HIP_CHECK(hipGetDevice(&dev));
          SPAI_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
          const int per_cu = FusedBlocksPerCu(method_ == SPAI_GLOBAL_MR);
          if (per_cu < 1) throw std::runtime_error("fused pass: no block fits on a CU");
          resident = std::min<int64_t>(int64_t(per_cu) * cus, kMaxFusedGrid);
        }
        gridf_ = (unsigned)std::min<int64_t>(std::max<int64_t>((ntiles + kBlk / kWave - 1) / (kBlk / kWave), 1),
                                             resident);
      }
      av_ = c.Alloc<V>(A_.nnz());
      ConvertValues<V><<<Blocks1(A_.nnz()), kBlk, 0, c.stream>>>(A_.nnz(), A_.v(), av_.get());
    }
    if (method_ >= SPAI_INNER_OUTER_MR) {
      if (!use_plan_) throw std::runtime_error("inner-outer fixed_pattern needs S2 columns of at most 65535 rows");
      // S^T with positions (rows of M for the sampled product M_outer R)
      DMat P;
      P.pat = s_pat_;
      P.val = c.Alloc<double>(nnz_s_);
      IotaD<<<Blocks1(nnz_s_), kBlk, 0, c.stream>>>(nnz_s_, P.val.get());
      DMat T = Transpose(c, P);
      st_pat_ = T.pat;
      stperm_ = c.Alloc<int>(nnz_s_);
      ToIntD<<<Blocks1(nnz_s_), kBlk, 0, c.stream>>>(nnz_s_, T.val.get(), stperm_.get());
      scol_ = c.Alloc<int>(nnz_s_);
      ColumnOfEntry<<<Blocks1(n_), kBlk, 0, c.stream>>>((int)n_, s_pat_->ptr.get(), scol_.get());
      Mout_ = c.Alloc<V>(nnz_s_);
      Q_ = c.Alloc<V>(nnz_s_);
      io_cols_ = c.Alloc<double>(3 * n_);
    }
    // value arrays
    atv_ = c.Alloc<V>(at_pat_->nnz);
    ConvertValues<V><<<Blocks1(at_pat_->nnz), kBlk, 0, c.stream>>>(at_pat_->nnz, at_d_.get(), atv_.get());
    prv_ = c.Alloc<V>(n_);
    ConvertValues<V><<<Blocks1(n_), kBlk, 0, c.stream>>>(n_, pr_d_.get(), prv_.get());
    if (fused_ok_ && (method_ == SPAI_GLOBAL_LOMR || method_ == SPAI_GLOBAL_MR)) {
      a0c_ = c.Alloc<int2>(nnz_s_);
      SEntryA<<<Blocks1(nnz_s_), kBlk, 0, c.stream>>>(nnz_s_, s_pat_->idx.get(), A_.ptr(), a0c_.get());
      stq_ = c.Alloc<uint16_t>(nnz_s_);
      s2d_ = c.Alloc<int>(n_);
      TileInfo<<<Blocks1(nnz_s2_), kBlk, 0, c.stream>>>(nnz_s2_, kWave / gl_, 16 / int(sizeof(V)), s2_to_s_.get(),
                                                       s2col_.get(),
                                                       s2_pat_->idx.get(), s2_pat_->ptr.get(), stq_.get(),
                                                       s2d_.get());
      prS_ = c.Alloc<V>(nnz_s_);
      GatherRows<V><<<Blocks1(nnz_s_), kBlk, 0, c.stream>>>(nnz_s_, s_pat_->idx.get(), prv_.get(), prS_.get());
      prS2_ = c.Alloc<V>(nnz_s2_ + 16);  // padding: the fused pass loads whole 16-byte chunks
      GatherRows<V><<<Blocks1(nnz_s2_), kBlk, 0, c.stream>>>(nnz_s2_, s2_pat_->idx.get(), prv_.get(), prS2_.get());
    }
    if (method_ == SPAI_GLOBAL_LOMR || method_ == SPAI_GLOBAL_MR) SetUpDiaPass();
    M_ = c.Alloc<V>(nnz_s_);
    Z_ = c.Alloc<V>(nnz_s_);
    P_ = c.Alloc<V>(nnz_s_);
    R_ = c.Alloc<V>(nnz_s2_ + 16);  // padding: the fused pass loads whole 16-byte chunks
    AZ_ = c.Alloc<V>(nnz_s2_);
    AP_ = c.Alloc<V>(nnz_s2_);
    grid2_ = GridFor(nnz_s2_);
    grid1_ = GridFor(nnz_s_);
    part_ = c.Alloc<double>(8 * kMaxGrid);
    partf_ = c.Alloc<double>(6 * size_t(kMaxFusedGrid));
    sc_ = c.Alloc<double>(kNumSlots);
    flags_ = c.Alloc<Flags>(1);
    hist_ = c.Alloc<double>(p.max_iterations + 1);
    c.Sync();
  }

  Ctx& context() override { return ctx_; }
  const DMat& MatrixA() const override { return A_; }

  // Fixed-order host sum of the G block partials of slot k.
  double HostPartialSum(int G, int k) {
    std::vector<double> h(G);
    ctx_.ReadBack(h.data(), part_.get() + size_t(k) * G, sizeof(double) * G);
    double t = 0.0;
    for (double v : h) t += v;
    return t;
  }

  // inner_outer::Run (src/inner_outer/inner_outer_common.hpp:14-107) on the
  // fixed pattern: Z = M_outer R sampled on S, per-column step sizes,
  // StabilizeResiduals (state_manager.cpp:81-95).  Host-driven (one norm
  // read-back per inner iteration).
  int64_t BuildInnerOuter() {
    Ctx& c = ctx_;
    hipStream_t st = c.stream;
    const bool lomr = method_ == SPAI_INNER_OUTER_LOMR;
    const int* map = s2_to_s_.get();
    SampleOnPattern<V><<<Blocks1(n_ * kSampleLanes), kBlk, 0, st>>>(
        (int)n_, s_pat_->ptr.get(), s_pat_->idx.get(), m0_.ptr(), m0_.idx(), m0_.v(), M_.get());
    SPAI_HIP_CHECK(hipMemsetAsync(flags_.get(), 0, sizeof(Flags), st));
    const double tol = prm_.tolerance;
    const double thr = std::max(10.0 * tol, 64.0 * 2.220446049250313e-16 * std::sqrt((double)n_));
    // FP32 data: pivots of the scaled per-column Gram below sqrt(FLT_EPSILON)
    // of the largest are noise (cos(az, aq) within ~1.5 degrees of +-1).
    const double rank_eps = std::is_same<V, float>::value ? 3.4526698300124393e-04 : 2.220446049250313e-16;
    auto residual = [&]() {
      Slot<1>(ResidualEpi<V>{map, prv_.get(), R_.get(), Z_.get(), P_.get(), 0}, M_.get(), nullptr);
      return std::sqrt(HostPartialSum((int)gridp_, 0));
    };
    io_hist_.clear();
    io_converged_ = false;
    double rn = residual();
    io_hist_.push_back(rn);
    if (rn < tol) {
      io_converged_ = true;
      return 0;
    }
    double* delta = io_cols_.get();
    double* gamma = delta + n_;
    double* norm = gamma + n_;
    const int* s2p = s2_pat_->ptr.get();
    for (int64_t it = 0; it < prm_.max_iterations;) {
      SPAI_HIP_CHECK(hipMemcpyAsync(Mout_.get(), M_.get(), sizeof(V) * nnz_s_, hipMemcpyDeviceToDevice, st));
      residual();
      FillZero<V><<<Blocks1(nnz_s_), kBlk, 0, st>>>(nnz_s_, Q_.get());
      const int64_t inner_count = std::min(prm_.inner_iterations, prm_.max_iterations - it);
      for (int64_t inner = 0; inner < inner_count; ++inner, ++it) {
        SampledOuterProduct<V><<<GridFor(nnz_s_), kBlk, 0, st>>>(
            nnz_s_, s_pat_->idx.get(), scol_.get(), st_pat_->ptr.get(), st_pat_->idx.get(), stperm_.get(),
            Mout_.get(), s2p, s2_pat_->idx.get(), R_.get(), Z_.get());
        if (lomr)
          Slot<2>(StoreTwoEpi<V>{AZ_.get(), AP_.get()}, Z_.get(), Q_.get());
        else
          Slot<1>(StoreTwoEpi<V>{AZ_.get(), nullptr}, Z_.get(), nullptr);
        IoColumnSteps<V><<<Blocks1(n_), kBlk, 0, st>>>((int)n_, s2p, R_.get(), AZ_.get(), AP_.get(), lomr ? 1 : 0,
                                                       (lomr && inner > 0) ? 1 : 0, rank_eps, delta, gamma, norm);
        IoDirection<V><<<GridFor(nnz_s_), kBlk, 0, st>>>(nnz_s_, scol_.get(), Z_.get(), Q_.get(), delta, gamma,
                                                         lomr ? 1 : 0, P_.get());
        Slot<1>(IoUpdateEpi<V>{map, nullptr, P_.get(), norm, R_.get(), M_.get(), Q_.get(), lomr ? 1 : 0}, P_.get(),
                nullptr);
        rn = std::sqrt(HostPartialSum((int)gridp_, 0));
        if (!(std::isfinite(rn) && rn > thr)) rn = residual();
        io_hist_.push_back(rn);
        if (rn < tol) {
          io_converged_ = true;
          return (int64_t)io_hist_.size() - 1;
        }
      }
    }
    return (int64_t)io_hist_.size() - 1;
  }

  ~FixedImpl() override {
    for (hipEvent_t e : events_)
      if (e) (void)hipEventDestroy(e);
    if (build_exec_) (void)hipGraphExecDestroy(build_exec_);
    if (build_graph_) (void)hipGraphDestroy(build_graph_);
  }

  void SetKernelTiming(bool on) override { timing_ = on; }

  int64_t FinishFusedBuild() {
    SPAI_HIP_CHECK(hipGetLastError());
    if (enqueue_after_) {
      const std::function<void()>* f = enqueue_after_;
      enqueue_after_ = nullptr;  // called here: BuildThen does not call it again
      (*f)();
    }
    Flags f;
    ctx_.ReadBack(&f, flags_.get(), sizeof(Flags));
    last_flags_ = f;
    return f.iterations;
  }

  double KernelMicros(int64_t* launches) override {
    if (launches) *launches = timed_launches_;
    return timed_launches_ > 0 ? timed_us_ : -1.0;
  }

  void FusedInfo(int64_t out[kFixedInfoLen]) const override {
    const bool fused = (method_ == SPAI_GLOBAL_LOMR || method_ == SPAI_GLOBAL_MR) && use_plan_ && fused_ok_;
    const int64_t w = dia_ ? kWave : kWave / gl_;
    out[0] = dia_ ? 2 : fused ? 1 : 0;
    out[1] = dia_ ? 1 : gl_;
    out[2] = w;
    out[3] = (n_ + w - 1) / w;
    out[4] = dia_ ? gridd_ : gridf_;
    out[5] = kBlk / kWave;
    out[6] = nnz_s_;
    out[7] = nnz_s2_;
    out[8] = dia_ ? dnd_ : 0;
    out[9] = dia_ ? dnd_ : 0;
    out[10] = dia_ ? dnd2_ : 0;
  }

  // The diagonal pass (DiaPass) for Global LOMR when A and S lie on the
  // diagonals of a 1-, 2- or 3-dimensional Laplacian stencil, and for Global
  // MR on those or a 2-dimensional 9-point box stencil's, with S2 on all their
  // sums, in the order of MakeDiaShape, each layout with at most 1.5 nnz
  // slots.  The environment variable SPAI_FIXED_LAYOUT ("csr", or "auto" the
  // default) can keep the general fused pass.
  void SetUpDiaPass() {
    const char* env = std::getenv("SPAI_FIXED_LAYOUT");
    if (env && *env && std::strcmp(env, "auto") != 0) {
      if (std::strcmp(env, "csr") == 0) return;
      throw std::invalid_argument("SPAI_FIXED_LAYOUT must be csr or auto");
    }
    Ctx& c = ctx_;
    ColDia a = MakeColDia(c, n_, *A_.pat, kDpS);
    if (!a.nd) return;
    ColDia s = MakeColDia(c, n_, *s_pat_, kDpS);
    if (s.nd != a.nd || s.off != a.off) return;
    ColDia s2 = MakeColDia(c, n_, *s2_pat_, kDpS2, false);
    if (!s2.nd) return;
    const bool mr = method_ == SPAI_GLOBAL_MR;
    int kind = 0;
    for (int k : {1, 2, 3, 5}) {
      if (k > 3 && !mr) break;  // box stencils: MR only
      if (MatchesShape(MakeDiaShape(k), s.off, s2.off)) {
        kind = k;
        break;
      }
    }
    if (!kind) return;
    for (int e = 0; e < s.nd; ++e) dargs_.offS[e] = s.off[e];
    for (int t = 0; t < s2.nd; ++t) dargs_.offS2[t] = s2.off[t];
    const int64_t la = int64_t(a.nd) * n_;
    dld_ = DiaStride(n_);
    dav_ = c.Alloc<V>(int64_t(a.nd) * dld_);
    ToDiagonals<double, V><<<Blocks1(la), kBlk, 0, c.stream>>>(la, (int)n_, dld_, a.src.get(), A_.v(), dav_.get());
    dM_ = c.Alloc<V>(int64_t(s.nd) * dld_);
    dZ_ = c.Alloc<V>(int64_t(s.nd) * dld_);
    if (!mr) dP_ = c.Alloc<V>(int64_t(s.nd) * dld_);
    dmask_s_ = std::move(s.mask);
    dsrc_s_ = std::move(s.src);
    partf2_ = c.Alloc<double>(6 * size_t(kMaxFusedGrid));
    ddims_ = kind;
    dnd_ = s.nd;
    dnd2_ = s2.nd;
    gridd_ = (unsigned)std::min<int64_t>(std::max<int64_t>((n_ + kBlk - 1) / kBlk, 1), kDiaMaxGrid);
    dia_ = true;
  }

  // Whether the offsets of S (ascending, 0 in the middle) and S2 pair up as
  // in shape sh: off[es] + off[ea] is S2 diagonal t for every pair (t, es,
  // ea) of the shape.
  static bool MatchesShape(const DiaShape& sh, const std::vector<int>& off, const std::vector<int>& off2) {
    if ((int)off.size() != sh.nd || (int)off2.size() != sh.nd2 || off[sh.nd / 2] != 0) return false;
    for (int q = 0; q < sh.np; ++q)
      if (off2[sh.pt[q]] != int64_t(off[sh.pes[q]]) + off[sh.pea[q]]) return false;
    return true;
  }

  // Global LOMR or MR with one fused pass per iteration (FusedPass, or
  // DiaPass on the diagonal layout).
  int64_t BuildFused() {
    Ctx& c = ctx_;
    hipStream_t st = c.stream;
    const bool mr = method_ == SPAI_GLOBAL_MR;
    timed_launches_ = 0;
    if (timing_ && (int64_t)events_.size() < 2 * prm_.max_iterations) {
      for (hipEvent_t e : events_) (void)hipEventDestroy(e);
      events_.assign(2 * prm_.max_iterations, nullptr);
      for (auto& e : events_) SPAI_HIP_CHECK(hipEventCreate(&e));
    }
    // untimed builds replay one recorded graph of the whole launch sequence
    const bool record = !timing_ && !build_exec_;
    if (!timing_ && build_exec_) {
      SPAI_HIP_CHECK(hipGraphLaunch(build_exec_, st));
      return FinishFusedBuild();
    }
    if (record) SPAI_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    SampleOnPattern<V><<<Blocks1(n_ * kSampleLanes), kBlk, 0, st>>>(
        (int)n_, s_pat_->ptr.get(), s_pat_->idx.get(), m0_.ptr(), m0_.idx(), m0_.v(), M_.get());
    const int64_t ls = int64_t(dnd_) * n_;
    if (dia_)
      ToDiagonals<V, V><<<Blocks1(ls), kBlk, 0, st>>>(ls, (int)n_, dld_, dsrc_s_.get(), M_.get(), dM_.get());
    ResetFlags<<<1, 1, 0, st>>>(flags_.get());
    double* part = partf_.get();
    const double tol = prm_.tolerance;
    const int G = int(dia_ ? gridd_ : gridf_);
    // the diagonal pass computes the step itself, from the partials of the
    // pass before (two buffers, written in turn)
    double* const pbuf[2] = {partf_.get(), partf2_.get()};
    for (int64_t it = 0; it <= prm_.max_iterations; ++it) {
      if (it > 0 && mr && !dia_)
        MrFusedStep<<<1, kStepBlk, 0, st>>>(part, G, (int)it, tol, hist_.get(), sc_.get(), flags_.get());
      else if (it > 0 && !dia_)
        LomrFusedStep<<<1, kStepBlk, 0, st>>>(part, G, (int)it, tol, hist_.get(), sc_.get(), flags_.get());
#define SPAI_FUSED_STASH(SlotT, plan, GLV, INIT, MR, STASH)                                                   \
  FusedPass<V, SlotT, GLV, INIT, MR, STASH><<<gridf_, kBlk, smemf_, st>>>(                                     \
      (int)n_, s_pat_->ptr.get(), a0c_.get(), stq_.get(), s2d_.get(), s2_pat_->ptr.get(), av_.get(),           \
      plan_ptr_.get(), plan, lmax_, smax_, prS_.get(), prS2_.get(), R_.get(), M_.get(), Z_.get(), P_.get(),    \
      sc_.get(), part, flags_.get(), maxcol_ <= gl_ ? 1 : 0)
#define SPAI_FUSED_ONE(SlotT, plan, GLV, INIT, MR)                          \
  do {                                                                      \
    if constexpr (sizeof(V) == 4 && sizeof(SlotT) == 1 && GLV <= 32) {     \
      if (terms_ == kTermsRegs) {                                           \
        SPAI_FUSED_STASH(SlotT, plan, GLV, INIT, MR, kTermsRegs);           \
        break;                                                              \
      }                                                                     \
    }                                                                       \
    if (terms_ == kTermsStash)                                              \
      SPAI_FUSED_STASH(SlotT, plan, GLV, INIT, MR, kTermsStash);            \
    else                                                                    \
      SPAI_FUSED_STASH(SlotT, plan, GLV, INIT, MR, kTermsReload);           \
  } while (0)
#define SPAI_FUSED_LAUNCH(SlotT, plan, GLV)                     \
  do {                                                          \
    if (mr) {                                                   \
      if (it == 0)                                              \
        SPAI_FUSED_ONE(SlotT, plan, GLV, true, true);           \
      else                                                      \
        SPAI_FUSED_ONE(SlotT, plan, GLV, false, true);          \
    } else {                                                    \
      if (it == 0)                                              \
        SPAI_FUSED_ONE(SlotT, plan, GLV, true, false);          \
      else                                                      \
        SPAI_FUSED_ONE(SlotT, plan, GLV, false, false);         \
    }                                                           \
  } while (0)
#define SPAI_FUSED_GL(SlotT, plan)                          \
  switch (gl_) {                                            \
    case 8: SPAI_FUSED_LAUNCH(SlotT, plan, 8); break;       \
    case 16: SPAI_FUSED_LAUNCH(SlotT, plan, 16); break;     \
    case 32: SPAI_FUSED_LAUNCH(SlotT, plan, 32); break;     \
    default: SPAI_FUSED_LAUNCH(SlotT, plan, 64); break;     \
  }
      const bool timed = timing_ && it > 0;
      if (timed) SPAI_HIP_CHECK(hipEventRecord(events_[2 * (it - 1)], st));
      if (dia_) {
        if (it == 0)
          LaunchDiaPass<true>(0, nullptr, pbuf[0]);
        else
          LaunchDiaPass<false>((int)it, pbuf[(it - 1) & 1], pbuf[it & 1]);
      } else if (slot16_) {
        SPAI_FUSED_GL(uint16_t, plan16_.get());
      } else {
        SPAI_FUSED_GL(uint8_t, plan8_.get());
      }
      if (timed) SPAI_HIP_CHECK(hipEventRecord(events_[2 * (it - 1) + 1], st));
#undef SPAI_FUSED_GL
#undef SPAI_FUSED_LAUNCH
#undef SPAI_FUSED_ONE
#undef SPAI_FUSED_STASH
    }
    LomrFusedFinish<<<1, kStepBlk, 0, st>>>(dia_ ? pbuf[prm_.max_iterations & 1] : part, G,
                                            (int)prm_.max_iterations, tol, hist_.get(), flags_.get());
    if (dia_) FromDiagonals<V><<<Blocks1(ls), kBlk, 0, st>>>(ls, (int)n_, dld_, dsrc_s_.get(), dM_.get(), M_.get());
    if (record) {
      SPAI_HIP_CHECK(hipStreamEndCapture(st, &build_graph_));
      SPAI_HIP_CHECK(hipGraphInstantiate(&build_exec_, build_graph_, nullptr, nullptr, 0));
      SPAI_HIP_CHECK(hipGraphLaunch(build_exec_, st));
      return FinishFusedBuild();
    }
    SPAI_HIP_CHECK(hipGetLastError());
    Flags f;
    c.ReadBack(&f, flags_.get(), sizeof(Flags));
    last_flags_ = f;
    if (timing_) {
      // passes 1..iterations ran in full; later ones returned at once
      double total = 0.0;
      for (int64_t it = 1; it <= f.iterations; ++it) {
        float ms = 0.f;
        SPAI_HIP_CHECK(hipEventElapsedTime(&ms, events_[2 * (it - 1)], events_[2 * (it - 1) + 1]));
        total += ms;
      }
      timed_launches_ = f.iterations;
      timed_us_ = f.iterations > 0 ? total * 1e3 / double(f.iterations) : 0.0;
    }
    return f.iterations;
  }

  template <bool kInit>
  void LaunchDiaPass(int it, const double* part_in, double* part_out) {
#define SPAI_DIA_PASS(KIND, ONE, MR)                                                                          \
  DiaPass<V, kInit, KIND, ONE, MR><<<gridd_, kBlk, 0, ctx_.stream>>>(                                         \
      (int)n_, dld_, dargs_, dmask_s_.get(), dav_.get(), prv_.get(), dM_.get(), dZ_.get(), dP_.get(), part_in, \
      (int)gridd_, it, prm_.tolerance, hist_.get(), part_out, flags_.get())
#define SPAI_DIA_PASS_LAPLACE(ONE, MR)          \
  switch (ddims_) {                             \
    case 1: SPAI_DIA_PASS(1, ONE, MR); break;   \
    case 2: SPAI_DIA_PASS(2, ONE, MR); break;   \
    default: SPAI_DIA_PASS(3, ONE, MR); break;  \
  }
    const bool one = int64_t(gridd_) * kBlk >= n_;
    if (ddims_ > 3) {
      // the 9-point box (MR): one instance, any n
      SPAI_DIA_PASS(5, false, true);
    } else if (method_ == SPAI_GLOBAL_MR) {
      if (one) {
        SPAI_DIA_PASS_LAPLACE(true, true);
      } else {
        SPAI_DIA_PASS_LAPLACE(false, true);
      }
    } else if (one) {
      SPAI_DIA_PASS_LAPLACE(true, false);
    } else {
      SPAI_DIA_PASS_LAPLACE(false, false);
    }
#undef SPAI_DIA_PASS_LAPLACE
#undef SPAI_DIA_PASS
  }

  // Blocks of the fused iteration kernel (the configuration the build
  // launches: slot width, GL, stash, method) resident per CU at once.
  int FusedBlocksPerCu(bool mr) const {
    int nb = 0;
#define SPAI_OCC(SlotT, GLV)                                                                                     \
  do {                                                                                                           \
    const void* f = nullptr;                                                                                     \
    if constexpr (sizeof(V) == 4 && sizeof(SlotT) == 1 && GLV <= 32) {                                          \
      if (terms_ == kTermsRegs)                                                                                  \
        f = mr ? reinterpret_cast<const void*>(&FusedPass<V, SlotT, GLV, false, true, kTermsRegs>)              \
               : reinterpret_cast<const void*>(&FusedPass<V, SlotT, GLV, false, false, kTermsRegs>);            \
    }                                                                                                            \
    if (f == nullptr) {                                                                                          \
      if (mr)                                                                                                    \
        f = terms_ == kTermsStash                                                                                \
                ? reinterpret_cast<const void*>(&FusedPass<V, SlotT, GLV, false, true, kTermsStash>)            \
                : reinterpret_cast<const void*>(&FusedPass<V, SlotT, GLV, false, true, kTermsReload>);          \
      else                                                                                                       \
        f = terms_ == kTermsStash                                                                                \
                ? reinterpret_cast<const void*>(&FusedPass<V, SlotT, GLV, false, false, kTermsStash>)           \
                : reinterpret_cast<const void*>(&FusedPass<V, SlotT, GLV, false, false, kTermsReload>);         \
    }                                                                                                            \
    SPAI_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kBlk, smemf_));                         \
  } while (0)
#define SPAI_OCC_GL(SlotT)                  \
  switch (gl_) {                            \
    case 8: SPAI_OCC(SlotT, 8); break;      \
    case 16: SPAI_OCC(SlotT, 16); break;    \
    case 32: SPAI_OCC(SlotT, 32); break;    \
    default: SPAI_OCC(SlotT, 64); break;    \
  }
    if (slot16_) {
      SPAI_OCC_GL(uint16_t);
    } else {
      SPAI_OCC_GL(uint8_t);
    }
#undef SPAI_OCC_GL
#undef SPAI_OCC
    return nb;
  }

  int64_t BuildThen(const std::function<void()>& enqueue_after) override {
    // the fused builds call it before their read-back; the others after
    enqueue_after_ = &enqueue_after;
    int64_t it = 0;
    try {
      it = Build();
    } catch (...) {
      enqueue_after_ = nullptr;
      throw;
    }
    const bool done = enqueue_after_ == nullptr;
    enqueue_after_ = nullptr;
    if (!done) enqueue_after();
    return it;
  }

  int64_t Build() override {
    timed_launches_ = 0;
    if (method_ >= SPAI_INNER_OUTER_MR) return BuildInnerOuter();
    if (dia_ || ((method_ == SPAI_GLOBAL_LOMR || method_ == SPAI_GLOBAL_MR) && use_plan_ && fused_ok_))
      return BuildFused();
    Ctx& c = ctx_;
    hipStream_t st = c.stream;
    // M = M0 on S, P = 0
    SampleOnPattern<V><<<Blocks1(n_ * kSampleLanes), kBlk, 0, st>>>(
        (int)n_, s_pat_->ptr.get(), s_pat_->idx.get(), m0_.ptr(), m0_.idx(), m0_.v(), M_.get());
    FillZero<V><<<Blocks1(nnz_s_), kBlk, 0, st>>>(nnz_s_, P_.get());
    SPAI_HIP_CHECK(hipMemsetAsync(flags_.get(), 0, sizeof(Flags), st));
    const int* s2r = s2_pat_->idx.get();
    const int* s2c = s2col_.get();
    const int* map = s2_to_s_.get();
    const int* atp = at_pat_->ptr.get();
    const int* ati = at_pat_->idx.get();
    const int* sp = s_pat_->ptr.get();
    const int* si = s_pat_->idx.get();
    const double tol = prm_.tolerance;
    const int G2 = (int)grid2_;
    const int Gp = use_plan_ ? (int)gridp_ : G2;  // partials of the product kernels
    auto residual = [&](int mode) {
      if (use_plan_) {
        Slot<1>(ResidualEpi<V>{map, prv_.get(), R_.get(), Z_.get(), P_.get(), mode}, M_.get(), nullptr);
      } else {
        ResidualKernel<V><<<grid2_, kBlk, 0, st>>>(nnz_s2_, s2r, s2c, map, atp, ati, atv_.get(), sp, si, M_.get(),
                                                   prv_.get(), R_.get(), Z_.get(), P_.get(), mode, part_.get(),
                                                   flags_.get());
      }
    };
    residual(method_ == SPAI_GLOBAL_CG ? 1 : 0);
    AppendInitial<<<1, kBlk, 0, st>>>(part_.get(), Gp, tol, hist_.get(), sc_.get(), flags_.get());
    for (int64_t it = 1; it <= prm_.max_iterations; ++it) {
      if (method_ == SPAI_GLOBAL_LOMR) {
        if (use_plan_) {
          // products are formed twice (project, then update) instead of storing A Z and A P
          Slot<2>(LomrProjectEpi<V>{R_.get(), prv_.get()}, Z_.get(), P_.get());
          LomrSteps<<<1, kBlk, 0, st>>>(part_.get(), Gp, it == 1 ? 1 : 0, sc_.get(), flags_.get());
          LomrUpdateEpi<V> ue{};
          ue.s2_to_s = map;
          ue.pr = prv_.get();
          ue.R = R_.get();
          ue.M = M_.get();
          ue.Z = Z_.get();
          ue.P = P_.get();
          Slot<2>(ue, Z_.get(), P_.get());
          AppendResidual<<<1, kBlk, 0, st>>>(part_.get(), Gp, tol, hist_.get(), flags_.get());
        } else {
          LomrProject<V><<<grid2_, kBlk, 0, st>>>(nnz_s2_, s2r, s2c, atp, ati, atv_.get(), sp, si, Z_.get(),
                                                  P_.get(), R_.get(), prv_.get(), AZ_.get(), AP_.get(), part_.get(),
                                                  flags_.get());
          LomrSteps<<<1, kBlk, 0, st>>>(part_.get(), G2, it == 1 ? 1 : 0, sc_.get(), flags_.get());
          LomrUpdate<V><<<grid2_, kBlk, 0, st>>>(nnz_s2_, s2r, map, AZ_.get(), AP_.get(), prv_.get(), R_.get(),
                                                 M_.get(), Z_.get(), P_.get(), sc_.get(), part_.get(), flags_.get());
          AppendResidual<<<1, kBlk, 0, st>>>(part_.get(), G2, tol, hist_.get(), flags_.get());
        }
      } else if (method_ == SPAI_GLOBAL_MR) {
        if (use_plan_) {
          Slot<1>(MrProjectEpi<V>{R_.get(), prv_.get()}, Z_.get(), nullptr);
        } else {
          MrProject<V><<<grid2_, kBlk, 0, st>>>(nnz_s2_, s2r, s2c, atp, ati, atv_.get(), sp, si, Z_.get(), R_.get(),
                                                prv_.get(), part_.get(), flags_.get());
        }
        MrStep<<<1, kBlk, 0, st>>>(part_.get(), Gp, sc_.get(), flags_.get());
        AxpyS<V><<<grid1_, kBlk, 0, st>>>(nnz_s_, M_.get(), Z_.get(), sc_.get(), flags_.get());
        residual(0);
        AppendResidual<<<1, kBlk, 0, st>>>(part_.get(), Gp, tol, hist_.get(), flags_.get());
      } else {
        if (use_plan_) {
          Slot<1>(CgProjectEpi<V>{map, P_.get(), AP_.get()}, P_.get(), nullptr);
        } else {
          CgProject<V><<<grid2_, kBlk, 0, st>>>(nnz_s2_, s2r, s2c, map, atp, ati, atv_.get(), sp, si, P_.get(),
                                                AP_.get(), part_.get(), flags_.get());
        }
        CgAlpha<<<1, kBlk, 0, st>>>(part_.get(), Gp, sc_.get(), flags_.get());
        CgUpdate<V><<<grid2_, kBlk, 0, st>>>(nnz_s2_, s2r, map, AP_.get(), prv_.get(), R_.get(), M_.get(), P_.get(),
                                             Z_.get(), sc_.get(), part_.get(), flags_.get());
        CgAppendBeta<<<1, kBlk, 0, st>>>(part_.get(), G2, tol, hist_.get(), sc_.get(), flags_.get());
        CgDirection<V><<<grid1_, kBlk, 0, st>>>(nnz_s_, P_.get(), Z_.get(), sc_.get(), flags_.get());
      }
    }
    SPAI_HIP_CHECK(hipGetLastError());
    Flags f;
    c.ReadBack(&f, flags_.get(), sizeof(Flags));
    last_flags_ = f;
    return f.iterations;
  }

  void Output(EngineOutput& out) override {
    Ctx& c = ctx_;
    std::vector<double> h;
    if (method_ >= SPAI_INNER_OUTER_MR) {
      h = io_hist_;
      out.iterations = (int64_t)h.size() - 1;
      out.converged = io_converged_;
    } else {
      const Flags f = last_flags_;
      out.iterations = f.iterations;
      out.converged = f.stop == 2;
      h.resize(f.iterations + 1);
      c.ReadBack(h.data(), hist_.get(), sizeof(double) * h.size());
    }
    const double entries = (double)n_ * (double)n_;
    const double dens = (double)nnz_s_ / entries;
    out.history.clear();
    for (int64_t i = 0; i <= out.iterations; ++i) out.history.push_back(SingleIterationData{i, h[i], dens, dens});
    DMat Md = MatrixM();
    out.M.rows = n_;
    out.M.cols = n_;
    out.M.ptr.resize(n_ + 1);
    out.M.idx.resize(nnz_s_);
    out.M.val.resize(nnz_s_);
    Download(c, Md, out.M.ptr.data(), out.M.idx.data(), out.M.val.data());
  }

  DMat MatrixM() override {
    Ctx& c = ctx_;
    DMat Md;
    Md.pat = s_pat_;
    Md.val = c.Alloc<double>(nnz_s_);
    ToDouble<V><<<Blocks1(nnz_s_), kBlk, 0, c.stream>>>(nnz_s_, M_.get(), Md.val.get());
    return Md;
  }

  void GatherM(const int* perm, int64_t len, double* out_d, float* out_f) override {
    if (len > 0) GatherValues<V><<<Blocks1(len), kBlk, 0, ctx_.stream>>>(len, perm, M_.get(), out_d, out_f);
  }

 private:
  Ctx ctx_;
  int64_t method_;
  EngineParams prm_;
  int64_t n_ = 0, nnz_s_ = 0, nnz_s2_ = 0;
  DMat A_, m0_;
  PatternPtr at_pat_, s_pat_, s2_pat_;
  DBuf<double> at_d_, pr_d_;
  DBuf<V> atv_, av_, prv_, M_, Z_, P_, R_, AZ_, AP_;
  // slot plan
  // inner-outer
  PatternPtr st_pat_;
  DBuf<int> stperm_, scol_;
  DBuf<V> Mout_, Q_;
  DBuf<double> io_cols_;
  std::vector<double> io_hist_;
  bool io_converged_ = false;
  DBuf<int64_t> plan_ptr_;
  DBuf<uint8_t> plan8_;
  DBuf<uint16_t> plan16_;
  int maxcol_ = 1, lmax_ = 1, gl_ = 8, smax_ = 1;
  size_t smemf_ = 0;
  bool fused_ok_ = false;
  bool timing_ = false;
  hipGraph_t build_graph_ = nullptr;
  hipGraphExec_t build_exec_ = nullptr;
  std::vector<hipEvent_t> events_;
  double timed_us_ = 0.0;
  int64_t timed_launches_ = 0;
  const std::function<void()>* enqueue_after_ = nullptr;  // see BuildThen
  unsigned gridf_ = 1;
  DBuf<double> partf_;
  DBuf<int2> a0c_;
  int terms_ = kTermsReload;  // how the fused pass's product 2 gets its terms
  DBuf<uint16_t> stq_;
  DBuf<int> s2d_;
  DBuf<V> prS_, prS2_;
  size_t smem_ = 0;
  bool slot16_ = false, use_plan_ = false;
  unsigned gridp_ = 1;
  // diagonal pass: stencil dimensions, diagonals of S and S2, their
  // offsets, S's column masks and slot sources, values on diagonals
  bool dia_ = false;
  unsigned gridd_ = 1;
  int ddims_ = 0, dnd_ = 0, dnd2_ = 0;  // ddims_: the shape kind of MakeDiaShape
  int64_t dld_ = 0;  // leading dimension of the diagonal arrays
  DiaPassArgs dargs_{};
  DBuf<unsigned> dmask_s_;
  DBuf<int> dsrc_s_;
  DBuf<double> partf2_;  // the diagonal pass's second partials buffer
  DBuf<V> dav_, dM_, dZ_, dP_;

  // Launches the slot product with epilogue e over Y = (y0, y1).
  template <int NY, class Epi>
  void Slot(const Epi& e, const V* y0, const V* y1) {
    hipStream_t st = ctx_.stream;
    const int* sp = s_pat_->ptr.get();
    const int* si = s_pat_->idx.get();
    const int* s2p = s2_pat_->ptr.get();
    const int* s2i = s2_pat_->idx.get();
#define SPAI_SLOT_LAUNCH(SlotT, plan, GLV)                                                                     \
  SlotProduct<V, SlotT, GLV, NY, Epi><<<gridp_, kBlk, smem_, st>>>((int)n_, sp, si, y0, y1, A_.ptr(), av_.get(), \
                                                                   plan_ptr_.get(), plan, s2p, s2i, lmax_, e,  \
                                                                   sc_.get(), part_.get(), flags_.get(),       \
                                                                   maxcol_ <= gl_ ? 1 : 0)
#define SPAI_SLOT_GL(SlotT, plan)             \
  switch (gl_) {                              \
    case 8: SPAI_SLOT_LAUNCH(SlotT, plan, 8); break;   \
    case 16: SPAI_SLOT_LAUNCH(SlotT, plan, 16); break; \
    case 32: SPAI_SLOT_LAUNCH(SlotT, plan, 32); break; \
    default: SPAI_SLOT_LAUNCH(SlotT, plan, 64); break; \
  }
    if (slot16_) {
      SPAI_SLOT_GL(uint16_t, plan16_.get());
    } else {
      SPAI_SLOT_GL(uint8_t, plan8_.get());
    }
#undef SPAI_SLOT_GL
#undef SPAI_SLOT_LAUNCH
  }
  DBuf<int> s2col_, s2_to_s_;
  DBuf<double> part_, sc_, hist_;
  DBuf<Flags> flags_;
  unsigned grid1_ = 1, grid2_ = 1;
  Flags last_flags_{0, 0};
};

}  // namespace

std::unique_ptr<FixedProblem> MakeFixedProblem(int64_t method, const MatrixData& A, const MatrixData* M0,
                                               const MatrixData* Pr, const EngineParams& params) {
  if (params.fp32) return std::make_unique<FixedImpl<float>>(method, A, M0, Pr, params);
  return std::make_unique<FixedImpl<double>>(method, A, M0, Pr, params);
}

EngineOutput RunFixedEngine(int64_t method, const MatrixData& A, const MatrixData* M0, const MatrixData* Pr,
                            const EngineParams& params) {
  auto prob = MakeFixedProblem(method, A, M0, Pr, params);
  Ctx& c = prob->context();
  hipEvent_t e0, e1;
  SPAI_HIP_CHECK(hipEventCreate(&e0));
  SPAI_HIP_CHECK(hipEventCreate(&e1));
  SPAI_HIP_CHECK(hipEventRecord(e0, c.stream));
  prob->Build();
  EngineOutput out;
  prob->Output(out);
  SPAI_HIP_CHECK(hipEventRecord(e1, c.stream));
  SPAI_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  SPAI_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  out.device_seconds = ms * 1e-3;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return out;
}

}  // namespace spai
