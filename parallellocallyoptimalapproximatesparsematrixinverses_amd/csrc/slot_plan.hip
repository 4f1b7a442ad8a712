// Slot plan of the fixed-pattern products (declared in slot_plan.hpp).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <stdexcept>

#include "dev_common.hpp"
#include "slot_plan.hpp"

namespace spai {
namespace {

constexpr int kPlanBlock = 256;

__device__ __forceinline__ int WaveMax(int v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, kWave));
  return v;
}

// terms[j] = sum over S(:, j) of |A(:, k)|, one thread per column (eight S
// entries' A column bounds loaded together); ext = the longest column of A,
// of S2 and of S (one atomic per block).
__global__ __launch_bounds__(kPlanBlock) void PlanTerms(int n, const int* __restrict__ sp, const int* __restrict__ si,
                                                        const int* __restrict__ ap, const int* __restrict__ s2p,
                                                        int64_t* __restrict__ terms, int* __restrict__ ext) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  int la = 0, ls2 = 0, ls = 0;
  if (j < n) {
    const int p0 = sp[j], p1 = sp[j + 1];
    int64_t u = 0;
    for (int p = p0; p < p1; p += 8) {
      int k[8], lo[8], hi[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) k[t] = p + t < p1 ? si[p + t] : -1;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        lo[t] = k[t] >= 0 ? ap[k[t]] : 0;
        hi[t] = k[t] >= 0 ? ap[k[t] + 1] : 0;
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) u += hi[t] - lo[t];
    }
    terms[j] = u;
    la = ap[j + 1] - ap[j];
    ls2 = s2p[j + 1] - s2p[j];
    ls = p1 - p0;
  }
  // block maxima in LDS, then one atomic per block and value
  __shared__ int bmax[3];
  if (threadIdx.x < 3) bmax[threadIdx.x] = 0;
  __syncthreads();
  la = WaveMax(la);
  ls2 = WaveMax(ls2);
  ls = WaveMax(ls);
  if (LaneId() == 0) {
    atomicMax(&bmax[0], la);
    atomicMax(&bmax[1], ls2);
    atomicMax(&bmax[2], ls);
  }
  __syncthreads();
  if (threadIdx.x < 3) atomicMax(&ext[threadIdx.x], bmax[threadIdx.x]);
}

// ---- int64 exclusive scan: tile sums, one block over the tile sums, apply
constexpr int kScanPer = 4;
constexpr int kScanTile64 = kPlanBlock * kScanPer;

__device__ __forceinline__ int64_t BlockExclusive64(int64_t v, int64_t& total) {
  __shared__ int64_t wsum[kPlanBlock / kWave];
  const int64_t incl = WaveInclusiveScan64(v);
  if (LaneId() == kWave - 1) wsum[WaveInBlock()] = incl;
  __syncthreads();
  int64_t before = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < kPlanBlock / kWave; ++w) {
    before += w < WaveInBlock() ? wsum[w] : 0;
    total += wsum[w];
  }
  __syncthreads();
  return before + incl - v;
}

__global__ __launch_bounds__(kPlanBlock) void Scan64TileSums(const int64_t* __restrict__ in, int64_t n,
                                                             int64_t* __restrict__ sums) {
  const int64_t base = int64_t(blockIdx.x) * kScanTile64 + int64_t(threadIdx.x) * kScanPer;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) s += base + i < n ? in[base + i] : 0;
  int64_t total = 0;
  BlockExclusive64(s, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the tile sums in place (one block); out_total = the sum.
__global__ __launch_bounds__(kPlanBlock) void Scan64Tiles(int64_t* __restrict__ sums, int64_t ntiles,
                                                          int64_t* __restrict__ out_total) {
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < ntiles; b0 += kScanTile64) {
    const int64_t base = b0 + int64_t(threadIdx.x) * kScanPer;
    int64_t v[kScanPer], s = 0;
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
      v[i] = base + i < ntiles ? sums[base + i] : 0;
      s += v[i];
    }
    int64_t total = 0;
    int64_t run = carry + BlockExclusive64(s, total);
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
      if (base + i < ntiles) sums[base + i] = run;
      run += v[i];
    }
    carry += total;
  }
  if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(kPlanBlock) void Scan64Apply(const int64_t* __restrict__ in, int64_t n,
                                                          const int64_t* __restrict__ sums,
                                                          const int64_t* __restrict__ total,
                                                          int64_t* __restrict__ out) {
  const int64_t base = int64_t(blockIdx.x) * kScanTile64 + int64_t(threadIdx.x) * kScanPer;
  int64_t v[kScanPer], s = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    v[i] = base + i < n ? in[base + i] : 0;
    s += v[i];
  }
  int64_t t = 0;
  int64_t run = sums[blockIdx.x] + BlockExclusive64(s, t);
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = *total;
}

// The plan itself: GL lanes per column of S2 (a power of two, 64 / GL columns
// per wavefront).  With kLds the group first copies S2(:, j) into LDS (lmax
// entries a group); then, for each S entry k of the column in order (lane
// gl fetches entry gl of every GL-entry chunk and the steps broadcast it),
// lanes u < |A(:, k)| (striding by GL) find row A(u, k) in S2(:, j) by binary
// search and store its position at plan[ptr[j] + offset of the term].
template <class SlotT, int GL, bool kLds>
__global__ __launch_bounds__(kPlanBlock) void BuildPlanGroups(int n, const int* __restrict__ sp,
                                                              const int* __restrict__ si,
                                                              const int* __restrict__ ap,
                                                              const int* __restrict__ ai,
                                                              const int* __restrict__ s2p,
                                                              const int* __restrict__ s2i,
                                                              const int64_t* __restrict__ plan_ptr, int lmax,
                                                              SlotT* __restrict__ plan, int* __restrict__ error) {
  extern __shared__ int s2_lds[];
  constexpr int kGroups = kPlanBlock / GL;
  const int g = threadIdx.x / GL, gl = threadIdx.x % GL;
  const int j = blockIdx.x * kGroups + g;
  if (j >= n) return;  // whole groups leave together
  const int qb = s2p[j], L = s2p[j + 1] - qb;
  const int* rows = s2i + qb;
  if constexpr (kLds) {
    int* const R = s2_lds + size_t(g) * lmax;
    for (int t = gl; t < L; t += GL) R[t] = rows[t];
    WaveFence();
    rows = R;
  }
  int64_t pp = plan_ptr[j];
  const int e1 = sp[j + 1];
  for (int eb = sp[j]; eb < e1; eb += GL) {
    int my_a0 = 0, my_c = 0;
    if (eb + gl < e1) {
      const int k = si[eb + gl];
      my_a0 = ap[k];
      my_c = ap[k + 1] - my_a0;
    }
    const int ns = min(GL, e1 - eb);
    for (int st = 0; st < ns; ++st) {
      const int a0 = __shfl(my_a0, st, GL), c = __shfl(my_c, st, GL);
      for (int u = gl; u < c; u += GL) {
        const int row = ai[a0 + u];
        int lo = 0, hi = L;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (rows[mid] < row)
            lo = mid + 1;
          else
            hi = mid;
        }
        const bool found = lo < L && rows[lo] == row;
        if (!found) *error = 2;  // the context's error code (mapped host memory)
        plan[pp + u] = SlotT(found ? lo : 0);
      }
      pp += c;
    }
  }
}

unsigned Grid1(int64_t n, int per) {
  const int64_t b = std::max<int64_t>(1, (n + per - 1) / per);
  if (b > (int64_t(1) << 31) - 1) throw std::runtime_error("grid too large");
  return unsigned(b);
}

template <class SlotT, int GL>
void LaunchBuild(Ctx& c, int64_t n, const int* sp, const int* si, const int* ap, const int* ai, const int* s2p,
                 const int* s2i, const int64_t* ptr, int lmax, SlotT* plan) {
  constexpr int kGroups = kPlanBlock / GL;
  const size_t lds = size_t(kGroups) * size_t(lmax) * sizeof(int);
  const unsigned g = Grid1(n, kGroups);
  if (lds <= 48 * 1024)
    BuildPlanGroups<SlotT, GL, true><<<g, kPlanBlock, lds, c.stream>>>((int)n, sp, si, ap, ai, s2p, s2i, ptr, lmax,
                                                                        plan, c.err_dev);
  else
    BuildPlanGroups<SlotT, GL, false><<<g, kPlanBlock, 0, c.stream>>>((int)n, sp, si, ap, ai, s2p, s2i, ptr, lmax,
                                                                       plan, c.err_dev);
}

template <class SlotT>
void LaunchBuildGl(Ctx& c, int gl, int64_t n, const int* sp, const int* si, const int* ap, const int* ai,
                   const int* s2p, const int* s2i, const int64_t* ptr, int lmax, SlotT* plan) {
  switch (gl) {
    case 8: LaunchBuild<SlotT, 8>(c, n, sp, si, ap, ai, s2p, s2i, ptr, lmax, plan); break;
    case 16: LaunchBuild<SlotT, 16>(c, n, sp, si, ap, ai, s2p, s2i, ptr, lmax, plan); break;
    case 32: LaunchBuild<SlotT, 32>(c, n, sp, si, ap, ai, s2p, s2i, ptr, lmax, plan); break;
    default: LaunchBuild<SlotT, 64>(c, n, sp, si, ap, ai, s2p, s2i, ptr, lmax, plan); break;
  }
}

}  // namespace

void ScanInt64(Ctx& c, const int64_t* in, int64_t n, int64_t* out) {
  const int64_t tiles = std::max<int64_t>(1, (n + kScanTile64 - 1) / kScanTile64);
  auto sums = c.Alloc<int64_t>(size_t(tiles));
  auto total = c.Alloc<int64_t>(1);
  Scan64TileSums<<<unsigned(tiles), kPlanBlock, 0, c.stream>>>(in, n, sums.get());
  Scan64Tiles<<<1, kPlanBlock, 0, c.stream>>>(sums.get(), tiles, total.get());
  Scan64Apply<<<unsigned(tiles), kPlanBlock, 0, c.stream>>>(in, n, sums.get(), total.get(), out);
}

SlotPlan BuildSlotPlan(Ctx& c, int64_t n, const int* sp, const int* si, const int* ap, const int* ai,
                       const int* s2p, const int* s2i) {
  SlotPlan P;
  auto terms = c.Alloc<int64_t>(size_t(std::max<int64_t>(n, 1)));
  auto ext = c.Alloc<int>(4);
  SPAI_HIP_CHECK(hipMemsetAsync(ext.get(), 0, 4 * sizeof(int), c.stream));
  P.ptr = c.Alloc<int64_t>(size_t(n) + 1);
  if (n > 0) PlanTerms<<<Grid1(n, kPlanBlock), kPlanBlock, 0, c.stream>>>((int)n, sp, si, ap, s2p, terms.get(),
                                                                         ext.get());
  ScanInt64(c, terms.get(), n, P.ptr.get());
  int h[4] = {0, 0, 0, 0};
  int64_t total = 0;
  c.ReadBack2(h, ext.get(), sizeof(h), &total, P.ptr.get() + n, sizeof(total));
  P.terms = total;
  P.max_a_col = h[0];
  P.max_s2_col = h[1];
  P.max_s_col = h[2];
  if (P.max_s2_col > 65535) return P;
  P.built = true;
  P.slot16 = P.max_s2_col > 256;
  const int gl = P.max_a_col <= 8 ? 8 : P.max_a_col <= 16 ? 16 : P.max_a_col <= 32 ? 32 : 64;
  const int lmax = std::max(P.max_s2_col, 1);
  if (P.slot16) {
    P.plan16 = c.Alloc<uint16_t>(size_t(std::max<int64_t>(total, 1)));
    if (n > 0) LaunchBuildGl<uint16_t>(c, gl, n, sp, si, ap, ai, s2p, s2i, P.ptr.get(), lmax, P.plan16.get());
  } else {
    P.plan8 = c.Alloc<uint8_t>(size_t(std::max<int64_t>(total, 1)));
    if (n > 0) LaunchBuildGl<uint8_t>(c, gl, n, sp, si, ap, ai, s2p, s2i, P.ptr.get(), lmax, P.plan8.get());
  }
  SPAI_HIP_CHECK(hipGetLastError());
  return P;
}

}  // namespace spai
