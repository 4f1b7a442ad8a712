// Device-side helpers shared by the kernels: wavefront/block reductions and
// small search routines.  gfx950 runs 64-wide wavefronts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spai {

constexpr int kWave = 64;

__device__ __forceinline__ int LaneId() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int WaveInBlock() { return threadIdx.x / kWave; }

// Orders the LDS/global accesses of one wavefront between two phases of a
// per-column loop (all lanes of the wave participate).
__device__ __forceinline__ void WaveFence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

template <class T>
__device__ __forceinline__ T WaveSum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ int WaveSumInt(int v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Inclusive prefix sum over the lanes of a wavefront.
__device__ __forceinline__ int WaveInclusiveScan(int v) {
  const int lane = LaneId();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int t = __shfl_up(v, off, kWave);
    if (lane >= off) v += t;
  }
  return v;
}

__device__ __forceinline__ int64_t WaveInclusiveScan64(int64_t v) {
  const int lane = LaneId();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int64_t t = __shfl_up(v, off, kWave);
    if (lane >= off) v += t;
  }
  return v;
}

// Number of set predicate bits of lanes strictly below this one.
__device__ __forceinline__ int LanePrefixCount(uint64_t mask) {
  const int lane = LaneId();
  const uint64_t below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
  return __popcll(below);
}

// Block-wide sum; every thread receives the total.  blockDim.x <= 1024.
template <class T>
__device__ T BlockSum(T v, T* scratch /* >= 16 entries in LDS */) {
  v = WaveSum(v);
  const int w = WaveInBlock();
  const int nw = (blockDim.x + kWave - 1) / kWave;
  __syncthreads();
  if (LaneId() == 0) scratch[w] = v;
  __syncthreads();
  T t = 0;
  for (int i = 0; i < nw; ++i) t += scratch[i];
  __syncthreads();
  return t;
}

// Two block sums in one pass: each value is summed in BlockSum's order (so
// the results equal two BlockSum calls) with one set of barriers.
template <class T>
__device__ void BlockSum2(T& a, T& b, T* scratch /* >= 32 entries in LDS */) {
  a = WaveSum(a);
  b = WaveSum(b);
  const int w = WaveInBlock();
  const int nw = (blockDim.x + kWave - 1) / kWave;
  __syncthreads();
  if (LaneId() == 0) {
    scratch[w] = a;
    scratch[16 + w] = b;
  }
  __syncthreads();
  T ta = 0, tb = 0;
  for (int i = 0; i < nw; ++i) {
    ta += scratch[i];
    tb += scratch[16 + i];
  }
  __syncthreads();
  a = ta;
  b = tb;
}

// First index in [lo, hi) of sorted a[] with a[i] >= key.
template <class I>
__device__ __forceinline__ int LowerBound(const I* a, int lo, int hi, I key) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Position of key in sorted a[lo, hi), or -1.
template <class I>
__device__ __forceinline__ int FindSorted(const I* a, int lo, int hi, I key) {
  const int p = LowerBound(a, lo, hi, key);
  return (p < hi && a[p] == key) ? p : -1;
}

// Intersection dot of two sorted sparse columns, added to s in row order
// (sequential, as Eigen's cwiseProduct().sum()).
__device__ __forceinline__ double ColumnDot(const int* xi, const double* xv, int xb, int xe, const int* yi,
                                            const double* yv, int yb, int ye, double s = 0.0) {
  while (xb < xe && yb < ye) {
    const int a = xi[xb], b = yi[yb];
    if (a == b) {
      s += xv[xb] * yv[yb];
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

// Order-preserving map of a double onto uint64 (+0.0 and -0.0 coincide).
__device__ __forceinline__ uint64_t OrderedKey(double x) {
  if (x == 0.0) x = 0.0;
  const uint64_t b = __double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

}  // namespace spai
