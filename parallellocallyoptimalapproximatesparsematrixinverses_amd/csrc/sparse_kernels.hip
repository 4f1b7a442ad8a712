// Sparse kernels of the reference-semantics (dynamic pattern) path: scan,
// two-phase SpGEMM, union merges, compaction, transposes and reductions.
//
// Values follow Eigen's operation order exactly where the reference's result
// depends on it (products, unions, scalings: no contraction, -ffp-contract=off),
// so the patterns these kernels produce are bit-identical to the reference's.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "dev_common.hpp"
#include "ls2x2.hpp"
#include "ops.hpp"

namespace spai {

namespace {

constexpr int kBlock = 256;

inline unsigned Blocks(int64_t n, int per_block = kBlock) {
  int64_t b = (n + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > (int64_t(1) << 30)) throw std::runtime_error("grid too large");
  return static_cast<unsigned>(b);
}

// ------------------------------------------------------------------ scan
constexpr int kScanItems = 4;  // per thread
constexpr int kScanTile = kBlock * kScanItems;

__global__ __launch_bounds__(kBlock) void ScanTileSums(const int* __restrict__ in, int64_t n,
                                                       int64_t* __restrict__ tile_sums) {
  __shared__ int64_t red[16];
  const int64_t base = int64_t(blockIdx.x) * kScanTile;
  int64_t s = 0;
  for (int t = 0; t < kScanItems; ++t) {
    const int64_t i = base + int64_t(t) * kBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  s = BlockSum<int64_t>(s, red);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

// Single block: exclusive scan of the tile sums in place; total to *total.
__global__ __launch_bounds__(1024) void ScanTiles(int64_t* __restrict__ sums, int64_t ntiles,
                                                  int64_t* __restrict__ total) {
  __shared__ int64_t wsum[16];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < ntiles; b0 += blockDim.x) {
    const int64_t i = b0 + threadIdx.x;
    const int64_t v = i < ntiles ? sums[i] : 0;
    int64_t incl = WaveInclusiveScan64(v);
    if (LaneId() == kWave - 1) wsum[WaveInBlock()] = incl;
    __syncthreads();
    int64_t wbase = 0;
    for (int w = 0; w < WaveInBlock(); ++w) wbase += wsum[w];
    const int64_t excl = carry + wbase + incl - v;
    __syncthreads();
    if (i < ntiles) sums[i] = excl;
    if (threadIdx.x == blockDim.x - 1) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kBlock) void ScanApply(const int* __restrict__ in, int64_t n,
                                                    const int64_t* __restrict__ tile_offsets,
                                                    int* __restrict__ out) {
  __shared__ int64_t wsum[16];
  __shared__ int64_t carry;
  const int64_t base = int64_t(blockIdx.x) * kScanTile;
  if (threadIdx.x == 0) carry = tile_offsets[blockIdx.x];
  __syncthreads();
  for (int t = 0; t < kScanItems; ++t) {
    const int64_t i = base + int64_t(t) * kBlock + threadIdx.x;
    const int64_t v = i < n ? in[i] : 0;
    const int64_t incl = WaveInclusiveScan64(v);
    if (LaneId() == kWave - 1) wsum[WaveInBlock()] = incl;
    __syncthreads();
    int64_t wbase = 0;
    for (int w = 0; w < WaveInBlock(); ++w) wbase += wsum[w];
    const int64_t excl = carry + wbase + incl - v;
    if (i < n) out[i] = static_cast<int>(excl);
    __syncthreads();
    if (threadIdx.x == kBlock - 1) carry = excl + v;
    __syncthreads();
  }
  // the last tile ends at n - 1, so its final carry is the total
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = static_cast<int>(carry);
}

__global__ void WriteLast(int* __restrict__ out, int64_t n, const int64_t* __restrict__ total) {
  out[n] = static_cast<int>(*total);
}

// The whole scan in one block for short inputs.  Wavefront w owns the
// contiguous segment [w seg, (w + 1) seg) and reads it in coalesced 64-entry
// chunks, kScanBatch chunks in flight: first the segment sums (the block adds
// the earlier segments' sums), then the scan of each chunk with a running
// carry.  out[n] and *total receive the sum.  in may be out: a chunk is read
// before any entry at or after it is written (each lane writes only its own
// entries, and later chunks are loaded ahead of the current chunk's writes).
constexpr int64_t kScanSmallMax = 65536;
constexpr int kScanBatch = 8;
__global__ __launch_bounds__(1024) void ScanSmall(const int* in, int64_t n, int* out, int64_t* __restrict__ total) {
  __shared__ int64_t wsum[16];
  const int w = WaveInBlock(), lane = LaneId();
  const int64_t seg = (n + 15) / 16;
  const int64_t b = min(n, int64_t(w) * seg), e = min(n, b + seg);
  int64_t s = 0;
  for (int64_t i0 = b; i0 < e; i0 += kScanBatch * kWave) {
    int v[kScanBatch];
#pragma unroll
    for (int t = 0; t < kScanBatch; ++t) {
      const int64_t i = i0 + t * kWave + lane;
      v[t] = i < e ? in[i] : 0;
    }
#pragma unroll
    for (int t = 0; t < kScanBatch; ++t) s += v[t];
  }
  s = WaveSum<int64_t>(s);
  if (lane == 0) wsum[w] = s;
  __syncthreads();
  int64_t run = 0, all = 0;
  for (int k = 0; k < 16; ++k) {
    if (k < w) run += wsum[k];
    all += wsum[k];
  }
  for (int64_t i0 = b; i0 < e; i0 += kScanBatch * kWave) {
    int v[kScanBatch];
#pragma unroll
    for (int t = 0; t < kScanBatch; ++t) {
      const int64_t i = i0 + t * kWave + lane;
      v[t] = i < e ? in[i] : 0;
    }
#pragma unroll
    for (int t = 0; t < kScanBatch; ++t) {
      const int64_t i = i0 + t * kWave + lane;
      const int64_t incl = WaveInclusiveScan64(v[t]);
      if (i < e) out[i] = static_cast<int>(run + incl - v[t]);
      run += __shfl(incl, kWave - 1, kWave);
    }
  }
  if (threadIdx.x == 0) {
    out[n] = static_cast<int>(all);
    *total = all;
  }
}

// ------------------------------------------------------------------ SpGEMM
constexpr int kEscCap = 4096;     // expansion keys per wavefront (LDS tiers)
constexpr int kSymWaves = 2;      // wavefronts per block in the symbolic kernel
constexpr int kStashMidColumns = 4096;
constexpr int kEscRegs = 4;  // expansion entries a lane gathers at once  // LDS-tier columns from which their count pass stashes
constexpr int kNumCap = 1024;     // output rows per column held in LDS (numeric)
constexpr int kNumWaves = 4;

// Short columns (at most kTmLists entries in Y(:, j) and an expansion of at
// most kTmCap rows, counting the identity's) are merged by one thread each;
// the other columns take the wavefront-per-column tiers.
constexpr int kTmLists = 16;
constexpr int kTmCap = 64;

__device__ __forceinline__ bool IsShortColumn(int64_t upper, int ny) {
  return upper <= kTmCap && ny <= kTmLists;
}

// A column's rows fit the LDS bitmap of the symbolic tier (one bit per row
// of its range, kEscCap words).
__device__ __forceinline__ bool BitmapRange(int lo_r, int hi_r) { return hi_r >= lo_r && hi_r - lo_r < 32 * kEscCap; }

__device__ __forceinline__ int WaveMinInt(int v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
  return v;
}
__device__ __forceinline__ int WaveMaxInt(int v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, kWave));
  return v;
}

// Block-wide ordered compaction helper: returns this thread's exclusive rank
// among flagged threads and writes the block total to *total.
__device__ int BlockRank(bool flag, int* wcount /* LDS >= 16 */, int* total) {
  const uint64_t mask = __ballot(flag);
  const int w = WaveInBlock();
  const int nw = blockDim.x / kWave;
  if (LaneId() == 0) wcount[w] = __popcll(mask);
  __syncthreads();
  int base = 0, t = 0;
  for (int i = 0; i < nw; ++i) {
    if (i < w) base += wcount[i];
    t += wcount[i];
  }
  *total = t;
  __syncthreads();
  return base + LanePrefixCount(mask);
}

// Appends `value` of every flagged thread to list[*count ...] with one atomic
// per block; threads keep their order within the block.  Called by all
// threads of the block.
__device__ void BlockAppend(bool flag, int value, int* __restrict__ list, int* __restrict__ count,
                            int* wcount /* LDS >= 16 */, int* base_s /* LDS */) {
  int tot = 0;
  const int rank = BlockRank(flag, wcount, &tot);
  if (threadIdx.x == 0 && tot > 0) *base_s = atomicAdd(count, tot);
  __syncthreads();
  if (flag) list[*base_s + rank] = value;
  __syncthreads();
}

// Sizes of the column lists of one product, the largest Y column among the
// short columns, and the summed expansion of the columns in the LDS tiers.
struct TierCounts {
  int n[4];  // short, long (LDS tiers and stamp), big (expansion over kEscCap), largest short-column |Y(:, j)|
  unsigned long long mid_expansion;
};

// k-way merge of one short column by one thread: the union, in ascending
// row order, of the sorted row lists X(:, k) over the entries k of Y(:, j)
// (list b is X(:, k_b), at most NL of them) and, with add_identity, of {j}.
// Each list keeps its next unread position, its end and a queue of its next
// D rows (and values), loaded D merge steps ahead of their use so that a step
// does not wait on the loads of the step before it (INT_MAX past the end).
template <int NL>
__device__ __forceinline__ void OpenLists(int y0, int ny, const int* __restrict__ yidx, const int* __restrict__ xptr,
                                          int (&pos)[NL], int (&end)[NL]) {
  int k[NL];
#pragma unroll
  for (int b = 0; b < NL; ++b) k[b] = b < ny ? yidx[y0 + b] : -1;
#pragma unroll
  for (int b = 0; b < NL; ++b) {
    pos[b] = k[b] >= 0 ? xptr[k[b]] : 0;
    end[b] = k[b] >= 0 ? xptr[k[b] + 1] : 0;
  }
}

template <int NL, int D, bool kValues>
struct MergeQueues {
  int pos[NL], end[NL];
  int row[D][NL];
  double val[D][NL];
  __device__ __forceinline__ void Fill(const int* __restrict__ xidx, const double* __restrict__ xval) {
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int b = 0; b < NL; ++b) {
        const bool h = pos[b] + d < end[b];
        row[d][b] = h ? xidx[pos[b] + d] : INT_MAX;
        if (kValues) val[d][b] = h ? xval[pos[b] + d] : 0.0;
      }
  }
  // Smallest head row (ident included).
  __device__ __forceinline__ int Min(int ident) const {
    int m = ident;
#pragma unroll
    for (int b = 0; b < NL; ++b) m = min(m, row[0][b]);
    return m;
  }
  // Pops every head equal to m, adding its value times y[b] to acc in list
  // order when kValues.
  __device__ __forceinline__ void Pop(int m, const int* __restrict__ xidx, const double* __restrict__ xval,
                                      const double (&y)[NL], double& acc) {
#pragma unroll
    for (int b = 0; b < NL; ++b) {
      if (row[0][b] == m) {
        if (kValues) acc += val[0][b] * y[b];
        ++pos[b];
#pragma unroll
        for (int d = 0; d + 1 < D; ++d) {
          row[d][b] = row[d + 1][b];
          if (kValues) val[d][b] = val[d + 1][b];
        }
        const bool h = pos[b] + D - 1 < end[b];
        row[D - 1][b] = h ? xidx[pos[b] + D - 1] : INT_MAX;
        if (kValues) val[D - 1][b] = h ? xval[pos[b] + D - 1] : 0.0;
      }
    }
  }
};

// u_j = sum over Y(:, j) of nnz(X(:, k)) (+1 for the identity), one thread
// per column, and the column's tier: short columns are counted (and their
// largest |Y(:, j)| kept), the others go to long_list (the
// wavefront-per-column kernels) and, when the expansion exceeds the LDS
// tier, also to big_list (the stamp kernel).
constexpr int kExpBlock = 1024;
__global__ __launch_bounds__(kExpBlock) void ExpansionSizes(int ncols, const int* __restrict__ xptr,
                                                            const int* __restrict__ yptr,
                                                            const int* __restrict__ yidx, int add_identity,
                                                            int64_t* __restrict__ upper, int* __restrict__ long_list,
                                                            int* __restrict__ big_list,
                                                            TierCounts* __restrict__ tiers) {
  __shared__ int wcount[kExpBlock / kWave];
  __shared__ int base_s;
  __shared__ unsigned long long mid_s;
  __shared__ int short_s, ny_s;
  if (threadIdx.x == 0) {
    mid_s = 0;
    short_s = 0;
    ny_s = 0;
  }
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = j < ncols;
  int64_t u = add_identity ? 1 : 0;
  int ny = 0;
  if (live) {
    const int p0 = yptr[j], p1 = yptr[j + 1];
    ny = p1 - p0;
    // eight Y entries at a time: their X column bounds are loaded together
    // rather than one dependent pair per entry (k = -1 past the column's end)
    for (int p = p0; p < p1; p += 8) {
      int k[8], lo[8], hi[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) k[t] = p + t < p1 ? yidx[p + t] : -1;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        lo[t] = k[t] >= 0 ? xptr[k[t]] : 0;
        hi[t] = k[t] >= 0 ? xptr[k[t] + 1] : 0;
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) u += hi[t] - lo[t];
    }
    upper[j] = u;
  }
  const bool is_short = live && IsShortColumn(u, ny);
  // short-column count and largest short |Y(:, j)|: one LDS atomic per
  // wavefront, one global atomic per block
  const uint64_t sm = __ballot(is_short);
  const int wmax = WaveMaxInt(is_short ? ny : 0);
  if (LaneId() == 0 && sm) {
    atomicAdd(&short_s, __popcll(sm));
    atomicMax(&ny_s, wmax);
  }
  BlockAppend(live && !is_short, j, long_list, tiers->n + 1, wcount, &base_s);
  BlockAppend(live && u > kEscCap, j, big_list, tiers->n + 2, wcount, &base_s);
  const int64_t mid = live && !is_short && u <= kEscCap ? u : 0;
  const int64_t wsum = WaveInclusiveScan64(mid);
  if (LaneId() == kWave - 1 && wsum > 0) atomicAdd(&mid_s, (unsigned long long)wsum);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (mid_s > 0) atomicAdd(&tiers->mid_expansion, mid_s);
    if (short_s > 0) {
      atomicAdd(&tiers->n[0], short_s);
      atomicMax(&tiers->n[3], ny_s);
    }
  }
}

constexpr int kMergeDepth = 3;     // rows loaded ahead per list (count)
constexpr int kMergeDepthNum = 2;  // rows and values loaded ahead per list (numeric)

// Distinct rows of the k-way merge of column j's NL (or fewer) lists.
template <int NL>
__device__ __forceinline__ int CountMerge(int j, int y0, int ny, const int* __restrict__ xptr,
                                          const int* __restrict__ xidx, const int* __restrict__ yidx,
                                          int add_identity) {
  MergeQueues<NL, kMergeDepth, false> q;
  OpenLists<NL>(y0, ny, yidx, xptr, q.pos, q.end);
  q.Fill(xidx, nullptr);
  const double y[NL] = {};
  double acc = 0.0;
  int ident = add_identity ? j : INT_MAX;
  int cnt = 0;
  for (;;) {
    const int m = q.Min(ident);
    if (m == INT_MAX) break;
    ++cnt;
    if (ident == m) ident = INT_MAX;
    q.Pop(m, xidx, nullptr, y, acc);
  }
  return cnt;
}

// Symbolic phase of the short columns, one thread per column: counts[j] =
// the number of distinct rows of the merge (0 for the other columns).  Each
// wavefront merges with the smallest list count (4, 8 or NL) that holds its
// short columns, so a stencil product's 7-list columns do not walk 16 lists.
template <int NL>
__global__ __launch_bounds__(256) void MergeCount(int ncols, const int* __restrict__ xptr,
                                                  const int* __restrict__ xidx, const int* __restrict__ yptr,
                                                  const int* __restrict__ yidx, const int64_t* __restrict__ upper,
                                                  int add_identity, int* __restrict__ counts) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = j < ncols;
  const int y0 = live ? yptr[j] : 0, ny = live ? yptr[j + 1] - y0 : 0;
  const bool active = live && ny <= NL && IsShortColumn(upper[j], ny);
  const int wmax = WaveMaxInt(active ? ny : 0);
  if (!active) {
    if (live) counts[j] = 0;  // a long column: its count comes from the wavefront tiers
    return;
  }
  int cnt;
  if (wmax <= 4)
    cnt = CountMerge<4>(j, y0, ny, xptr, xidx, yidx, add_identity);
  else if (wmax <= 8 || NL <= 8)
    cnt = CountMerge<(NL < 8 ? NL : 8)>(j, y0, ny, xptr, xidx, yidx, add_identity);
  else
    cnt = CountMerge<NL>(j, y0, ny, xptr, xidx, yidx, add_identity);
  counts[j] = cnt;
}

// Numeric phase of the short columns, one thread per column: the same merge,
// writing each distinct row and its value to C in row order.  A row's value
// is its products x(i, k) y(k, j) added to 0 in ascending k -- the lists in
// order -- as NumericGustavson adds them.  residual: -(sum) (+1 on the
// diagonal, which the identity list inserts when the product lacks it).
// NL >= the largest short |Y(:, j)|; long columns produce nothing here.
// Outputs go through LDS, kMergeChunk merge steps at a time: each lane
// stages its next entries, then eight lanes write one column's chunk, so a
// store instruction covers eight contiguous runs instead of 64 scattered
// entries.
constexpr int kMergeChunk = 8;
constexpr int kNumMergeWaves = 4;
template <int NL>
__global__ __launch_bounds__(kNumMergeWaves* kWave) void MergeNumeric(
    int ncols, const int* __restrict__ xptr, const int* __restrict__ xidx, const double* __restrict__ xval,
    const int* __restrict__ yptr, const int* __restrict__ yidx, const double* __restrict__ yval,
    const int64_t* __restrict__ upper, int residual, const int* __restrict__ cptr, int* __restrict__ cidx,
    double* __restrict__ cval) {
  __shared__ int sidx_all[kNumMergeWaves][kWave * kMergeChunk];
  __shared__ double sval_all[kNumMergeWaves][kWave * kMergeChunk];
  int* const sidx = sidx_all[WaveInBlock()];
  double* const sval = sval_all[WaveInBlock()];
  const int lane = LaneId();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  bool active = false;
  int y0 = 0, ny = 0;
  if (j < ncols) {
    y0 = yptr[j];
    ny = yptr[j + 1] - y0;
    active = ny <= NL && IsShortColumn(upper[j], ny);
  }
  if (!__ballot(active)) return;  // whole wavefront
  MergeQueues<NL, kMergeDepthNum, true> q;
  double y[NL];
  OpenLists<NL>(y0, active ? ny : 0, yidx, xptr, q.pos, q.end);
#pragma unroll
  for (int b = 0; b < NL; ++b) y[b] = b < ny && active ? yval[y0 + b] : 0.0;
  q.Fill(xidx, xval);
  int ident = residual && active ? j : INT_MAX;
  int out = active ? cptr[j] : 0;
  for (;;) {
    // up to kMergeChunk merge steps into this lane's LDS slots
    int cnt = 0;
    for (; cnt < kMergeChunk; ++cnt) {
      const int m = q.Min(ident);
      if (m == INT_MAX) break;
      double acc = 0.0;
      q.Pop(m, xidx, xval, y, acc);
      if (ident == m) ident = INT_MAX;
      double v = acc;
      if (residual) {
        v = -v;
        if (m == j) v += 1.0;
      }
      sidx[lane * kMergeChunk + cnt] = m;
      sval[lane * kMergeChunk + cnt] = v;
    }
    WaveFence();
    // eight lanes per column chunk: lane 8 r + e writes entry e of the chunk
    // of column lane (8 r' + lane / 8) in round r'
#pragma unroll
    for (int r = 0; r < kWave / 8; ++r) {
      const int src = r * 8 + lane / 8, e = lane & 7;
      const int c = __shfl(cnt, src, kWave), o = __shfl(out, src, kWave);
      if (e < c) {
        cidx[o + e] = sidx[src * kMergeChunk + e];
        cval[o + e] = sval[src * kMergeChunk + e];
      }
    }
    out += cnt;
    if (!__ballot(cnt == kMergeChunk && q.Min(ident) != INT_MAX)) break;
    WaveFence();
  }
}

__global__ void ZeroTiers(TierCounts* __restrict__ t) { *t = TierCounts{}; }

// cnt[g] = the expansion of long-list column g when it is in the LDS tiers
// (the stash offsets of the stashing count pass), else 0.
__global__ void MidExpansions(int nlist, const int* __restrict__ list, const int64_t* __restrict__ upper,
                              int* __restrict__ cnt) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nlist) return;
  const int64_t u = upper[list[g]];
  cnt[g] = u <= kEscCap ? int(u) : 0;
}

// Symbolic phase, one wavefront per output column whose expansion fits the
// LDS tier: counts (fill == 0) or writes in row order (fill == 1) the
// distinct rows of X(:, k) over k in Y(:, j) (and j itself when add_identity);
// with fill == 0 and a stash, also writes them to stash[soff[g] ...].
// X's columns are sorted, so their first and last rows bound the column's row
// range.  When that range spans fewer than 32 kEscCap rows, the rows are
// marked in an LDS bitmap of the range and read back in order (popcounts per
// lane, a wave scan, then the set bits); otherwise the expansion is gathered
// into LDS, bitonic-sorted and compressed.

__global__ __launch_bounds__(kSymWaves* kWave) void SymbolicEsc(
    int nlist, const int* __restrict__ list, const int* __restrict__ xptr, const int* __restrict__ xidx, const int* __restrict__ yptr,
    const int* __restrict__ yidx, const int64_t* __restrict__ upper, int add_identity, int fill,
    int* __restrict__ counts, const int* __restrict__ cptr, int* __restrict__ cidx, const int* __restrict__ soff,
    int* __restrict__ stash) {
  __shared__ __align__(16) int keys[kSymWaves][kEscCap];
  __shared__ int pre[kSymWaves][kWave];
  __shared__ int beg[kSymWaves][kWave];
  const int w = WaveInBlock(), lane = LaneId();
  const int g = blockIdx.x * kSymWaves + w;
  if (g >= nlist) return;
  const int j = list[g];
  // an expansion over kEscCap is not stashed (the fill pass writes its rows)
  const bool wide = upper[j] > kEscCap;
  if (fill && stash != nullptr && !wide) return;  // its rows are in the stash already
  if (!fill && wide) stash = nullptr;
  int* K = keys[w];
  const int y0 = yptr[j], ye = yptr[j + 1];
  // the first 64 Y entries (most columns have no more) stay in registers:
  // lane l holds entry y0 + l's X column start and length
  int b1 = 0, c1 = 0;
  if (y0 + lane < ye) {
    const int k = yidx[y0 + lane];
    b1 = xptr[k];
    c1 = xptr[k + 1] - b1;
  }
  // row range of the expansion
  int lo_r = add_identity ? j : INT_MAX, hi_r = add_identity ? j : -1;
  if (c1 > 0) {
    lo_r = min(lo_r, xidx[b1]);
    hi_r = max(hi_r, xidx[b1 + c1 - 1]);
  }
  for (int p = y0 + kWave + lane; p < ye; p += kWave) {
    const int k = yidx[p];
    const int b = xptr[k], e = xptr[k + 1];
    if (e > b) {
      lo_r = min(lo_r, xidx[b]);
      hi_r = max(hi_r, xidx[e - 1]);
    }
  }
  lo_r = WaveMinInt(lo_r);
  hi_r = WaveMaxInt(hi_r);
  const bool bitmap = BitmapRange(lo_r, hi_r);
  if (wide && !bitmap) return;  // the stamp kernel's column
  const int nwords = bitmap ? ((hi_r - lo_r) >> 5) + 1 : 0;
  if (bitmap) {
    for (int i = lane; i < (nwords + 3) / 4; i += kWave) reinterpret_cast<int4*>(K)[i] = make_int4(0, 0, 0, 0);
    WaveFence();
  }
  // gather: the expansion entries of 64 Y entries at a time, kEscRegs per
  // lane at once (e = e0 + lane + 64 r): the source entry by binary search,
  // then the loads, then the marks
  int off = 0;
  for (int p0 = y0; p0 < ye; p0 += kWave) {
    int b = b1, c = c1;
    if (p0 != y0) {
      b = 0;
      c = 0;
      if (p0 + lane < ye) {
        const int k = yidx[p0 + lane];
        b = xptr[k];
        c = xptr[k + 1] - b;
      }
    }
    const int nyc = min(kWave, ye - p0);
    const int incl = WaveInclusiveScan(c);
    const int tot = __shfl(incl, kWave - 1, kWave);
    pre[w][lane] = incl - c;
    beg[w][lane] = b;
    WaveFence();
    for (int e0 = 0; e0 < tot; e0 += kEscRegs * kWave) {
      int q[kEscRegs];
#pragma unroll
      for (int r = 0; r < kEscRegs; ++r) {
        const int e = e0 + lane + kWave * r;
        q[r] = -1;
        if (e < tot) {
          int lo = 0, hi = nyc;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pre[w][mid] <= e)
              lo = mid;
            else
              hi = mid;
          }
          q[r] = beg[w][lo] + (e - pre[w][lo]);
        }
      }
      int row[kEscRegs];
#pragma unroll
      for (int r = 0; r < kEscRegs; ++r) row[r] = q[r] >= 0 ? xidx[q[r]] : 0;
#pragma unroll
      for (int r = 0; r < kEscRegs; ++r) {
        if (q[r] < 0) continue;
        if (bitmap) {
          const int rr = row[r] - lo_r;
          atomicOr(&K[rr >> 5], int(1u << (rr & 31)));
        } else {
          K[off + e0 + lane + kWave * r] = row[r];
        }
      }
    }
    off += tot;
    WaveFence();
  }
  if (bitmap) {
    if (add_identity && lane == 0) atomicOr(&K[(j - lo_r) >> 5], int(1u << ((j - lo_r) & 31)));
    WaveFence();
    // lane ranges of a multiple of 4 words, read 16 bytes at a time (the
    // words past nwords up to the next multiple of 4 were zeroed)
    const int per = ((nwords + kWave - 1) / kWave + 3) & ~3;
    const int w0 = min(nwords, lane * per), w1 = min(nwords, w0 + per);
    int cnt = 0;
    for (int i = w0; i < w1; i += 4) {
      const int4 q = *reinterpret_cast<const int4*>(K + i);
      cnt += __popc(uint32_t(q.x)) + __popc(uint32_t(q.y)) + __popc(uint32_t(q.z)) + __popc(uint32_t(q.w));
    }
    if (!fill && stash == nullptr) {
      cnt = WaveSumInt(cnt);
      if (lane == 0) counts[j] = cnt;
    } else {
      const int incl = WaveInclusiveScan(cnt);
      if (!fill && lane == kWave - 1) counts[j] = incl;
      int* const out = fill ? cidx + cptr[j] : stash + soff[g];
      int o = incl - cnt;
      for (int i = w0; i < w1; i += 4) {
        const int4 q = *reinterpret_cast<const int4*>(K + i);
        const int qw[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          uint32_t word = uint32_t(qw[t]);
          while (word) {
            const int bit = __ffs(word) - 1;
            out[o++] = lo_r + (i + t) * 32 + bit;
            word &= word - 1u;
          }
        }
      }
    }
    return;
  }
  if (add_identity) {
    if (lane == 0) K[off] = j;
    off += 1;
  }
  int n2 = 1;
  while (n2 < off) n2 <<= 1;
  for (int e = off + lane; e < n2; e += kWave) K[e] = INT_MAX;
  WaveFence();
  for (int k = 2; k <= n2; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = lane; i < n2; i += kWave) {
        const int ixj = i ^ jj;
        if (ixj > i) {
          const int a = K[i], b = K[ixj];
          const bool asc = (i & k) == 0;
          if (asc ? (a > b) : (a < b)) {
            K[i] = b;
            K[ixj] = a;
          }
        }
      }
      WaveFence();
    }
  }
  if (!fill && stash == nullptr) {
    int cnt = 0;
    for (int i = lane; i < off; i += kWave) cnt += (i == 0 || K[i] != K[i - 1]) ? 1 : 0;
    cnt = WaveSumInt(cnt);
    if (lane == 0) counts[j] = cnt;
  } else {
    int* const out = fill ? cidx + cptr[j] : stash + soff[g];
    int base = 0;
    for (int i0 = 0; i0 < off; i0 += kWave) {
      const int i = i0 + lane;
      const bool flag = i < off && (i == 0 || K[i] != K[i - 1]);
      const uint64_t mask = __ballot(flag);
      if (flag) out[base + LanePrefixCount(mask)] = K[i];
      base += __popcll(mask);
    }
    if (!fill && lane == 0) counts[j] = base;
  }
}

// The columns whose expansion exceeds kEscCap, one wavefront each: those
// whose row range fits the LDS bitmap stay with SymbolicEsc; the others are
// appended to stamp_list (count in *nstamp) for the stamp kernel.
__global__ __launch_bounds__(kSymWaves* kWave) void BigColumnRanges(int nbig, const int* __restrict__ big_list,
                                                                    const int* __restrict__ xptr,
                                                                    const int* __restrict__ xidx,
                                                                    const int* __restrict__ yptr,
                                                                    const int* __restrict__ yidx, int add_identity,
                                                                    int* __restrict__ stamp_list,
                                                                    int* __restrict__ nstamp) {
  const int lane = LaneId();
  const int g = blockIdx.x * kSymWaves + WaveInBlock();
  if (g >= nbig) return;
  const int j = big_list[g];
  int lo_r = add_identity ? j : INT_MAX, hi_r = add_identity ? j : -1;
  for (int p = yptr[j] + lane; p < yptr[j + 1]; p += kWave) {
    const int k = yidx[p];
    const int b = xptr[k], e = xptr[k + 1];
    if (e > b) {
      lo_r = min(lo_r, xidx[b]);
      hi_r = max(hi_r, xidx[e - 1]);
    }
  }
  lo_r = WaveMinInt(lo_r);
  hi_r = WaveMaxInt(hi_r);
  if (lane == 0 && !BitmapRange(lo_r, hi_r)) stamp_list[atomicAdd(nstamp, 1)] = j;
}

// Symbolic phase for columns whose expansion exceeds the LDS tier: a block per
// column marks rows in a per-block stamp array (length nrows, initialised to
// -1) and compacts them in row order.
__global__ __launch_bounds__(kBlock) void SymbolicStamp(int nlist, const int* __restrict__ list, int nrows,
                                                        const int* __restrict__ xptr, const int* __restrict__ xidx,
                                                        const int* __restrict__ yptr, const int* __restrict__ yidx,
                                                        int add_identity, int fill, int* __restrict__ stamp_ws,
                                                        int* __restrict__ counts, const int* __restrict__ cptr,
                                                        int* __restrict__ cidx) {
  __shared__ int wcount[16];
  int* stamp = stamp_ws + size_t(blockIdx.x) * size_t(nrows);
  for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
    const int j = list[li];
    for (int p = yptr[j]; p < yptr[j + 1]; ++p) {
      const int k = yidx[p];
      for (int q = xptr[k] + threadIdx.x; q < xptr[k + 1]; q += blockDim.x) stamp[xidx[q]] = j;
    }
    if (add_identity && threadIdx.x == 0) stamp[j] = j;
    __syncthreads();
    int base = fill ? cptr[j] : 0;
    for (int r0 = 0; r0 < nrows; r0 += blockDim.x) {
      const int r = r0 + threadIdx.x;
      const bool flag = r < nrows && stamp[r] == j;
      int tot = 0;
      const int rank = BlockRank(flag, wcount, &tot);
      if (fill && flag) cidx[base + rank] = r;
      base += tot;
    }
    if (!fill && threadIdx.x == 0) counts[j] = base;
    __syncthreads();
  }
}

// Numeric phase, one wavefront per output column.  Contributions are added in
// Eigen's order: Y(:, j)'s entries k ascending, each scattering X(:, k) into
// the column's accumulator (distinct rows inside one k, so the lanes never
// collide).  residual: value = -(sum) (+1 on the diagonal).
//
// Columns of at most kNumCap rows accumulate in LDS.  Y's entries are taken
// 64 at a time (lane l holds entry l: X column start, length, y), and their
// expansion kNumRegs * 64 entries at a time: every lane loads and places its
// entries (e = lane + 64 r) up front, then register by register adds them
// grouped by Y entry in ascending order, so the loads and searches of a chunk
// overlap and only the additions are ordered.  Longer columns accumulate in
// global memory, one Y entry at a time.
//
// Two instances cover the long columns: kCap = kNumCap with kNumWaves
// wavefronts per block, and kCap = kNumCapWide with one (48 KiB of LDS) for
// the columns of kNumCap < L <= kNumCapWide rows.  Each is launched over the
// whole list and skips the columns of the other (L <= min_len: the smaller
// instance's; L > kCap: the global path, run by the instance with
// global_long).
constexpr int kNumRegs = 4;
constexpr int kNumCapWide = 4096;
template <int kCap, int kWaves>
__global__ __launch_bounds__(kWaves* kWave) void NumericGustavson(
    int nlist, const int* __restrict__ list, const int* __restrict__ xptr, const int* __restrict__ xidx,
    const double* __restrict__ xval, const int* __restrict__ yptr, const int* __restrict__ yidx,
    const double* __restrict__ yval, const int* __restrict__ cptr, int* __restrict__ cidx,
    double* __restrict__ cval, int residual, int* __restrict__ error_flag, const int64_t* __restrict__ upper,
    const int* __restrict__ soff, const int* __restrict__ stash, int min_len, int global_long) {
  __shared__ int rows_s[kWaves][kCap];
  __shared__ double acc_s[kWaves][kCap];
  __shared__ int pre_s[kWaves][kWave];
  __shared__ int beg_s[kWaves][kWave];
  __shared__ double y_s[kWaves][kWave];
  const int w = WaveInBlock(), lane = LaneId();
  const int g = blockIdx.x * kWaves + w;
  if (g >= nlist) return;
  const int j = list[g];
  const int cb = cptr[j];
  const int L = cptr[j + 1] - cb;
  if (L <= min_len || (L > kCap && !global_long)) return;  // the other instance's column
  const int y0 = yptr[j], ye = yptr[j + 1];
  // rows of an LDS-tier column whose count pass stashed them (the fill pass
  // skipped it): read from the stash and placed in C here
  const bool stashed = stash != nullptr && upper[j] <= kEscCap;
  const int* const rows = stashed ? stash + soff[g] : cidx + cb;
  if (L > kCap) {
    for (int e = lane; e < L; e += kWave) cval[cb + e] = 0.0;
    WaveFence();
    const int* R = rows;
    for (int p = y0; p < ye; ++p) {
      const int k = yidx[p];
      const double y = yval[p];
      for (int q = xptr[k] + lane; q < xptr[k + 1]; q += kWave) {
        const int i = xidx[q];
        const double x = xval[q];
        const int pos = LowerBound(R, 0, L, i);
        if (pos >= L || R[pos] != i) {
          *error_flag = 1;
          continue;
        }
        atomicAdd(&cval[cb + pos], x * y);
      }
      WaveFence();
    }
    for (int e = lane; e < L; e += kWave) {
      double v = __hip_atomic_load(&cval[cb + e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (residual) {
        v = -v;
        if (R[e] == j) v += 1.0;
      }
      if (stashed) cidx[cb + e] = R[e];
      cval[cb + e] = v;
    }
    return;
  }
  int* const R = rows_s[w];
  double* const acc = acc_s[w];
  for (int e = lane; e < L; e += kWave) {
    R[e] = rows[e];
    acc[e] = 0.0;
  }
  for (int p0 = y0; p0 < ye; p0 += kWave) {
    const int p = p0 + lane;
    const int nyc = min(kWave, ye - p0);
    int b = 0, c = 0;
    double y = 0.0;
    if (p < ye) {
      const int k = yidx[p];
      y = yval[p];
      b = xptr[k];
      c = xptr[k + 1] - b;
    }
    const int incl = WaveInclusiveScan(c);
    const int tot = __shfl(incl, kWave - 1, kWave);
    pre_s[w][lane] = incl - c;
    beg_s[w][lane] = b;
    y_s[w][lane] = y;
    WaveFence();
    for (int e0 = 0; e0 < tot; e0 += kNumRegs * kWave) {
      int src[kNumRegs], pos[kNumRegs];
      double prod[kNumRegs];
#pragma unroll
      for (int r = 0; r < kNumRegs; ++r) {
        const int e = e0 + lane + kWave * r;
        src[r] = -1;
        pos[r] = -1;
        prod[r] = 0.0;
        if (e < tot) {
          int lo = 0, hi = nyc;  // the last Y entry whose first expansion entry is <= e
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pre_s[w][mid] <= e)
              lo = mid;
            else
              hi = mid;
          }
          const int q = beg_s[w][lo] + (e - pre_s[w][lo]);
          const int i = xidx[q];
          prod[r] = xval[q] * y_s[w][lo];
          src[r] = lo;
          const int at = LowerBound(R, 0, L, i);
          if (at < L && R[at] == i)
            pos[r] = at;
          else
            *error_flag = 1;
        }
      }
#pragma unroll
      for (int r = 0; r < kNumRegs; ++r) {
        const int first = e0 + kWave * r;
        if (first >= tot) break;
        const int s_lo = __shfl(src[r], 0, kWave);
        const int s_hi = __shfl(src[r], min(tot - 1 - first, kWave - 1), kWave);
        for (int sv = s_lo; sv <= s_hi; ++sv) {
          if (src[r] == sv && pos[r] >= 0) acc[pos[r]] += prod[r];
          WaveFence();
        }
      }
    }
    WaveFence();
  }
  for (int e = lane; e < L; e += kWave) {
    double v = acc[e];
    if (residual) {
      v = -v;
      if (R[e] == j) v += 1.0;
    }
    if (stashed) cidx[cb + e] = R[e];
    cval[cb + e] = v;
  }
}

// ------------------------------------------------------------------ filters
// Keep flags -> per-column counts.
__global__ void KeepToInt(int64_t nnz, const unsigned char* __restrict__ keep, int* __restrict__ flags) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nnz) flags[i] = keep[i];
}

// out[j] = pos[ptr[j]] for j <= ncols.
__global__ void GatherPtr(int64_t m, const int* __restrict__ ptr, const int* __restrict__ pos, int* __restrict__ out) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < m) out[j] = pos[ptr[j] - ptr[0]];
}

__global__ void ScatterKept(int64_t nnz, const int* __restrict__ idx, const double* __restrict__ val,
                            const unsigned char* __restrict__ keep, const int* __restrict__ pos,
                            int* __restrict__ oidx, double* __restrict__ oval) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nnz || !keep[i]) return;
  const int o = pos[i];
  oidx[o] = idx[i];
  oval[o] = val[i];
}

// dropped (nullable) counts the entries not kept.
__global__ void FlagNonZero(int64_t nnz, const double* __restrict__ val, unsigned char* __restrict__ keep,
                            int* __restrict__ dropped) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const double v = val[i];
  const bool k = !(fabs(v) <= 0.0);  // prune(0.0): keep |v| > 0 and NaN
  keep[i] = k ? 1 : 0;
  if (!k && dropped) atomicAdd(dropped, 1);
}

// Up to this many entries, PruneZeros filters at once instead of first
// reading back whether any entry is dropped (one host round trip less).
constexpr int64_t kPruneFilterAlways = int64_t(1) << 20;

// ------------------------------------------------------------------ elementwise
// out = d(idx) .* yv over the stored entries ptr[0] + [0, nnz).
__global__ void DiagScaleKernel(int64_t nnz, const int* __restrict__ ptr, const int* __restrict__ idx,
                                const double* __restrict__ d, const double* __restrict__ yv,
                                double* __restrict__ out) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nnz) return;
  const int64_t p = ptr[0] + t;
  out[p] = d[idx[p]] * yv[p];
}

__global__ void ColScaleKernel(int ncols, const int* __restrict__ ptr, const double* __restrict__ f,
                               const double* __restrict__ xv, double* __restrict__ out) {
  const int j = blockIdx.x;
  const double s = f[j];
  for (int p = ptr[j] + threadIdx.x; p < ptr[j + 1]; p += blockDim.x) out[p] = xv[p] * s;
}

// ------------------------------------------------------------------ merges
__device__ __forceinline__ double MergeBoth(int op, double l, double r, double a, double b) {
  switch (op) {
    case 0: return l + a * r;
    case 1: return l * b + a * r;
    case 2: return 0.5 * (l + r);
    default: return l;
  }
}
__device__ __forceinline__ double MergeLeft(int op, double l, double a, double b) {
  switch (op) {
    case 0: return l + 0.0;
    case 1: return l * b + 0.0;
    case 2: return 0.5 * (l + 0.0);
    default: return l;
  }
}
__device__ __forceinline__ double MergeRight(int op, double r, double a, double b) {
  switch (op) {
    case 0: return 0.0 + a * r;
    case 1: return 0.0 + a * r;
    case 2: return 0.5 * (0.0 + r);
    default: return r;
  }
}

// Merge path of two sorted columns with ties taken from the left first: the
// number of left entries among the first d entries of the merged sequence.
__device__ __forceinline__ int MergePathSplit(const int* __restrict__ li, int L, const int* __restrict__ ri, int R,
                                              int d) {
  int lo = max(0, d - R), hi = min(d, L);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ri[d - mid - 1] < li[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Union of two sorted columns, one wavefront per column: lane l walks
// positions [l c, (l + 1) c) of the merged sequence (c = ceil((L + R) / 64)),
// found by a merge-path search.  An entry present in both columns is one
// union entry, handled by the lane that holds its left copy; the lane after
// skips the right copy.  The union entries and their values are those of
// the sequential merge (Eigen's binary iterator order), so fill == 1 writes
// the same column as a thread-per-column merge would.  fill == 0 counts kept
// entries.
//
// Columns with L + R <= kMergeStage first copy both index lists into LDS and
// do the search and the walk there; the walk then only records, for each
// union entry, where it comes from, and the values are loaded, combined and
// written by consecutive lanes for consecutive entries.
constexpr int kMergeWaves = 4;
constexpr int kMergeStage = 1024;
// union entry record: kind (0 left only, 1 both, 2 right only), x, y
__device__ __forceinline__ int MergeRecord(int kind, int x, int y) { return (kind << 20) | (x << 10) | y; }

__global__ __launch_bounds__(kMergeWaves* kWave) void MergeWave(
    int ncols, const int* __restrict__ lp, const int* __restrict__ li, const double* __restrict__ lv,
    const int* __restrict__ rp, const int* __restrict__ ri, const double* __restrict__ rv, int op, double a, double b,
    int prune, int fill, int* __restrict__ counts, const int* __restrict__ optr, int* __restrict__ oidx,
    double* __restrict__ oval) {
  __shared__ int stage_s[kMergeWaves][kMergeStage];
  __shared__ int rec_s[kMergeWaves][kMergeStage];
  const int lane = LaneId(), w = WaveInBlock();
  const int j = blockIdx.x * kMergeWaves + w;
  if (j >= ncols) return;
  const int xb = lp[j], L = lp[j + 1] - xb, yb = rp[j], R = rp[j + 1] - yb;
  const int T = L + R;
  if (T <= kMergeStage) {
    int* const S = stage_s[w];
    int* const rec = rec_s[w];
    for (int t = lane; t < L; t += kWave) S[t] = li[xb + t];
    for (int t = lane; t < R; t += kWave) S[L + t] = ri[yb + t];
    WaveFence();
    const int* const lx = S;
    const int* const ry = S + L;
    const int c = (T + kWave - 1) / kWave;
    const int d0 = min(T, lane * c), d1 = min(T, d0 + c);
    int x0 = MergePathSplit(lx, L, ry, R, d0), y0 = d0 - x0;
    if (d0 < d1 && x0 > 0 && y0 < R && ry[y0] == lx[x0 - 1]) ++y0;
    auto walk = [&](bool write, int base) {
      int x = x0, y = y0, u = 0;
      while (x + y < d1) {
        int kind;
        if (x < L && (y >= R || lx[x] <= ry[y])) {
          kind = (y < R && lx[x] == ry[y]) ? 1 : 0;
          if (write) rec[base + u] = MergeRecord(kind, x, kind ? y : 0);
          if (kind == 1) ++y;
          ++x;
        } else {
          if (write) rec[base + u] = MergeRecord(2, 0, y);
          ++y;
        }
        ++u;
      }
      return u;
    };
    const int u = walk(false, 0);
    const int incl = WaveInclusiveScan(u);
    const int U = __shfl(incl, kWave - 1, kWave);  // union entries
    if (!fill && !prune) {
      if (lane == 0) counts[j] = U;
      return;
    }
    walk(true, incl - u);
    WaveFence();
    int base = fill ? optr[j] : 0, kept = 0;
    for (int s0 = 0; s0 < U; s0 += kWave) {
      const int slot = s0 + lane;
      int row = 0;
      double v = 0.0;
      bool keep_it = false;
      if (slot < U) {
        const int rc = rec[slot];
        const int kind = rc >> 20, x = (rc >> 10) & 1023, y = rc & 1023;
        if (kind == 0) {
          row = lx[x];
          v = MergeLeft(op, lv[xb + x], a, b);
        } else if (kind == 1) {
          row = lx[x];
          v = MergeBoth(op, lv[xb + x], rv[yb + y], a, b);
        } else {
          row = ry[y];
          v = MergeRight(op, rv[yb + y], a, b);
        }
        keep_it = !(prune && !(fabs(v) > 0.0) && !isnan(v));
      }
      const uint64_t mask = __ballot(keep_it);
      if (fill && keep_it) {
        const int o = base + LanePrefixCount(mask);
        oidx[o] = row;
        oval[o] = v;
      }
      base += __popcll(mask);
      kept += __popcll(mask);
    }
    if (!fill && lane == 0) counts[j] = kept;
    return;
  }
  const int* const lx = li + xb;
  const int* const ry = ri + yb;
  const int c = (T + kWave - 1) / kWave;
  const int d0 = min(T, lane * c), d1 = min(T, d0 + c);
  int x0 = MergePathSplit(lx, L, ry, R, d0), y0 = d0 - x0;
  if (d0 < d1 && x0 > 0 && y0 < R && ry[y0] == lx[x0 - 1]) ++y0;  // right copy of the previous lane's last entry
  auto walk = [&](bool write, int base) {
    int x = x0, y = y0, kept = 0;
    while (x + y < d1) {
      int row;
      double v;
      if (x < L && (y >= R || lx[x] <= ry[y])) {
        row = lx[x];
        if (y < R && lx[x] == ry[y]) {
          v = MergeBoth(op, lv[xb + x], rv[yb + y], a, b);
          ++y;
        } else {
          v = MergeLeft(op, lv[xb + x], a, b);
        }
        ++x;
      } else {
        row = ry[y];
        v = MergeRight(op, rv[yb + y], a, b);
        ++y;
      }
      if (prune && !(fabs(v) > 0.0) && !isnan(v)) continue;
      if (write) {
        oidx[base + kept] = row;
        oval[base + kept] = v;
      }
      ++kept;
    }
    return kept;
  };
  const int kept = walk(false, 0);
  if (!fill) {
    const int total = WaveSumInt(kept);
    if (lane == 0) counts[j] = total;
  } else {
    const int excl = WaveInclusiveScan(kept) - kept;
    walk(true, optr[j] + excl);
  }
}

// ------------------------------------------------------------------ reductions
// Per-column intersection dot (same == 0) or squared norm (same == 1) into
// out[j], one wavefront per column.  The products are formed in parallel (the
// matches of an intersection by merge-path segments, as in MergeWave), placed
// in LDS in the column's row order, and added by one lane in that order: the
// order of the sequential loop (ColumnDot), so out[j] is bit-identical to it.
// An intersection with more than kDotCap matches takes ColumnDot in lane 0.
// When L + R <= kDotCap, both index lists are first staged in the upper half
// of the product buffer, the search and walk run there and record the
// matches, and consecutive lanes form consecutive products.
constexpr int kDotWaves = 4;
constexpr int kDotCap = 1024;

// s + P[0] + P[1] + ... + P[n-1], added in that order by the calling lane;
// P is 16-byte aligned and read 8 values ahead of the additions.
__device__ __forceinline__ double OrderedSumLds(const double* P, int n, double s) {
  const double2* P2 = reinterpret_cast<const double2*>(P);
  int e = 0;
  for (; e + 8 <= n; e += 8) {
    const double2 a = P2[e / 2], b = P2[e / 2 + 1], c = P2[e / 2 + 2], d = P2[e / 2 + 3];
    s += a.x;
    s += a.y;
    s += b.x;
    s += b.y;
    s += c.x;
    s += c.y;
    s += d.x;
    s += d.y;
  }
  for (; e < n; ++e) s += P[e];
  return s;
}
// The matched products of column j of X .* Y, in the column's row order,
// into P (kDotCap doubles of the calling wavefront's LDS); returns their
// count, or -1 when there are more than kDotCap (nothing usable in P).
__device__ __forceinline__ int ColumnProducts(int j, int lane, const int* __restrict__ xp, const int* __restrict__ xi,
                                              const double* __restrict__ xv, const int* __restrict__ yp,
                                              const int* __restrict__ yi, const double* __restrict__ yv,
                                              double* P) {
  const int xb = xp[j], L = xp[j + 1] - xb;
  const int yb = yp[j], R = yp[j + 1] - yb;
  const int T = L + R;
  if (T <= kDotCap) {
    // ints [kDotCap, 2 kDotCap) of P: the index lists; ints [0, 512): the
    // match records (x << 10 | y), read back before products [0, 512) land
    int* const S = reinterpret_cast<int*>(P) + kDotCap;
    int* const rec = reinterpret_cast<int*>(P);
    for (int t = lane; t < L; t += kWave) S[t] = xi[xb + t];
    for (int t = lane; t < R; t += kWave) S[L + t] = yi[yb + t];
    WaveFence();
    const int* const lx = S;
    const int* const ry = S + L;
    const int c = (T + kWave - 1) / kWave;
    const int d0 = min(T, lane * c), d1 = min(T, d0 + c);
    int x0 = MergePathSplit(lx, L, ry, R, d0), y0 = d0 - x0;
    if (d0 < d1 && x0 > 0 && y0 < R && ry[y0] == lx[x0 - 1]) ++y0;  // matched by the previous lane
    auto walk = [&](bool write, int base) {
      int x = x0, y = y0, k = 0;
      while (x + y < d1) {
        if (x < L && (y >= R || lx[x] <= ry[y])) {
          if (y < R && lx[x] == ry[y]) {
            if (write) rec[base + k] = (x << 10) | y;
            ++k;
            ++y;
          }
          ++x;
        } else {
          ++y;
        }
      }
      return k;
    };
    const int k = walk(false, 0);
    const int incl = WaveInclusiveScan(k);
    const int total = __shfl(incl, kWave - 1, kWave);  // <= min(L, R) <= kDotCap / 2
    walk(true, incl - k);
    WaveFence();
    constexpr int kRecs = kDotCap / 2 / kWave;
    int rc[kRecs];
#pragma unroll
    for (int t = 0; t < kRecs; ++t) rc[t] = lane + kWave * t < total ? rec[lane + kWave * t] : -1;
    WaveFence();
#pragma unroll
    for (int t = 0; t < kRecs; ++t)
      if (rc[t] >= 0) P[lane + kWave * t] = xv[xb + (rc[t] >> 10)] * yv[yb + (rc[t] & 1023)];
    WaveFence();
    return total;
  }
  const int* const lx = xi + xb;
  const int* const ry = yi + yb;
  const int c = (T + kWave - 1) / kWave;
  const int d0 = min(T, lane * c), d1 = min(T, d0 + c);
  int x0 = MergePathSplit(lx, L, ry, R, d0), y0 = d0 - x0;
  if (d0 < d1 && x0 > 0 && y0 < R && ry[y0] == lx[x0 - 1]) ++y0;  // matched by the previous lane
  auto walk = [&](bool write, int base) {
    int x = x0, y = y0, k = 0;
    while (x + y < d1) {
      if (x < L && (y >= R || lx[x] <= ry[y])) {
        if (y < R && lx[x] == ry[y]) {
          if (write) P[base + k] = xv[xb + x] * yv[yb + y];
          ++k;
          ++y;
        }
        ++x;
      } else {
        ++y;
      }
    }
    return k;
  };
  const int k = walk(false, 0);
  const int incl = WaveInclusiveScan(k);
  const int total = __shfl(incl, kWave - 1, kWave);
  if (total > kDotCap) return -1;
  walk(true, incl - k);
  WaveFence();
  return total;
}

__global__ __launch_bounds__(kDotWaves* kWave) void ColumnDotWave(int ncols, const int* __restrict__ xp,
                                                                  const int* __restrict__ xi,
                                                                  const double* __restrict__ xv,
                                                                  const int* __restrict__ yp,
                                                                  const int* __restrict__ yi,
                                                                  const double* __restrict__ yv, int same,
                                                                  double* __restrict__ out) {
  __shared__ __align__(16) double prod[kDotWaves][kDotCap];
  const int lane = LaneId(), w = WaveInBlock();
  const int j = blockIdx.x * kDotWaves + w;
  if (j >= ncols) return;
  double* const P = prod[w];
  if (same) {
    const int xb = xp[j], L = xp[j + 1] - xb;
    double s = 0.0;
    for (int e0 = 0; e0 < L; e0 += kDotCap) {
      const int m = min(kDotCap, L - e0);
      for (int e = lane; e < m; e += kWave) {
        const double v = xv[xb + e0 + e];
        P[e] = v * v;
      }
      WaveFence();
      if (lane == 0) s = OrderedSumLds(P, m, s);
      WaveFence();
    }
    if (lane == 0) out[j] = s;
    return;
  }
  const int total = ColumnProducts(j, lane, xp, xi, xv, yp, yi, yv, P);
  if (lane == 0)
    out[j] = total >= 0 ? OrderedSumLds(P, total, 0.0)
                        : ColumnDot(xi, xv, xp[j], xp[j + 1], yi, yv, yp[j], yp[j + 1]);
}

// Batch-ordered sums, one wavefront per column batch b (columns
// [first[b], first[b + 1])): the batch's terms added one after another in
// CSC order from 0.0, as the reference's sequential per-batch loop does, so
// out[b] is bit-identical to it.  The terms are formed by all lanes and
// staged in LDS; lane 0 adds them (the only sequential part).
//
// Squared norm: the batch's entries are one contiguous range, taken in
// chunks of kDotCap values.  Each lane holds kSqPer values of a chunk in
// registers; the next chunk is loaded before lane 0 adds the current one, so
// its load latency hides behind the additions.
constexpr int kSqPer = kDotCap / kWave;
__global__ __launch_bounds__(kWave) void BatchSqWave(const int* __restrict__ first, const int* __restrict__ ptr,
                                                     const double* __restrict__ v, double* __restrict__ out) {
  __shared__ __align__(16) double P[kDotCap];
  const int lane = LaneId(), b = blockIdx.x;
  const int p0 = ptr[first[b]], p1 = ptr[first[b + 1]];
  double r[kSqPer];
  auto load = [&](int c0) {
#pragma unroll
    for (int t = 0; t < kSqPer; ++t) {
      const int p = c0 + lane + kWave * t;
      r[t] = p < p1 ? v[p] : 0.0;
    }
  };
  double s = 0.0;
  load(p0);
  for (int c0 = p0; c0 < p1; c0 += kDotCap) {
    const int m = min(kDotCap, p1 - c0);
#pragma unroll
    for (int t = 0; t < kSqPer; ++t) P[lane + kWave * t] = r[t] * r[t];
    WaveFence();
    if (c0 + kDotCap < p1) load(c0 + kDotCap);
    if (lane == 0) s = OrderedSumLds(P, m, s);
    WaveFence();
  }
  if (lane == 0) out[b] = s;
}

// Intersection dot: column by column, each column's matched products formed
// in parallel (ColumnProducts) and added by lane 0 to the running sum; a
// column with more than kDotCap matches is added by lane 0 sequentially.
__global__ __launch_bounds__(kWave) void BatchDotWave(const int* __restrict__ first, const int* __restrict__ xp,
                                                      const int* __restrict__ xi, const double* __restrict__ xv,
                                                      const int* __restrict__ yp, const int* __restrict__ yi,
                                                      const double* __restrict__ yv, double* __restrict__ out) {
  __shared__ __align__(16) double P[kDotCap];
  const int lane = LaneId(), b = blockIdx.x;
  double s = 0.0;
  for (int j = first[b]; j < first[b + 1]; ++j) {
    const int total = ColumnProducts(j, lane, xp, xi, xv, yp, yi, yv, P);
    if (lane == 0)
      s = total >= 0 ? OrderedSumLds(P, total, s) : ColumnDot(xi, xv, xp[j], xp[j + 1], yi, yv, yp[j], yp[j + 1], s);
    WaveFence();
  }
  if (lane == 0) out[b] = s;
}

// Deterministic sum of v[0..n): per-block partials, then one block.
__global__ __launch_bounds__(kBlock) void SumPartials(const double* __restrict__ v, int64_t n,
                                                      double* __restrict__ partial) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    s += v[i];
  s = BlockSum<double>(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Per-batch ||X_b||^2 with one block per batch: strided per-thread sums,
// then the block's fixed tree (deterministic, not the sequential order).
// Tree form of the per-batch squared norms: each batch's entries split into
// kBatchSplit equal ranges, one block each (block sums), then the ranges'
// sums added in range order (BatchSqNormJoin).
constexpr int kBatchSplit = 16;
__global__ __launch_bounds__(256) void BatchSqNormTreeKernel(const int* __restrict__ first,
                                                             const int* __restrict__ ptr,
                                                             const double* __restrict__ v,
                                                             double* __restrict__ part) {
  __shared__ double red[16];
  const int b = blockIdx.x / kBatchSplit, piece = blockIdx.x % kBatchSplit;
  const int64_t p0 = ptr[first[b]], p1 = ptr[first[b + 1]];
  const int64_t len = (p1 - p0 + kBatchSplit - 1) / kBatchSplit;
  const int64_t q0 = min(p1, p0 + piece * len), q1 = min(p1, q0 + len);
  double s = 0.0;
  for (int64_t p = q0 + threadIdx.x; p < q1; p += blockDim.x) s += v[p] * v[p];
  s = BlockSum<double>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void BatchSqNormJoin(int nbatches, const double* __restrict__ part, double* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatches) return;
  double s = 0.0;
  for (int i = 0; i < kBatchSplit; ++i) s += part[b * kBatchSplit + i];
  out[b] = s;
}

__global__ void SequentialSum(int n, const double* __restrict__ v, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += v[i];
  *out = s;
}

// ------------------------------------------------------------------ dropping

// ApplyDroppingStrategy threshold (dropping.cpp:124-133): sequential column
// norm, keep v != 0 && (diag || |v| >= eps * norm).
__global__ void ThresholdColumnsKernel(int ncols, const int* __restrict__ ptr, const int* __restrict__ idx,
                                       const double* __restrict__ v, double eps, unsigned char* __restrict__ keep) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  double s = 0.0;
  for (int p = ptr[j]; p < ptr[j + 1]; ++p) s += v[p] * v[p];
  const double thr = eps * sqrt(s);
  for (int p = ptr[j]; p < ptr[j + 1]; ++p) {
    const double x = v[p];
    keep[p] = (x != 0.0 && (idx[p] == j || fabs(x) >= thr)) ? 1 : 0;
  }
}

// DropForMatrixType per-batch prune (dropping.cpp:182-185): keep |v| > norm*eps.
__global__ void ThresholdBatchesKernel(int64_t nnz_cols, int ncols, const int* __restrict__ ptr,
                                       const double* __restrict__ v, const double* __restrict__ batch_sq,
                                       const int* __restrict__ batch_of_col, double eps,
                                       unsigned char* __restrict__ keep) {
  const int j = blockIdx.x;
  if (j >= ncols) return;
  const double thr = sqrt(batch_sq[batch_of_col[j]]) * eps;
  for (int p = ptr[j] + threadIdx.x; p < ptr[j + 1]; p += blockDim.x) keep[p] = !(fabs(v[p]) <= thr) ? 1 : 0;
}

// One step of the 8-bit radix select: H holds the 256-bin histogram of the
// candidate keys that match `prefix` on `mask`; finds the bin that holds the
// target-th largest of them and narrows prefix, mask and target to it.
__device__ __forceinline__ void TopKPickBin(const unsigned* H, int shift, uint64_t& prefix, uint64_t& mask,
                                            int& target) {
  const int lane = LaneId();
  int c[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) c[i] = int(H[4 * lane + i]);
  const int mine = c[0] + c[1] + c[2] + c[3];
  const int incl = WaveInclusiveScan(mine);
  const int above = __shfl(incl, kWave - 1, kWave) - incl;  // keys in the bins of higher lanes
  const bool here = above < target && target <= above + mine;
  int bin = 0, rank = 0;
  if (here) {
    int t = target - above;
    for (int i = 3; i >= 0; --i) {
      if (c[i] >= t) {
        bin = 4 * lane + i;
        rank = t;
        break;
      }
      t -= c[i];
    }
  }
  const int src = __ffsll((unsigned long long)__ballot(here)) - 1;
  bin = __shfl(bin, src, kWave);
  target = __shfl(rank, src, kWave);
  prefix |= uint64_t(bin) << shift;
  mask |= uint64_t(255) << shift;
}

// Whether the entry at p (row `row`, key `key` equal to the threshold key T)
// is among the first `take` off-diagonal entries of column j with key T in
// row order.
__device__ __forceinline__ unsigned char TopKTie(const int* __restrict__ idx, const double* __restrict__ v, int b,
                                                 int e, int j, int row, uint64_t T, int take) {
  int before = 0;
  for (int q = b; q < e; ++q)
    if (idx[q] != j && idx[q] < row && OrderedKey(fabs(v[q])) == T) ++before;
  return before < take ? 1 : 0;
}

// Per-column top-k by |v| (ties: smaller row first) with the diagonal always
// kept, one wavefront per column: an 8-bit radix select over the
// order-preserving keys of |v| (8 passes, a 256-bin LDS histogram per pass,
// the crossing bin found by a wave scan over 4 bins per lane), then one pass
// that keeps keys above the threshold key and the first `take` keys equal to
// it in row order.  Columns of at most kTopKRegs * 64 entries keep their keys
// in registers across the passes; longer ones re-read them each pass.
constexpr int kTopKWaves = 4;
constexpr int kTopKRegs = 16;
__global__ __launch_bounds__(kTopKWaves* kWave) void TopKWave(int ncols, const int* __restrict__ ptr,
                                                              const int* __restrict__ idx,
                                                              const double* __restrict__ v, int64_t k,
                                                              unsigned char* __restrict__ keep) {
  __shared__ unsigned hist[kTopKWaves][256];
  const int lane = LaneId(), w = WaveInBlock();
  const int j = blockIdx.x * kTopKWaves + w;
  if (j >= ncols) return;
  unsigned* const H = hist[w];
  const int b = ptr[j], e = ptr[j + 1];
  uint64_t prefix = 0, mask = 0;
  int target = int(k);  // rank of the wanted key counted from the largest
  if (e - b <= kTopKRegs * kWave) {
    uint64_t key[kTopKRegs];
    int row[kTopKRegs];
    int offd = 0;
#pragma unroll
    for (int r = 0; r < kTopKRegs; ++r) {
      const int p = b + lane + kWave * r;
      row[r] = -1;
      key[r] = 0;
      if (p < e) {
        row[r] = idx[p];
        key[r] = OrderedKey(fabs(v[p]));
        offd += row[r] != j ? 1 : 0;
      }
    }
    offd = WaveSumInt(offd);
    if (offd <= k) {
#pragma unroll
      for (int r = 0; r < kTopKRegs; ++r)
        if (row[r] >= 0) keep[b + lane + kWave * r] = 1;
      return;
    }
#pragma unroll 1
    for (int shift = 56; shift >= 0; shift -= 8) {
      for (int t = lane; t < 256; t += kWave) H[t] = 0;
      WaveFence();
#pragma unroll
      for (int r = 0; r < kTopKRegs; ++r)
        if (row[r] >= 0 && row[r] != j && (key[r] & mask) == prefix) atomicAdd(&H[(key[r] >> shift) & 255u], 1u);
      WaveFence();
      TopKPickBin(H, shift, prefix, mask, target);
      WaveFence();
    }
    // entries with the threshold key: all kept when there are at most
    // `target` of them (the usual single one), else ranked by row
    bool tie[kTopKRegs];
    int nties = 0;
#pragma unroll
    for (int r = 0; r < kTopKRegs; ++r) {
      tie[r] = row[r] >= 0 && row[r] != j && key[r] == prefix;
      nties += tie[r] ? 1 : 0;
    }
    nties = WaveSumInt(nties);
    int before[kTopKRegs];
#pragma unroll
    for (int r = 0; r < kTopKRegs; ++r) before[r] = 0;
    if (nties > target) {
#pragma unroll
      for (int rr = 0; rr < kTopKRegs; ++rr) {
        uint64_t m = __ballot(tie[rr]);
        while (m) {
          const int src = __ffsll((unsigned long long)m) - 1;
          m &= m - 1;
          const int trow = __shfl(row[rr], src, kWave);
#pragma unroll
          for (int r = 0; r < kTopKRegs; ++r) before[r] += tie[r] && trow < row[r] ? 1 : 0;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kTopKRegs; ++r) {
      if (row[r] < 0) continue;
      const int p = b + lane + kWave * r;
      unsigned char kp = 1;
      if (row[r] != j) kp = key[r] > prefix || (tie[r] && before[r] < target) ? 1 : 0;
      keep[p] = kp;
    }
    return;
  }
  int offd = 0;
  for (int p = b + lane; p < e; p += kWave) offd += idx[p] != j ? 1 : 0;
  offd = WaveSumInt(offd);
  if (offd <= k) {
    for (int p = b + lane; p < e; p += kWave) keep[p] = 1;
    return;
  }
  for (int shift = 56; shift >= 0; shift -= 8) {
    for (int t = lane; t < 256; t += kWave) H[t] = 0;
    WaveFence();
    for (int p = b + lane; p < e; p += kWave) {
      if (idx[p] == j) continue;
      const uint64_t key = OrderedKey(fabs(v[p]));
      if ((key & mask) == prefix) atomicAdd(&H[(key >> shift) & 255u], 1u);
    }
    WaveFence();
    TopKPickBin(H, shift, prefix, mask, target);
    WaveFence();
  }
  const uint64_t T = prefix;
  const int take = target;  // how many entries with key == T are kept
  int nties = 0;
  for (int p = b + lane; p < e; p += kWave) nties += idx[p] != j && OrderedKey(fabs(v[p])) == T ? 1 : 0;
  nties = WaveSumInt(nties);
  for (int p = b + lane; p < e; p += kWave) {
    const int row = idx[p];
    if (row == j) {
      keep[p] = 1;
      continue;
    }
    const uint64_t key = OrderedKey(fabs(v[p]));
    unsigned char kp = key > T ? 1 : 0;
    if (key == T) kp = nties <= take ? 1 : TopKTie(idx, v, b, e, j, row, T, take);
    keep[p] = kp;
  }
}

// Candidate keys in CSC order.  mode 0: scores of off-diagonal entries of M
// (m^2 * ||A e_k||^2 + 2 m (A R)_kl, -inf when non-finite), diagonal entries
// are not candidates (key = UINT64_MAX).  mode 1: |v| of every entry.
__global__ void CandidateKeys(int ncols, const int* __restrict__ ptr, const int* __restrict__ idx,
                              const double* __restrict__ v, int mode, const int* __restrict__ arp,
                              const int* __restrict__ ari, const double* __restrict__ arv,
                              const double* __restrict__ col_sq, uint64_t* __restrict__ keys) {
  const int j = blockIdx.x;
  for (int p = ptr[j] + threadIdx.x; p < ptr[j + 1]; p += blockDim.x) {
    const int row = idx[p];
    if (mode == 1) {
      keys[p] = OrderedKey(fabs(v[p]));
      continue;
    }
    if (row == j) {
      keys[p] = ~0ull;
      continue;
    }
    const int q = FindSorted(ari, arp[j], arp[j + 1], row);
    const double t = q >= 0 ? arv[q] : 0.0;
    const double m = v[p];
    double s = m * m * col_sq[row] + 2.0 * m * t;
    if (!isfinite(s)) s = -INFINITY;
    keys[p] = OrderedKey(s);
  }
}

// Radix select state: [0] prefix, [1] mask, [2] remaining rank (1-based).
__global__ __launch_bounds__(kBlock) void RadixHistogram(int64_t n, const uint64_t* __restrict__ keys,
                                                         const uint64_t* __restrict__ state, int shift,
                                                         unsigned* __restrict__ hist) {
  __shared__ unsigned h[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const uint64_t prefix = state[0], mask = state[1];
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    if (k == ~0ull || (k & mask) != prefix) continue;
    atomicAdd(&h[(k >> shift) & 0xFF], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

__global__ void RadixPick(uint64_t* __restrict__ state, int shift, unsigned* __restrict__ hist) {
  if (threadIdx.x != 0) return;
  uint64_t rem = state[2];
  uint64_t cum = 0;
  int bin = 255;
  for (int b = 0; b < 256; ++b) {
    if (cum + hist[b] >= rem) {
      bin = b;
      break;
    }
    cum += hist[b];
  }
  state[2] = rem - cum;
  state[0] |= uint64_t(bin) << shift;
  state[1] |= uint64_t(0xFF) << shift;
  for (int b = 0; b < 256; ++b) hist[b] = 0;
}

__global__ void EqualFlags(int64_t n, const uint64_t* __restrict__ keys, const uint64_t* __restrict__ state,
                           int* __restrict__ flags) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = keys[i] == state[0] ? 1 : 0;
}

// Finds the position of the r-th (1-based) key equal to the cutoff.
__global__ void FindTiePosition(int64_t n, const int* __restrict__ flags, const int* __restrict__ scan,
                                const uint64_t* __restrict__ state, int64_t* __restrict__ pos) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flags[i] && uint64_t(scan[i]) + 1 == state[2]) *pos = i;
}

// Keep the entry unless it is a candidate at or below the cutoff (key, pos).
__global__ void KeepAboveCutoff(int64_t n, const uint64_t* __restrict__ keys, const uint64_t* __restrict__ state,
                                const int64_t* __restrict__ cut_pos, unsigned char* __restrict__ keep) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const uint64_t c = state[0];
  bool drop = false;
  if (k != ~0ull) drop = k < c || (k == c && i <= *cut_pos);
  keep[i] = drop ? 0 : 1;
}

// ------------------------------------------------------------------ transpose
__global__ void RowCounts(int64_t nnz, const int* __restrict__ idx, int* __restrict__ counts) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  atomicAdd(&counts[idx[i]], 1);
}

__global__ void ScatterTranspose(int ncols, const int* __restrict__ ptr, const int* __restrict__ idx,
                                 const double* __restrict__ v, const int* __restrict__ tptr,
                                 int* __restrict__ cursor, int* __restrict__ tidx, double* __restrict__ tval) {
  const int j = blockIdx.x;
  for (int p = ptr[j] + threadIdx.x; p < ptr[j + 1]; p += blockDim.x) {
    const int r = idx[p];
    const int d = tptr[r] + atomicAdd(&cursor[r], 1);
    tidx[d] = j;
    tval[d] = v[p];
  }
}

constexpr int kSortCap = 2048;

// Sorts each segment [tptr[c], tptr[c+1]) by index (values follow).  Short
// segments (at most 32 entries): one thread, a 32-element bitonic network in
// registers (indices distinct, padding INT_MAX); medium: wavefront bitonic
// sort of (index << 11 | slot) keys in LDS; long segments are listed for the
// stamp pass.
__global__ __launch_bounds__(kBlock) void SortSegmentsSmall(int nseg, const int* __restrict__ tptr, int* __restrict__ tidx,
                                  double* __restrict__ tval, int* __restrict__ long_list,
                                  int* __restrict__ long_count, int* __restrict__ mid_list,
                                  int* __restrict__ mid_count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nseg) return;
  const int b = tptr[c], e = tptr[c + 1];
  const int len = e - b;
  if (len > kSortCap) {
    long_list[atomicAdd(long_count, 1)] = c;
    return;
  }
  if (len > 32) {
    mid_list[atomicAdd(mid_count, 1)] = c;
    return;
  }
  if (len <= 1) return;
  int key[32];
  double val[32];
#pragma unroll
  for (int t = 0; t < 32; ++t) {
    key[t] = t < len ? tidx[b + t] : INT_MAX;
    val[t] = t < len ? tval[b + t] : 0.0;
  }
#pragma unroll
  for (int k = 2; k <= 32; k <<= 1) {
#pragma unroll
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const int l = i ^ jj;
        if (l > i) {
          const bool asc = (i & k) == 0;
          const bool swap = asc ? key[i] > key[l] : key[i] < key[l];
          const int ki = key[i], kl = key[l];
          const double vi = val[i], vl = val[l];
          key[i] = swap ? kl : ki;
          key[l] = swap ? ki : kl;
          val[i] = swap ? vl : vi;
          val[l] = swap ? vi : vl;
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 32; ++t) {
    if (t < len) {
      tidx[b + t] = key[t];
      tval[b + t] = val[t];
    }
  }
}

__global__ __launch_bounds__(kWave) void SortSegmentsWave(int nlist, const int* __restrict__ list,
                                                          const int* __restrict__ tptr, int* __restrict__ tidx,
                                                          double* __restrict__ tval, double* __restrict__ vtmp) {
  __shared__ uint64_t K[kSortCap];
  const int lane = LaneId();
  for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
    const int c = list[li];
    const int b = tptr[c], len = tptr[c + 1] - b;
    int n2 = 1;
    while (n2 < len) n2 <<= 1;
    for (int i = lane; i < n2; i += kWave)
      K[i] = i < len ? ((uint64_t(uint32_t(tidx[b + i])) << 11) | uint64_t(i)) : ~0ull;
    for (int i = lane; i < len; i += kWave) vtmp[b + i] = tval[b + i];
    WaveFence();
    for (int k = 2; k <= n2; k <<= 1) {
      for (int jj = k >> 1; jj > 0; jj >>= 1) {
        for (int i = lane; i < n2; i += kWave) {
          const int ixj = i ^ jj;
          if (ixj > i) {
            const uint64_t x = K[i], y = K[ixj];
            const bool asc = (i & k) == 0;
            if (asc ? (x > y) : (x < y)) {
              K[i] = y;
              K[ixj] = x;
            }
          }
        }
        WaveFence();
      }
    }
    for (int i = lane; i < len; i += kWave) {
      const uint64_t key = K[i];
      tidx[b + i] = int(key >> 11);
      tval[b + i] = vtmp[b + int(key & 2047)];
    }
    WaveFence();
  }
}

// Long segments: a block per segment, stamp[index] = slot, ordered compaction
// over all indices (ws length = ncols per block, initialised to -1).
__global__ __launch_bounds__(kBlock) void SortSegmentsStamp(int nlist, const int* __restrict__ list,
                                                            const int* __restrict__ tptr, int* __restrict__ tidx,
                                                            double* __restrict__ tval, int nidx,
                                                            int* __restrict__ ws, int* __restrict__ itmp,
                                                            double* __restrict__ vtmp) {
  __shared__ int wcount[16];
  int* stamp = ws + size_t(blockIdx.x) * size_t(nidx);
  for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
    const int c = list[li];
    const int b = tptr[c], len = tptr[c + 1] - b;
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
      stamp[tidx[b + i]] = i;
      vtmp[b + i] = tval[b + i];
      itmp[b + i] = tidx[b + i];
    }
    __syncthreads();
    int base = 0;
    for (int r0 = 0; r0 < nidx; r0 += blockDim.x) {
      const int r = r0 + threadIdx.x;
      int slot = -1;
      if (r < nidx) {
        slot = stamp[r];
        // a stamp left by another segment is valid only if it points back here
        if (slot >= 0 && (slot >= len || itmp[b + slot] != r)) slot = -1;
      }
      int tot = 0;
      const int rank = BlockRank(slot >= 0, wcount, &tot);
      if (slot >= 0) {
        tidx[b + base + rank] = r;
        tval[b + base + rank] = vtmp[b + slot];
      }
      base += tot;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < len; i += blockDim.x) stamp[itmp[b + i]] = -1;
    __syncthreads();
  }
}

__global__ void ExtractDiagonalKernel(int ncols, const int* __restrict__ ptr, const int* __restrict__ idx,
                                      const double* __restrict__ v, double* __restrict__ d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  const int q = FindSorted(idx, ptr[j], ptr[j + 1], j);
  d[j] = q >= 0 ? v[q] : 0.0;
}

__global__ void FullDiagonalCheck(int ncols, const int* __restrict__ ptr, const int* __restrict__ idx,
                                  int* __restrict__ bad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  if (ptr[j + 1] - ptr[j] != 1 || idx[ptr[j]] != j) atomicAdd(bad, 1);
}

__global__ void IdentityKernel(int n, int* __restrict__ ptr, int* __restrict__ idx, double* __restrict__ v) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n) return;
  ptr[j] = j;
  if (j < n) {
    idx[j] = j;
    v[j] = 1.0;
  }
}

__global__ void FillKernel(int64_t n, double* __restrict__ p, double v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ double SafeQuotientDev(double num, double den) {
  if (!(isfinite(den) && fabs(den) > 0.0)) return 0.0;
  const double q = num / den;
  return isfinite(q) ? q : 0.0;
}

__global__ void InnerOuterStepsKernel(int64_t n, int lomr, int two_term, const double* __restrict__ r_az,
                                      const double* __restrict__ az_sq, const double* __restrict__ az_aq,
                                      const double* __restrict__ aq_sq, const double* __restrict__ r_aq,
                                      double* __restrict__ delta, double* __restrict__ gamma,
                                      double* __restrict__ normalization) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double d, g;
  if (!two_term) {
    d = SafeQuotientDev(r_az[j], az_sq[j]);
    g = lomr ? 1.0 : 0.0;
  } else {
    SolveLeastSquares2x2(az_sq[j], az_aq[j], az_aq[j], aq_sq[j], r_az[j], r_aq[j], d, g);
  }
  delta[j] = d;
  gamma[j] = g;
  const double inv = d != 0.0 ? 1.0 / d : 1.0;
  normalization[j] = isfinite(inv) ? inv : 1.0;
}

__global__ void SafeQuotientKernel(int64_t n, const double* __restrict__ num, const double* __restrict__ den,
                                   double* __restrict__ out) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n) out[j] = SafeQuotientDev(num[j], den[j]);
}

}  // namespace

// ====================================================================== host

void FillValue(Ctx& ctx, double* p, int64_t n, double v) {
  if (n > 0) FillKernel<<<Blocks(n), kBlock, 0, ctx.stream>>>(n, p, v);
}

void InnerOuterSteps(Ctx& ctx, int64_t n, bool lomr, bool two_term, const double* r_az, const double* az_sq,
                     const double* az_aq, const double* aq_sq, const double* r_aq, double* delta, double* gamma,
                     double* normalization) {
  if (n > 0)
    InnerOuterStepsKernel<<<Blocks(n), kBlock, 0, ctx.stream>>>(n, lomr ? 1 : 0, two_term ? 1 : 0, r_az, az_sq,
                                                                az_aq, aq_sq, r_aq, delta, gamma, normalization);
}

void SafeQuotients(Ctx& ctx, int64_t n, const double* num, const double* den, double* out) {
  if (n > 0) SafeQuotientKernel<<<Blocks(n), kBlock, 0, ctx.stream>>>(n, num, den, out);
}

Ctx::Ctx() {
  SPAI_HIP_CHECK(hipGetDevice(&device));
  hipDeviceProp_t prop;
  SPAI_HIP_CHECK(hipGetDeviceProperties(&prop, device));
  num_cus = prop.multiProcessorCount;
  pool.SetDevice(device);
  SPAI_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  InitHostBuffers();
}

Ctx::Ctx(hipStream_t external) {
  SPAI_HIP_CHECK(hipGetDevice(&device));
  hipDeviceProp_t prop;
  SPAI_HIP_CHECK(hipGetDeviceProperties(&prop, device));
  num_cus = prop.multiProcessorCount;
  pool.SetDevice(device);
  stream = external;
  owns_stream = false;
  InitHostBuffers();
}

// out[0 .. na) = a, out[64 .. 64 + nb) = b (words; na, nb <= 64).
__global__ void GatherWords(const unsigned* __restrict__ a, int na, const unsigned* __restrict__ b, int nb,
                            unsigned* __restrict__ out) {
  const int t = threadIdx.x;
  if (t < na) out[t] = a[t];
  if (t >= 64 && t - 64 < nb) out[t] = b[t - 64];
}

// One pinned allocation: the staging area for read-backs, then the error code.
void Ctx::InitHostBuffers() {
  pinned_bytes = 1 << 16;
  SPAI_HIP_CHECK(hipHostMalloc(&pinned, pinned_bytes + 256, hipHostMallocMapped | hipHostMallocCoherent));
  err_host = reinterpret_cast<int*>(static_cast<char*>(pinned) + pinned_bytes);
  *err_host = 0;
  SPAI_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&err_dev), err_host, 0));
  SPAI_HIP_CHECK(hipHostGetDevicePointer(&pinned_dev, pinned, 0));
}

void Ctx::CheckErrorCode() {
  const int e = *static_cast<volatile int*>(err_host);
  if (e == 0) return;
  *err_host = 0;
  if (e == 1) throw std::runtime_error("SpGEMM numeric phase found a row outside the symbolic pattern");
  if (e == 2) throw std::runtime_error("product term outside pattern(A S)");
  if (e == 5) throw std::runtime_error("fixed pattern S is not contained in pattern(A S)");
  throw std::runtime_error("device consistency check failed (code " + std::to_string(e) + ")");
}

Ctx::~Ctx() {
  (void)hipStreamSynchronize(stream);
  if (owns_stream && stream) (void)hipStreamDestroy(stream);
  if (pinned) (void)hipHostFree(pinned);
}

void Ctx::ReadBack(void* host, const void* dev, size_t bytes) {
  if (bytes <= pinned_bytes) {
    SPAI_HIP_CHECK(hipMemcpyAsync(pinned, dev, bytes, hipMemcpyDeviceToHost, stream));
    SPAI_HIP_CHECK(hipStreamSynchronize(stream));
    std::memcpy(host, pinned, bytes);
  } else {
    SPAI_HIP_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream));
    SPAI_HIP_CHECK(hipStreamSynchronize(stream));
  }
  CheckErrorCode();
}

void Ctx::ReadBack2(void* h1, const void* d1, size_t b1, void* h2, const void* d2, size_t b2) {
  if (b1 > 256 || b2 > 256 || b1 % 4 || b2 % 4) throw std::invalid_argument("ReadBack2 reads whole words, at most 256 bytes a piece");
  char* const p = static_cast<char*>(pinned);
  // one small kernel stores both pieces into the mapped staging area
  GatherWords<<<1, 128, 0, stream>>>(static_cast<const unsigned*>(d1), int(b1 / 4), static_cast<const unsigned*>(d2),
                                     int(b2 / 4), reinterpret_cast<unsigned*>(pinned_dev));
  SPAI_HIP_CHECK(hipStreamSynchronize(stream));
  std::memcpy(h1, p, b1);
  std::memcpy(h2, p + 256, b2);
  CheckErrorCode();
}

void Ctx::Sync() {
  SPAI_HIP_CHECK(hipStreamSynchronize(stream));
  CheckErrorCode();
}

PatternPtr MakePattern(Ctx& ctx, int64_t rows, int64_t cols, int64_t nnz) {
  if (nnz > kMaxDeviceNnz) throw std::runtime_error("matrix exceeds 2^31-1 stored entries on the device");
  if (rows > kMaxDeviceNnz || cols > kMaxDeviceNnz) throw std::runtime_error("matrix dimension exceeds int32");
  auto p = std::make_shared<Pattern>();
  p->rows = rows;
  p->cols = cols;
  p->nnz = nnz;
  p->ptr = ctx.Alloc<int>(cols + 1);
  p->idx = ctx.Alloc<int>(nnz > 0 ? nnz : 1);
  return p;
}

int64_t ScanCounts(Ctx& ctx, const int* counts, int64_t n, int* ptr, bool read_total, int64_t* total_dev) {
  if (n > 0 && n <= kScanSmallMax) {
    auto total = ctx.Alloc<int64_t>(1);
    ScanSmall<<<1, 1024, 0, ctx.stream>>>(counts, n, ptr, total_dev ? total_dev : total.get());
    if (!read_total) return -1;
    int64_t t = 0;
    ctx.ReadBack(&t, total.get(), sizeof(t));
    if (t > kMaxDeviceNnz) throw std::runtime_error("result exceeds 2^31-1 stored entries on the device");
    return t;
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  auto sums = ctx.Alloc<int64_t>(tiles > 0 ? tiles : 1);
  auto total_buf = ctx.Alloc<int64_t>(1);
  int64_t* const total = total_dev ? total_dev : total_buf.get();
  if (n > 0) {
    ScanTileSums<<<Blocks(n, kScanTile), kBlock, 0, ctx.stream>>>(counts, n, sums.get());
    ScanTiles<<<1, 1024, 0, ctx.stream>>>(sums.get(), tiles, total);
    ScanApply<<<Blocks(n, kScanTile), kBlock, 0, ctx.stream>>>(counts, n, sums.get(), ptr);  // and ptr[n]
  } else {
    SPAI_HIP_CHECK(hipMemsetAsync(total, 0, sizeof(int64_t), ctx.stream));
    WriteLast<<<1, 1, 0, ctx.stream>>>(ptr, n, total);
  }
  if (!read_total) return -1;
  int64_t t = 0;
  ctx.ReadBack(&t, total, sizeof(t));
  if (t > kMaxDeviceNnz) throw std::runtime_error("result exceeds 2^31-1 stored entries on the device");
  return t;
}

DMat Upload(Ctx& ctx, int64_t rows, int64_t cols, const int64_t* ptr, const int64_t* idx, const double* val) {
  const int64_t base = ptr[0];
  const int64_t nnz = ptr[cols] - base;
  DMat m;
  m.pat = MakePattern(ctx, rows, cols, nnz);
  m.val = ctx.Alloc<double>(nnz > 0 ? nnz : 1);
  NarrowToDevice(ctx, m.pat->ptr.get(), ptr, cols + 1, base);
  if (nnz > 0) {
    NarrowToDevice(ctx, m.pat->idx.get(), idx + base, nnz, 0);
    CopyToDevice(ctx, m.val.get(), val + base, sizeof(double) * size_t(nnz));
  }
  return m;
}

void Download(Ctx& ctx, const DMat& m, int64_t* ptr, int64_t* idx, double* val) {
  const int64_t cols = m.cols(), nnz = m.nnz();
  WidenToHost(ctx, ptr, m.ptr(), cols + 1);
  if (nnz > 0) {
    WidenToHost(ctx, idx, m.idx(), nnz);
    CopyToHost(ctx, val, m.v(), sizeof(double) * size_t(nnz));
  }
  ctx.Sync();
}

DMat EmptyMatrix(Ctx& ctx, int64_t rows, int64_t cols) {
  DMat m;
  m.pat = MakePattern(ctx, rows, cols, 0);
  SPAI_HIP_CHECK(hipMemsetAsync(m.pat->ptr.get(), 0, sizeof(int) * (cols + 1), ctx.stream));
  m.val = ctx.Alloc<double>(1);
  return m;
}

DMat Identity(Ctx& ctx, int64_t n) {
  DMat m;
  m.pat = MakePattern(ctx, n, n, n);
  m.val = ctx.Alloc<double>(n > 0 ? n : 1);
  IdentityKernel<<<Blocks(n + 1), kBlock, 0, ctx.stream>>>((int)n, m.pat->ptr.get(), m.pat->idx.get(), m.val.get());
  return m;
}

DMat Clone(Ctx& ctx, const DMat& m) {
  DMat c;
  c.pat = MakePattern(ctx, m.rows(), m.cols(), m.nnz());
  SPAI_HIP_CHECK(hipMemcpyAsync(c.pat->ptr.get(), m.ptr(), sizeof(int) * (m.cols() + 1), hipMemcpyDeviceToDevice,
                                ctx.stream));
  if (m.nnz() > 0)
    SPAI_HIP_CHECK(
        hipMemcpyAsync(c.pat->idx.get(), m.idx(), sizeof(int) * m.nnz(), hipMemcpyDeviceToDevice, ctx.stream));
  c.val = ctx.Alloc<double>(m.nnz() > 0 ? m.nnz() : 1);
  if (m.nnz() > 0)
    SPAI_HIP_CHECK(hipMemcpyAsync(c.val.get(), m.v(), sizeof(double) * m.nnz(), hipMemcpyDeviceToDevice, ctx.stream));
  return c;
}

DMat ShareClone(Ctx& ctx, const DMat& m) {
  DMat c;
  c.pat = m.pat;
  c.val = ctx.Alloc<double>(m.nnz() > 0 ? m.nnz() : 1);
  if (m.nnz() > 0)
    SPAI_HIP_CHECK(hipMemcpyAsync(c.val.get(), m.v(), sizeof(double) * m.nnz(), hipMemcpyDeviceToDevice, ctx.stream));
  return c;
}

namespace {

// New matrix made of the entries of X whose keep flag is set.
DMat Filter(Ctx& ctx, const DMat& X, const unsigned char* keep) {
  // stream compaction over all entries: exclusive scan of the keep flags
  // gives every kept entry its output position, and the scan at the column
  // starts gives the output column pointers
  const int64_t n = X.cols(), nnz = X.nnz();
  DMat out;
  out.pat = MakePattern(ctx, X.rows(), n, 0);
  auto flags = ctx.Alloc<int>(nnz > 0 ? nnz : 1);
  auto pos = ctx.Alloc<int>(nnz + 1);
  if (nnz > 0) KeepToInt<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, keep, flags.get());
  const int64_t total = ScanCounts(ctx, flags.get(), nnz, pos.get());
  out.pat->nnz = total;
  GatherPtr<<<Blocks(n + 1), kBlock, 0, ctx.stream>>>(n + 1, X.ptr(), pos.get(), out.pat->ptr.get());
  out.pat->idx = ctx.Alloc<int>(total > 0 ? total : 1);
  out.val = ctx.Alloc<double>(total > 0 ? total : 1);
  if (nnz > 0)
    ScatterKept<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, X.idx(), X.v(), keep, pos.get(), out.pat->idx.get(),
                                                        out.val.get());
  return out;
}

}  // namespace

DMat PruneZeros(Ctx& ctx, DMat&& X) {
  const int64_t nnz = X.nnz();
  if (nnz == 0) return std::move(X);
  auto keep = ctx.Alloc<unsigned char>(nnz);
  if (nnz <= kPruneFilterAlways) {
    // small: filter without first reading back whether anything is dropped
    FlagNonZero<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, X.v(), keep.get(), nullptr);
    return Filter(ctx, X, keep.get());
  }
  auto dropped = ctx.Alloc<int>(1);
  SPAI_HIP_CHECK(hipMemsetAsync(dropped.get(), 0, sizeof(int), ctx.stream));
  FlagNonZero<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, X.v(), keep.get(), dropped.get());
  int d = 0;
  ctx.ReadBack(&d, dropped.get(), sizeof(int));
  if (d == 0) return std::move(X);
  return Filter(ctx, X, keep.get());
}

DMat SpGemm(Ctx& ctx, const DMat& X, const DMat& Y, bool prune_zero, bool residual) {
  if (X.cols() != Y.rows()) throw std::invalid_argument("Sparse matrix dimensions do not match");
  const int64_t n = Y.cols();
  const int64_t nrows = X.rows();
  if (residual && nrows != n) throw std::invalid_argument("residual product needs a square result");
  auto upper = ctx.Alloc<int64_t>(n > 0 ? n : 1);
  auto counts = ctx.Alloc<int>(n > 0 ? n : 1);
  auto lists = ctx.Alloc<int>(2 * (n > 0 ? n : 1));  // long, big column lists
  int* const long_list = lists.get();
  int* const big_list = long_list + (n > 0 ? n : 1);
  auto tiers_d = ctx.Alloc<TierCounts>(1);
  ZeroTiers<<<1, 1, 0, ctx.stream>>>(tiers_d.get());
  if (n > 0)
    ExpansionSizes<<<Blocks(n, kExpBlock), kExpBlock, 0, ctx.stream>>>((int)n, X.ptr(), Y.ptr(), Y.idx(),
                                                                       residual ? 1 : 0, upper.get(), long_list,
                                                                       big_list, tiers_d.get());
  // symbolic phase of the short columns, then the scan, before reading
  // anything back: when every column is short (the usual case) the tiers and
  // the total come back together in one round trip
  if (n > 0)
    MergeCount<kTmLists><<<Blocks(n, 256), 256, 0, ctx.stream>>>((int)n, X.ptr(), X.idx(), Y.ptr(), Y.idx(),
                                                                 upper.get(), residual ? 1 : 0, counts.get());
  DMat C;
  C.pat = MakePattern(ctx, nrows, n, 0);
  auto total_d = ctx.Alloc<int64_t>(1);
  ScanCounts(ctx, counts.get(), n, C.pat->ptr.get(), false, total_d.get());
  TierCounts tiers{};
  int64_t total = 0;
  ctx.ReadBack2(&tiers, tiers_d.get(), sizeof(tiers), &total, total_d.get(), sizeof(total));
  const int nshort = tiers.n[0], nlong = tiers.n[1], nbig = tiers.n[2], short_ny = tiers.n[3];
  // the numeric merge's list count: the smallest that holds every short column
  const int merge_lists = short_ny <= 4 ? 4 : short_ny <= 8 ? 8 : 16;
  // Many LDS-tier columns: their count pass stashes the sorted rows (at the
  // offsets of a scan of their expansions, when those stay under 1 GiB) and
  // the numeric phase places them, instead of a fill pass repeating the
  // gather.
  const bool stash_mid = nlong - nbig >= kStashMidColumns && tiers.mid_expansion <= (1ull << 28);
  DBuf<int> mid_off, mid_stash;
  if (stash_mid) {
    auto mid_cnt = ctx.Alloc<int>(size_t(nlong));
    mid_off = ctx.Alloc<int>(size_t(nlong) + 1);
    MidExpansions<<<Blocks(nlong), kBlock, 0, ctx.stream>>>(nlong, long_list, upper.get(), mid_cnt.get());
    const int64_t mid_total = ScanCounts(ctx, mid_cnt.get(), nlong, mid_off.get());
    mid_stash = ctx.Alloc<int>(mid_total > 0 ? size_t(mid_total) : 1);
  }

  // The large-expansion columns whose rows fit the LDS bitmap stay in the
  // wavefront tier; only the others take the stamp kernel.  Their order in
  // stamp_list does not matter: each column's result depends on it alone.
  int nstamp = 0;
  DBuf<int> stamp_list;
  if (nbig > 0) {
    stamp_list = ctx.Alloc<int>(size_t(nbig) + 1);
    int* const nstamp_d = stamp_list.get() + nbig;
    SPAI_HIP_CHECK(hipMemsetAsync(nstamp_d, 0, sizeof(int), ctx.stream));
    BigColumnRanges<<<Blocks(nbig, kSymWaves), kSymWaves * kWave, 0, ctx.stream>>>(
        nbig, big_list, X.ptr(), X.idx(), Y.ptr(), Y.idx(), residual ? 1 : 0, stamp_list.get(), nstamp_d);
    ctx.ReadBack(&nstamp, nstamp_d, sizeof(int));
  }
  const int nwide = nbig - nstamp;
  // Each long tier is launched over its own column list, so a tier with no
  // columns costs nothing.
  auto symbolic = [&](int fill, int* cnt, const int* cptr, int* cidx, DBuf<int>& stamp, unsigned stamp_blocks) {
    // with the stash, the numeric phase places the LDS-tier columns' rows
    // and the fill pass only writes the wide (bitmap, unstashed) columns
    if (nlong > nstamp && !(fill && stash_mid && nwide == 0))
      SymbolicEsc<<<Blocks(nlong, kSymWaves), kSymWaves * kWave, 0, ctx.stream>>>(
          nlong, long_list, X.ptr(), X.idx(), Y.ptr(), Y.idx(), upper.get(), residual ? 1 : 0, fill, cnt, cptr,
          cidx, stash_mid ? mid_off.get() : nullptr, stash_mid ? mid_stash.get() : nullptr);
    if (nstamp > 0) {
      SPAI_HIP_CHECK(
          hipMemsetAsync(stamp.get(), 0xFF, sizeof(int) * size_t(stamp_blocks) * size_t(nrows), ctx.stream));
      SymbolicStamp<<<stamp_blocks, kBlock, 0, ctx.stream>>>(nstamp, stamp_list.get(), (int)nrows, X.ptr(), X.idx(), Y.ptr(),
                                                             Y.idx(), residual ? 1 : 0, fill, stamp.get(), cnt,
                                                             cptr, cidx);
    }
  };
  DBuf<int> stamp;
  unsigned stamp_blocks = 0;
  if (nstamp > 0) {
    const int64_t max_ws = (int64_t(1) << 28) / std::max<int64_t>(nrows, 1);  // <= 1 GiB of stamps
    stamp_blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>({(int64_t)nstamp, max_ws, 1024}));
    stamp = ctx.Alloc<int>(size_t(stamp_blocks) * size_t(nrows));
  }
  if (nlong > 0) {
    // the long columns' counts, then the scan again
    symbolic(0, counts.get(), nullptr, nullptr, stamp, stamp_blocks);
    total = ScanCounts(ctx, counts.get(), n, C.pat->ptr.get());
  } else if (total > kMaxDeviceNnz) {
    throw std::runtime_error("result exceeds 2^31-1 stored entries on the device");
  }
  C.pat->nnz = total;
  C.pat->idx = ctx.Alloc<int>(total > 0 ? total : 1);
  C.val = ctx.Alloc<double>(total > 0 ? total : 1);
  if (nlong > 0) symbolic(1, nullptr, C.pat->ptr.get(), C.pat->idx.get(), stamp, stamp_blocks);
  // numeric phase: the short columns write their rows and values in place;
  // the long kernels check their rows against the symbolic pattern and
  // report a miss through the context's device error code (raised at its
  // next synchronisation)
  if (nshort > 0) {
#define SPAI_MERGE_NUMERIC(NL)                                                                                 \
  MergeNumeric<NL><<<Blocks(n, kNumMergeWaves * kWave), kNumMergeWaves * kWave, 0, ctx.stream>>>(                    \
      (int)n, X.ptr(), X.idx(), X.v(), Y.ptr(), Y.idx(), Y.v(), upper.get(), residual ? 1 : 0, C.pat->ptr.get(), \
      C.pat->idx.get(), C.val.get())
    if (merge_lists == 4)
      SPAI_MERGE_NUMERIC(4);
    else if (merge_lists == 8)
      SPAI_MERGE_NUMERIC(8);
    else
      SPAI_MERGE_NUMERIC(16);
#undef SPAI_MERGE_NUMERIC
  }
  if (nlong > 0) {
    NumericGustavson<kNumCap, kNumWaves><<<Blocks(nlong, kNumWaves), kNumWaves * kWave, 0, ctx.stream>>>(
        nlong, long_list, X.ptr(), X.idx(), X.v(), Y.ptr(), Y.idx(), Y.v(), C.pat->ptr.get(), C.pat->idx.get(),
        C.val.get(), residual ? 1 : 0, ctx.err_dev, upper.get(), stash_mid ? mid_off.get() : nullptr,
        stash_mid ? mid_stash.get() : nullptr, -1, 0);
    // the columns of more than kNumCap rows (an expansion over kNumCap)
    if (nbig > 0 || total > int64_t(kNumCap))
      NumericGustavson<kNumCapWide, 1><<<(unsigned)nlong, kWave, 0, ctx.stream>>>(
          nlong, long_list, X.ptr(), X.idx(), X.v(), Y.ptr(), Y.idx(), Y.v(), C.pat->ptr.get(), C.pat->idx.get(),
          C.val.get(), residual ? 1 : 0, ctx.err_dev, upper.get(), stash_mid ? mid_off.get() : nullptr,
          stash_mid ? mid_stash.get() : nullptr, kNumCap, 1);
  }
  if (prune_zero) return PruneZeros(ctx, std::move(C));
  return C;
}

DMat DiagScale(Ctx& ctx, const double* d, const DMat& Y, bool prune_zero) {
  DMat C;
  C.pat = Y.pat;
  C.val = ctx.Alloc<double>(Y.nnz() > 0 ? Y.nnz() : 1);
  if (Y.cols() > 0 && Y.nnz() > 0)
    DiagScaleKernel<<<Blocks(Y.nnz()), kBlock, 0, ctx.stream>>>(Y.nnz(), Y.ptr(), Y.idx(), d, Y.v(), C.val.get());
  if (prune_zero) return PruneZeros(ctx, std::move(C));
  return C;
}

DMat ColScale(Ctx& ctx, const DMat& X, const double* f, bool prune_zero) {
  DMat C;
  C.pat = X.pat;
  C.val = ctx.Alloc<double>(X.nnz() > 0 ? X.nnz() : 1);
  if (X.cols() > 0 && X.nnz() > 0)
    ColScaleKernel<<<(unsigned)X.cols(), 64, 0, ctx.stream>>>((int)X.cols(), X.ptr(), f, X.v(), C.val.get());
  if (prune_zero) return PruneZeros(ctx, std::move(C));
  return C;
}

DMat Merge(Ctx& ctx, const DMat& L, const DMat& R, MergeOp op, double a, double b, bool prune_zero) {
  if (L.rows() != R.rows() || L.cols() != R.cols()) throw std::invalid_argument("Sparse matrices must have the same shape");
  const int64_t n = L.cols();
  auto counts = ctx.Alloc<int>(n > 0 ? n : 1);
  const unsigned mg = Blocks(n, kMergeWaves);
  MergeWave<<<mg, kMergeWaves * kWave, 0, ctx.stream>>>((int)n, L.ptr(), L.idx(), L.v(), R.ptr(), R.idx(), R.v(),
                                                        (int)op, a, b, prune_zero ? 1 : 0, 0, counts.get(), nullptr,
                                                        nullptr, nullptr);
  DMat C;
  C.pat = MakePattern(ctx, L.rows(), n, 0);
  const int64_t total = ScanCounts(ctx, counts.get(), n, C.pat->ptr.get());
  C.pat->nnz = total;
  C.pat->idx = ctx.Alloc<int>(total > 0 ? total : 1);
  C.val = ctx.Alloc<double>(total > 0 ? total : 1);
  MergeWave<<<mg, kMergeWaves * kWave, 0, ctx.stream>>>((int)n, L.ptr(), L.idx(), L.v(), R.ptr(), R.idx(), R.v(),
                                                        (int)op, a, b, prune_zero ? 1 : 0, 1, nullptr,
                                                        C.pat->ptr.get(), C.pat->idx.get(), C.val.get());
  return C;
}

DMat Transpose(Ctx& ctx, const DMat& X) {
  const int64_t rows = X.rows(), cols = X.cols(), nnz = X.nnz();
  auto counts = ctx.Alloc<int>(rows > 0 ? rows : 1);
  SPAI_HIP_CHECK(hipMemsetAsync(counts.get(), 0, sizeof(int) * rows, ctx.stream));
  if (nnz > 0) RowCounts<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, X.idx(), counts.get());
  DMat T;
  T.pat = MakePattern(ctx, cols, rows, nnz);
  T.val = ctx.Alloc<double>(nnz > 0 ? nnz : 1);
  ScanCounts(ctx, counts.get(), rows, T.pat->ptr.get(), false);  // the sum is nnz
  if (nnz == 0) return T;
  SPAI_HIP_CHECK(hipMemsetAsync(counts.get(), 0, sizeof(int) * rows, ctx.stream));
  ScatterTranspose<<<(unsigned)cols, 64, 0, ctx.stream>>>((int)cols, X.ptr(), X.idx(), X.v(), T.pat->ptr.get(),
                                                          counts.get(), T.pat->idx.get(), T.val.get());
  auto lists = ctx.Alloc<int>(2 * rows + 2);
  int* long_list = lists.get();
  int* mid_list = lists.get() + rows;
  int* cnts = lists.get() + 2 * rows;
  SPAI_HIP_CHECK(hipMemsetAsync(cnts, 0, 2 * sizeof(int), ctx.stream));
  SortSegmentsSmall<<<Blocks(rows), kBlock, 0, ctx.stream>>>((int)rows, T.pat->ptr.get(), T.pat->idx.get(),
                                                             T.val.get(), long_list, cnts, mid_list, cnts + 1);
  int nl[2] = {0, 0};
  ctx.ReadBack(nl, cnts, sizeof(nl));
  auto vtmp = ctx.Alloc<double>(nnz);
  if (nl[1] > 0) {
    const unsigned g = (unsigned)std::min(nl[1], 65535);
    SortSegmentsWave<<<g, kWave, 0, ctx.stream>>>(nl[1], mid_list, T.pat->ptr.get(), T.pat->idx.get(), T.val.get(),
                                                  vtmp.get());
  }
  if (nl[0] > 0) {
    const int64_t max_ws = (int64_t(1) << 28) / std::max<int64_t>(cols, 1);
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>({(int64_t)nl[0], max_ws, 1024}));
    auto ws = ctx.Alloc<int>(size_t(g) * size_t(cols));
    auto itmp = ctx.Alloc<int>(nnz);
    SPAI_HIP_CHECK(hipMemsetAsync(ws.get(), 0xFF, sizeof(int) * size_t(g) * size_t(cols), ctx.stream));
    SortSegmentsStamp<<<g, kBlock, 0, ctx.stream>>>(nl[0], long_list, T.pat->ptr.get(), T.pat->idx.get(),
                                                    T.val.get(), (int)cols, ws.get(), itmp.get(), vtmp.get());
  }
  return T;
}

namespace {
__global__ void IotaValues(int64_t n, double* __restrict__ v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = double(i);
}

__global__ void ValuesToPositions(int64_t n, const double* __restrict__ v, int* __restrict__ pos) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) pos[i] = int(v[i]);
}
}  // namespace

DBuf<int> TransposePositions(Ctx& ctx, const DMat& X) {
  const int64_t nnz = X.nnz();
  DMat I;  // X's pattern with each entry's position as its value (exact in FP64)
  I.pat = X.pat;
  I.val = ctx.Alloc<double>(nnz > 0 ? nnz : 1);
  if (nnz > 0) IotaValues<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, I.val.get());
  const DMat T = Transpose(ctx, I);
  auto pos = ctx.Alloc<int>(nnz > 0 ? nnz : 1);
  if (nnz > 0) ValuesToPositions<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, T.v(), pos.get());
  return pos;
}

namespace {
double SumDevice(Ctx& ctx, const double* v, int64_t n) {
  const unsigned g = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (n + kBlock - 1) / kBlock), 1024);
  auto part = ctx.Alloc<double>(g);
  SumPartials<<<g, kBlock, 0, ctx.stream>>>(v, n, part.get());
  auto fin = ctx.Alloc<double>(1);
  SumPartials<<<1, kBlock, 0, ctx.stream>>>(part.get(), g, fin.get());
  double s = 0.0;
  ctx.ReadBack(&s, fin.get(), sizeof(double));
  return s;
}
}  // namespace

void ColumnDots(Ctx& ctx, const DMat& X, const DMat& Y, double* out) {
  const int64_t n = X.cols();
  ColumnDotWave<<<Blocks(n, kDotWaves), kDotWaves * kWave, 0, ctx.stream>>>((int)n, X.ptr(), X.idx(), X.v(), Y.ptr(),
                                                                            Y.idx(), Y.v(), 0, out);
}

void ColumnSqNorms(Ctx& ctx, const DMat& X, double* out) {
  const int64_t n = X.cols();
  ColumnDotWave<<<Blocks(n, kDotWaves), kDotWaves * kWave, 0, ctx.stream>>>((int)n, X.ptr(), X.idx(), X.v(), X.ptr(),
                                                                            X.idx(), X.v(), 1, out);
}

double FrobDot(Ctx& ctx, const DMat& X, const DMat& Y) {
  if (X.rows() != Y.rows() || X.cols() != Y.cols()) throw std::invalid_argument("Sparse matrices must have the same shape");
  const int64_t n = X.cols();
  auto cd = ctx.Alloc<double>(n > 0 ? n : 1);
  ColumnDots(ctx, X, Y, cd.get());
  return SumDevice(ctx, cd.get(), n);
}

double SqNorm(Ctx& ctx, const DMat& X) {
  const int64_t n = X.cols();
  auto cd = ctx.Alloc<double>(n > 0 ? n : 1);
  ColumnSqNorms(ctx, X, cd.get());
  return SumDevice(ctx, cd.get(), n);
}

double BatchOrderedDot(Ctx& ctx, const DMat& X, const DMat& Y, const int* batch_first, int nbatches, bool same) {
  if (X.rows() != Y.rows() || X.cols() != Y.cols()) throw std::invalid_argument("Sparse matrices must have the same shape");
  auto part = ctx.Alloc<double>(nbatches + 1);
  if (nbatches > 0) {
    if (same)
      BatchSqWave<<<(unsigned)nbatches, kWave, 0, ctx.stream>>>(batch_first, X.ptr(), X.v(), part.get());
    else
      BatchDotWave<<<(unsigned)nbatches, kWave, 0, ctx.stream>>>(batch_first, X.ptr(), X.idx(), X.v(), Y.ptr(),
                                                                  Y.idx(), Y.v(), part.get());
  }
  SequentialSum<<<1, 64, 0, ctx.stream>>>(nbatches, part.get(), part.get() + nbatches);
  double s = 0.0;
  ctx.ReadBack(&s, part.get() + nbatches, sizeof(double));
  return s;
}

void BatchSqNorms(Ctx& ctx, const DMat& X, const int* batch_first, int nbatches, double* out, bool ordered) {
  if (ordered) {
    if (nbatches > 0) BatchSqWave<<<(unsigned)nbatches, kWave, 0, ctx.stream>>>(batch_first, X.ptr(), X.v(), out);
  } else if (nbatches > 0) {
    auto part = ctx.Alloc<double>(size_t(nbatches) * kBatchSplit);
    BatchSqNormTreeKernel<<<(unsigned)nbatches * kBatchSplit, 256, 0, ctx.stream>>>(batch_first, X.ptr(), X.v(),
                                                                                    part.get());
    BatchSqNormJoin<<<Blocks(nbatches), kBlock, 0, ctx.stream>>>(nbatches, part.get(), out);
  }
}

DMat ThresholdColumns(Ctx& ctx, const DMat& M, double eps) {
  const int64_t n = M.cols(), nnz = M.nnz();
  auto keep = ctx.Alloc<unsigned char>(nnz > 0 ? nnz : 1);
  ThresholdColumnsKernel<<<Blocks(n), kBlock, 0, ctx.stream>>>((int)n, M.ptr(), M.idx(), M.v(), eps, keep.get());
  return Filter(ctx, M, keep.get());
}

DMat ThresholdBatches(Ctx& ctx, const DMat& X, const double* batch_sq, const int* batch_of_col, double eps) {
  const int64_t n = X.cols(), nnz = X.nnz();
  auto keep = ctx.Alloc<unsigned char>(nnz > 0 ? nnz : 1);
  if (n > 0)
    ThresholdBatchesKernel<<<(unsigned)n, 64, 0, ctx.stream>>>(nnz, (int)n, X.ptr(), X.v(), batch_sq, batch_of_col,
                                                              eps, keep.get());
  return Filter(ctx, X, keep.get());
}

DMat ColumnTopK(Ctx& ctx, const DMat& X, int64_t k) {
  const int64_t n = X.cols(), nnz = X.nnz();
  auto keep = ctx.Alloc<unsigned char>(nnz > 0 ? nnz : 1);
  if (n > 0)
    TopKWave<<<Blocks(n, kTopKWaves), kTopKWaves * kWave, 0, ctx.stream>>>((int)n, X.ptr(), X.idx(), X.v(), k,
                                                                          keep.get());
  return Filter(ctx, X, keep.get());
}

namespace {

// Drops the `drop` smallest candidates under (key, position).
DMat DropByKeys(Ctx& ctx, const DMat& X, const uint64_t* keys, int64_t drop) {
  const int64_t nnz = X.nnz();
  auto state = ctx.Alloc<uint64_t>(3);
  const uint64_t init[3] = {0, 0, (uint64_t)drop};
  SPAI_HIP_CHECK(hipMemcpyAsync(state.get(), init, sizeof(init), hipMemcpyHostToDevice, ctx.stream));
  auto hist = ctx.Alloc<unsigned>(256);
  SPAI_HIP_CHECK(hipMemsetAsync(hist.get(), 0, 256 * sizeof(unsigned), ctx.stream));
  const unsigned g = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (nnz + kBlock - 1) / kBlock), 2048);
  for (int shift = 56; shift >= 0; shift -= 8) {
    RadixHistogram<<<g, kBlock, 0, ctx.stream>>>(nnz, keys, state.get(), shift, hist.get());
    RadixPick<<<1, 64, 0, ctx.stream>>>(state.get(), shift, hist.get());
  }
  auto flags = ctx.Alloc<int>(nnz);
  auto scan = ctx.Alloc<int>(nnz + 1);
  EqualFlags<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, keys, state.get(), flags.get());
  ScanCounts(ctx, flags.get(), nnz, scan.get(), false);
  auto pos = ctx.Alloc<int64_t>(1);
  SPAI_HIP_CHECK(hipMemsetAsync(pos.get(), 0xFF, sizeof(int64_t), ctx.stream));
  FindTiePosition<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, flags.get(), scan.get(), state.get(), pos.get());
  auto keep = ctx.Alloc<unsigned char>(nnz);
  KeepAboveCutoff<<<Blocks(nnz), kBlock, 0, ctx.stream>>>(nnz, keys, state.get(), pos.get(), keep.get());
  return Filter(ctx, X, keep.get());
}

}  // namespace

DMat DropSmallestByScore(Ctx& ctx, const DMat& M, const DMat& AR, const double* a_col_sq, int64_t drop) {
  const int64_t n = M.cols(), nnz = M.nnz();
  auto keys = ctx.Alloc<uint64_t>(nnz > 0 ? nnz : 1);
  if (n > 0)
    CandidateKeys<<<(unsigned)n, 64, 0, ctx.stream>>>((int)n, M.ptr(), M.idx(), M.v(), 0, AR.ptr(), AR.idx(), AR.v(),
                                                      a_col_sq, keys.get());
  return DropByKeys(ctx, M, keys.get(), drop);
}

DMat DropSmallestByAbs(Ctx& ctx, const DMat& X, int64_t drop) {
  const int64_t n = X.cols(), nnz = X.nnz();
  auto keys = ctx.Alloc<uint64_t>(nnz > 0 ? nnz : 1);
  if (n > 0)
    CandidateKeys<<<(unsigned)n, 64, 0, ctx.stream>>>((int)n, X.ptr(), X.idx(), X.v(), 1, nullptr, nullptr, nullptr,
                                                      nullptr, keys.get());
  return DropByKeys(ctx, X, keys.get(), drop);
}

void ExtractDiagonal(Ctx& ctx, const DMat& A, double* d) {
  const int64_t n = A.cols();
  ExtractDiagonalKernel<<<Blocks(n), kBlock, 0, ctx.stream>>>((int)n, A.ptr(), A.idx(), A.v(), d);
}

bool IsFullDiagonal(Ctx& ctx, const DMat& X) {
  if (X.rows() != X.cols() || X.nnz() != X.cols()) return false;
  auto bad = ctx.Alloc<int>(1);
  SPAI_HIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(int), ctx.stream));
  FullDiagonalCheck<<<Blocks(X.cols()), kBlock, 0, ctx.stream>>>((int)X.cols(), X.ptr(), X.idx(), bad.get());
  int b = 0;
  ctx.ReadBack(&b, bad.get(), sizeof(int));
  return b == 0;
}

}  // namespace spai

// ------------------------------------------------- no-allocation top-k
// Device memory of the no-allocation per-column top-k: the keep flag of every
// input entry, then the per-tile sums of the column-count scan.  One copy per
// GPU, held until the process ends; calls on one GPU share it.
static __device__ __align__(256) unsigned char scratch[spai::kTopKNoAllocScratchBytes];

namespace spai {

namespace {

constexpr int kNaTile = 1024;  // column counts per scan tile (256 threads x 4)
constexpr int64_t kNaFlagBytes = kTopKNoAllocMaxNnz;
static_assert(kNaFlagBytes % 256 == 0, "tile sums must stay aligned");
static_assert(kTopKNoAllocScratchBytes >= kNaFlagBytes + sizeof(int) * (kTopKNoAllocMaxCols / kNaTile + 1),
              "scratch too small for the tile sums");

// out_ptr[j + 1] = kept entries of column j (one wavefront per column);
// out_ptr[0] = 0.
__global__ __launch_bounds__(kBlock) void NaKeptCounts(int ncols, const int* __restrict__ ptr,
                                                       const unsigned char* __restrict__ keep,
                                                       int* __restrict__ out_ptr) {
  const int j = blockIdx.x * (kBlock / kWave) + WaveInBlock();
  if (blockIdx.x == 0 && threadIdx.x == 0) out_ptr[0] = 0;
  if (j >= ncols) return;
  int c = 0;
  for (int p = ptr[j] + LaneId(); p < ptr[j + 1]; p += kWave) c += keep[p];
  c = WaveSumInt(c);
  if (LaneId() == 0) out_ptr[j + 1] = c;
}

// Inclusive block scan of 4 values per thread (256 threads); returns the
// exclusive prefix of this thread's first value and sets total.
__device__ __forceinline__ int NaBlockScan4(int (&v)[4], int& total) {
  __shared__ int wsum[kBlock / kWave];
  int t = v[0] + v[1] + v[2] + v[3];
  // inclusive scan over the wavefront
  const int lane = LaneId(), w = WaveInBlock();
  int x = t;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  if (lane == kWave - 1) wsum[w] = x;
  __syncthreads();
  int before = 0;
  total = 0;
#pragma unroll
  for (int q = 0; q < kBlock / kWave; ++q) {
    before += q < w ? wsum[q] : 0;
    total += wsum[q];
  }
  __syncthreads();
  return before + x - t;
}

// tiles[t] = sum of out_ptr[1 + t * kNaTile, ...) over the tile.
__global__ __launch_bounds__(kBlock) void NaTileSums(int ncols, const int* __restrict__ out_ptr,
                                                     int* __restrict__ tiles) {
  const int base = blockIdx.x * kNaTile + threadIdx.x * 4;
  int v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = base + i < ncols ? out_ptr[1 + base + i] : 0;
  int total = 0;
  NaBlockScan4(v, total);
  if (threadIdx.x == 0) tiles[blockIdx.x] = total;
}

// Exclusive scan of the tile sums in place (one block, ntiles <= 4097).
__global__ __launch_bounds__(kBlock) void NaScanTiles(int ntiles, int* __restrict__ tiles) {
  int carry = 0;
  for (int base = 0; base < ntiles; base += kNaTile) {
    const int i0 = base + threadIdx.x * 4;
    int v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = i0 + i < ntiles ? tiles[i0 + i] : 0;
    int total = 0;
    int run = carry + NaBlockScan4(v, total);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i0 + i < ntiles) tiles[i0 + i] = run;
      run += v[i];
    }
    carry += total;
  }
}

// out_ptr[1 + i] becomes the inclusive prefix sum of the counts.
__global__ __launch_bounds__(kBlock) void NaApplyScan(int ncols, int* __restrict__ out_ptr,
                                                      const int* __restrict__ tiles) {
  const int base = blockIdx.x * kNaTile + threadIdx.x * 4;
  int v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = base + i < ncols ? out_ptr[1 + base + i] : 0;
  int total = 0;
  int run = tiles[blockIdx.x] + NaBlockScan4(v, total);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    run += v[i];
    if (base + i < ncols) out_ptr[1 + base + i] = run;
  }
}

// Kept entries of column j to out[out_ptr[j] ...) in input order (one
// wavefront per column, 64 entries per ballot).
__global__ __launch_bounds__(kBlock) void NaScatterKept(int ncols, const int* __restrict__ ptr,
                                                        const int* __restrict__ idx, const double* __restrict__ v,
                                                        const unsigned char* __restrict__ keep,
                                                        const int* __restrict__ out_ptr, int* __restrict__ out_idx,
                                                        double* __restrict__ out_v) {
  const int j = blockIdx.x * (kBlock / kWave) + WaveInBlock();
  if (j >= ncols) return;
  const int lane = LaneId();
  const uint64_t below = (uint64_t(1) << lane) - 1;
  int dst = out_ptr[j];
  for (int p0 = ptr[j]; p0 < ptr[j + 1]; p0 += kWave) {
    const int p = p0 + lane;
    const bool k = p < ptr[j + 1] && keep[p];
    const uint64_t m = __ballot(k);
    if (k) {
      const int q = dst + __popcll(m & below);
      out_idx[q] = idx[p];
      out_v[q] = v[p];
    }
    dst += __popcll(m);
  }
}

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (each GPU holds its own copy at its own address).
std::mutex g_na_mu;
void* g_na_ptr[64] = {nullptr};

unsigned char* NoAllocScratch() {
  int dev = 0;
  SPAI_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_na_ptr[dev] = p;
  }
  return static_cast<unsigned char*>(g_na_ptr[dev]);
}

}  // namespace

void CheckTopKNoAllocSizes(int64_t rows, int64_t cols, int64_t nnz, int64_t k) {
  if (rows < 0 || cols < 0 || nnz < 0) throw std::invalid_argument("rows, cols and nnz must be non-negative");
  if (k < 0) throw std::invalid_argument("k must be non-negative");
  if (rows > INT_MAX) throw std::invalid_argument("rows must fit a 32-bit row index");
  if (cols > kTopKNoAllocMaxCols)
    throw std::invalid_argument("no-allocation top-k: cols " + std::to_string(cols) + " exceeds the maximum " +
                                std::to_string(kTopKNoAllocMaxCols));
  if (nnz > kTopKNoAllocMaxNnz)
    throw std::invalid_argument("no-allocation top-k: nnz " + std::to_string(nnz) + " exceeds the maximum " +
                                std::to_string(kTopKNoAllocMaxNnz));
}

void ColumnTopKNoAlloc(hipStream_t stream, int64_t rows, int64_t cols, int64_t nnz, const int* ptr, const int* idx,
                       const double* v, int64_t k, int* out_ptr, int* out_idx, double* out_v) {
  CheckTopKNoAllocSizes(rows, cols, nnz, k);
  unsigned char* const base = NoAllocScratch();
  unsigned char* const keep = base;
  int* const tiles = reinterpret_cast<int*>(base + kNaFlagBytes);
  const int n = int(cols);
  if (n == 0) {
    // an empty matrix still gets its one column offset
    NaKeptCounts<<<1, kBlock, 0, stream>>>(0, ptr, keep, out_ptr);
    SPAI_HIP_CHECK(hipGetLastError());
    return;
  }
  // every flag of the input's entries is written by TopKWave, every tile sum
  // by NaTileSums: nothing left from an earlier call is read
  TopKWave<<<Blocks(n, kTopKWaves), kTopKWaves * kWave, 0, stream>>>(n, ptr, idx, v, k, keep);
  const unsigned wave_blocks = Blocks(n, kBlock / kWave);
  NaKeptCounts<<<wave_blocks, kBlock, 0, stream>>>(n, ptr, keep, out_ptr);
  const int ntiles = (n + kNaTile - 1) / kNaTile;
  NaTileSums<<<ntiles, kBlock, 0, stream>>>(n, out_ptr, tiles);
  NaScanTiles<<<1, kBlock, 0, stream>>>(ntiles, tiles);
  NaApplyScan<<<ntiles, kBlock, 0, stream>>>(n, out_ptr, tiles);
  NaScatterKept<<<wave_blocks, kBlock, 0, stream>>>(n, ptr, idx, v, keep, out_ptr, out_idx, out_v);
  SPAI_HIP_CHECK(hipGetLastError());
}

}  // namespace spai
