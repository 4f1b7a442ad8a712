// Achievable HBM bandwidth on this device: streaming read (sum), copy
// (y = x) and triad (a = b + s c) over arrays far larger than the last-level
// caches, timed with HIP events.  The fractions the benchmark reports for the
// sparse kernels are stated against these figures as well as against the
// 8 TB/s peak.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "dev_common.hpp"
#include "ops.hpp"

namespace spai {
namespace {

constexpr int kBwBlock = 256;
constexpr int kBwUnroll = 4;  // 16-byte accesses in flight per thread

// native 16-byte vector (the nontemporal builtins take clang vector types)
typedef double D2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(kBwBlock) void BwRead(int64_t n2, const D2* __restrict__ x, double* __restrict__ out) {
  const int64_t base = (int64_t(blockIdx.x) * kBwUnroll) * kBwBlock + threadIdx.x;
  D2 v[kBwUnroll];
#pragma unroll
  for (int u = 0; u < kBwUnroll; ++u) {
    const int64_t i = base + int64_t(u) * kBwBlock;
    v[u] = 0.0;
    if (i < n2) v[u] = __builtin_nontemporal_load(&x[i]);
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < kBwUnroll; ++u) s += v[u].x + v[u].y;
  // a store the compiler cannot drop, taken by almost no thread
  if (s == 1.2345e300) out[threadIdx.x] = s;
}

__global__ __launch_bounds__(kBwBlock) void BwCopy(int64_t n2, const D2* __restrict__ x, D2* __restrict__ y) {
  const int64_t base = (int64_t(blockIdx.x) * kBwUnroll) * kBwBlock + threadIdx.x;
  D2 v[kBwUnroll];
#pragma unroll
  for (int u = 0; u < kBwUnroll; ++u) {
    const int64_t i = base + int64_t(u) * kBwBlock;
    if (i < n2) v[u] = __builtin_nontemporal_load(&x[i]);
  }
#pragma unroll
  for (int u = 0; u < kBwUnroll; ++u) {
    const int64_t i = base + int64_t(u) * kBwBlock;
    if (i < n2) __builtin_nontemporal_store(v[u], &y[i]);
  }
}

__global__ __launch_bounds__(kBwBlock) void BwTriad(int64_t n2, D2* __restrict__ a, const D2* __restrict__ b,
                                                    const D2* __restrict__ c, double s) {
  const int64_t base = (int64_t(blockIdx.x) * kBwUnroll) * kBwBlock + threadIdx.x;
  D2 vb[kBwUnroll], vc[kBwUnroll];
#pragma unroll
  for (int u = 0; u < kBwUnroll; ++u) {
    const int64_t i = base + int64_t(u) * kBwBlock;
    if (i < n2) {
      vb[u] = __builtin_nontemporal_load(&b[i]);
      vc[u] = __builtin_nontemporal_load(&c[i]);
    }
  }
#pragma unroll
  for (int u = 0; u < kBwUnroll; ++u) {
    const int64_t i = base + int64_t(u) * kBwBlock;
    if (i < n2) {
      D2 r;
      r.x = vb[u].x + s * vc[u].x;
      r.y = vb[u].y + s * vc[u].y;
      __builtin_nontemporal_store(r, &a[i]);
    }
  }
}

// Device array straight from hipMalloc (not the library's block cache, which
// would keep gigabytes idle after the measurement).
struct RawArray {
  double* p = nullptr;
  explicit RawArray(int64_t n) { SPAI_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), sizeof(double) * n)); }
  ~RawArray() {
    if (p) (void)hipFree(p);
  }
  RawArray(const RawArray&) = delete;
  RawArray& operator=(const RawArray&) = delete;
  double* get() const { return p; }
};

__global__ void BwFill(int64_t n, double* __restrict__ p, double v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

}  // namespace

// Median GB/s over `repeats` launches of each kernel on arrays of
// bytes_per_array bytes: out[0] read, out[1] copy, out[2] triad (bytes moved
// = 1, 2 and 3 arrays per launch).
void StreamBandwidth(int64_t bytes_per_array, int64_t repeats, double out[3]) {
  if (bytes_per_array < (int64_t(1) << 20) || bytes_per_array > (int64_t(1) << 36))
    throw std::invalid_argument("bytes_per_array must lie in [1 MiB, 64 GiB]");
  if (repeats < 1 || repeats > 1000) throw std::invalid_argument("repeats must lie in [1, 1000]");
  Ctx c;
  const int64_t n = bytes_per_array / 16 * 2;  // doubles, a whole number of double2
  const int64_t n2 = n / 2;
  RawArray a(n), b(n), x(n), sink(kBwBlock);
  BwFill<<<4096, 256, 0, c.stream>>>(n, a.get(), 0.0);
  BwFill<<<4096, 256, 0, c.stream>>>(n, b.get(), 1.0);
  BwFill<<<4096, 256, 0, c.stream>>>(n, x.get(), 2.0);
  const int64_t blocks = (n2 + int64_t(kBwBlock) * kBwUnroll - 1) / (int64_t(kBwBlock) * kBwUnroll);
  if (blocks > (int64_t(1) << 31) - 1) throw std::invalid_argument("array too large for one grid");
  const unsigned g = unsigned(blocks);
  std::vector<hipEvent_t> ev(2 * repeats);
  for (auto& e : ev) SPAI_HIP_CHECK(hipEventCreate(&e));
  const D2* xs = reinterpret_cast<const D2*>(x.get());
  const D2* bs = reinterpret_cast<const D2*>(b.get());
  D2* as = reinterpret_cast<D2*>(a.get());
  for (int kind = 0; kind < 3; ++kind) {
    // one untimed launch first
    for (int64_t r = -1; r < repeats; ++r) {
      if (r >= 0) SPAI_HIP_CHECK(hipEventRecord(ev[2 * r], c.stream));
      if (kind == 0)
        BwRead<<<g, kBwBlock, 0, c.stream>>>(n2, xs, sink.get());
      else if (kind == 1)
        BwCopy<<<g, kBwBlock, 0, c.stream>>>(n2, xs, as);
      else
        BwTriad<<<g, kBwBlock, 0, c.stream>>>(n2, as, bs, xs, 0.5);
      if (r >= 0) SPAI_HIP_CHECK(hipEventRecord(ev[2 * r + 1], c.stream));
    }
    SPAI_HIP_CHECK(hipGetLastError());
    SPAI_HIP_CHECK(hipStreamSynchronize(c.stream));
    std::vector<double> ms(repeats);
    for (int64_t r = 0; r < repeats; ++r) {
      float t = 0.f;
      SPAI_HIP_CHECK(hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]));
      ms[r] = t;
    }
    std::sort(ms.begin(), ms.end());
    const double med = ms[ms.size() / 2];
    const double bytes = double(n) * 8.0 * (kind + 1);
    out[kind] = bytes / (med * 1e-3) / 1e9;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  c.Sync();
}

}  // namespace spai
