// Device plumbing for libspai_hip: error checks, a caching allocator, device
// CSC matrices and the per-call execution context (one HIP stream).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace spai {

#define SPAI_HIP_CHECK(expr)                                                                          \
  do {                                                                                                \
    hipError_t _e = (expr);                                                                           \
    if (_e != hipSuccess) {                                                                           \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " __FILE__ \
                               ":" + std::to_string(__LINE__) + " (" #expr ")");                      \
    }                                                                                                 \
  } while (0)

// Largest index the device kernels address with int32 offsets.
constexpr int64_t kMaxDeviceNnz = (int64_t(1) << 31) - 1;

// Process-wide cache of idle device blocks, per device, shared by the pools
// of successive contexts.  A block enters it only when its pool is destroyed,
// after the owning context has synchronised its stream, so no queued work
// still uses it.  It holds at most kCacheBytes (blocks of at most
// kCacheBlockBytes); the rest is freed.  Never destroyed: freeing at process
// exit could run after the HIP runtime has shut down.
class BlockCache {
 public:
  static constexpr size_t kCacheBytes = size_t(2) << 30;
  static constexpr size_t kCacheBlockBytes = size_t(512) << 20;

  static BlockCache& Get() {
    static BlockCache* cache = new BlockCache();
    return *cache;
  }
  // A cached block of [sz, 2 sz) bytes on `device` (its size in *got), or null.
  void* Take(int device, size_t sz, size_t* got) {
    std::lock_guard<std::mutex> lock(mu_);
    auto& m = blocks_[device];
    auto it = m.lower_bound(sz);
    if (it == m.end() || it->first >= 2 * sz) return nullptr;
    void* p = it->second;
    *got = it->first;
    bytes_ -= it->first;
    m.erase(it);
    return p;
  }
  // Keeps an idle block, or frees it when the cache is full.
  void Give(int device, size_t sz, void* p) {
    {
      std::lock_guard<std::mutex> lock(mu_);
      if (sz <= kCacheBlockBytes && bytes_ + sz <= kCacheBytes) {
        blocks_[device].emplace(sz, p);
        bytes_ += sz;
        return;
      }
    }
    (void)hipFree(p);
  }
  // Frees every cached block of `device` (to make room for an allocation).
  void Drain(int device) {
    std::multimap<size_t, void*> m;
    {
      std::lock_guard<std::mutex> lock(mu_);
      m.swap(blocks_[device]);
      for (auto& kv : m) bytes_ -= kv.first;
    }
    for (auto& kv : m) (void)hipFree(kv.second);
  }

 private:
  std::mutex mu_;
  std::map<int, std::multimap<size_t, void*>> blocks_;
  size_t bytes_ = 0;
};

// Caching device allocator.  Every kernel of one build runs on one stream, so
// a block released after its last use can be handed to a later launch on the
// same stream without synchronisation (stream order makes the reuse safe).
// Blocks come from the pool's own idle list, then from the process-wide
// cache, then from hipMalloc; when the pool is destroyed (its context has
// synchronised) its idle blocks go back to the cache.
class Pool {
 public:
  Pool() = default;
  Pool(const Pool&) = delete;
  Pool& operator=(const Pool&) = delete;
  ~Pool() {
    for (auto& kv : free_) BlockCache::Get().Give(device_, kv.first, kv.second);
    free_.clear();
  }
  void SetDevice(int device) { device_ = device; }

  void* Alloc(size_t bytes) {
    if (bytes == 0) bytes = 1;
    const size_t sz = RoundUp(bytes);
    auto it = free_.lower_bound(sz);
    if (it != free_.end() && it->first <= 2 * sz) {
      void* p = it->second;
      live_[p] = it->first;
      free_.erase(it);
      return p;
    }
    size_t got = 0;
    if (void* c = BlockCache::Get().Take(device_, sz, &got)) {
      live_[c] = got;
      bytes_reserved_ += got;
      return c;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      Trim();
      SPAI_HIP_CHECK(hipMalloc(&p, sz));
    }
    live_[p] = sz;
    bytes_reserved_ += sz;
    return p;
  }

  void Release(void* p) {
    if (!p) return;
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_.emplace(it->second, p);
    live_.erase(it);
  }

  // Frees the pool's idle blocks and the cache's blocks of this device.
  void Trim() {
    for (auto& kv : free_) {
      (void)hipFree(kv.second);
      bytes_reserved_ -= kv.first;
    }
    free_.clear();
    BlockCache::Get().Drain(device_);
  }

  size_t BytesReserved() const { return bytes_reserved_; }

 private:
  static size_t RoundUp(size_t b) {
    const size_t g = 256;
    b = (b + g - 1) / g * g;
    if (b <= (size_t(1) << 20)) {
      size_t p = 256;
      while (p < b) p <<= 1;
      return p;
    }
    // above 1 MiB: multiples of 1/8 of the next power of two
    size_t p = size_t(1) << 20;
    while (p < b) p <<= 1;
    const size_t step = p / 8;
    return (b + step - 1) / step * step;
  }

  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> live_;
  size_t bytes_reserved_ = 0;
  int device_ = 0;
};

// Move-only device array drawn from a Pool.
template <class T>
class DBuf {
 public:
  DBuf() = default;
  DBuf(Pool* pool, size_t n) : pool_(pool), n_(n) {
    if (n > 0) p_ = static_cast<T*>(pool->Alloc(n * sizeof(T)));
  }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : pool_(o.pool_), p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      Reset();
      pool_ = o.pool_;
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  ~DBuf() { Reset(); }
  void Reset() {
    if (p_ && pool_) pool_->Release(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  Pool* pool_ = nullptr;
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Sparsity pattern of a CSC matrix on the device (int32 offsets and rows,
// rows sorted inside each column).  Shared between matrices whose patterns
// coincide (e.g. Z = Pr R keeps R's pattern).
struct Pattern {
  int64_t rows = 0, cols = 0, nnz = 0;
  DBuf<int> ptr;  // cols + 1
  DBuf<int> idx;  // nnz
};

using PatternPtr = std::shared_ptr<Pattern>;

// Device CSC matrix with FP64 values.
struct DMat {
  PatternPtr pat;
  DBuf<double> val;
  int64_t rows() const { return pat ? pat->rows : 0; }
  int64_t cols() const { return pat ? pat->cols : 0; }
  int64_t nnz() const { return pat ? pat->nnz : 0; }
  const int* ptr() const { return pat->ptr.get(); }
  const int* idx() const { return pat->idx.get(); }
  double* v() const { return val.get(); }
};

}  // namespace spai
