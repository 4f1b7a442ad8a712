// Copies between large pageable host arrays (the C ABI's inputs and Result
// arrays) and device memory.
//
// A copy from pageable memory is staged by the runtime through pinned memory
// one piece at a time, with the host side of each piece (a memcpy, and the
// first-touch page faults of a freshly allocated destination) on one thread:
// about 4-10 GB/s.  Here the data moves through two pinned chunks of this
// library in turn: while the DMA engine copies one chunk, worker threads fill
// (or drain) the other, each taking a slice of it.  Index arrays cross the
// bus as int32 (the device's type): the host narrows int64 indices while it
// fills a chunk and widens them while it drains one, which halves their PCIe
// bytes.
#include <pthread.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "ops.hpp"

namespace spai {
namespace {

constexpr size_t kChunkBytes = size_t(8) << 20;  // pinned bytes per chunk
constexpr int64_t kMinPart = int64_t(1) << 15;   // elements per worker slice, at least
constexpr int kMaxWorkers = 8;

// Host worker threads that run the slices of one loop at a time; the calling
// thread takes slices too.
class HostPool {
 public:
  HostPool() {
    cpu_set_t set;
    int cpus = 1;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = CPU_COUNT(&set);
    const int workers = std::max(0, std::min(kMaxWorkers, cpus) - 1);
    for (int w = 0; w < workers; ++w) threads_.emplace_back([this] { Work(); });
  }
  // The pool is never destroyed (see Get).
  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;

  // f(begin, end) over [0, n) in slices of at least kMinPart elements.
  void For(int64_t n, const std::function<void(int64_t, int64_t)>& f) {
    const int64_t parts = std::min<int64_t>(int64_t(threads_.size()) + 1, std::max<int64_t>(1, n / kMinPart));
    if (parts <= 1) {
      if (n > 0) f(0, n);
      return;
    }
    std::lock_guard<std::mutex> call(call_mu_);
    {
      std::lock_guard<std::mutex> lock(mu_);
      job_ = &f;
      n_ = n;
      parts_ = parts;
      next_.store(0);
      busy_ = int(threads_.size());
      ++gen_;
    }
    cv_.notify_all();
    RunSlices();
    std::unique_lock<std::mutex> lock(mu_);
    done_cv_.wait(lock, [this] { return busy_ == 0; });
    job_ = nullptr;
  }

  static HostPool& Get();

 private:
  void RunSlices() {
    for (;;) {
      const int64_t p = next_.fetch_add(1);
      if (p >= parts_) return;
      const int64_t b = n_ * p / parts_, e = n_ * (p + 1) / parts_;
      (*job_)(b, e);
    }
  }
  void Work() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return gen_ != seen; });
        seen = gen_;
      }
      RunSlices();
      {
        std::lock_guard<std::mutex> lock(mu_);
        --busy_;
      }
      done_cv_.notify_one();
    }
  }

  std::vector<std::thread> threads_;
  std::mutex call_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int64_t, int64_t)>* job_ = nullptr;
  int64_t n_ = 0, parts_ = 0;
  std::atomic<int64_t> next_{0};
  int busy_ = 0;
  uint64_t gen_ = 0;
};

// The pinned chunks, shared by every context of the process (one transfer at
// a time; pinned host memory serves every device).
struct Staging {
  std::mutex mu;
  char* buf[2] = {nullptr, nullptr};
};

// Process-wide singletons, created on first use and never destroyed (their
// threads and pinned memory live until the process ends).  A child created
// by fork() has neither the parent's threads nor its HIP state, so it drops
// both pointers and builds its own on first use.
HostPool* g_pool = nullptr;
Staging* g_staging = nullptr;
std::mutex* g_init_mu = new std::mutex();
std::once_flag g_atfork_once;

// (the parent's objects are left as they are: their mutexes may have been
// held by threads that the child does not have)
void ChildAfterFork() {
  g_pool = nullptr;
  g_staging = nullptr;
  g_init_mu = new std::mutex();
}

void RegisterAtFork() { std::call_once(g_atfork_once, [] { pthread_atfork(nullptr, nullptr, ChildAfterFork); }); }

HostPool& HostPool::Get() {
  RegisterAtFork();
  std::lock_guard<std::mutex> lock(*g_init_mu);
  if (!g_pool) g_pool = new HostPool();
  return *g_pool;
}

Staging& GetStaging() {
  RegisterAtFork();
  std::lock_guard<std::mutex> lock(*g_init_mu);
  if (!g_staging) {
    auto* s = new Staging();
    for (auto& b : s->buf) SPAI_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&b), kChunkBytes, hipHostMallocDefault));
    g_staging = s;
  }
  return *g_staging;
}

// Two events of the calling device, destroyed with the object.
struct EventPair {
  hipEvent_t ev[2] = {nullptr, nullptr};
  EventPair() {
    for (auto& e : ev) SPAI_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  ~EventPair() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// When a transfer leaves by an exception, copies it queued may still use the
// pinned chunks: wait for them before the staging lock is released.
struct QuiesceOnThrow {
  hipStream_t stream;
  int pending = std::uncaught_exceptions();
  ~QuiesceOnThrow() {
    if (std::uncaught_exceptions() > pending) (void)hipStreamSynchronize(stream);
  }
};

// n device elements of `elem` bytes at dst, chunk by chunk: fill(b, e, out)
// writes elements [b, e) into out (pinned); returns once the device holds
// everything.
void StageIn(Ctx& ctx, char* dst, size_t elem, int64_t n,
             const std::function<void(int64_t, int64_t, char*)>& fill) {
  if (n <= 0) return;
  Staging& st = GetStaging();
  HostPool& pool = HostPool::Get();
  std::lock_guard<std::mutex> lock(st.mu);
  EventPair evs;
  QuiesceOnThrow guard{ctx.stream};
  const int64_t per = int64_t(kChunkBytes / elem);
  const int64_t chunks = (n + per - 1) / per;
  for (int64_t c = 0; c < chunks; ++c) {
    const int b = int(c & 1);
    if (c >= 2) SPAI_HIP_CHECK(hipEventSynchronize(evs.ev[b]));  // chunk c - 2 has left this buffer
    const int64_t off = c * per, len = std::min(per, n - off);
    char* const buf = st.buf[b];
    pool.For(len, [&](int64_t s, int64_t e) { fill(off + s, off + e, buf + size_t(s) * elem); });
    SPAI_HIP_CHECK(hipMemcpyAsync(dst + size_t(off) * elem, buf, size_t(len) * elem, hipMemcpyHostToDevice, ctx.stream));
    SPAI_HIP_CHECK(hipEventRecord(evs.ev[b], ctx.stream));
  }
  for (int b = 0; b < 2 && b < chunks; ++b) SPAI_HIP_CHECK(hipEventSynchronize(evs.ev[b]));
}

// n device elements of `elem` bytes from src, chunk by chunk: drain(b, e, in)
// consumes elements [b, e) from in (pinned).  The copies are queued after the
// work already on the context's stream.
void StageOut(Ctx& ctx, const char* src, size_t elem, int64_t n,
              const std::function<void(int64_t, int64_t, const char*)>& drain) {
  if (n <= 0) return;
  Staging& st = GetStaging();
  HostPool& pool = HostPool::Get();
  std::lock_guard<std::mutex> lock(st.mu);
  EventPair evs;
  QuiesceOnThrow guard{ctx.stream};
  const int64_t per = int64_t(kChunkBytes / elem);
  const int64_t chunks = (n + per - 1) / per;
  auto issue = [&](int64_t c) {
    const int b = int(c & 1);
    const int64_t off = c * per, len = std::min(per, n - off);
    SPAI_HIP_CHECK(hipMemcpyAsync(st.buf[b], src + size_t(off) * elem, size_t(len) * elem, hipMemcpyDeviceToHost,
                                  ctx.stream));
    SPAI_HIP_CHECK(hipEventRecord(evs.ev[b], ctx.stream));
  };
  issue(0);
  if (chunks > 1) issue(1);
  for (int64_t c = 0; c < chunks; ++c) {
    const int b = int(c & 1);
    SPAI_HIP_CHECK(hipEventSynchronize(evs.ev[b]));
    const int64_t off = c * per, len = std::min(per, n - off);
    const char* const buf = st.buf[b];
    pool.For(len, [&](int64_t s, int64_t e) { drain(off + s, off + e, buf + size_t(s) * elem); });
    if (c + 2 < chunks) issue(c + 2);
  }
}

}  // namespace

void HostParallelFor(int64_t n, const std::function<void(int64_t, int64_t)>& f) { HostPool::Get().For(n, f); }

void CopyToDevice(Ctx& ctx, void* dst, const void* src, size_t bytes) {
  const char* const in = static_cast<const char*>(src);
  StageIn(ctx, static_cast<char*>(dst), 1, int64_t(bytes),
          [&](int64_t b, int64_t e, char* out) { std::memcpy(out, in + b, size_t(e - b)); });
}

void NarrowToDevice(Ctx& ctx, int* dst, const int64_t* src, int64_t n, int64_t sub) {
  std::atomic<bool> bad{false};
  StageIn(ctx, reinterpret_cast<char*>(dst), sizeof(int), n, [&](int64_t b, int64_t e, char* out) {
    int* const o = reinterpret_cast<int*>(out);
    int64_t lo = 0, hi = 0;
    for (int64_t i = b; i < e; ++i) {
      const int64_t v = src[i] - sub;
      lo = std::min(lo, v);
      hi = std::max(hi, v);
      o[i - b] = int(v);
    }
    if (lo < 0 || hi > kMaxDeviceNnz) bad.store(true);
  });
  if (bad.load()) throw std::runtime_error("index outside the int32 range of the device arrays");
}

void CopyToHost(Ctx& ctx, void* dst, const void* src, size_t bytes) {
  char* const out = static_cast<char*>(dst);
  StageOut(ctx, static_cast<const char*>(src), 1, int64_t(bytes),
           [&](int64_t b, int64_t e, const char* in) { std::memcpy(out + b, in, size_t(e - b)); });
}

void WidenToHost(Ctx& ctx, int64_t* dst, const int* src, int64_t n) {
  StageOut(ctx, reinterpret_cast<const char*>(src), sizeof(int), n, [&](int64_t b, int64_t e, const char* in) {
    const int* const v = reinterpret_cast<const int*>(in);
    for (int64_t i = b; i < e; ++i) dst[i] = v[i - b];
  });
}

}  // namespace spai
