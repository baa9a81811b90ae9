// Fake HIP runtime for the host-code harness (see hip/hip_runtime.h).
// Test infrastructure only.
//
// Semantics modelled after HIP's documentation:
//   - a stream runs its operations in order on a thread of its own, each
//     after a random delay; streams created non-blocking do not order with
//     the null stream or with each other except through events
//   - hipMemcpyAsync from pageable host memory reads the source when called
//     (staging) and writes the device in stream order; into pageable host
//     memory it returns once the copy has run; pinned memory (hipHostMalloc)
//     is read or written when the copy runs
//   - hipFree and hipHostFree wait for all work on the device first
//   - hipStreamWaitEvent waits for the event's latest record at call time
//     (none: no wait); hipEventSynchronize likewise
//   - a capturing stream records its operations into the graph instead of
//     running them; hipGraphLaunch enqueues the recorded operations
#include "fake_device.hpp"

#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <thread>

struct FakeStream {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::function<void()>> q;
  uint64_t enq = 0, done = 0;
  bool stop = false, capturing = false;
  std::vector<std::function<void()>> captured;
  std::thread th;
};

struct FakeEvent {
  std::mutex mu;
  std::condition_variable cv;
  uint64_t recorded = 0, completed = 0;
  std::chrono::steady_clock::time_point t;
};

struct FakeGraph {
  std::vector<std::function<void()>> ops;
};
struct FakeGraphExec {
  std::vector<std::function<void()>> ops;
};

namespace {

using Clock = std::chrono::steady_clock;

struct Write {
  size_t len;
  uint64_t seq;
};
struct Alloc {
  size_t size;
  bool mapped;
  std::map<size_t, Write> writes;  // offset -> latest copy that wrote there
};

std::mutex g_mu;  // registries, errors, hooks, statistics
std::map<uintptr_t, Alloc> g_dev;
std::map<uintptr_t, size_t> g_pinned;
std::vector<std::string> g_errors;
std::vector<std::function<void(const void *, size_t)>> g_free_hooks;
std::vector<FakeStream *> g_streams;
std::map<FakeEvent *, std::shared_ptr<FakeEvent>> g_events;
fakehip::Stats g_stats;
std::atomic<uint64_t> g_seq{0};
std::atomic<int> g_max_delay{40};
thread_local hipError_t t_last = hipSuccess;
constexpr size_t kMapThreshold = 64ull << 20;  // larger buffers: untouched mmap

hipError_t ret(hipError_t e) {
  if (e != hipSuccess) t_last = e;
  return e;
}

void delay() {
  thread_local std::mt19937 rng(
      (uint32_t)std::hash<std::thread::id>()(std::this_thread::get_id()));
  const int m = g_max_delay.load();
  if (m <= 0) return;
  const int us = std::uniform_int_distribution<int>(0, m)(rng);
  if (us > m / 2) std::this_thread::sleep_for(std::chrono::microseconds(us));
}

void worker(FakeStream *s) {
  for (;;) {
    std::function<void()> f;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->q.empty(); });
      if (s->q.empty()) return;
      f = std::move(s->q.front());
      s->q.pop_front();
    }
    delay();
    f();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      ++s->done;
    }
    s->cv.notify_all();
  }
}

FakeStream *new_stream() {
  auto *s = new FakeStream();
  s->th = std::thread(worker, s);
  std::lock_guard<std::mutex> g(g_mu);
  g_streams.push_back(s);
  g_stats.streams++;
  return s;
}

FakeStream *null_stream() {
  static FakeStream *s = new_stream();
  return s;
}

FakeStream *S(hipStream_t s) { return s ? s : null_stream(); }

void sync_stream(FakeStream *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] { return s->done == s->enq; });
}

void sync_all() {
  std::vector<FakeStream *> all;
  {
    std::lock_guard<std::mutex> g(g_mu);
    all = g_streams;
  }
  for (FakeStream *s : all) sync_stream(s);
}

// allocation holding [p, p + n), under g_mu
Alloc *find(const void *p, size_t n, uintptr_t *base) {
  const uintptr_t a = (uintptr_t)p;
  auto it = g_dev.upper_bound(a);
  if (it == g_dev.begin()) return nullptr;
  --it;
  if (a + n > it->first + it->second.size || a < it->first) return nullptr;
  *base = it->first;
  return &it->second;
}

bool pinned(const void *p, size_t n) {
  std::lock_guard<std::mutex> g(g_mu);
  const uintptr_t a = (uintptr_t)p;
  auto it = g_pinned.upper_bound(a);
  if (it == g_pinned.begin()) return false;
  --it;
  return a + n <= it->first + it->second;
}

// a copy or memset of [p, p + n) on the device: checked, and its sequence
// number recorded
bool device_write(void *p, size_t n, const char *what) {
  std::lock_guard<std::mutex> g(g_mu);
  uintptr_t base = 0;
  Alloc *a = find(p, n, &base);
  if (!a) {
    g_errors.push_back(std::string(what) + ": device range not inside a live allocation");
    return false;
  }
  const size_t off = (uintptr_t)p - base;
  auto lo = a->writes.lower_bound(off);
  while (lo != a->writes.end() && lo->first + lo->second.len <= off + n && lo->first < off + n)
    lo = a->writes.erase(lo);
  a->writes[off] = Write{n, ++g_seq};
  return true;
}

bool device_read(const void *p, size_t n, const char *what) {
  std::lock_guard<std::mutex> g(g_mu);
  uintptr_t base = 0;
  if (!find(p, n, &base)) {
    g_errors.push_back(std::string(what) + ": device range not inside a live allocation");
    return false;
  }
  return true;
}

}  // namespace

namespace fakehip {

void enqueue(hipStream_t sh, std::function<void()> fn) {
  FakeStream *s = S(sh);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->capturing) {
      s->captured.push_back(std::move(fn));
      return;
    }
    s->q.push_back(std::move(fn));
    ++s->enq;
  }
  s->cv.notify_all();
  std::lock_guard<std::mutex> g(g_mu);
  g_stats.ops++;
}

bool live(const void *p, size_t bytes, std::string *why) {
  std::lock_guard<std::mutex> g(g_mu);
  uintptr_t base = 0;
  if (find(p, bytes ? bytes : 1, &base)) return true;
  if (why) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%p + %zu bytes is not inside a live device allocation", p,
                  bytes);
    *why = buf;
  }
  return false;
}

uint64_t bytes_from(const void *p) {
  std::lock_guard<std::mutex> g(g_mu);
  uintptr_t base = 0;
  Alloc *a = find(p, 1, &base);
  return a ? base + a->size - (uintptr_t)p : 0;
}

uint64_t last_write(const void *p, size_t bytes) {
  std::lock_guard<std::mutex> g(g_mu);
  uintptr_t base = 0;
  Alloc *a = find(p, bytes, &base);
  if (!a) return 0;
  const size_t off = (uintptr_t)p - base;
  uint64_t m = 0;
  for (const auto &w : a->writes)
    if (w.first < off + bytes && w.first + w.second.len > off) m = std::max(m, w.second.seq);
  return m;
}

void error(const std::string &msg) {
  std::lock_guard<std::mutex> g(g_mu);
  g_errors.push_back(msg);
}

std::vector<std::string> errors() {
  std::lock_guard<std::mutex> g(g_mu);
  return g_errors;
}

void on_free(std::function<void(const void *, size_t)> fn) {
  std::lock_guard<std::mutex> g(g_mu);
  g_free_hooks.push_back(std::move(fn));
}

void set_max_delay_us(int us) { g_max_delay = us; }

Stats stats() {
  std::lock_guard<std::mutex> g(g_mu);
  Stats s = g_stats;
  s.live_allocs = g_dev.size();
  return s;
}

void count_kernel() {
  std::lock_guard<std::mutex> g(g_mu);
  g_stats.kernels++;
}

}  // namespace fakehip

// ------------------------------------------------------------ device -----
hipError_t hipGetDeviceCount(int *n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return ret(d == 0 ? hipSuccess : hipErrorInvalidValue); }
hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
  std::memset(p, 0, sizeof(*p));
  std::snprintf(p->name, sizeof(p->name), "fake MI355X");
  std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "gfx950:sramecc+:xnack-");
  p->multiProcessorCount = 256;
  p->totalGlobalMem = 288ull << 30;
  p->pciBusID = 5;
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) {
  *v = 256;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() {
  sync_all();
  return hipSuccess;
}
hipError_t hipGetLastError() {
  const hipError_t e = t_last;
  t_last = hipSuccess;
  return e;
}
const char *hipGetErrorString(hipError_t e) {
  switch (e) {
    case hipSuccess: return "no error";
    case hipErrorInvalidValue: return "invalid argument";
    case hipErrorOutOfMemory: return "out of memory";
    default: return "fake device error";
  }
}

// ------------------------------------------------------------ memory -----
hipError_t hipMalloc(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return hipSuccess;
  const bool mapped = bytes >= kMapThreshold;
  void *q = mapped ? mmap(nullptr, bytes, PROT_READ | PROT_WRITE,
                          MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0)
                   : std::malloc(bytes);
  if (!q || q == MAP_FAILED) return ret(hipErrorOutOfMemory);
  std::lock_guard<std::mutex> g(g_mu);
  g_dev[(uintptr_t)q] = Alloc{bytes, mapped, {}};
  g_stats.mallocs++;
  *p = q;
  return hipSuccess;
}

// FAKEHIP_FREE_NOSYNC=1: hipFree and hipHostFree return without waiting for
// queued work.  On the MI355X a hipFree issued behind 358 ms of work queued on
// another stream returned after 143 ms, so a free does not order itself after
// all earlier work; code that is correct without the wait is correct there.
static bool free_waits() {
  static const bool nosync = [] {
    const char *e = std::getenv("FAKEHIP_FREE_NOSYNC");
    return e && e[0] == '1';
  }();
  return !nosync;
}

hipError_t hipFree(void *p) {
  if (!p) return hipSuccess;
  if (free_waits()) sync_all();
  size_t size = 0;
  bool mapped = false;
  std::vector<std::function<void(const void *, size_t)>> hooks;
  {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_dev.find((uintptr_t)p);
    if (it == g_dev.end()) {
      g_errors.push_back("hipFree of a pointer that is not a live device allocation");
      return ret(hipErrorInvalidValue);
    }
    size = it->second.size;
    mapped = it->second.mapped;
    g_dev.erase(it);
    hooks = g_free_hooks;
    g_stats.frees++;
  }
  for (auto &h : hooks) h(p, size);
  if (mapped) munmap(p, size);
  else std::free(p);
  return hipSuccess;
}

hipError_t hipGetSymbolAddress(void **p, const void *symbol, size_t bytes) {
  std::lock_guard<std::mutex> g(g_mu);
  g_dev.emplace((uintptr_t)symbol, Alloc{bytes, true, {}});  // static: never freed
  *p = const_cast<void *>(symbol);
  return hipSuccess;
}

hipError_t hipHostMalloc(void **p, size_t bytes, unsigned) {
  *p = std::malloc(bytes ? bytes : 1);
  if (!*p) return ret(hipErrorOutOfMemory);
  std::lock_guard<std::mutex> g(g_mu);
  g_pinned[(uintptr_t)*p] = bytes;
  return hipSuccess;
}

hipError_t hipHostFree(void *p) {
  if (!p) return hipSuccess;
  if (free_waits()) sync_all();
  {
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_pinned.erase((uintptr_t)p)) {
      g_errors.push_back("hipHostFree of a pointer that is not pinned host memory");
      return ret(hipErrorInvalidValue);
    }
  }
  std::free(p);
  return hipSuccess;
}

static hipError_t copy(void *dst, const void *src, size_t n, hipMemcpyKind k, hipStream_t s,
                       bool sync) {
  if (n == 0) return hipSuccess;
  {
    std::lock_guard<std::mutex> g(g_mu);
    g_stats.copies++;
  }
  if (k == hipMemcpyHostToDevice) {
    if (pinned(src, n)) {
      fakehip::enqueue(s, [dst, src, n] {
        if (device_write(dst, n, "host-to-device copy")) std::memcpy(dst, src, n);
      });
    } else {
      auto staged = std::make_shared<std::vector<uint8_t>>((const uint8_t *)src,
                                                            (const uint8_t *)src + n);
      fakehip::enqueue(s, [dst, staged, n] {
        if (device_write(dst, n, "host-to-device copy")) std::memcpy(dst, staged->data(), n);
      });
    }
    if (sync) sync_stream(S(s));
  } else if (k == hipMemcpyDeviceToHost) {
    fakehip::enqueue(s, [dst, src, n] {
      if (device_read(src, n, "device-to-host copy")) std::memcpy(dst, src, n);
    });
    if (sync || !pinned(dst, n)) sync_stream(S(s));
  } else if (k == hipMemcpyDeviceToDevice) {
    fakehip::enqueue(s, [dst, src, n] {
      if (device_read(src, n, "device-to-device copy") &&
          device_write(dst, n, "device-to-device copy"))
        std::memmove(dst, src, n);
    });
    if (sync) sync_stream(S(s));
  } else {
    return ret(hipErrorInvalidValue);
  }
  return hipSuccess;
}

hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind k) {
  return copy(dst, src, n, k, nullptr, true);
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind k, hipStream_t s) {
  return copy(dst, src, n, k, s, false);
}
hipError_t hipMemsetAsync(void *dst, int v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  fakehip::enqueue(s, [dst, v, n] {
    if (device_write(dst, n, "memset")) std::memset(dst, v, n);
  });
  return hipSuccess;
}
hipError_t hipMemset(void *dst, int v, size_t n) {
  hipError_t e = hipMemsetAsync(dst, v, n, nullptr);
  sync_stream(null_stream());
  return e;
}

// ----------------------------------------------------------- streams -----
hipError_t hipStreamCreate(hipStream_t *s) {
  *s = new_stream();
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) {
  if (!s) return ret(hipErrorInvalidValue);
  sync_stream(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  s->th.join();
  {
    std::lock_guard<std::mutex> g(g_mu);
    g_streams.erase(std::find(g_streams.begin(), g_streams.end(), s));
  }
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  FakeStream *f = S(s);
  {
    std::lock_guard<std::mutex> lk(f->mu);
    if (f->capturing) {
      fakehip::error("hipStreamSynchronize on a capturing stream");
      return ret(hipErrorInvalidValue);
    }
  }
  sync_stream(f);
  return hipSuccess;
}

static std::shared_ptr<FakeEvent> event_ref(hipEvent_t e) {
  std::lock_guard<std::mutex> g(g_mu);
  auto it = g_events.find(e);
  return it == g_events.end() ? nullptr : it->second;
}

hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  auto ev = event_ref(e);
  if (!ev) {
    fakehip::error("hipStreamWaitEvent on a destroyed event");
    return ret(hipErrorInvalidValue);
  }
  uint64_t gen;
  {
    std::lock_guard<std::mutex> lk(ev->mu);
    gen = ev->recorded;
  }
  if (gen == 0) return hipSuccess;
  fakehip::enqueue(s, [ev, gen] {
    std::unique_lock<std::mutex> lk(ev->mu);
    ev->cv.wait(lk, [&] { return ev->completed >= gen; });
  });
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (!s) return ret(hipErrorInvalidValue);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->capturing) return ret(hipErrorInvalidValue);
  s->capturing = true;
  s->captured.clear();
  return hipSuccess;
}

hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (!s) return ret(hipErrorInvalidValue);
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->capturing) return ret(hipErrorInvalidValue);
  s->capturing = false;
  *g = new FakeGraph{std::move(s->captured)};
  s->captured.clear();
  return hipSuccess;
}

// ------------------------------------------------------------ events -----
hipError_t hipEventCreate(hipEvent_t *e) {
  auto ev = std::make_shared<FakeEvent>();
  std::lock_guard<std::mutex> g(g_mu);
  g_events[ev.get()] = ev;
  *e = ev.get();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  std::lock_guard<std::mutex> g(g_mu);
  return ret(g_events.erase(e) ? hipSuccess : hipErrorInvalidValue);
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  auto ev = event_ref(e);
  if (!ev) return ret(hipErrorInvalidValue);
  uint64_t gen;
  {
    std::lock_guard<std::mutex> lk(ev->mu);
    gen = ++ev->recorded;
  }
  fakehip::enqueue(s, [ev, gen] {
    {
      std::lock_guard<std::mutex> lk(ev->mu);
      if (gen > ev->completed) ev->completed = gen;
      ev->t = Clock::now();
    }
    ev->cv.notify_all();
  });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  auto ev = event_ref(e);
  if (!ev) return ret(hipErrorInvalidValue);
  std::unique_lock<std::mutex> lk(ev->mu);
  const uint64_t gen = ev->recorded;
  ev->cv.wait(lk, [&] { return ev->completed >= gen; });
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  auto ea = event_ref(a), eb = event_ref(b);
  if (!ea || !eb) return ret(hipErrorInvalidValue);
  std::lock_guard<std::mutex> la(ea->mu);
  std::lock_guard<std::mutex> lb(eb->mu);
  *ms = std::chrono::duration<float, std::milli>(eb->t - ea->t).count();
  return hipSuccess;
}

// ------------------------------------------------------------ graphs -----
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  if (!g) return ret(hipErrorInvalidValue);
  *x = new FakeGraphExec{g->ops};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
  if (!x) return ret(hipErrorInvalidValue);
  for (const auto &op : x->ops) fakehip::enqueue(s, op);
  std::lock_guard<std::mutex> g(g_mu);
  g_stats.graph_launches++;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  delete x;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
