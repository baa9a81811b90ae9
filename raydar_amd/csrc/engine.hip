// Wavefront engine (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <cstring>

namespace rdh {

Engine::~Engine() {
  if (external_) return;  // nothing here is owned
  release_buffers();
  for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
  if (arena_free_) (void)hipEventDestroy(arena_free_);
  if (tickets_pin_) (void)hipHostFree(tickets_pin_);
  if (counters) (void)hipFree(counters);
  if (stream) (void)hipStreamDestroy(stream);
}

bool Engine::check(hipError_t e, const char *what, std::string *err) {
  if (e == hipSuccess) return true;
  *err = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

bool Engine::init(std::string *err) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    *err = std::string("no HIP device available (") + hipGetErrorString(e) +
           "); the HIP renderer needs an MI355X (gfx950) GPU";
    (void)hipGetLastError();
    return false;
  }
  if (!check(hipSetDevice(device), "hipSetDevice", err)) return false;
  hipDeviceProp_t prop;
  if (!check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties", err)) return false;
  num_cus = prop.multiProcessorCount;
  if (!check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate", err))
    return false;
  if (!check(hipMalloc(&counters, sizeof(DevCounters)), "hipMalloc(counters)", err)) return false;
  if (!check(hipMemsetAsync(counters, 0, sizeof(DevCounters), stream), "hipMemset", err))
    return false;
  if (!check(hipHostMalloc(&ctl_host, sizeof(QueueCtl), hipHostMallocDefault), "hipHostMalloc",
             err))
    return false;
  if (!check(hipMalloc(&ctl, sizeof(QueueCtl)), "hipMalloc(ctl)", err)) return false;
  return check(hipStreamSynchronize(stream), "hipStreamSynchronize", err);
}

void Engine::bind_external(const ExternalBuffers &b, const Engine &owner) {
  external_ = true;
  device = owner.device;
  num_cus = owner.num_cus;
  stream = owner.stream;
  ctl_host = owner.ctl_host;
  q[0] = b.q[0];
  q[1] = b.q[1];
  hit_t = b.hit_t;
  hit_id = b.hit_id;
  ctl = b.ctl;
  L = b.L;
  d_tickets = b.tickets;
  d_ticket_done = b.ticket_done;
  counters = b.counters;
  q_alloc = b.queue;
  seg_cap = b.seg_cap;
  l_alloc = b.l_slots;
  ticket_alloc = b.tickets_n;
  opt.queue = b.queue;
  opt.l_capacity = b.l_slots;
  opt.pipes = 1;
}

void Engine::release_buffers() {
  if (external_) return;
  for (int p = 0; p < 2; ++p) {
    void *f[] = {q[p].ray, q[p].dir, q[p].word, q[p].att, q[p].light};
    for (void *x : f)
      if (x) (void)hipFree(x);
    q[p] = PathQueue{};
  }
  if (hit_t) { (void)hipFree(hit_t); hit_t = nullptr; }
  if (hit_id) { (void)hipFree(hit_id); hit_id = nullptr; }
  if (L) { (void)hipFree(L); L = nullptr; }
  if (first_hit_) { (void)hipFree(first_hit_); first_hit_ = nullptr; }
  first_hit_px_ = 0;
  first_hit_frame_ = 0;
  q_alloc = 0;
  l_alloc = 0;
  if (d_tickets) { (void)hipFree(d_tickets); d_tickets = nullptr; }
  if (d_ticket_done) { (void)hipFree(d_ticket_done); d_ticket_done = nullptr; }
  ticket_alloc = 0;
  if (ctl) { (void)hipFree(ctl); ctl = nullptr; }
  if (ctl_host) { (void)hipHostFree(ctl_host); ctl_host = nullptr; }
}

// Slots per segment so that the shade workgroups sharing a segment can never
// overfill it: workgroup b handles at most ceil(Q / (G*B)) tiles of B paths
// and ceil(G / kSegments) workgroups share a segment.
uint32_t Engine::segment_slots() const {
  const LaunchShape sh = shape(2);
  const uint64_t tile = (uint64_t)sh.grid * sh.block;
  const uint64_t per_block = ((opt.queue + tile - 1) / tile) * sh.block;
  const uint64_t blocks = (sh.grid + kSegments - 1) / kSegments;
  const uint64_t even = (opt.queue + kSegments - 1) / kSegments;
  return (uint32_t)std::max<uint64_t>(blocks * per_block, even);
}

uint32_t Engine::queued(const QueueCtl &c, int parity) {
  uint32_t n = 0;
  for (int s = 0; s < kSegments; ++s) n += c.count[parity][s * kSegStride];
  return n;
}

bool Engine::ensure_buffers(std::string *err) {
  if (opt.queue < 1024) opt.queue = 1024;
  const uint32_t need = segment_slots();
  if (external_) {
    // borrowed buffers: they must already fit these options
    if (opt.queue != q_alloc || need > seg_cap || opt.l_capacity > l_alloc) {
      *err = "launch shape needs larger borrowed buffers (internal error)";
      return false;
    }
    return true;
  }
  if (q_alloc != opt.queue || need > seg_cap) {
    const uint64_t slots = (uint64_t)kSegments * need;
    for (int p = 0; p < 2; ++p) {
      void *f[] = {q[p].ray, q[p].dir, q[p].word, q[p].att, q[p].light};
      for (void *x : f)
        if (x) (void)hipFree(x);
      q[p] = PathQueue{};
      if (!check(hipMalloc(&q[p].ray, sizeof(QRay) * slots), "hipMalloc(queue)", err) ||
          !check(hipMalloc(&q[p].dir, sizeof(QDir) * slots), "hipMalloc(queue)", err) ||
          !check(hipMalloc(&q[p].word, sizeof(QWord) * slots), "hipMalloc(queue)", err) ||
          !check(hipMalloc(&q[p].att, sizeof(QAtt) * slots), "hipMalloc(queue)", err) ||
          !check(hipMalloc(&q[p].light, sizeof(QLight) * slots), "hipMalloc(queue)", err))
        return false;
    }
    if (hit_t) (void)hipFree(hit_t);
    if (hit_id) (void)hipFree(hit_id);
    hit_t = nullptr;
    hit_id = nullptr;
    if (!check(hipMalloc(&hit_t, 4ull * slots), "hipMalloc(hits)", err)) return false;
    if (!check(hipMalloc(&hit_id, 4ull * slots), "hipMalloc(hits)", err)) return false;
    q_alloc = opt.queue;
    seg_cap = need;
    ++generation;
  }
  if (l_alloc < opt.l_capacity) {
    if (L) (void)hipFree(L);
    L = nullptr;
    if (!check(hipMalloc(&L, sizeof(float4) * opt.l_capacity), "hipMalloc(radiance arena)", err))
      return false;
    l_alloc = opt.l_capacity;
    ++generation;
  }
  if (!ctl && !check(hipMalloc(&ctl, sizeof(QueueCtl)), "hipMalloc(ctl)", err)) return false;
  if (!ctl_host &&
      !check(hipHostMalloc(&ctl_host, sizeof(QueueCtl), hipHostMallocDefault), "hipHostMalloc", err))
    return false;
  return true;
}

uint64_t Engine::device_bytes() const {
  // two queues of 14 words per slot plus the 8-byte hit record
  uint64_t b = (uint64_t)kSegments * seg_cap * (2 * 56 + 8) + l_alloc * sizeof(float4) +
               (uint64_t)ticket_alloc * (sizeof(DevTicket) + 8) + first_hit_px_ * sizeof(HitWord);
  for (const auto &e : aux_) b += e->device_bytes();
  return b;
}

bool Engine::ensure_tickets(uint32_t n, std::string *err) {
  if (n <= ticket_alloc) return true;
  if (external_) {
    *err = "more tickets than the borrowed table holds (internal error)";
    return false;
  }
  uint32_t cap = std::max<uint32_t>(n, 2 * ticket_alloc);
  if (d_tickets) (void)hipFree(d_tickets);
  if (d_ticket_done) (void)hipFree(d_ticket_done);
  d_tickets = nullptr;
  d_ticket_done = nullptr;
  if (!check(hipMalloc(&d_tickets, sizeof(DevTicket) * cap), "hipMalloc(tickets)", err)) return false;
  if (!check(hipMalloc(&d_ticket_done, 8ull * cap), "hipMalloc(tickets)", err)) return false;
  ticket_alloc = cap;
  ++generation;  // a captured graph holds the old table pointers: recapture
  return true;
}

size_t DeviceCache::size_class(size_t bytes) {
  if (bytes <= 4096) return 4096;
  int e = 63 - __builtin_clzll((unsigned long long)bytes);
  const size_t step = (size_t)1 << (e - 3);
  return (bytes + step - 1) / step * step;
}

hipError_t DeviceCache::alloc(void **p, size_t bytes) {
  const size_t c = size_class(bytes);
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = kept_.find(c);
    if (it != kept_.end()) {
      *p = it->second;
      kept_.erase(it);
      kept_bytes_ -= c;
      ++hits;
      return hipSuccess;
    }
    ++misses;
  }
  hipError_t e = hipMalloc(p, c);
  if (e == hipErrorOutOfMemory && kept_bytes() > 0) {
    // the kept buffers may be what fills the device: give them back, retry
    (void)hipGetLastError();
    trim();
    e = hipMalloc(p, c);
  }
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(mu_);
  class_of_[*p] = c;
  return hipSuccess;
}

void DeviceCache::free(void *p) {
  if (!p) return;
  std::vector<void *> drop;
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = class_of_.find(p);
    if (it == class_of_.end()) return;
    kept_.emplace(it->second, p);
    kept_bytes_ += it->second;
    // over the limit: give back the largest kept buffers down to half of it
    if (kept_bytes_ > keep_max)
      while (kept_bytes_ > keep_max / 2 && !kept_.empty()) {
        auto last = std::prev(kept_.end());
        kept_bytes_ -= last->first;
        drop.push_back(last->second);
        class_of_.erase(last->second);
        kept_.erase(last);
      }
  }
  for (void *q : drop) (void)hipFree(q);
}

void DeviceCache::trim() {
  std::vector<void *> drop;
  {
    std::lock_guard<std::mutex> g(mu_);
    for (auto &kv : kept_) {
      drop.push_back(kv.second);
      class_of_.erase(kv.second);
    }
    kept_.clear();
    kept_bytes_ = 0;
  }
  for (void *q : drop) (void)hipFree(q);
}

uint64_t DeviceCache::kept_bytes() const {
  std::lock_guard<std::mutex> g(mu_);
  return kept_bytes_;
}

hipError_t Engine::dmalloc(void **p, size_t bytes) {
  return cache ? cache->alloc(p, bytes) : hipMalloc(p, bytes);
}

void Engine::dfree(void *p) {
  if (!p) return;
  if (cache) cache->free(p);
  else (void)hipFree(p);
}

bool Engine::upload_scene(const BuiltScene &b, DeviceScene *out, std::string *err) {
  free_scene(out);
  out->n_nodes = (uint32_t)b.nodes.size();
  out->n_nodes4 = (uint32_t)b.nodes4.size();
  out->stack4 = b.stack4;
  out->max_leaf = b.max_leaf;
  out->n_prims = (uint32_t)b.prims.size();
  out->r0 = b.r0;
  out->depth = b.depth;
  const size_t nb = sizeof(DevNode) * b.nodes.size(), pb = sizeof(DevPrim) * b.prims.size(),
               ob = sizeof(DevObject) * b.objects.size(), n4 = sizeof(DevNode4) * b.nodes4.size();
  out->bytes = nb + pb + ob + n4;
  if (nb && !check(dmalloc((void **)&out->nodes, nb), "hipMalloc(bvh)", err)) return false;
  if (n4 && !check(dmalloc((void **)&out->nodes4, n4), "hipMalloc(bvh4)", err)) return false;
  if (pb && !check(dmalloc((void **)&out->prims, pb), "hipMalloc(prims)", err)) return false;
  if (ob && !check(dmalloc((void **)&out->objects, ob), "hipMalloc(objects)", err)) return false;
  if (nb && !check(hipMemcpyAsync(out->nodes, b.nodes.data(), nb, hipMemcpyHostToDevice, stream),
                   "upload bvh", err))
    return false;
  if (n4 && !check(hipMemcpyAsync(out->nodes4, b.nodes4.data(), n4, hipMemcpyHostToDevice, stream),
                   "upload bvh4", err))
    return false;
  if (pb && !check(hipMemcpyAsync(out->prims, b.prims.data(), pb, hipMemcpyHostToDevice, stream),
                   "upload prims", err))
    return false;
  if (ob && !check(hipMemcpyAsync(out->objects, b.objects.data(), ob, hipMemcpyHostToDevice,
                                  stream),
                   "upload objects", err))
    return false;
  return check(hipStreamSynchronize(stream), "upload scene", err);
}

void Engine::free_scene(DeviceScene *s) {
  dfree(s->nodes);
  dfree(s->nodes4);
  dfree(s->prims);
  dfree(s->objects);
  *s = DeviceScene();
}

IterArgs Engine::iter_args(const DevJob *d_jobs) const {
  IterArgs a;
  a.q[0] = q[0];
  a.q[1] = q[1];
  a.hit_t = hit_t;
  a.hit_id = hit_id;
  a.ctl = ctl;
  a.jobs = d_jobs;
  a.tickets = d_tickets;
  a.capacity = q_alloc;
  a.seg_cap = seg_cap;
  a.L = L;
  a.ticket_done = d_ticket_done;
  a.counters = counters;
  a.brute_force = opt.brute_force;
  a.refill = opt.refill;
  a.lds_bvh_bytes = 0;
  a.stack_rows = 0;
  a.lds_obj_bytes = 0;
  a.bvh4 = 0;
  a.leaf1 = 0;
  a.vk = opt.vk;
  a.first_hit = nullptr;
  a.first_hit_mode = 0;
  a.chunk_log2 = opt.lane_chunk_log2;
  a.slots = (uint32_t)kSegments * seg_cap;
  a.ticket_cap = ticket_alloc;
  a.job_cap = 1;
  a.l_cap = l_alloc;
  a.single_job = 0;
  a.track_done = 1;
  a.wide_stack = 0;
  return a;
}

LaunchShape Engine::shape(int which, const IterArgs *a) const {
  LaunchShape s;
  s.block = opt.block;
  if (opt.grid > 0) {
    s.grid = opt.grid;
  } else {
    const bool lanes = which == 1 && opt.refill > 0;
    const uint32_t dyn = (which == 1 && a) ? extend_dynamic_lds(*a, s.block) : 0;
    const bool w4 = which == 1 && a && a->bvh4;
    const bool single = a ? a->single_job != 0 : true;
    const uint32_t dyn_lds = which == 2 && a ? shade_dynamic_lds(*a) : dyn;
    int per_cu = std::max(1, occupancy_blocks(which, s.block, lanes, dyn ? a->lds_bvh_bytes : 0, dyn_lds, w4, single));
    if (which == 2 && opt.shade_blocks > 0) {
      per_cu = std::min(per_cu, opt.shade_blocks);
    } else if (which == 2 && single) {
      // single-job shade fits 8 waves/SIMD but runs faster at 7 (plain_1080p
      // frames -0.6%, shade alone -8%, the closest-hit kernel +0.5%)
      per_cu = std::min(per_cu, std::max(1, 7 * 4 / (s.block / 64)));
    }
    s.grid = num_cus * per_cu;
  }
  return s;
}

// Grid for an iteration known to hold at most `bound` paths: one lane per
// path for extend and shade (never more than the occupancy grid).  A smaller
// shade grid changes how many workgroups share a queue segment and how many
// tiles each handles, so it is used only if the segment capacity still
// covers the worst case, ceil(G / kSegments) * ceil(bound / (G * B)) * B.
LaunchShape Engine::bounded_shape(int which, const IterArgs &a, uint32_t bound) const {
  LaunchShape s = shape(which, &a);
  if (bound == kNoBound || opt.grid > 0 || which == 0) return s;
  const uint64_t b = (uint64_t)s.block;
  const int g = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)s.grid, (bound + b - 1) / b));
  if (which == 2) {
    const uint64_t tiles = ((uint64_t)bound + (uint64_t)g * b - 1) / ((uint64_t)g * b);
    const uint64_t need = (uint64_t)((g + kSegments - 1) / kSegments) * tiles * b;
    if (need > seg_cap) return s;
  }
  s.grid = g;
  return s;
}

bool Engine::timed(int which, std::string *err,
                   hipError_t (*fn)(const IterArgs &, int, const LaunchShape &, bool, hipStream_t),
                   const IterArgs &a, int parity, uint32_t bound) {
  const bool count = opt.count != 0;
  const LaunchShape sh = bounded_shape(which, a, bound);
  if (!opt.kernel_timing || count) return check(fn(a, parity, sh, count, stream), "kernel launch", err);
  hipEvent_t e0, e1;
  if (!check(hipEventCreate(&e0), "hipEventCreate", err)) return false;
  if (!check(hipEventCreate(&e1), "hipEventCreate", err)) return false;
  if (!check(hipEventRecord(e0, stream), "hipEventRecord", err)) return false;
  if (!check(fn(a, parity, sh, false, stream), "kernel launch", err)) return false;
  if (!check(hipEventRecord(e1, stream), "hipEventRecord", err)) return false;
  if (!check(hipEventSynchronize(e1), "hipEventSynchronize", err)) return false;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  kernel_ms[which] += ms;
  kernel_launches[which] += 1;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return true;
}

bool Engine::launch_iteration(const IterArgs &a, int parity, std::string *err, uint32_t bound) {
  ++iterations;
  // with a bound every path has been generated: generate would add nothing
  return (bound != kNoBound || timed(0, err, launch_generate, a, parity, bound)) &&
         timed(1, err, launch_extend, a, parity, bound) && timed(2, err, launch_shade, a, parity, bound);
}

bool Engine::run_tickets(const DevJob *d_jobs, std::vector<DevTicket> tickets,
                         const std::vector<DevJob> &host_jobs, std::string *err) {
  const uint64_t frame = hit_frame;  // set by the caller for this pass only
  hit_frame = 0;
  if (tickets.empty()) return true;
  if (!ensure_buffers(err)) return false;
  uint64_t w = 0;
  uint32_t max_bounces = 1;
  std::vector<DevTicket> traced;
  for (auto &t : tickets) {
    t.work_begin = w;
    t.l_offset = w;
    const DevJob &j = host_jobs[t.job];
    if (j.max_bounces > 0) {
      traced.push_back(t);
      w += t.paths;
      max_bounces = std::max(max_bounces, j.max_bounces);
    }
  }
  if (w > l_alloc) {
    *err = "pass needs " + std::to_string(w) + " radiance slots, arena holds " +
           std::to_string(l_alloc);
    return false;
  }
  if (!traced.empty()) {
    // Pipes: a one-job pass is divided by samples (pipe k traces samples
    // [s_k, s_k+1) into their own radiance slots), anything else runs whole.
    // Each pipe must still get 4M paths (or half its queue, if smaller):
    // below that the extra launches cost more than the overlap gains
    // (measured break-even between 4M and 8M paths per pass).
    std::vector<std::vector<DevTicket>> parts;
    int np = 1;
    if (traced.size() == 1 && host_jobs.size() == 1) {
      // automatic: two pipes when the scene is traversed from LDS (extend
      // bound by VALU issue), three otherwise (extend also waits on node
      // fetches, so a third stream hides more of it)
      const DevJob &j = host_jobs[0];
      const uint64_t tree = opt.bvh4 && j.n_nodes4 ? 128ull * j.n_nodes4 : 64ull * j.n_nodes;
      const bool lds = tree + 32ull * j.n_prims <= opt.lds_bvh;
      const int want = opt.pipes > 0 ? opt.pipes : (lds ? 2 : 3);
      const uint64_t per_pipe = std::min<uint64_t>(q_alloc / 2, 1ull << 22);
      const uint64_t by_size = traced[0].paths / std::max<uint64_t>(1, per_pipe);
      np = (int)std::min<uint64_t>({(uint64_t)want, traced[0].spp, by_size});
      np = std::max(np, 1);
    }
    pipes_used = np;
    if (np <= 1) {
      parts.push_back(traced);
    } else {
      const DevTicket &t = traced[0];
      const uint64_t npix = t.paths / t.spp;
      uint32_t s0 = 0;
      for (int k = 0; k < np; ++k) {
        const uint32_t s1 = (uint32_t)((uint64_t)t.spp * (k + 1) / np);
        DevTicket p = t;
        p.work_begin = 0;
        p.l_offset = t.l_offset + s0 * npix;
        p.sample_base = t.sample_base + s0;
        p.spp = s1 - s0;
        p.paths = (uint32_t)(p.spp * npix);
        parts.push_back({p});
        s0 = s1;
      }
    }
    while ((int)aux_.size() < np - 1) {
      auto e = std::make_unique<Engine>();
      e->device = device;
      if (!e->init(err)) return false;
      aux_.push_back(std::move(e));
    }
    // Pipes 1.. write this engine's radiance arena from their own streams:
    // they start after the work already queued here (the previous pass's
    // accumulation still reads the arena).
    if (np > 1) {
      if (!arena_free_ &&
          !check(hipEventCreateWithFlags(&arena_free_, hipEventDisableTiming), "hipEventCreate", err))
        return false;
      if (!check(hipEventRecord(arena_free_, stream), "hipEventRecord", err)) return false;
      for (int k = 1; k < np; ++k)
        if (!check(hipStreamWaitEvent(aux_[k - 1]->stream, arena_free_, 0), "hipStreamWaitEvent", err))
          return false;
    }
    struct Pipe {
      Engine *e;
      IterArgs a;
      uint64_t w;
      int parity;
      uint64_t it, cap_iters;
      bool done;
      uint32_t bound;  // paths in flight at most (kNoBound until all are generated)
    };
    std::vector<Pipe> pipes;
    // On an early return (error), every pipe's stream is drained first: their
    // kernels write this engine's radiance arena, which the caller may free.
    struct DrainOnError {
      std::vector<Pipe> &p;
      bool ok = false;
      ~DrainOnError() {
        if (!ok)
          for (Pipe &x : p) (void)hipStreamSynchronize(x.e->stream);
      }
    } drain{pipes};
    for (int k = 0; k < np; ++k) {
      Engine *e = k == 0 ? this : aux_[k - 1].get();
      if (k > 0) {
        e->opt = opt;
        e->opt.l_capacity = 0;  // radiance goes to this engine's arena
        e->opt.pipes = 1;
        if (!e->ensure_buffers(err)) return false;
      }
      const std::vector<DevTicket> &tk = parts[k];
      const uint32_t n = (uint32_t)tk.size();
      uint64_t wk = 0;
      for (const DevTicket &t : tk) wk += t.paths;
      if (!e->ensure_tickets(n, err)) return false;
      if (e->external_) {
        // no pinned copy to allocate: the runtime stages the pageable source
        if (!check(hipMemcpyAsync(e->d_tickets, tk.data(), sizeof(DevTicket) * n,
                                  hipMemcpyHostToDevice, e->stream),
                   "upload tickets", err))
          return false;
      } else {
        // from pinned memory (no staging copy); the previous call's upload
        // from it has completed, as every pipe was synchronised at its end
        if (n > e->tickets_pin_cap_) {
          if (e->tickets_pin_) (void)hipHostFree(e->tickets_pin_);
          e->tickets_pin_ = nullptr;
          e->tickets_pin_cap_ = 0;
          if (!check(hipHostMalloc((void **)&e->tickets_pin_, sizeof(DevTicket) * n,
                                   hipHostMallocDefault),
                     "hipHostMalloc(tickets)", err))
            return false;
          e->tickets_pin_cap_ = n;
        }
        std::memcpy(e->tickets_pin_, tk.data(), sizeof(DevTicket) * n);
        if (!check(hipMemcpyAsync(e->d_tickets, e->tickets_pin_, sizeof(DevTicket) * n,
                                  hipMemcpyHostToDevice, e->stream),
                   "upload tickets", err))
          return false;
      }
      // (no per-ticket completion counts here: track_done is off below)
      if (!check(launch_set_ctl(e->ctl, wk, n, e->stream), "set ctl", err)) return false;
      IterArgs a = e->iter_args(d_jobs);
      a.L = L;
      a.l_cap = l_alloc;
      a.counters = counters;
      // the single-job kernels take every path as job 0 / ticket 0
      a.single_job = host_jobs.size() == 1 && n == 1 ? 1 : 0;
      a.job_cap = (uint32_t)host_jobs.size();
      a.track_done = 0;  // completion is known from the drained queue
      for (const DevJob &j : host_jobs)
        if (j.n_nodes > kNarrowStackNodes) a.wide_stack = 1;
      if (a.single_job && !a.wide_stack && !opt.brute_force && opt.refill > 0) {
        // One scene: the 4-wide tree when its stack bound fits the kernel's
        // stack; a small scene is traversed from LDS, with a stack as deep as
        // the traversal can need.
        const DevJob &j = host_jobs[0];
        a.bvh4 = opt.bvh4 && j.n_nodes4 > 0 && j.stack4 + 3 <= (uint32_t)kStackDepth ? 1 : 0;
        // (the Vulkan-semantics kernels have no one-primitive-leaf variant)
        a.leaf1 = a.bvh4 && j.max_leaf == 1 && !a.vk ? 1 : 0;
        const uint64_t nb = a.bvh4 ? (uint64_t)sizeof(DevNode4) * j.n_nodes4 : (uint64_t)sizeof(DevNode) * j.n_nodes;
        const uint64_t bytes = nb + (uint64_t)sizeof(DevPrim) * j.n_prims;
        const uint32_t rows = a.bvh4 ? j.stack4 + 3 : j.depth + 1;
        if (j.n_nodes > 0 && bytes <= opt.lds_bvh && rows <= (uint32_t)kStackDepth) {
          a.lds_bvh_bytes = (uint32_t)bytes;
          a.stack_rows = rows;
        }
      }
      if (a.single_job) {
        // a small scene's objects are read by shade from LDS: the object load
        // depends on the hit id, so from L2 it is a second dependent fetch
        const uint64_t ob = (uint64_t)sizeof(DevObject) * host_jobs[0].n_prims;
        if (ob > 0 && ob <= opt.lds_objects) a.lds_obj_bytes = (uint32_t)ob;
      }
      const uint64_t cap_iters = (wk / e->q_alloc + 2) * (uint64_t)(max_bounces + 1) + 64;
      pipes.push_back(Pipe{e, a, wk, 0, 0, cap_iters, false, kNoBound});
    }
    // Camera-ray hits: one scene, CPU semantics and an owned queue that holds
    // a whole sample, so that each pipe's first iteration generates every
    // pixel's camera path at least once and its shade records each pixel's
    // hit; from the second iteration on, fresh paths take their pixel's hit
    // and extend traces only the survivors (same hits, same bits: the camera
    // ray of a pixel does not depend on the sample).
    bool first_hits = opt.hit_reuse && !external_ && host_jobs.size() == 1 &&
                      pipes[0].a.single_job && !opt.brute_force;
    uint64_t npix_fh = first_hits ? (uint64_t)host_jobs[0].width * host_jobs[0].height : 0;
    for (const Pipe &p : pipes) first_hits = first_hits && p.e->q_alloc >= npix_fh && p.w >= npix_fh;
    if (first_hits && npix_fh > first_hit_px_) {
      if (first_hit_) (void)hipFree(first_hit_);
      first_hit_ = nullptr;
      first_hit_px_ = 0;
      first_hit_frame_ = 0;
      if (!check(hipMalloc(&first_hit_, sizeof(HitWord) * npix_fh), "hipMalloc(first hits)", err))
        return false;
      first_hit_px_ = npix_fh;
    }
    // an earlier pass of the same frame recorded them: every iteration reuses
    const bool carried = first_hits && frame != 0 && frame == first_hit_frame_;
    if (!carried) first_hit_frame_ = 0;  // rewritten by this pass's first iterations
    for (Pipe &p : pipes) p.a.first_hit = first_hits ? first_hit_ : nullptr;
    const int group = std::max(1, opt.sync_every);
    for (;;) {
      for (int g = 0; g < group; ++g)
        for (Pipe &p : pipes) {
          if (p.done) continue;
          const bool record = first_hits && !carried && p.it == 0;
          p.a.first_hit_mode = first_hits ? (record ? 1 : 2) : 0;
          // the recording iteration generates one sample only, so that the
          // pipe's other camera paths all take their hits from it, when the
          // pass needs several fills of the queue anyway; a pass that fits
          // in one fill is generated whole (an extra iteration would cost
          // more than the queries it saves: plain_small 0.50 -> 0.56 ms)
          p.a.capacity = record && p.w > p.e->q_alloc ? (uint32_t)npix_fh : p.e->q_alloc;
          if (!p.e->launch_iteration(p.a, p.parity, err, p.bound)) return false;
          p.parity ^= 1;
          ++p.it;
        }
      for (Pipe &p : pipes)
        if (!p.done && !check(hipMemcpyAsync(p.e->ctl_host, p.e->ctl, sizeof(QueueCtl),
                                             hipMemcpyDeviceToHost, p.e->stream),
                              "read queue state", err))
          return false;
      bool all = true;
      for (Pipe &p : pipes) {
        if (p.done) continue;
        if (!check(hipStreamSynchronize(p.e->stream), "iteration", err)) return false;
        const QueueCtl &c = *p.e->ctl_host;
        if (c.overflow) {
          *err = "path queue segment overflow (internal error)";
          return false;
        }
        if (c.violation) {
          *err = "device bounds check failed, mask " + std::to_string(c.violation) +
                 " (internal error)";
          return false;
        }
        if (queued(c, p.parity) == 0 && c.cursor[p.parity] == p.w) {
          p.done = true;
        } else if (c.cursor[p.parity] == p.w) {
          // every path generated: later iterations only shrink the queue
          p.bound = queued(c, p.parity);
        }
        if (!p.done && p.it > p.cap_iters) {
          *err = "wavefront loop made no progress (internal error)";
          return false;
        }
        all = all && p.done;
      }
      if (all) break;
    }
    drain.ok = true;
    // the table now holds this frame's hits (every pipe ran its recording
    // iteration, or they were already there)
    if (first_hits) first_hit_frame_ = frame;
    for (int k = 1; k < np; ++k) {  // the pipes' kernel statistics are this engine's
      Engine &e = *aux_[k - 1];
      for (int i = 0; i < 5; ++i) {
        kernel_ms[i] += e.kernel_ms[i];
        kernel_launches[i] += e.kernel_launches[i];
        e.kernel_ms[i] = 0;
        e.kernel_launches[i] = 0;
      }
      iterations += e.iterations;
      e.iterations = 0;
    }
  }
  // Accumulate every ticket in order: per pixel, samples in sample order.
  for (const auto &t : tickets) {
    const DevJob &j = host_jobs[t.job];
    const float4 *src = j.max_bounces > 0 ? L : nullptr;
    DevTicket tt = t;
    if (j.max_bounces == 0) tt.l_offset = 0;
    const int grid = std::min<int>((int)((j.width * j.height + 255) / 256), num_cus * 8);
    if (opt.kernel_timing && !opt.count) {
      hipEvent_t e0, e1;
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      (void)hipEventRecord(e0, stream);
      if (!check(launch_accumulate(d_jobs + t.job, tt, src, counters, false, std::max(grid, 1), stream),
                 "accumulate", err))
        return false;
      (void)hipEventRecord(e1, stream);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      kernel_ms[3] += ms;
      kernel_launches[3] += 1;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    } else if (!check(launch_accumulate(d_jobs + t.job, tt, src, counters, opt.count != 0,
                                        std::max(grid, 1), stream),
                      "accumulate", err)) {
      return false;
    }
  }
  return true;
}

void fill_job(const rd_scene &s, const DeviceScene &ds, float *acc, uint32_t max_bounces,
              uint64_t seed, DevJob *j) {
  std::memset(j, 0, sizeof(*j));
  j->nodes = ds.nodes;
  j->prims = ds.prims;
  j->objects = ds.objects;
  j->acc = acc;
  j->n_prims = ds.n_prims;
  j->n_nodes = ds.n_nodes;
  j->width = s.camera.resolution_x;
  j->height = s.camera.resolution_y;
  j->max_bounces = max_bounces;
  j->world_kind = s.world.kind;
  j->seed = seed;
  std::memcpy(j->inv_proj, s.camera.inverse_proj_matrix, sizeof(j->inv_proj));
  std::memcpy(j->inv_view, s.camera.inverse_view_matrix, sizeof(j->inv_view));
  for (int k = 0; k < 3; ++k) {
    j->position[k] = s.camera.position[k];
    j->top[k] = s.world.top_color[k];
    j->bottom[k] = s.world.bottom_color[k];
    j->solid[k] = s.world.solid_color[k];
  }
  j->r0_tol = ds.r0;
  j->depth = ds.depth;
  j->nodes4 = ds.nodes4;
  j->n_nodes4 = ds.n_nodes4;
  j->stack4 = ds.stack4;
  j->max_leaf = ds.max_leaf;
}

}  // namespace rdh
