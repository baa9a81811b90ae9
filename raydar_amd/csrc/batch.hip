// Streaming wavefront engine behind the batched renderer and the render
// server (include/raydar_hip_batch.h).
//
// Every admitted job owns device-resident state (BVH, primitives, materials,
// an RGBA32F accumulator, an RGBA8 image) and a slot in a fixed-capacity
// device job table.  Its samples are cut into passes; each pass becomes a
// ticket appended to one global work list.  The wavefront kernels consume the
// work list through a cursor, so paths of many jobs share the same queues:
// new camera paths fill whatever room the survivors of earlier jobs leave.
// The host loop ("pump") admits jobs under the HBM budget, appends tickets
// while the radiance ring has room, replays a captured hipGraph of two
// iterations, then reads back the per-ticket finished-path counts and
// accumulates every finished pass in pass order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/raydar_hip_batch.h"
#include "bvh.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "ring.hpp"

namespace rdh {

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint32_t kJobSlots = 4096;        // device job table entries
constexpr uint32_t kTicketCap = 1u << 20;   // default device ticket table entries

int64_t ns_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count();
}

struct SJob {
  uint64_t id = 0;
  uint32_t slot = 0;
  rd_scene scene{};                // header copy; objects point into `objects`
  std::vector<rd_object> objects;
  BuiltScene built;
  DeviceScene dev;
  float *acc = nullptr;
  uchar4 *rgba_d = nullptr;
  uint32_t W = 0, H = 0, spp = 0, bounces = 0, pass_spp = 0;
  uint64_t seed = 0;
  uint32_t passes_total = 0, passes_issued = 0, passes_done = 0;
  std::deque<uint32_t> tickets;    // issued, not yet accumulated (in pass order)
  uint8_t *out_rgba = nullptr;
  float *out_acc = nullptr;
  uint8_t *progress = nullptr;
  uint64_t bytes = 0;
  int state = 0;                   // 0 queued, 1 resident, 2 done, 3 failed
  int32_t status = RD_OK;
  Clock::time_point t_submit, t_admit, t_done;
};

// Device-memory account of one or more streamers (and the batch renderer's
// solo lane, which admits jobs from a thread of its own): one budget for jobs
// plus the engines' own buffers.
struct HbmPool {
  uint64_t budget = 200ull << 30;
  uint64_t fixed = 0;  // engine buffers of every ready streamer (and the solo lane)
  uint64_t in_use = 0, peak = 0;
  uint32_t resident = 0, max_resident = 0;
  mutable std::mutex mu;

  // Takes `bytes` for one job unless that would pass the budget while other
  // jobs are resident (a job always runs when nothing else is resident).
  bool reserve(uint64_t bytes) {
    std::lock_guard<std::mutex> g(mu);
    if (in_use + bytes + fixed > budget && resident > 0) return false;
    in_use += bytes;
    peak = std::max(peak, in_use);
    resident++;
    max_resident = std::max(max_resident, resident);
    return true;
  }
  void release(uint64_t bytes) {
    std::lock_guard<std::mutex> g(mu);
    in_use -= bytes;
    resident--;
  }
  void add_fixed(int64_t delta) {
    std::lock_guard<std::mutex> g(mu);
    fixed += (uint64_t)delta;
  }
  bool fits_alone(uint64_t bytes) const {
    std::lock_guard<std::mutex> g(mu);
    return bytes + fixed <= budget;
  }
  template <class T> T read(T HbmPool::*field) const {
    std::lock_guard<std::mutex> g(mu);
    return this->*field;
  }
};

}  // namespace

class Streamer {
 public:
  Engine eng;
  std::string err;
  HbmPool own_pool;
  HbmPool *pool = &own_pool;  // shared when the streamer is a group member
  int use_graph = 1;
  int iters_per_pump = 2;
  uint32_t default_pass = 0;
  uint32_t ticket_cap = kTicketCap;  // ticket table entries (fixed once ready)
  // statistics
  uint64_t paths_done = 0, graph_launches = 0, pumps = 0, rewinds = 0;

  ~Streamer() { teardown(); }

  // A streamer's radiance ring holds 128 Mi slots (2 GB): passes take at
  // most half of it, and the batch and the server run two or three streamers.
  Streamer() { eng.opt.l_capacity = 1ull << 27; }

  bool init() {
    if (!eng.init(&err)) return false;
    return true;
  }

  // Fixed-capacity tables and engine buffers (pointers stay stable so the
  // captured graph's kernel arguments never change).
  bool ensure_ready() {
    if (ready_) return true;
    if (!eng.ensure_buffers(&err)) return false;
    if (!eng.ensure_tickets(ticket_cap, &err)) return false;
    if (!eng.check(hipMalloc(&d_jobs_, sizeof(DevJob) * kJobSlots), "hipMalloc(job table)", &err))
      return false;
    if (!eng.check(hipMemsetAsync(eng.d_ticket_done, 0, 8ull * ticket_cap, eng.stream), "hipMemset", &err))
      return false;
    if (!eng.check(launch_set_ctl(eng.ctl, 0, 0, eng.stream), "set ctl", &err)) return false;
    if (!eng.check(hipHostMalloc(&done_host_, 8ull * ticket_cap, hipHostMallocDefault), "hipHostMalloc", &err))
      return false;
    for (uint32_t s = kJobSlots; s-- > 0;) free_slots_.push_back(s);
    h_jobs_.assign(kJobSlots, DevJob{});
    tickets_.reserve(ticket_cap);  // never reallocates: async uploads read from it
    ring_.reset(eng.l_alloc);
    fixed_bytes_ = engine_bytes();
    pool->add_fixed((int64_t)fixed_bytes_);
    ready_ = true;
    return eng.check(hipStreamSynchronize(eng.stream), "init", &err);
  }

  bool ready() const { return ready_; }
  uint64_t engine_bytes_if_ready() const { return ready_ ? fixed_bytes_ : 0; }

  uint64_t engine_bytes() const {
    return (uint64_t)kSegments * eng.seg_cap * (2 * 56 + 8) + eng.l_alloc * 16 + ticket_cap * 40ull +
           sizeof(DevJob) * (uint64_t)kJobSlots;
  }

  // Validates and prepares a job on the caller's thread (scene copy + BVH).
  std::unique_ptr<SJob> make_job(const rd_job_desc &d, int leaf, std::string *e) {
    if (!d.scene) { *e = "job scene is NULL"; return nullptr; }
    const rd_scene &s = *d.scene;
    if (s.camera.resolution_x == 0 || s.camera.resolution_y == 0) { *e = "zero-size image"; return nullptr; }
    if (s.world.kind == RD_WORLD_TRANSPARENT) {
      *e = "transparent world is not supported (raydar: todo!(\"transparent world support\"))";
      return nullptr;
    }
    if (d.samples == 0 || d.max_bounces == 0) { *e = "samples and max_bounces must be >= 1"; return nullptr; }
    if (d.max_bounces > 255) { *e = "max_bounces above 255"; return nullptr; }
    auto j = std::make_unique<SJob>();
    j->scene = s;
    j->objects.assign(s.objects, s.objects + s.object_count);
    j->scene.objects = j->objects.data();
    if (!build_scene(j->scene, leaf, &j->built, e) || !check_bvh(j->built, e)) return nullptr;
    j->W = s.camera.resolution_x;
    j->H = s.camera.resolution_y;
    j->spp = d.samples;
    j->bounces = d.max_bounces;
    j->seed = d.seed;
    const uint64_t npix = (uint64_t)j->W * j->H;
    // A pass of one sample must fit the radiance ring, or its ticket could
    // never be issued and the job would wait forever.
    if (npix > eng.opt.l_capacity) {
      *e = "image of " + std::to_string(npix) + " pixels is larger than the radiance ring (" +
           std::to_string(eng.opt.l_capacity) + " slots)";
      return nullptr;
    }
    uint32_t pass = d.pass_samples ? d.pass_samples : (default_pass ? default_pass : d.samples);
    pass = std::min(pass, d.samples);
    const uint64_t fit = std::max<uint64_t>(1, eng.opt.l_capacity / 2 / npix);
    pass = (uint32_t)std::min<uint64_t>({(uint64_t)pass, fit, 0xffffffffull / npix});
    if (pass == 0) { *e = "image too large for the radiance arena"; return nullptr; }
    j->pass_spp = pass;
    j->passes_total = (d.samples + pass - 1) / pass;
    j->out_rgba = d.rgba;
    j->out_acc = d.accumulator;
    j->progress = d.progress_rgba;
    j->bytes = j->built.nodes.size() * sizeof(DevNode) + j->built.prims.size() * sizeof(DevPrim) +
               j->built.objects.size() * sizeof(DevObject) + npix * 28;
    j->t_submit = Clock::now();
    return j;
  }

  bool fits_alone(const SJob &j) const { return pool->fits_alone(j.bytes); }

  void enqueue(std::unique_ptr<SJob> j) {
    queued_.push_back(j.get());
    owned_[j->id] = std::move(j);
  }

  bool idle() const { return queued_.empty() && resident_.empty(); }
  uint32_t n_resident() const { return (uint32_t)resident_.size(); }
  uint32_t n_queued() const { return (uint32_t)queued_.size(); }
  SJob *job(uint64_t id) {
    auto it = owned_.find(id);
    return it == owned_.end() ? nullptr : it->second.get();
  }
  void forget(uint64_t id) { owned_.erase(id); }

  // One round of the host loop.  Returns false on a device error.
  bool pump() { return launch() && harvest(); }

  // First half of a round: admit jobs, append tickets and enqueue
  // iterations on this streamer's stream without waiting for them.
  bool launch() {
    if (!ensure_ready()) return false;
    ++pumps;
    // Launch-shape options may have changed since the queues were sized;
    // resize only while no path is in flight.
    if (last_count_ == 0 && ctl_total_ == last_cursor_ && !eng.ensure_buffers(&err)) return false;
    if (!rewind()) return false;
    if (!admit()) return false;
    if (!issue()) return false;
    const bool work = ctl_total_ > last_cursor_ || last_count_ > 0;
    if (work) {
      if (!run_iterations()) return false;
    }
    return true;
  }

  // Once every issued ticket is accumulated and no path is in flight, the
  // ticket list starts over at index 0, so the table never fills for good
  // (at the cap, issue() stops until the work in flight drains).
  bool rewind() {
    if (uploaded_ == 0 || ticket_lo_ != uploaded_ || last_count_ != 0 || last_cursor_ != ctl_total_)
      return true;
    if (!eng.check(hipMemsetAsync(eng.d_ticket_done, 0, 8ull * uploaded_, eng.stream), "hipMemset", &err))
      return false;
    if (!eng.check(launch_set_ctl(eng.ctl, 0, 0, eng.stream), "set ctl", &err)) return false;
    tickets_.clear();
    ticket_job_.clear();
    ticket_fin_.clear();
    uploaded_ = ticket_lo_ = 0;
    ctl_total_ = last_cursor_ = 0;
    stalls_ = 0;
    ++rewinds;
    return true;
  }

  // Paths of queued and resident jobs not yet accumulated.
  uint64_t pending_paths() const {
    uint64_t n = 0;
    for (const SJob *j : queued_) n += (uint64_t)j->W * j->H * j->spp;
    for (const SJob *j : resident_)
      n += (uint64_t)j->W * j->H * (j->spp - std::min(j->spp, j->passes_done * j->pass_spp));
    return n;
  }

  uint64_t iterations() const { return eng.iterations; }

 private:
  bool ready_ = false;
  DevJob *d_jobs_ = nullptr;
  unsigned long long *done_host_ = nullptr;
  std::vector<uint32_t> free_slots_;
  std::deque<SJob *> queued_;
  std::vector<SJob *> resident_;
  std::map<uint64_t, std::unique_ptr<SJob>> owned_;
  std::vector<DevTicket> tickets_;
  std::vector<DevJob> h_jobs_;
  std::vector<SJob *> ticket_job_;
  std::vector<uint8_t> ticket_fin_;
  uint32_t uploaded_ = 0, ticket_lo_ = 0;
  uint64_t ctl_total_ = 0, last_cursor_ = 0;
  uint32_t last_count_ = 0;
  uint32_t stalls_ = 0;
  int parity_ = 0;
  RadianceRing ring_;
  uint64_t fixed_bytes_ = 0;
  hipGraphExec_t graph_ = nullptr;
  hipGraph_t graph_def_ = nullptr;
  int graph_block_ = 0, graph_grid_ = 0, graph_brute_ = 0, graph_refill_ = -1, graph_shade_ = -1;

  void teardown() {
    if (graph_) (void)hipGraphExecDestroy(graph_);
    if (graph_def_) (void)hipGraphDestroy(graph_def_);
    graph_ = nullptr;
    graph_def_ = nullptr;
    for (auto &kv : owned_) release_job(kv.second.get());
    if (d_jobs_) (void)hipFree(d_jobs_);
    d_jobs_ = nullptr;
    if (done_host_) (void)hipHostFree(done_host_);
    done_host_ = nullptr;
  }

  void release_job(SJob *j) {
    if (j->state != 1 && !j->acc && !j->dev.nodes && !j->dev.prims) return;
    eng.free_scene(&j->dev);
    eng.dfree(j->acc);
    eng.dfree(j->rgba_d);
    j->acc = nullptr;
    j->rgba_d = nullptr;
  }

  bool admit() {
    while (!queued_.empty()) {
      SJob *j = queued_.front();
      if (free_slots_.empty()) break;
      // wait for room unless nothing is resident (a job that fits alone always runs)
      if (!pool->reserve(j->bytes)) break;
      queued_.pop_front();
      // until the job is resident a failure gives its reservation back
      struct Unreserve {
        HbmPool *p;
        uint64_t b;
        bool keep = false;
        ~Unreserve() {
          if (!keep) p->release(b);
        }
      } unreserve{pool, j->bytes};
      j->slot = free_slots_.back();
      free_slots_.pop_back();
      const uint64_t npix = (uint64_t)j->W * j->H;
      if (!eng.upload_scene(j->built, &j->dev, &err)) return false;
      // the accumulator (16 B per pixel) is followed by the job's camera-ray
      // hit table (8 B per pixel), zero (not recorded) like the accumulator
      if (!eng.check(eng.dmalloc((void **)&j->acc, 24 * npix), "hipMalloc(accumulator)", &err)) return false;
      if (!eng.check(eng.dmalloc((void **)&j->rgba_d, 4 * npix), "hipMalloc(image)", &err)) return false;
      if (!eng.check(launch_zero_acc(j->acc, 6 * npix, eng.stream), "zero accumulator", &err)) return false;
      // the host mirror entry outlives the asynchronous copy
      DevJob &dj = h_jobs_[j->slot];
      fill_job(j->scene, j->dev, j->acc, j->bounces, j->seed, &dj);
      dj.first_hit = reinterpret_cast<HitWord *>(j->acc + 4 * npix);
      if (!eng.check(hipMemcpyAsync(d_jobs_ + j->slot, &dj, sizeof(DevJob), hipMemcpyHostToDevice,
                                    eng.stream),
                     "upload job", &err))
        return false;
      if (j->built.nodes.size() > kNarrowStackNodes) wide_ = true;
      j->built = BuiltScene();  // host copy no longer needed
      j->state = 1;
      j->t_admit = Clock::now();
      resident_.push_back(j);
      unreserve.keep = true;
    }
    return true;
  }

  bool issue() {
    // Keep enough work ahead of the cursor that no iteration of the next
    // pump finds the work list empty (each can admit a queue load).
    const uint64_t ahead = (uint64_t)(iters_per_pump + 1) * eng.q_alloc;
    bool added = false;
    for (SJob *j : resident_) {
      while (j->passes_issued < j->passes_total && ctl_total_ - last_cursor_ < ahead) {
        if (tickets_.size() >= ticket_cap) break;  // resumes after the next rewind
        const uint64_t npix = (uint64_t)j->W * j->H;
        const uint32_t s0 = j->passes_issued * j->pass_spp;
        const uint32_t spp = std::min(j->pass_spp, j->spp - s0);
        const uint32_t ti = (uint32_t)tickets_.size();
        uint64_t off;
        if (!ring_.alloc(spp * npix, ti, &off)) break;
        DevTicket t{};
        t.work_begin = ctl_total_;
        t.l_offset = off;
        t.job = j->slot;
        t.sample_base = s0;
        t.spp = spp;
        t.paths = (uint32_t)(spp * npix);
        tickets_.push_back(t);
        ticket_job_.push_back(j);
        ticket_fin_.push_back(0);
        j->tickets.push_back(ti);
        j->passes_issued++;
        ctl_total_ += t.paths;
        added = true;
      }
    }
    if (!added) return true;
    const uint32_t n = (uint32_t)tickets_.size();
    if (!eng.check(hipMemcpyAsync(eng.d_tickets + uploaded_, tickets_.data() + uploaded_,
                                  sizeof(DevTicket) * (n - uploaded_), hipMemcpyHostToDevice,
                                  eng.stream),
                   "upload tickets", &err))
      return false;
    uploaded_ = n;
    return eng.check(launch_extend_work(eng.ctl, ctl_total_, n, ticket_lo_, eng.stream),
                     "append work", &err);
  }

  bool wide_ = false;

  // The 4-wide tree while every resident job's walk fits the kernel's stack
  // (both trees of every job stay resident, so this may change between any
  // two iterations).
  bool want_w4() const {
    if (!eng.opt.bvh4 || wide_ || eng.opt.brute_force || eng.opt.refill <= 0) return false;
    for (const SJob *j : resident_)
      if (j->dev.n_nodes > 0 && (j->dev.n_nodes4 == 0 || j->dev.stack4 + 3 > (uint32_t)kStackDepth))
        return false;
    return true;
  }

  IterArgs args() const {
    IterArgs a = eng.iter_args(d_jobs_);
    a.single_job = 0;
    a.job_cap = kJobSlots;
    a.track_done = 1;
    a.wide_stack = wide_ ? 1 : 0;
    a.bvh4 = want_w4() ? 1 : 0;
    return a;
  }

  bool capture() {
    if (graph_) {
      (void)hipGraphExecDestroy(graph_);
      (void)hipGraphDestroy(graph_def_);
      graph_ = nullptr;
      graph_def_ = nullptr;
    }
    const IterArgs a = args();
    if (!eng.check(hipStreamBeginCapture(eng.stream, hipStreamCaptureModeThreadLocal), "begin capture", &err))
      return false;
    hipError_t e = hipSuccess;
    for (int p = 0; p < 2 && e == hipSuccess; ++p) {
      e = launch_generate(a, p, eng.shape(0), false, eng.stream);
      if (e == hipSuccess) e = launch_extend(a, p, eng.shape(1, &a), false, eng.stream);
      if (e == hipSuccess) e = launch_shade(a, p, eng.shape(2), false, eng.stream);
    }
    hipError_t e2 = hipStreamEndCapture(eng.stream, &graph_def_);
    if (!eng.check(e, "capture", &err) || !eng.check(e2, "end capture", &err)) return false;
    if (!eng.check(hipGraphInstantiate(&graph_, graph_def_, nullptr, nullptr, 0), "graph instantiate", &err))
      return false;
    graph_block_ = eng.opt.block;
    graph_grid_ = eng.opt.grid;
    graph_brute_ = eng.opt.brute_force;
    graph_refill_ = eng.opt.refill;
    graph_chunk_ = eng.opt.lane_chunk_log2;
    graph_shade_ = eng.opt.shade_blocks;
    graph_wide_ = wide_;
    graph_w4_ = a.bvh4 != 0;
    graph_gen_ = eng.generation;
    return true;
  }
  bool graph_wide_ = false, graph_w4_ = false;
  uint32_t graph_chunk_ = 0;
  uint64_t graph_gen_ = 0;

  bool run_iterations() {
    const int n = std::max(2, iters_per_pump & ~1);
    const bool direct = !use_graph || eng.opt.count || eng.opt.kernel_timing;
    if (direct) {
      const IterArgs a = args();
      for (int i = 0; i < n; ++i) {
        if (!eng.launch_iteration(a, parity_, &err)) return false;
        parity_ ^= 1;
      }
      return true;
    }
    if (!graph_ || graph_block_ != eng.opt.block || graph_grid_ != eng.opt.grid ||
        graph_brute_ != eng.opt.brute_force || graph_refill_ != eng.opt.refill || graph_shade_ != eng.opt.shade_blocks ||
        graph_chunk_ != eng.opt.lane_chunk_log2 ||
        graph_wide_ != wide_ || graph_w4_ != want_w4() ||
        graph_gen_ != eng.generation) {
      if (!capture()) return false;
    }
    for (int i = 0; i < n; i += 2) {
      if (!eng.check(hipGraphLaunch(graph_, eng.stream), "graph launch", &err)) return false;
      ++graph_launches;
      eng.iterations += 2;
    }
    return true;
  }

  // Second half of a round: wait for the stream, read the ticket progress
  // and accumulate every finished pass in pass order.
 public:
  bool harvest() {
    const uint32_t n = uploaded_;
    if (!eng.check(hipMemcpyAsync(eng.ctl_host, eng.ctl, sizeof(QueueCtl), hipMemcpyDeviceToHost,
                                  eng.stream),
                   "read queue state", &err))
      return false;
    if (n > ticket_lo_ &&
        !eng.check(hipMemcpyAsync(done_host_ + ticket_lo_, eng.d_ticket_done + ticket_lo_,
                                  8ull * (n - ticket_lo_), hipMemcpyDeviceToHost, eng.stream),
                   "read ticket progress", &err))
      return false;
    if (!eng.check(hipStreamSynchronize(eng.stream), "iteration", &err)) return false;
    if (eng.ctl_host->overflow) {
      err = "path queue segment overflow (internal error)";
      return false;
    }
    if (eng.ctl_host->violation) {
      err = "device bounds check failed, mask " + std::to_string(eng.ctl_host->violation) +
            " (internal error)";
      return false;
    }
    const uint32_t count = Engine::queued(*eng.ctl_host, parity_);
    const uint64_t cursor = eng.ctl_host->cursor[parity_];
    // every iteration either generates paths or advances every path in flight
    // by one bounce, so the state cannot stay frozen while work is pending
    if (count == last_count_ && cursor == last_cursor_ && (count > 0 || cursor < ctl_total_)) {
      if (++stalls_ > 1000) {
        err = "wavefront loop made no progress (internal error)";
        return false;
      }
    } else {
      stalls_ = 0;
    }
    last_count_ = count;
    last_cursor_ = cursor;
    for (uint32_t t = ticket_lo_; t < n; ++t)
      if (!ticket_fin_[t] && done_host_[t] == tickets_[t].paths) ticket_fin_[t] = 1;
    // accumulate finished passes in pass order, per job
    std::vector<SJob *> finished;
    for (SJob *j : resident_) {
      bool progressed = false;
      while (!j->tickets.empty() && ticket_fin_[j->tickets.front()] == 1) {
        const uint32_t ti = j->tickets.front();
        j->tickets.pop_front();
        const DevTicket &t = tickets_[ti];
        const int grid = (int)std::min<uint64_t>(((uint64_t)j->W * j->H + 255) / 256, 8192);
        if (!eng.check(launch_accumulate(d_jobs_ + j->slot, t, eng.L, eng.counters, eng.opt.count != 0,
                                         std::max(grid, 1), eng.stream),
                       "accumulate", &err))
          return false;
        ticket_fin_[ti] = 2;
        ring_.release(ti);
        j->passes_done++;
        paths_done += t.paths;
        progressed = true;
      }
      if (progressed && (j->progress || j->passes_done == j->passes_total)) {
        const uint64_t npix = (uint64_t)j->W * j->H;
        const int grid = (int)std::min<uint64_t>((npix + 255) / 256, 8192);
        const uint32_t samples = std::min(j->spp, j->passes_done * j->pass_spp);
        if (!eng.check(launch_tonemap(d_jobs_ + j->slot, (uint32_t)npix, (float)samples, j->rgba_d,
                                      eng.counters, false, std::max(grid, 1), eng.stream),
                       "tonemap", &err))
          return false;
        uint8_t *dst = j->passes_done == j->passes_total ? j->out_rgba : j->progress;
        if (j->passes_done == j->passes_total && j->progress &&
            !eng.check(hipMemcpyAsync(j->progress, j->rgba_d, 4 * npix, hipMemcpyDeviceToHost, eng.stream),
                       "read image", &err))
          return false;
        if (dst && !eng.check(hipMemcpyAsync(dst, j->rgba_d, 4 * npix, hipMemcpyDeviceToHost, eng.stream),
                              "read image", &err))
          return false;
      }
      if (j->passes_done == j->passes_total) {
        const uint64_t npix = (uint64_t)j->W * j->H;
        if (j->out_acc && !eng.check(hipMemcpyAsync(j->out_acc, j->acc, 16 * npix, hipMemcpyDeviceToHost, eng.stream),
                                     "read accumulator", &err))
          return false;
        finished.push_back(j);
      }
    }
    while (ticket_lo_ < n && ticket_fin_[ticket_lo_] == 2) ++ticket_lo_;
    if (!finished.empty()) {
      if (!eng.check(hipStreamSynchronize(eng.stream), "read results", &err)) return false;
      const auto now = Clock::now();
      for (SJob *j : finished) {
        j->t_done = now;
        j->state = 2;
        release_job(j);
        free_slots_.push_back(j->slot);
        pool->release(j->bytes);
        resident_.erase(std::find(resident_.begin(), resident_.end(), j));
      }
    }
    return true;
  }
};

// Streamers ("pipes") pumped in turn by one host thread.  Each owns its
// stream, queues, work list, radiance ring and job table; every job is given
// whole to the member with the fewest pending paths.  A round harvests a
// member's previous iterations and launches its next ones before moving to
// the next member, so while the host waits on one stream the other streams
// have iterations queued, and their closest-hit (ALU-bound) and shading
// (memory-bound) kernels share the CUs.
class StreamerGroup {
 public:
  DeviceCache cache;  // job buffers of every member (declared first: outlives them)
  std::vector<std::unique_ptr<Streamer>> m;
  std::string err;
  uint64_t pumps = 0;

  bool init(int n) {
    m.clear();
    launched_.clear();
    return grow(n);
  }

  // Members can be added or removed only before any of them has rendered.
  bool unused() const {
    for (const auto &s : m)
      if (s->eng.q_alloc || !s->idle()) return false;
    return true;
  }
  bool resize(int n) {
    if ((int)m.size() > n) {
      for (size_t i = n; i < m.size(); ++i) pool_.add_fixed(-(int64_t)m[i]->engine_bytes_if_ready());
      m.resize(n);
      launched_.resize(n);
      return true;
    }
    return grow(n);
  }

  Streamer &first() { return *m[0]; }
  const Streamer &first() const { return *m[0]; }
  template <class F> void each(F f) {
    for (auto &s : m) f(*s);
  }

  // One HBM budget for all members (jobs plus every member's own buffers).
  void set_budget(uint64_t total) {
    std::lock_guard<std::mutex> g(pool_.mu);
    pool_.budget = total;
  }
  uint64_t budget() const { return pool_.read(&HbmPool::budget); }
  HbmPool &pool() { return pool_; }

  std::unique_ptr<SJob> make_job(const rd_job_desc &d, int leaf, std::string *e) {
    return m[0]->make_job(d, leaf, e);
  }
  bool fits_alone(const SJob &j) const { return m[0]->fits_alone(j); }

  void enqueue(std::unique_ptr<SJob> j) {
    size_t best = 0;
    uint64_t least = ~0ull;
    for (size_t i = 0; i < m.size(); ++i) {
      const uint64_t p = m[i]->pending_paths();
      if (p < least) {
        least = p;
        best = i;
      }
    }
    m[best]->enqueue(std::move(j));
  }

  bool idle() const {
    for (const auto &s : m)
      if (!s->idle()) return false;
    return true;
  }

  SJob *job(uint64_t id) {
    for (auto &s : m)
      if (SJob *j = s->job(id)) return j;
    return nullptr;
  }
  void forget(uint64_t id) {
    for (auto &s : m) s->forget(id);
  }

  bool ensure_ready() {
    for (auto &s : m)
      if (!s->ensure_ready()) {
        err = s->err;
        return false;
      }
    return true;
  }

  bool pump() {
    ++pumps;
    for (size_t i = 0; i < m.size(); ++i) {
      Streamer &s = *m[i];
      if (launched_[i]) {
        launched_[i] = 0;
        if (!s.harvest()) {
          err = s.err;
          return false;
        }
      }
      if (!s.idle()) {
        if (!s.launch()) {
          err = s.err;
          return false;
        }
        launched_[i] = 1;
      }
    }
    return true;
  }

  uint64_t iterations() const { return sum([](const Streamer &s) { return s.iterations(); }); }
  uint64_t graph_launches() const { return sum([](const Streamer &s) { return s.graph_launches; }); }
  uint64_t paths_done() const { return sum([](const Streamer &s) { return s.paths_done; }); }
  uint64_t hbm_in_use() const { return pool_.read(&HbmPool::in_use); }
  uint64_t hbm_peak() const { return pool_.read(&HbmPool::peak); }
  uint64_t engine_bytes() const { return sum([](const Streamer &s) { return s.engine_bytes(); }); }
  uint32_t max_resident() const { return pool_.read(&HbmPool::max_resident); }
  uint32_t n_resident() const { return (uint32_t)sum([](const Streamer &s) { return s.n_resident(); }); }
  uint32_t n_queued() const { return (uint32_t)sum([](const Streamer &s) { return s.n_queued(); }); }

 private:
  std::vector<uint8_t> launched_;  // member has iterations not yet harvested
  HbmPool pool_;

  template <class F> uint64_t sum(F f) const {
    uint64_t n = 0;
    for (const auto &s : m) n += (uint64_t)f(*s);
    return n;
  }
  bool grow(int n) {
    while ((int)m.size() < n) {
      auto s = std::make_unique<Streamer>();
      if (!m.empty()) {  // same settings as the first member
        s->eng.opt = m[0]->eng.opt;
        s->use_graph = m[0]->use_graph;
        s->iters_per_pump = m[0]->iters_per_pump;
        s->default_pass = m[0]->default_pass;
        s->ticket_cap = m[0]->ticket_cap;
      }
      if (!s->init()) {
        err = s->err;
        return false;
      }
      s->pool = &pool_;
      s->eng.cache = &cache;
      m.push_back(std::move(s));
      launched_.push_back(0);
    }
    return true;
  }
};

// The batch renderer's solo lane.  A job of at least `min_paths` paths fills
// the GPU by itself, and the single-scene kernels render it faster than the
// shared queues do: the job's fields are launch constants, a tree of up to
// 14 KB is traversed from LDS, and the samples are split over pipes.  Such
// jobs are handed to this lane, which renders them one after another on a
// host thread of its own with a per-scene engine (the plain renderer's code
// path), while the streamers' shared queues take the smaller jobs; the two
// sets of kernels share the CUs.  A job's image does not depend on which
// side renders it: each pixel's samples are drawn from the same counters and
// accumulated in sample order either way.
constexpr uint64_t kSoloPaths = 64ull << 20;

class SoloLane {
 public:
  Engine eng;
  uint64_t min_paths = kSoloPaths;  // 0: every job goes to the streamers
  uint64_t jobs_done = 0, paths_done = 0;

  ~SoloLane() {
    stop();
    if (d_job_) (void)hipFree(d_job_);
  }

  // Whether a job goes to this lane: large enough, no progress image (the
  // lane renders a job in one go) and samples that fit the radiance arena.
  bool wants(const rd_job_desc &d) const {
    if (min_paths == 0 || d.progress_rgba || !d.scene) return false;
    const uint64_t npix = (uint64_t)d.scene->camera.resolution_x * d.scene->camera.resolution_y;
    return npix * d.samples >= min_paths && npix <= EngineOptions().l_capacity;
  }

  // Queues a job, starting the lane's thread for this batch call if needed.
  // The job stays owned by the caller until finish() or stop() returns.
  bool push(SJob *j, HbmPool *pool, DeviceCache *cache, const EngineOptions &group_opt,
            std::string *err) {
    if (!th_.joinable()) {
      if (!inited_) {
        if (!eng.init(err)) return false;
        inited_ = true;
      }
      // launch-shape and traversal options follow the batch; the queue and
      // the radiance arena keep the per-scene renderer's sizes
      const EngineOptions d;
      eng.opt = group_opt;
      eng.opt.queue = d.queue;
      eng.opt.l_capacity = d.l_capacity;
      eng.opt.lds_bvh = d.lds_bvh;
      eng.opt.pipes = 0;
      eng.cache = cache;
      pool_ = pool;
      closed_ = false;
      stop_ = false;
      failed_ = false;
      err_.clear();
      th_ = std::thread(&SoloLane::run, this);
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      q_.push_back(j);
    }
    cv_.notify_one();
    return true;
  }

  bool failed() const { return failed_.load(); }

  // Waits until every queued job is rendered.  False (with the message) if a
  // job failed; jobs queued behind a failed one are not rendered.
  bool finish(std::string *err) {
    join();
    if (failed_) *err = err_;
    return !failed_;
  }

  // Drops the jobs not yet started and waits for the current one.
  void stop() {
    {
      std::lock_guard<std::mutex> g(mu_);
      q_.clear();
      stop_ = true;
    }
    join();
  }

 private:
  std::thread th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<SJob *> q_;
  bool closed_ = false;
  std::atomic<bool> stop_{false}, failed_{false};
  std::string err_;
  HbmPool *pool_ = nullptr;
  DevJob *d_job_ = nullptr;
  uint64_t frame_serial_ = 0;  // one per solo job (Engine::hit_frame)
  bool inited_ = false;
  uint64_t accounted_ = 0;  // engine bytes already counted in the pool

  void join() {
    {
      std::lock_guard<std::mutex> g(mu_);
      closed_ = true;
    }
    cv_.notify_all();
    if (th_.joinable()) th_.join();
  }

  void run() {
    if (!eng.check(hipSetDevice(eng.device), "hipSetDevice", &err_)) failed_ = true;
    for (;;) {
      SJob *j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !q_.empty() || closed_; });
        if (q_.empty()) return;
        j = q_.front();
        q_.pop_front();
      }
      if (failed_ || stop_) continue;
      if (!render(*j)) failed_ = true;
    }
  }

  bool render(SJob &j) {
    // admission under the group's budget, as for the streamers' jobs
    while (!pool_->reserve(j.bytes)) {
      if (stop_) return true;
      std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    const uint64_t npix = (uint64_t)j.W * j.H;
    const uint64_t fit = std::min<uint64_t>(eng.opt.l_capacity / npix, 0xffffffffull / npix);
    std::string ignored;
    bool ok = fit > 0;
    if (!ok) err_ = "image larger than the radiance arena";
    j.t_admit = Clock::now();
    ok = ok && eng.upload_scene(j.built, &j.dev, &err_) &&
         eng.check(eng.dmalloc((void **)&j.acc, 16 * npix), "hipMalloc(accumulator)", &err_) &&
         eng.check(eng.dmalloc((void **)&j.rgba_d, 4 * npix), "hipMalloc(image)", &err_) &&
         eng.check(launch_zero_acc(j.acc, 4 * npix, eng.stream), "zero accumulator", &err_);
    if (ok && !d_job_) ok = eng.check(hipMalloc(&d_job_, sizeof(DevJob)), "hipMalloc(job)", &err_);
    DevJob h{};
    if (ok) {
      fill_job(j.scene, j.dev, j.acc, j.bounces, j.seed, &h);
      ok = eng.check(hipMemcpyAsync(d_job_, &h, sizeof(DevJob), hipMemcpyHostToDevice, eng.stream),
                     "upload job", &err_);
    }
    const std::vector<DevJob> host{h};
    ++frame_serial_;  // the job's passes share its camera-ray hits
    for (uint32_t done = 0; ok && done < j.spp;) {
      // passes as large as the radiance arena allows, in sample order
      DevTicket t{};
      t.job = 0;
      t.sample_base = done;
      t.spp = (uint32_t)std::min<uint64_t>(j.spp - done, fit);
      t.paths = (uint32_t)(t.spp * npix);
      eng.hit_frame = frame_serial_;
      ok = eng.run_tickets(d_job_, {t}, host, &err_);
      done += t.spp;
    }
    if (ok) {
      const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((npix + 255) / 256, 8192));
      ok = eng.check(launch_tonemap(d_job_, (uint32_t)npix, (float)j.spp, j.rgba_d, eng.counters,
                                    false, grid, eng.stream),
                     "tonemap", &err_) &&
           (!j.out_rgba || eng.check(hipMemcpyAsync(j.out_rgba, j.rgba_d, 4 * npix,
                                                    hipMemcpyDeviceToHost, eng.stream),
                                     "read image", &err_)) &&
           (!j.out_acc || eng.check(hipMemcpyAsync(j.out_acc, j.acc, 16 * npix,
                                                   hipMemcpyDeviceToHost, eng.stream),
                                    "read accumulator", &err_));
    }
    // nothing may be in flight when the job's buffers are freed
    const bool synced = eng.check(hipStreamSynchronize(eng.stream), "render", ok ? &err_ : &ignored);
    ok = ok && synced;
    eng.free_scene(&j.dev);
    eng.dfree(j.acc);
    eng.dfree(j.rgba_d);
    j.acc = nullptr;
    j.rgba_d = nullptr;
    j.built = BuiltScene();
    pool_->release(j.bytes);
    // the engine keeps its buffers between jobs: count them once
    const uint64_t held = eng.device_bytes();
    if (held > accounted_) {
      pool_->add_fixed((int64_t)(held - accounted_));
      accounted_ = held;
    }
    if (!ok) {
      j.state = 3;
      return false;
    }
    j.state = 2;
    j.t_done = Clock::now();
    ++jobs_done;
    paths_done += npix * j.spp;
    return true;
  }
};

}  // namespace rdh

// ------------------------------------------------------------- batch ABI ---
struct rd_batch {
  rdh::StreamerGroup st;
  rdh::SoloLane solo;  // declared after st: its thread stops before the group goes
  std::string last_error;
  bool failed = false;  // a device error left work behind: the engines are unusable
  rd_batch_stats stats{};
  uint64_t next_id = 1;
};

namespace {

rd_status batch_fail(rd_batch *b, rd_status s, const std::string &m) {
  b->last_error = m;
  return s;
}

// ---- autotune results kept per device and mix (rd_batch_autotune) --------
std::mutex g_tune_mu;
std::map<std::string, rd_tune_result> g_tune;  // guarded by g_tune_mu
bool g_tune_file_read = false;                 // guarded by g_tune_mu

const char *tune_file() {
  const char *f = std::getenv("RAYDAR_HIP_TUNE_CACHE");
  return f && *f ? f : nullptr;
}

// Device identity + batch shape + per-job work shape, as one line of text
// (no tabs or newlines: it is the first field of a cache-file line).
bool tune_key(const rd_batch *b, const rd_job_desc *jobs, uint32_t n, std::string *key) {
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  std::string k;
  for (const char *c = p.name; *c && k.size() < 64; ++c) k += (*c == '\t' || *c == '\n') ? ' ' : *c;
  k += "|" + std::string(p.gcnArchName) + "|cu" + std::to_string(p.multiProcessorCount) + "|mem" +
       std::to_string((unsigned long long)p.totalGlobalMem) + "|pci" + std::to_string(p.pciDomainID) +
       ":" + std::to_string(p.pciBusID) + ":" + std::to_string(p.pciDeviceID) + "|pipes" +
       std::to_string(b->st.m.size()) + "|solo" + std::to_string((unsigned long long)b->solo.min_paths) +
       "|graph" + std::to_string(b->st.first().use_graph) + "|jobs" + std::to_string(n);
  // FNV-1a over each job's work shape
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    for (int i = 0; i < 8; ++i) {
      h ^= (v >> (8 * i)) & 0xffu;
      h *= 1099511628211ull;
    }
  };
  for (uint32_t i = 0; i < n; ++i) {
    const rd_scene &sc = *jobs[i].scene;
    mix(sc.object_count);
    mix(((uint64_t)sc.camera.resolution_x << 32) | sc.camera.resolution_y);
    mix(((uint64_t)jobs[i].samples << 32) | jobs[i].max_bounces);
    mix(jobs[i].pass_samples);
  }
  char hex[17];
  std::snprintf(hex, sizeof hex, "%016llx", (unsigned long long)h);
  *key = k + "|mix" + hex;
  return true;
}

// Reads RAYDAR_HIP_TUNE_CACHE once per process: lines of
// key TAB block TAB leaf TAB queue TAB best_ms TAB default_ms TAB configs;
// malformed lines are skipped.  Caller holds g_tune_mu.
void tune_read_file() {
  if (g_tune_file_read) return;
  g_tune_file_read = true;
  const char *f = tune_file();
  if (!f) return;
  FILE *fp = std::fopen(f, "r");
  if (!fp) return;
  char line[1024];
  while (std::fgets(line, sizeof line, fp)) {
    char *tab = std::strchr(line, '\t');
    if (!tab) continue;
    *tab = 0;
    rd_tune_result r{};
    unsigned q = 0, cfg = 0;
    if (std::sscanf(tab + 1, "%d\t%d\t%u\t%lf\t%lf\t%u", &r.block, &r.leaf, &q, &r.best_ms,
                    &r.default_ms, &cfg) != 6)
      continue;
    if ((r.block != 64 && r.block != 128 && r.block != 256) || r.leaf < 1 || r.leaf > 8 ||
        q < 1024 || q > (1u << 28))
      continue;
    r.queue = q;
    r.configs_tried = cfg;
    g_tune[line] = r;
  }
  std::fclose(fp);
}

bool tune_lookup(const std::string &key, rd_tune_result *r) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_read_file();
  auto it = g_tune.find(key);
  if (it == g_tune.end()) return false;
  *r = it->second;
  return true;
}

void tune_store(const std::string &key, const rd_tune_result &r) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_read_file();
  g_tune[key] = r;
  if (const char *f = tune_file()) {
    if (FILE *fp = std::fopen(f, "a")) {
      std::fprintf(fp, "%s\t%d\t%d\t%u\t%.6f\t%.6f\t%u\n", key.c_str(), r.block, r.leaf, r.queue,
                   r.best_ms, r.default_ms, r.configs_tried);
      std::fclose(fp);
    }
  }
}

void apply_tuning(rd_batch *b, const rd_tune_result &r) {
  b->st.each([&](rdh::Streamer &s) {
    s.eng.opt.block = r.block;
    s.eng.opt.leaf = r.leaf;
    if (!s.eng.q_alloc) s.eng.opt.queue = r.queue;
  });
}

bool engine_option(rdh::Streamer &st, const std::string &k, int64_t v, std::string *err) {
  rdh::EngineOptions &o = st.eng.opt;
  if (k == "block") {
    if (v != 64 && v != 128 && v != 256) { *err = "block must be 64, 128 or 256"; return false; }
    o.block = (int)v;
  } else if (k == "queue") {
    if (v < 1024 || v > (1ll << 28)) { *err = "queue out of range"; return false; }
    if (st.eng.q_alloc && (uint32_t)v != st.eng.q_alloc) { *err = "queue is fixed once rendering started"; return false; }
    o.queue = (uint32_t)v;
  } else if (k == "leaf") {
    if (v < 1 || v > rdh::kMaxLeafSize) { *err = "leaf must be in 1..8"; return false; }
    o.leaf = (int)v;
  } else if (k == "grid") {
    if (v < 0 || v > (1 << 20)) { *err = "grid out of range"; return false; }
    o.grid = (int)v;
  } else if (k == "graph") {
    st.use_graph = v != 0;
  } else if (k == "ticket_cap") {
    if (v < 1 || v > (1 << 24)) { *err = "ticket_cap must be in 1..2^24"; return false; }
    if (st.ready() && (uint32_t)v != st.ticket_cap) { *err = "ticket_cap is fixed once rendering started"; return false; }
    st.ticket_cap = (uint32_t)v;
  } else if (k == "iters_per_pump") {
    if (v < 2 || v > 1024) { *err = "iters_per_pump out of range"; return false; }
    st.iters_per_pump = (int)v;
  } else if (k == "l_capacity") {
    if (v < (1 << 16) || v > (1ll << 30)) { *err = "l_capacity out of range"; return false; }
    if (st.eng.l_alloc && (uint64_t)v != st.eng.l_alloc) { *err = "l_capacity is fixed once rendering started"; return false; }
    o.l_capacity = (uint64_t)v;
  } else if (k == "traversal") {
    o.brute_force = v != 0;
  } else if (k == "refill") {
    if (v < 0 || v > 64) { *err = "refill must be in 0..64"; return false; }
    o.refill = (int)v;
  } else if (k == "lane_chunk") {
    if (v < 0 || (v > 0 && v < 64) || v > 65536 || (v & (v - 1)) != 0) {
      *err = "lane_chunk must be 0 or a power of two in 64..65536";
      return false;
    }
    o.lane_chunk_log2 = v ? (uint32_t)__builtin_ctzll((unsigned long long)v) : 0u;
  } else if (k == "shade_blocks") {
    if (v < 0 || v > 64) { *err = "shade_blocks must be in 0..64"; return false; }
    o.shade_blocks = (int)v;
  } else if (k == "bvh4") {
    if (v != 0 && v != 1) { *err = "bvh4 must be 0 or 1"; return false; }
    o.bvh4 = (int)v;
  } else if (k == "count") {
    o.count = v != 0;
  } else if (k == "kernel_timing") {
    o.kernel_timing = v != 0;
  } else {
    *err = "unknown option " + k;
    return false;
  }
  return true;
}

// Options of a group: "pipes" (members) while nothing has rendered, every
// other option on every member.
bool group_option(rdh::StreamerGroup &g, const std::string &k, int64_t v, std::string *err) {
  if (k == "pipes") {
    if (v < 1 || v > 4) { *err = "pipes must be in 1..4"; return false; }
    if ((int)v != (int)g.m.size() && !g.unused()) { *err = "pipes is fixed once rendering started"; return false; }
    if (!g.resize((int)v)) { *err = g.err; return false; }
    return true;
  }
  for (auto &s : g.m)
    if (!engine_option(*s, k, v, err)) return false;
  return true;
}

bool engine_option_get(const rdh::Streamer &st, const std::string &k, int64_t *v) {
  const rdh::EngineOptions &o = st.eng.opt;
  if (k == "block") *v = o.block;
  else if (k == "queue") *v = o.queue;
  else if (k == "leaf") *v = o.leaf;
  else if (k == "grid") *v = o.grid;
  else if (k == "refill") *v = o.refill;
  else if (k == "lane_chunk") *v = o.lane_chunk_log2 ? (1ll << o.lane_chunk_log2) : 0;
  else if (k == "shade_blocks") *v = o.shade_blocks;
  else if (k == "bvh4") *v = o.bvh4;
  else if (k == "graph") *v = st.use_graph;
  else if (k == "iters_per_pump") *v = st.iters_per_pump;
  else if (k == "ticket_cap") *v = st.ticket_cap;
  else if (k == "rewinds") *v = (int64_t)st.rewinds;
  else if (k == "l_capacity") *v = (int64_t)o.l_capacity;
  else if (k == "num_cus") *v = st.eng.num_cus;
  else return false;
  return true;
}

// Cheap checks of every job before any work starts, so a bad job fails the
// call before anything is rendered.
bool validate_job(const rd_job_desc &d, std::string *e, rd_status *st) {
  *st = RD_ERR_INVALID_ARGUMENT;
  if (!d.scene) { *e = "job scene is NULL"; return false; }
  const rd_scene &s = *d.scene;
  if (s.camera.resolution_x == 0 || s.camera.resolution_y == 0) { *e = "zero-size image"; return false; }
  if (s.world.kind == RD_WORLD_TRANSPARENT) {
    *st = RD_ERR_UNSUPPORTED;
    *e = "transparent world is not supported (raydar: todo!(\"transparent world support\"))";
    return false;
  }
  if (s.world.kind > RD_WORLD_TRANSPARENT) { *e = "unknown world kind"; return false; }
  if (d.samples == 0 || d.max_bounces == 0 || d.max_bounces > 255) {
    *e = "samples must be >= 1 and max_bounces in 1..255";
    return false;
  }
  if (s.object_count && !s.objects) { *e = "objects is NULL"; return false; }
  for (uint32_t i = 0; i < s.object_count; ++i) {
    const rd_object &o = s.objects[i];
    if ((o.kind != RD_SPHERE && o.kind != RD_CUBE) || !std::isfinite(o.center[0]) ||
        !std::isfinite(o.center[1]) || !std::isfinite(o.center[2]) || !std::isfinite(o.size)) {
      *e = "object " + std::to_string(i) + ": unknown kind or non-finite geometry";
      return false;
    }
  }
  return true;
}

rd_status run_batch(rd_batch *b, const rd_job_desc *jobs, uint32_t n) {
  rdh::StreamerGroup &st = b->st;
  if (b->failed) return batch_fail(b, RD_ERR_DEVICE, "the batch renderer failed earlier: " + b->last_error);
  for (uint32_t i = 0; i < n; ++i) {
    std::string e;
    rd_status s;
    if (!validate_job(jobs[i], &e, &s)) return batch_fail(b, s, "job " + std::to_string(i) + ": " + e);
  }
  const auto t0 = std::chrono::steady_clock::now();
  const uint64_t it0 = st.iterations(), g0 = st.graph_launches(), p0 = st.pumps, paths0 = st.paths_done();
  const uint64_t solo_it0 = b->solo.eng.iterations, solo_paths0 = b->solo.paths_done;
  // Host-side scene digestion (BVH builds) runs on a thread pool and
  // overlaps the GPU: jobs join the work list in submission order as soon as
  // they are built.
  std::vector<std::unique_ptr<rdh::SJob>> built(n);
  std::vector<std::string> errs(n);
  std::unique_ptr<std::atomic<int>[]> ready(new std::atomic<int>[n]);
  for (uint32_t i = 0; i < n; ++i) ready[i] = 0;
  std::atomic<uint32_t> next{0};
  std::atomic<int64_t> build_ns{0};
  const int leaf = st.first().eng.opt.leaf;
  auto worker = [&] {
    for (uint32_t i; (i = next.fetch_add(1)) < n;) {
      const auto b0 = std::chrono::steady_clock::now();
      built[i] = st.make_job(jobs[i], leaf, &errs[i]);
      build_ns += rdh::ns_between(b0, std::chrono::steady_clock::now());
      ready[i].store(built[i] ? 1 : 2, std::memory_order_release);
    }
  };
  const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < std::min(hw, n); ++t) pool.emplace_back(worker);
  std::vector<uint64_t> ids;
  std::vector<std::unique_ptr<rdh::SJob>> solo_jobs;  // owned here until the lane has finished
  uint32_t enq = 0;
  rd_status status = RD_OK;
  while (status == RD_OK) {
    while (enq < n && ready[enq].load(std::memory_order_acquire) != 0) {
      if (ready[enq] == 2) {
        status = batch_fail(b, RD_ERR_INVALID_ARGUMENT, "job " + std::to_string(enq) + ": " + errs[enq]);
        break;
      }
      built[enq]->id = b->next_id++;
      if (b->solo.wants(jobs[enq])) {
        solo_jobs.push_back(std::move(built[enq]));
        std::string e;
        if (!b->solo.push(solo_jobs.back().get(), &st.pool(), &st.cache, st.first().eng.opt, &e)) {
          status = batch_fail(b, RD_ERR_DEVICE, e);
          break;
        }
      } else {
        ids.push_back(built[enq]->id);
        st.enqueue(std::move(built[enq]));
      }
      ++enq;
    }
    if (status != RD_OK) break;
    if (b->solo.failed()) {
      status = RD_ERR_DEVICE;  // the lane's message is taken when it has stopped
      break;
    }
    if (st.idle()) {
      if (enq == n) break;
      std::this_thread::yield();  // waiting for the next scene to be built
      continue;
    }
    if (!st.pump())
      status = batch_fail(b, st.err.find("out of memory") != std::string::npos ? RD_ERR_OUT_OF_MEMORY : RD_ERR_DEVICE, st.err);
  }
  next = n;  // stop workers that have not started a job
  for (auto &t : pool) t.join();
  // the lane's remaining jobs: all of them on success, none after a failure
  bool lane_failed = false;
  if (status == RD_OK) {
    std::string e;
    if (!b->solo.finish(&e)) {
      status = batch_fail(b, RD_ERR_DEVICE, "solo lane: " + e);
      lane_failed = true;
    }
  } else {
    lane_failed = b->solo.failed();
    b->solo.stop();
    if (lane_failed) {
      std::string e;
      (void)b->solo.finish(&e);
      b->last_error = "solo lane: " + e;
    }
  }
  if (lane_failed) b->failed = true;  // the lane's engine may hold unfinished work
  // after a failure, drain what was already submitted so no job is left half-rendered
  while (status != RD_OK && !st.idle() && st.pump()) {
  }
  if (status != RD_OK) {
    // nothing may still be writing the caller's buffers when this returns
    st.each([](rdh::Streamer &s) { (void)hipStreamSynchronize(s.eng.stream); });
    if (!st.idle()) {
      // the drain itself failed: the jobs stay owned (destroy releases them)
      b->failed = true;
      return status;
    }
  }
  for (uint64_t id : ids) st.forget(id);
  if (status != RD_OK) return status;
  b->stats.build_ns = build_ns.load();
  b->stats.host_threads = std::min(hw, n);
  b->stats.elapsed_ns = rdh::ns_between(t0, std::chrono::steady_clock::now());
  b->stats.iterations = st.iterations() - it0 + (b->solo.eng.iterations - solo_it0);
  b->stats.graph_launches = st.graph_launches() - g0;
  b->stats.pumps = st.pumps - p0;
  b->stats.paths = st.paths_done() - paths0 + (b->solo.paths_done - solo_paths0);
  return RD_OK;
}

}  // namespace

extern "C" {

rd_status rd_batch_create(rd_batch **out) {
  if (!out) return RD_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto *b = new (std::nothrow) rd_batch();
  if (!b) return RD_ERR_OUT_OF_MEMORY;
  if (!b->st.init(3)) {  // three pipes: measured best on mixed batches (2: -7%, 4: no gain)
    rdh::set_global_error(b->st.err);
    delete b;
    return RD_ERR_DEVICE;
  }
  // throughput over latency: fewer host round trips
  b->st.each([](rdh::Streamer &s) { s.iters_per_pump = 4; });
  *out = b;
  return RD_OK;
}

void rd_batch_destroy(rd_batch *b) {
  if (!b) return;
  b->st.each([](rdh::Streamer &s) { (void)hipStreamSynchronize(s.eng.stream); });
  delete b;
}

rd_status rd_batch_set_option(rd_batch *b, const char *key, int64_t value) {
  if (!b || !key) return RD_ERR_INVALID_ARGUMENT;
  std::string e;
  if (std::string(key) == "solo_paths") {
    if (value < 0) return batch_fail(b, RD_ERR_INVALID_ARGUMENT, "solo_paths must be >= 0");
    b->solo.min_paths = (uint64_t)value;
    return RD_OK;
  }
  if (!group_option(b->st, key, value, &e)) return batch_fail(b, RD_ERR_INVALID_ARGUMENT, e);
  return RD_OK;
}

rd_status rd_batch_get_option(const rd_batch *b, const char *key, int64_t *value) {
  if (!b || !key || !value) return RD_ERR_INVALID_ARGUMENT;
  const std::string k = key;
  if (k == "pipes") *value = (int64_t)b->st.m.size();
  else if (k == "solo_paths") *value = (int64_t)b->solo.min_paths;
  else if (k == "solo_jobs") *value = (int64_t)b->solo.jobs_done;
  else return engine_option_get(b->st.first(), key, value) ? RD_OK : RD_ERR_INVALID_ARGUMENT;
  return RD_OK;
}

rd_status rd_batch_render(rd_batch *b, const rd_job_desc *jobs, uint32_t n) {
  if (!b || (n && !jobs)) return RD_ERR_INVALID_ARGUMENT;
  return run_batch(b, jobs, n);
}

rd_status rd_batch_stats_get(const rd_batch *b, rd_batch_stats *out) {
  if (!b || !out) return RD_ERR_INVALID_ARGUMENT;
  *out = b->stats;
  return RD_OK;
}

const char *rd_batch_last_error(const rd_batch *b) { return b ? b->last_error.c_str() : ""; }

rd_status rd_batch_autotune(rd_batch *b, const rd_job_desc *jobs, uint32_t n, rd_tune_result *out) {
  if (!b || !jobs || !n || !out) return RD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n; ++i)
    if (!jobs[i].scene) return batch_fail(b, RD_ERR_INVALID_ARGUMENT, "job without a scene");
  std::string key;
  if (!tune_key(b, jobs, n, &key)) return batch_fail(b, RD_ERR_DEVICE, "hipGetDeviceProperties failed");
  rd_tune_result kept{};
  if (tune_lookup(key, &kept)) {
    apply_tuning(b, kept);
    kept.cached = 1;
    *out = kept;
    return RD_OK;
  }
  // The queue capacity and radiance arena are fixed once allocated, so the
  // sweep runs on fresh batches and the winner is applied to `b`.
  const int blocks[] = {64, 128, 256};
  const uint32_t queues[] = {1u << 20, 1u << 22, 1u << 23, 1u << 24};
  const int leaves[] = {1, 2, 4, 8};
  double best = 1e300, dflt = 0.0;
  rd_tune_result r{};
  const rdh::EngineOptions base = b->st.first().eng.opt;
  r.block = base.block;
  r.leaf = base.leaf;
  r.queue = base.queue;
  for (uint32_t q : queues) {
    rd_batch *t = nullptr;
    rd_status s = rd_batch_create(&t);
    if (s != RD_OK) return batch_fail(b, s, rd_last_error(nullptr));
    std::unique_ptr<rd_batch, void (*)(rd_batch *)> guard(t, rd_batch_destroy);
    if (!t->st.resize((int)b->st.m.size())) return batch_fail(b, RD_ERR_DEVICE, t->st.err);
    t->solo.min_paths = b->solo.min_paths;
    const int use_graph = b->st.first().use_graph;
    t->st.each([&](rdh::Streamer &s) {
      s.eng.opt = base;
      s.eng.opt.queue = q;
      s.use_graph = use_graph;
    });
    for (int blk : blocks)
      for (int lf : leaves) {
        t->st.each([&](rdh::Streamer &s) {
          s.eng.opt.block = blk;
          s.eng.opt.leaf = lf;
        });
        // one untimed run per launch shape compiles / warms the kernels
        if ((s = run_batch(t, jobs, n)) != RD_OK) return batch_fail(b, s, t->last_error);
        if ((s = run_batch(t, jobs, n)) != RD_OK) return batch_fail(b, s, t->last_error);
        const double ms = t->stats.elapsed_ns / 1e6;
        r.configs_tried++;
        const rdh::EngineOptions d;  // the library's defaults, for comparison
        if (blk == d.block && lf == d.leaf && q == d.queue) dflt = ms;
        if (ms < best) {
          best = ms;
          r.block = blk;
          r.leaf = lf;
          r.queue = q;
        }
      }
  }
  r.best_ms = best;
  r.default_ms = dflt;
  apply_tuning(b, r);
  tune_store(key, r);
  *out = r;
  return RD_OK;
}

}  // extern "C"

// ------------------------------------------------------------ server ABI ---
struct rd_server {
  rdh::StreamerGroup st;
  std::string last_error;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::unique_ptr<rdh::SJob>> inbox;
  std::map<uint64_t, rd_job_result> results;
  uint64_t next_id = 1;
  bool stop = false, failed = false;
  std::thread worker;
  rd_server_stats stats{};
  std::chrono::steady_clock::time_point busy_since;
  bool busy = false;
};

namespace {

void server_fill_result(rdh::SJob *j, rd_job_result *r) {
  r->status = j->status;
  r->done = j->state == 2 ? 1 : 0;
  r->passes_done = j->passes_done;
  r->samples_done = std::min(j->spp, j->passes_done * j->pass_spp);
  r->device_bytes = j->bytes;
  r->queued_ns = j->state >= 1 ? rdh::ns_between(j->t_submit, j->t_admit) : -1;
  r->latency_ns = j->state == 2 ? rdh::ns_between(j->t_submit, j->t_done) : -1;
}

void server_loop(rd_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    while (!s->inbox.empty()) {
      auto j = std::move(s->inbox.front());
      s->inbox.pop_front();
      s->st.enqueue(std::move(j));
    }
    if (s->st.idle()) {
      if (s->busy) {
        s->stats.busy_ns += rdh::ns_between(s->busy_since, std::chrono::steady_clock::now());
        s->busy = false;
      }
      if (s->stop) break;
      s->cv_work.wait(lk, [s] { return s->stop || !s->inbox.empty(); });
      continue;
    }
    if (!s->busy) {
      s->busy = true;
      s->busy_since = std::chrono::steady_clock::now();
    }
    lk.unlock();
    const bool ok = s->st.pump();
    lk.lock();
    // publish progress of every known job
    for (auto &kv : s->results) {
      rdh::SJob *j = s->st.job(kv.first);
      if (!j) continue;
      server_fill_result(j, &kv.second);
      if (j->state == 2) {
        s->stats.jobs_done++;
        s->st.forget(kv.first);
      }
    }
    s->stats.paths = s->st.paths_done();
    s->stats.iterations = s->st.iterations();
    s->stats.hbm_in_use = s->st.hbm_in_use();
    s->stats.hbm_peak = s->st.hbm_peak();
    s->stats.engine_bytes = s->st.engine_bytes();
    s->stats.max_resident = s->st.max_resident();
    s->stats.jobs_resident = s->st.n_resident();
    s->stats.jobs_queued = s->st.n_queued() + (uint32_t)s->inbox.size();
    s->cv_done.notify_all();
    if (!ok) {
      // nothing may still be writing the callers' buffers once they see the error
      s->st.each([](rdh::Streamer &m) { (void)hipStreamSynchronize(m.eng.stream); });
      s->failed = true;
      s->last_error = s->st.err;
      for (auto &kv : s->results)
        if (!kv.second.done) {
          kv.second.status = RD_ERR_DEVICE;
          kv.second.done = 1;
        }
      s->cv_done.notify_all();
      break;
    }
  }
}

}  // namespace

extern "C" {

rd_status rd_server_create(const rd_server_config *cfg, rd_server **out) {
  if (!out) return RD_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto *s = new (std::nothrow) rd_server();
  if (!s) return RD_ERR_OUT_OF_MEMORY;
  if (!s->st.init(2)) {  // two pipes
    rdh::set_global_error(s->st.err);
    delete s;
    return RD_ERR_DEVICE;
  }
  s->st.each([](rdh::Streamer &m) { m.default_pass = 8; });
  if (cfg) {
    if (cfg->hbm_budget_bytes) s->st.set_budget(cfg->hbm_budget_bytes);
    if (cfg->pass_samples) s->st.each([&](rdh::Streamer &m) { m.default_pass = cfg->pass_samples; });
    std::string e;
    if ((cfg->block && !group_option(s->st, "block", cfg->block, &e)) ||
        (cfg->queue && !group_option(s->st, "queue", cfg->queue, &e)) ||
        (cfg->leaf && !group_option(s->st, "leaf", cfg->leaf, &e))) {
      rdh::set_global_error(e);
      delete s;
      return RD_ERR_INVALID_ARGUMENT;
    }
  }
  if (!s->st.ensure_ready()) {
    rdh::set_global_error(s->st.err);
    delete s;
    return RD_ERR_DEVICE;
  }
  if (s->st.engine_bytes() > s->st.budget()) {
    rdh::set_global_error("HBM budget smaller than the engine's own buffers");
    delete s;
    return RD_ERR_BUDGET;
  }
  s->stats.engine_bytes = s->st.engine_bytes();
  s->worker = std::thread(server_loop, s);
  *out = s;
  return RD_OK;
}

void rd_server_destroy(rd_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  delete s;
}

rd_status rd_server_submit(rd_server *s, const rd_job_desc *job, uint64_t *job_id) {
  if (!s || !job || !job_id) return RD_ERR_INVALID_ARGUMENT;
  std::string e;
  int leaf;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->failed) return RD_ERR_DEVICE;
    if (s->stop) return RD_ERR_INVALID_ARGUMENT;
    leaf = s->st.first().eng.opt.leaf;
  }
  auto j = s->st.make_job(*job, leaf, &e);  // BVH build on the caller's thread
  std::lock_guard<std::mutex> lk(s->mu);
  if (!j) {
    s->last_error = e;
    s->stats.jobs_rejected++;
    return e.find("transparent") != std::string::npos ? RD_ERR_UNSUPPORTED : RD_ERR_INVALID_ARGUMENT;
  }
  if (!s->st.fits_alone(*j)) {
    s->last_error = "job needs " + std::to_string(j->bytes) + " bytes of HBM, above the budget";
    s->stats.jobs_rejected++;
    return RD_ERR_BUDGET;
  }
  j->id = s->next_id++;
  *job_id = j->id;
  rd_job_result r{};
  r.status = RD_OK;
  r.queued_ns = -1;
  r.latency_ns = -1;
  r.device_bytes = j->bytes;
  s->results[j->id] = r;
  s->inbox.push_back(std::move(j));
  s->stats.jobs_submitted++;
  s->cv_work.notify_all();
  return RD_OK;
}

rd_status rd_server_poll(rd_server *s, uint64_t id, rd_job_result *out) {
  if (!s || !out) return RD_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->results.find(id);
  if (it == s->results.end()) return RD_ERR_NOT_FOUND;
  *out = it->second;
  return RD_OK;
}

rd_status rd_server_wait(rd_server *s, uint64_t id, int64_t timeout_ms, rd_job_result *out) {
  if (!s || !out) return RD_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->results.find(id);
  if (it == s->results.end()) return RD_ERR_NOT_FOUND;
  auto pred = [&] { return it->second.done != 0; };
  if (timeout_ms < 0) s->cv_done.wait(lk, pred);
  else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
  *out = it->second;
  return RD_OK;
}

rd_status rd_server_stats_get(rd_server *s, rd_server_stats *out) {
  if (!s || !out) return RD_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  if (s->busy) out->busy_ns += rdh::ns_between(s->busy_since, std::chrono::steady_clock::now());
  return RD_OK;
}

const char *rd_server_last_error(const rd_server *s) { return s ? s->last_error.c_str() : ""; }

}  // extern "C"
