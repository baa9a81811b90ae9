// Wavefront engine: owns the device path queues, the radiance arena, the
// work-ticket list and the library's HIP stream, and drives iterations of
// generate -> extend -> shade until every ticket's paths have finished.
//
// The same engine serves the per-scene renderer (one job, one ticket per
// pass), the batched renderer (many jobs, tickets for all of them in one
// work list) and the render server (tickets appended as requests arrive).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "bvh.hpp"
#include "kernels.hpp"
#include "layout.hpp"

namespace rdh {

// Device copy of a BuiltScene.
struct DeviceScene {
  DevNode *nodes = nullptr;
  DevNode4 *nodes4 = nullptr;
  uint32_t n_nodes4 = 0, stack4 = 0;
  uint32_t max_leaf = 0;
  DevPrim *prims = nullptr;
  DevObject *objects = nullptr;
  uint32_t n_nodes = 0, n_prims = 0;
  float r0 = 0.f;
  uint32_t depth = 0;
  size_t bytes = 0;
};

// Device buffers given back by finished jobs, kept for later jobs.  hipFree
// waits for the work queued on every stream of the device, so freeing a
// job's buffers while other pipes have kernels queued would stall the host
// thread and let those queues run dry.  Buffers are rounded up to size
// classes (each power of two split in eight steps) and reused by class; they
// are freed for good by trim(), when more than `keep_max` bytes are kept, when
// a new allocation runs out of device memory (then it is retried), and on
// destruction.  Thread-safe.
class DeviceCache {
 public:
  ~DeviceCache() { trim(); }
  hipError_t alloc(void **p, size_t bytes);
  void free(void *p);  // p from alloc(), or nullptr
  void trim();
  uint64_t kept_bytes() const;
  uint64_t hits = 0, misses = 0;
  uint64_t keep_max = 4ull << 30;
  static size_t size_class(size_t bytes);

 private:
  mutable std::mutex mu_;
  std::unordered_map<void *, size_t> class_of_;  // every buffer handed out
  std::multimap<size_t, void *> kept_;
  uint64_t kept_bytes_ = 0;
};

struct EngineOptions {
  int block = 256;
  uint32_t queue = 1u << 24;      // paths in flight per pipe (two ping-pong queues)
  uint64_t l_capacity = 1ull << 28;  // radiance slots (16 B each): 128 spp of 1080p per
                                     // pass (fewer pass boundaries, ~1% faster than 2^27)
  int grid = 0;                   // 0: occupancy-derived
  int brute_force = 0;
  int refill = 48;                // extend: persistent lanes refill at >= 48 idle (0: off)
  uint32_t lds_bvh = 14336;       // single scenes with at most this many BVH bytes
                                  // are traversed from LDS (0: never)
  int shade_blocks = 0;           // shade workgroups per CU (0: as many as fit)
  uint32_t lds_objects = 16384;   // single scenes with at most this many bytes of
                                  // objects (64 B each) are shaded from LDS (0: never)
  int bvh4 = 1;                   // single scenes: traverse the 4-wide tree
  int leaf = 1;                   // BVH leaf size; 1 measured fastest for one scene
  int count = 0;
  int kernel_timing = 0;
  int sync_every = 4;             // iterations between completion checks
  int pipes = 0;                  // one-scene passes: independent streams the
                                  // samples are divided over (1..4; 0 = 2 for
                                  // LDS-resident scenes, else 3)
  uint32_t lane_chunk_log2 = 6;   // closest-hit lanes: the queue's chunks of 64 paths
                                  // dealt round-robin to the waves (0: one contiguous
                                  // range per wave; lone extend launches 10-22% slower)
  int hit_reuse = 1;              // 1: camera-ray hits traced once per pixel and reused
  int vk = 0;                     // 1: the Vulkan renderer's semantics (per-scene
                                  // renderer only)
};

// Device buffers an engine borrows instead of allocating (the allocation-free
// entry point carves them from a static device array).
struct ExternalBuffers {
  PathQueue q[2];
  float *hit_t;
  uint32_t *hit_id;
  QueueCtl *ctl;
  float4 *L;
  DevTicket *tickets;
  unsigned long long *ticket_done;
  DevCounters *counters;
  uint32_t queue;       // nominal paths in flight
  uint32_t seg_cap;     // slots per queue segment
  uint64_t l_slots;     // radiance slots
  uint32_t tickets_n;   // ticket table entries
};

class Engine {
 public:
  ~Engine();
  // Selects the device and creates the stream; allocation is lazy.
  bool init(std::string *err);
  // Borrows every device buffer from `b`, the stream and pinned control-block
  // copy from `owner` (an initialised engine): no allocation, no pipes.  The
  // engine frees nothing it borrowed.
  void bind_external(const ExternalBuffers &b, const Engine &owner);
  bool external() const { return external_; }
  bool ensure_buffers(std::string *err);  // (re)allocates for current options
  void release_buffers();

  bool upload_scene(const BuiltScene &b, DeviceScene *out, std::string *err);
  void free_scene(DeviceScene *s);
  // Job buffers: from `cache` when one is set, else hipMalloc / hipFree.
  DeviceCache *cache = nullptr;
  hipError_t dmalloc(void **p, size_t bytes);
  void dfree(void *p);

  // Runs one batch of tickets to completion (all paths finished), then
  // accumulates every ticket into its job's accumulator in ticket order.
  // `jobs` is the device job table (already uploaded).  A single-job pass of
  // two or more samples is divided by samples over `opt.pipes` pipes: this
  // engine and auxiliary engines with their own stream, queues and tickets,
  // writing this engine's radiance arena.  The pipes' kernels run
  // concurrently, so one pipe's closest-hit work (ALU-bound) overlaps
  // another's shading (memory-bound).
  bool run_tickets(const DevJob *d_jobs, std::vector<DevTicket> tickets,
                   const std::vector<DevJob> &host_jobs, std::string *err);

  bool check(hipError_t e, const char *what, std::string *err);

  EngineOptions opt;
  hipStream_t stream = nullptr;
  int device = 0;
  int num_cus = 0;
  DevCounters *counters = nullptr;  // device
  double kernel_ms[5] = {0, 0, 0, 0, 0};
  uint64_t kernel_launches[5] = {0, 0, 0, 0, 0};
  uint64_t iterations = 0;
  int pipes_used = 0;  // pipes of the last traced pass
  // Frame of the next run_tickets pass (0: none): passes of one frame whose
  // scene and camera are unchanged reuse the camera-ray hits the frame's
  // first pass recorded instead of recording them again.
  uint64_t hit_frame = 0;

  // buffers (public so the server can drive iterations itself)
  PathQueue q[2] = {};
  float *hit_t = nullptr;
  uint32_t *hit_id = nullptr;
  QueueCtl *ctl = nullptr;
  QueueCtl *ctl_host = nullptr;  // pinned
  float4 *L = nullptr;
  uint64_t l_alloc = 0;
  uint32_t q_alloc = 0;   // nominal paths in flight (opt.queue when allocated)
  uint32_t seg_cap = 0;   // slots per queue segment (kSegments segments)
  uint64_t generation = 0;  // bumped whenever queue buffers are reallocated
  uint32_t segment_slots() const;
  static uint32_t queued(const QueueCtl &c, int parity);
  DevTicket *d_tickets = nullptr;
  uint32_t ticket_alloc = 0;
  unsigned long long *d_ticket_done = nullptr;

  IterArgs iter_args(const DevJob *d_jobs) const;
  static constexpr uint32_t kNoBound = 0xffffffffu;
  LaunchShape shape(int which, const IterArgs *a = nullptr) const;
  LaunchShape bounded_shape(int which, const IterArgs &a, uint32_t bound) const;
  // One generate/extend/shade round.  `bound`: the host knows every path has
  // been generated and at most `bound` are in flight (kNoBound otherwise).
  bool launch_iteration(const IterArgs &a, int parity, std::string *err, uint32_t bound = kNoBound);
  bool ensure_tickets(uint32_t n, std::string *err);
  // Device bytes held by the queues, hit buffers, radiance arena and ticket
  // tables of this engine and its auxiliary pipes.
  uint64_t device_bytes() const;

 private:
  bool timed(int which, std::string *err, hipError_t (*fn)(const IterArgs &, int,
                                                             const LaunchShape &, bool,
                                                             hipStream_t),
             const IterArgs &a, int parity, uint32_t bound);
  bool external_ = false;  // buffers, stream and control copy are borrowed
  std::vector<hipEvent_t> ev_pool_;
  std::vector<std::unique_ptr<Engine>> aux_;  // pipes 1.. of run_tickets
  hipEvent_t arena_free_ = nullptr;           // this stream's work before a pass
  DevTicket *tickets_pin_ = nullptr;          // pinned source of run_tickets' uploads
  uint32_t tickets_pin_cap_ = 0;
  HitWord *first_hit_ = nullptr;              // per-pixel camera-ray hits (IterArgs::first_hit)
  uint64_t first_hit_px_ = 0;                 // pixels first_hit_ holds
  uint64_t first_hit_frame_ = 0;              // hit_frame whose hits first_hit_ holds (0: none)
};

// Message returned by rd_last_error(NULL) (calls with no renderer handle).
void set_global_error(const std::string &msg);

// Fills the camera / world / render fields of a DevJob from a scene.
void fill_job(const rd_scene &s, const DeviceScene &ds, float *acc, uint32_t max_bounces,
              uint64_t seed, DevJob *j);

}  // namespace rdh
