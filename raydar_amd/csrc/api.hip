// C ABI of the per-scene HIP renderer (include/raydar_hip.h) and the parity
// hooks (include/raydar_hip_debug.h).
//
// rd_renderer mirrors CpuRenderer (src/renderer/cpu.rs:110-116): a config,
// a sample count, a profiler, and an accumulator that persists across
// render_sample calls — here a device-resident RGBA32F buffer next to the
// scene's device image (BVH, primitives, materials) snapshotted at new_frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/raydar_hip.h"
#include "../../include/raydar_hip_debug.h"
#include "bvh.hpp"
#include "engine.hpp"
#include "ring.hpp"
#include "kernels.hpp"
#include "pathmath.hpp"

namespace rdh {

namespace {
std::mutex g_err_mu;
std::string g_err;

// timing.rs:66-127 Timer: optional start, optional duration.
struct Timer {
  bool started = false, has_duration = false;
  std::chrono::steady_clock::time_point t0;
  int64_t ns = -1;
  void start() {
    started = true;
    t0 = std::chrono::steady_clock::now();
  }
  void start_if_not_started() {
    if (!started) start();
  }
  void end() {
    if (started) {
      has_duration = true;
      ns = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0)
               .count();
    }
  }
  void end_if_not_ended() {
    if (!has_duration) end();
  }
  // timing.rs:117-121: the average over `count` iterations timed together
  void end_multiple(uint32_t count) {
    if (started && count > 0) {
      end();
      ns /= count;
    }
  }
};
}  // namespace

void set_global_error(const std::string &msg) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = msg;
}

}  // namespace rdh

// ------------------------------------------- static scratch (no-alloc path) --
// Device memory of rd_render_frame_noalloc: one array per GPU, resolved at
// rd_create and held until the process ends.  Host code must never pass
// `scratch` itself to a kernel or a copy (it names a host-side stand-in): it
// uses the device address hipGetSymbolAddress returned for the current GPU.
namespace rdh {
constexpr uint64_t kScratchBytes = 192ull << 20;  // 192 MiB per GPU
constexpr uint32_t kScratchQueue = 1u << 19;      // paths in flight
constexpr int kScratchGrid = 512;                 // workgroups per launch: seg_cap = queue / 32
constexpr int kScratchDevices = 64;
}  // namespace rdh

static __device__ __align__(256) unsigned char scratch[rdh::kScratchBytes];

namespace {
std::mutex g_scratch_mu;                                  // guards g_scratch_addr
void *g_scratch_addr[rdh::kScratchDevices] = {};          // `scratch` on each GPU
std::mutex g_scratch_use[rdh::kScratchDevices];           // one no-alloc call per GPU at a time

// Once per device, with that device current.
bool scratch_setup(int device, std::string *err) {
  if (device < 0 || device >= rdh::kScratchDevices) {
    *err = "device index beyond the static scratch table";
    return false;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_addr[device]) return true;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    (void)hipGetLastError();
    return false;
  }
  g_scratch_addr[device] = p;
  return true;
}

void *scratch_addr(int device) {
  if (device < 0 || device >= rdh::kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return g_scratch_addr[device];
}
}  // namespace

struct rd_renderer {
  rd_config cfg{1024, 12};
  uint32_t sample_count = 0;
  uint32_t samples_per_pass = 1;
  uint64_t seed = 0;
  std::string last_error;
  rdh::DeviceCache cache;  // scene and image buffers reused across frames (outlives eng)
  rdh::Engine eng;
  rdh::Engine noalloc_eng;  // rd_render_frame_noalloc: borrows the static scratch
  rdh::DeviceScene scene;
  rdh::DevJob job_h{};
  rdh::DevJob *job_d = nullptr;
  // pinned copy the job record is uploaded from (a pageable source costs a
  // staging copy on every upload); uploads only when the record changed
  rdh::DevJob *job_pin = nullptr;
  bool job_uploaded = false;
  float *acc = nullptr;
  uchar4 *rgba_d = nullptr;
  uint32_t *vk_rng = nullptr;  // Vulkan semantics: per-pixel PCG state
  uint32_t W = 0, H = 0;
  bool has_frame = false;
  // bumped by new_frame and restart_frame: the passes of one frame share
  // their camera-ray hits (recorded by the frame's first pass)
  uint64_t frame_serial = 0;
  int frame_vk = 0;  // the semantics the current frame was set up with
  rdh::Timer frame_timer, sample_timer, prepare_timer, render_timer;

  rd_status fail(rd_status st, const std::string &msg) {
    last_error = msg;
    return st;
  }
};

namespace {

rd_status device_fail(rd_renderer *r, const std::string &msg) {
  r->last_error = msg;
  return msg.find("out of memory") != std::string::npos ? RD_ERR_OUT_OF_MEMORY : RD_ERR_DEVICE;
}

bool valid_scene(const rd_scene *s, std::string *err) {
  if (!s) { *err = "scene is NULL"; return false; }
  if (s->camera.resolution_x == 0 || s->camera.resolution_y == 0) {
    *err = "zero-size image";
    return false;
  }
  if ((uint64_t)s->camera.resolution_x * s->camera.resolution_y > (1ull << 28)) {
    *err = "image too large (more than 2^28 pixels)";
    return false;
  }
  if (s->world.kind == RD_WORLD_TRANSPARENT) {
    // World::sample -> todo!("transparent world support") (world.rs:31)
    *err = "transparent world is not supported (raydar: todo!(\"transparent world support\"))";
    return false;
  }
  if (s->world.kind > RD_WORLD_TRANSPARENT) { *err = "unknown world kind"; return false; }
  return true;
}

// Snapshot a scene into device memory and (re)allocate the accumulator.
rd_status snapshot(rd_renderer *r, const rd_scene *s, bool zero_acc) {
  ++r->frame_serial;  // a new scene or camera: its camera-ray hits are recorded anew
  std::string err;
  if (!valid_scene(s, &err)) return r->fail(s && s->world.kind == RD_WORLD_TRANSPARENT
                                                ? RD_ERR_UNSUPPORTED
                                                : RD_ERR_INVALID_ARGUMENT,
                                            err);
  rdh::BuiltScene built;
  if (!rdh::build_scene(*s, r->eng.opt.leaf, &built, &err)) return r->fail(RD_ERR_INVALID_ARGUMENT, err);
  if (!rdh::check_bvh(built, &err)) return r->fail(RD_ERR_INVALID_ARGUMENT, "BVH build: " + err);
  if (!r->eng.upload_scene(built, &r->scene, &err)) return device_fail(r, err);
  const uint32_t W = s->camera.resolution_x, H = s->camera.resolution_y;
  if (W != r->W || H != r->H || !r->acc) {
    r->eng.dfree(r->acc);
    r->eng.dfree(r->rgba_d);
    r->eng.dfree(r->vk_rng);
    r->acc = nullptr;
    r->rgba_d = nullptr;
    r->vk_rng = nullptr;
    if (!r->eng.check(r->eng.dmalloc((void **)&r->acc, 16ull * W * H), "hipMalloc(accumulator)", &err) ||
        !r->eng.check(r->eng.dmalloc((void **)&r->rgba_d, 4ull * W * H), "hipMalloc(image)", &err))
      return device_fail(r, err);
    r->W = W;
    r->H = H;
    zero_acc = true;
  }
  if (zero_acc && !r->eng.check(rdh::launch_zero_acc(r->acc, 4ull * W * H, r->eng.stream),
                                "zero accumulator", &err))
    return device_fail(r, err);
  rdh::fill_job(*s, r->scene, r->acc, r->cfg.max_bounces, r->seed, &r->job_h);
  r->frame_vk = r->eng.opt.vk;
  if (r->frame_vk) {
    // the Vulkan renderer's camera (vulkan.rs:251-272) and a PCG state per pixel
    if (rd_camera_vulkan_inverse_view(&s->camera, r->job_h.inv_view) != RD_OK)
      return r->fail(RD_ERR_INVALID_ARGUMENT, "singular Vulkan view matrix");
    if (!r->vk_rng && !r->eng.check(r->eng.dmalloc((void **)&r->vk_rng, 4ull * W * H),
                                    "hipMalloc(PCG states)", &err))
      return device_fail(r, err);
    r->job_h.vk_rng = r->vk_rng;
  }
  if (!r->job_d && !r->eng.check(hipMalloc(&r->job_d, sizeof(rdh::DevJob)), "hipMalloc(job)", &err))
    return device_fail(r, err);
  *r->job_pin = r->job_h;  // no upload from it is in flight: every call ends synchronised
  r->job_uploaded = true;
  if (!r->eng.check(hipMemcpyAsync(r->job_d, r->job_pin, sizeof(rdh::DevJob),
                                   hipMemcpyHostToDevice, r->eng.stream),
                    "upload job", &err) ||
      !r->eng.check(hipStreamSynchronize(r->eng.stream), "upload job", &err))
    return device_fail(r, err);
  r->has_frame = true;
  return RD_OK;
}

// Job fields that can change between passes without a new snapshot.
rd_status refresh_job(rd_renderer *r) {
  std::string err;
  if (r->job_uploaded && r->job_h.max_bounces == r->cfg.max_bounces && r->job_h.seed == r->seed)
    return RD_OK;
  r->job_h.max_bounces = r->cfg.max_bounces;
  r->job_h.seed = r->seed;
  *r->job_pin = r->job_h;
  r->job_uploaded = true;
  if (!r->eng.check(hipMemcpyAsync(r->job_d, r->job_pin, sizeof(rdh::DevJob),
                                   hipMemcpyHostToDevice, r->eng.stream),
                    "upload job", &err))
    return device_fail(r, err);
  return RD_OK;
}

// Renders `k` more samples (cpu.rs:193-219 repeated k times).  The k samples
// run as one pass, so the sample timer reports their average, as the Vulkan
// renderer does for its all-samples dispatch (vulkan.rs:175-177); with k = 1
// that is the last sample's time, as CpuRenderer reports (cpu.rs:196, 218).
//
// Vulkan semantics: a pixel's samples share one PCG stream whose position
// after a sample depends on that sample's draws, so each pass holds one
// sample (all pixels), and the next pass continues every pixel's stream.
rd_status render_samples(rd_renderer *r, uint32_t k) {
  const rdh::EngineOptions &o = r->eng.opt;
  if (o.vk != r->frame_vk)
    return r->fail(RD_ERR_INVALID_ARGUMENT, "\"semantics\" changed after new_frame: start a new frame");
  if (o.vk && (o.block != 256 || o.refill <= 0 || o.count))
    return r->fail(RD_ERR_INVALID_ARGUMENT,
                   "\"semantics\" = 1 runs with block 256, refill > 0 and no counting");
  r->prepare_timer.end_if_not_ended();
  r->render_timer.start_if_not_started();
  r->sample_timer.start();
  rd_status st = refresh_job(r);
  if (st != RD_OK) return st;
  std::string err;
  const uint64_t npix = (uint64_t)r->W * r->H;
  std::vector<rdh::DevJob> host_jobs{r->job_h};
  uint32_t done = 0;
  while (done < k) {
    uint64_t fit = r->eng.opt.l_capacity / npix;
    if (fit == 0) return r->fail(RD_ERR_INVALID_ARGUMENT, "image larger than the radiance arena");
    if (o.vk) fit = 1;
    uint32_t spp = (uint32_t)std::min<uint64_t>({(uint64_t)(k - done), fit, 0xffffffffull / npix});
    rdh::DevTicket t{};
    t.job = 0;
    t.sample_base = r->sample_count + done;
    t.spp = spp;
    t.paths = (uint32_t)(spp * npix);
    r->eng.hit_frame = r->frame_serial;
    if (!r->eng.run_tickets(r->job_d, {t}, host_jobs, &err)) return device_fail(r, err);
    done += spp;
  }
  if (!r->eng.check(hipStreamSynchronize(r->eng.stream), "render", &err)) return device_fail(r, err);
  r->sample_count += k;
  if (r->sample_count == r->cfg.max_sample_count) {
    r->render_timer.end();
    r->frame_timer.end();
  }
  r->sample_timer.end_multiple(k);
  return RD_OK;
}

// print_frame_buffer (cpu.rs:221-230) on the device, then copy out.
rd_status tonemap_out(rd_renderer *r, uint8_t *rgba, size_t cap) {
  std::string err;
  const uint64_t npix = (uint64_t)r->W * r->H;
  const int grid = (int)std::min<uint64_t>((npix + 255) / 256, 8192);
  if (!r->eng.check(rdh::launch_tonemap(r->job_d, (uint32_t)npix, (float)r->sample_count,
                                        r->rgba_d, r->eng.counters, r->eng.opt.count != 0, grid,
                                        r->eng.stream, r->frame_vk != 0),
                    "tonemap", &err) ||
      !r->eng.check(hipMemcpyAsync(rgba, r->rgba_d, 4 * npix, hipMemcpyDeviceToHost, r->eng.stream),
                    "read image", &err) ||
      !r->eng.check(hipStreamSynchronize(r->eng.stream), "read image", &err))
    return device_fail(r, err);
  (void)cap;
  return RD_OK;
}

bool option_value(const rd_renderer *r, const std::string &k, int64_t *v) {
  const rdh::EngineOptions &o = r->eng.opt;
  if (k == "block") *v = o.block;
  else if (k == "queue") *v = o.queue;
  else if (k == "leaf") *v = o.leaf;
  else if (k == "grid") *v = o.grid;
  else if (k == "traversal") *v = o.brute_force;
  else if (k == "refill") *v = o.refill;
  else if (k == "lds_bvh") *v = o.lds_bvh;
  else if (k == "lds_objects") *v = o.lds_objects;
  else if (k == "shade_blocks") *v = o.shade_blocks;
  else if (k == "hit_reuse") *v = o.hit_reuse;
  else if (k == "bvh4") *v = o.bvh4;
  else if (k == "pipes") *v = o.pipes;
  else if (k == "pipes_used") *v = r->eng.pipes_used;
  else if (k == "count") *v = o.count;
  else if (k == "kernel_timing") *v = o.kernel_timing;
  else if (k == "l_capacity") *v = (int64_t)o.l_capacity;
  else if (k == "sync_every") *v = o.sync_every;
  else if (k == "samples_per_pass") *v = r->samples_per_pass;
  else if (k == "bvh_depth") *v = r->scene.depth;
  else if (k == "bvh_nodes") *v = r->scene.n_nodes;
  else if (k == "iterations") *v = (int64_t)r->eng.iterations;
  else if (k == "num_cus") *v = r->eng.num_cus;
  else if (k == "semantics") *v = o.vk;
  else if (k == "lane_chunk") *v = o.lane_chunk_log2 ? (1ll << o.lane_chunk_log2) : 0;
  else return false;
  return true;
}

}  // namespace

extern "C" {

rd_status rd_config_default(rd_config *out) {
  if (!out) return RD_ERR_INVALID_ARGUMENT;
  out->max_sample_count = 1024;  // renderer/mod.rs:16-23
  out->max_bounces = 12;
  return RD_OK;
}

rd_status rd_create(const rd_config *config, rd_renderer **out) {
  if (!out) {
    rdh::set_global_error("rd_create: out is NULL");
    return RD_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  rd_renderer *r = new (std::nothrow) rd_renderer();
  if (!r) return RD_ERR_OUT_OF_MEMORY;
  if (config) r->cfg = *config;
  r->eng.cache = &r->cache;
  std::string err;
  if (!r->eng.init(&err) ||
      !r->eng.check(hipHostMalloc((void **)&r->job_pin, sizeof(rdh::DevJob), hipHostMallocDefault),
                    "hipHostMalloc(job)", &err) ||
      !scratch_setup(r->eng.device, &err)) {
    rdh::set_global_error(err);
    delete r;
    return RD_ERR_DEVICE;
  }
  *out = r;
  return RD_OK;
}

void rd_destroy(rd_renderer *r) {
  if (!r) return;
  (void)hipDeviceSynchronize();
  r->eng.free_scene(&r->scene);
  r->eng.dfree(r->acc);
  r->eng.dfree(r->rgba_d);
  r->eng.dfree(r->vk_rng);
  if (r->job_d) (void)hipFree(r->job_d);
  if (r->job_pin) (void)hipHostFree(r->job_pin);
  delete r;
}

rd_status rd_new_frame(rd_renderer *r, const rd_scene *scene) {
  if (!r) return RD_ERR_INVALID_ARGUMENT;
  r->frame_timer.start();    // cpu.rs:136-137
  r->prepare_timer.start();
  r->sample_count = 0;
  return snapshot(r, scene, true);
}

rd_status rd_restart_frame(rd_renderer *r) {
  if (!r) return RD_ERR_INVALID_ARGUMENT;
  if (!r->has_frame) return r->fail(RD_ERR_INVALID_ARGUMENT, "restart_frame before new_frame");
  ++r->frame_serial;
  r->frame_timer.start();
  r->prepare_timer.start();
  r->sample_count = 0;
  std::string err;
  if (!r->eng.check(rdh::launch_zero_acc(r->acc, 4ull * r->W * r->H, r->eng.stream),
                    "zero accumulator", &err))
    return device_fail(r, err);
  return RD_OK;
}

rd_status rd_render_sample(rd_renderer *r, const rd_scene *scene, uint8_t *rgba, size_t cap,
                           int *produced) {
  if (!r || !produced) return RD_ERR_INVALID_ARGUMENT;
  *produced = 0;
  if (r->sample_count >= r->cfg.max_sample_count) return RD_OK;  // cpu.rs:143-145 -> None
  if (!r->has_frame || (scene && (scene->camera.resolution_x != r->W ||
                                  scene->camera.resolution_y != r->H))) {
    // frame_buffer(): a fresh buffer when none exists or the size changed
    // (cpu.rs:400-417); sample_count is kept, as there.
    if (!scene) return r->fail(RD_ERR_INVALID_ARGUMENT, "render_sample before new_frame needs a scene");
    rd_status st = snapshot(r, scene, true);
    if (st != RD_OK) return st;
  }
  if (!rgba || cap < 4ull * r->W * r->H)
    return r->fail(RD_ERR_BUFFER_TOO_SMALL, "image buffer smaller than W*H*4 bytes");
  // the Vulkan renderer renders every sample in one dispatch (vulkan.rs:109-184)
  const uint32_t k = r->frame_vk ? r->cfg.max_sample_count - r->sample_count
                                 : std::min(r->samples_per_pass, r->cfg.max_sample_count - r->sample_count);
  rd_status st = render_samples(r, k);
  if (st != RD_OK) return st;
  st = tonemap_out(r, rgba, cap);
  if (st != RD_OK) return st;
  *produced = 1;
  return RD_OK;
}

rd_status rd_render_frame(rd_renderer *r, const rd_scene *scene, uint8_t *rgba, size_t cap) {
  if (!r) return RD_ERR_INVALID_ARGUMENT;
  if (scene && (!rgba || cap < 4ull * scene->camera.resolution_x * scene->camera.resolution_y))
    return r->fail(RD_ERR_BUFFER_TOO_SMALL, "image buffer smaller than W*H*4 bytes");
  rd_status st = rd_new_frame(r, scene);
  if (st != RD_OK) return st;
  if (r->sample_count < r->cfg.max_sample_count) {
    st = render_samples(r, r->cfg.max_sample_count - r->sample_count);
    if (st != RD_OK) return st;
  }
  return tonemap_out(r, rgba, cap);
}

rd_status rd_render_frame_noalloc(rd_renderer *r, const rd_scene *scene, uint8_t *rgba,
                                  size_t cap) {
  if (!r) return RD_ERR_INVALID_ARGUMENT;
  std::string err;
  if (!valid_scene(scene, &err))
    return r->fail(scene && scene->world.kind == RD_WORLD_TRANSPARENT ? RD_ERR_UNSUPPORTED
                                                                       : RD_ERR_INVALID_ARGUMENT,
                   err);
  const uint32_t W = scene->camera.resolution_x, H = scene->camera.resolution_y;
  const uint64_t npix = (uint64_t)W * H;
  if (!rgba || cap < 4 * npix) return r->fail(RD_ERR_BUFFER_TOO_SMALL, "image buffer smaller than W*H*4 bytes");
  if (r->eng.opt.vk)
    return r->fail(RD_ERR_UNSUPPORTED, "\"semantics\" = 1 is not available without allocation");
  const int dev = r->eng.device;
  uint8_t *base = static_cast<uint8_t *>(scratch_addr(dev));
  if (!base) return r->fail(RD_ERR_DEVICE, "static scratch not set up on this device");
  std::lock_guard<std::mutex> one_at_a_time(g_scratch_use[dev]);

  rdh::BuiltScene b;
  if (!rdh::build_scene(*scene, r->eng.opt.leaf, &b, &err)) return r->fail(RD_ERR_INVALID_ARGUMENT, err);
  if (!rdh::check_bvh(b, &err)) return r->fail(RD_ERR_INVALID_ARGUMENT, "BVH build: " + err);

  // The engine's launch shape decides the queue segments: a fixed grid keeps
  // them at queue / 32 slots for every block size.
  rdh::Engine &se = r->noalloc_eng;
  se.opt = r->eng.opt;
  se.opt.grid = rdh::kScratchGrid;
  se.opt.queue = rdh::kScratchQueue;
  se.opt.pipes = 1;
  se.opt.count = 0;
  se.opt.kernel_timing = 0;
  se.num_cus = r->eng.num_cus;
  const uint64_t seg_cap = se.segment_slots();
  const uint64_t slots = (uint64_t)rdh::kSegments * seg_cap;

  // Layout of the scratch array, 256-byte aligned pieces.
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    off = (off + 255) & ~255ull;
    const uint64_t at = off;
    off += bytes;
    return at;
  };
  const uint64_t o_ctl = take(sizeof(rdh::QueueCtl)), o_cnt = take(sizeof(rdh::DevCounters));
  const uint64_t o_job = take(sizeof(rdh::DevJob)), o_tk = take(sizeof(rdh::DevTicket));
  const uint64_t o_done = take(8);
  const uint64_t o_nodes = take(sizeof(rdh::DevNode) * b.nodes.size());
  const uint64_t o_nodes4 = take(sizeof(rdh::DevNode4) * b.nodes4.size());
  const uint64_t o_prims = take(sizeof(rdh::DevPrim) * b.prims.size());
  const uint64_t o_obj = take(sizeof(rdh::DevObject) * b.objects.size());
  const uint64_t o_acc = take(16 * npix), o_img = take(4 * npix);
  uint64_t o_q[2][5];
  for (int p = 0; p < 2; ++p) {
    o_q[p][0] = take(sizeof(rdh::QRay) * slots);
    o_q[p][1] = take(sizeof(rdh::QDir) * slots);
    o_q[p][2] = take(sizeof(rdh::QWord) * slots);
    o_q[p][3] = take(sizeof(rdh::QAtt) * slots);
    o_q[p][4] = take(sizeof(rdh::QLight) * slots);
  }
  const uint64_t o_hit_t = take(4 * slots), o_hit_id = take(4 * slots);
  const uint64_t o_L = take(0);
  // the radiance arena takes the rest: at least one sample of every pixel
  const uint64_t l_slots = off < rdh::kScratchBytes ? (rdh::kScratchBytes - off) / 16 : 0;
  if (l_slots < npix)
    return r->fail(RD_ERR_INVALID_ARGUMENT,
                   "scene needs " + std::to_string(off + 16 * npix) + " bytes of the " +
                       std::to_string(rdh::kScratchBytes) +
                       "-byte static scratch (image, BVH, queues and one sample of radiance slots); "
                       "rd_render_frame has no such limit");

  rdh::ExternalBuffers eb{};
  for (int p = 0; p < 2; ++p) {
    eb.q[p].ray = reinterpret_cast<rdh::QRay *>(base + o_q[p][0]);
    eb.q[p].dir = reinterpret_cast<rdh::QDir *>(base + o_q[p][1]);
    eb.q[p].word = reinterpret_cast<rdh::QWord *>(base + o_q[p][2]);
    eb.q[p].att = reinterpret_cast<rdh::QAtt *>(base + o_q[p][3]);
    eb.q[p].light = reinterpret_cast<rdh::QLight *>(base + o_q[p][4]);
  }
  eb.hit_t = reinterpret_cast<float *>(base + o_hit_t);
  eb.hit_id = reinterpret_cast<uint32_t *>(base + o_hit_id);
  eb.ctl = reinterpret_cast<rdh::QueueCtl *>(base + o_ctl);
  eb.L = reinterpret_cast<float4 *>(base + o_L);
  eb.tickets = reinterpret_cast<rdh::DevTicket *>(base + o_tk);
  eb.ticket_done = reinterpret_cast<unsigned long long *>(base + o_done);
  eb.counters = reinterpret_cast<rdh::DevCounters *>(base + o_cnt);
  eb.queue = rdh::kScratchQueue;
  eb.seg_cap = (uint32_t)seg_cap;
  eb.l_slots = l_slots;
  eb.tickets_n = 1;
  se.bind_external(eb, r->eng);

  // new_frame (cpu.rs:135-140): timers, then the snapshot into the scratch
  r->frame_timer.start();
  r->prepare_timer.start();
  r->sample_count = 0;
  rdh::DeviceScene ds;
  ds.nodes = reinterpret_cast<rdh::DevNode *>(base + o_nodes);
  ds.nodes4 = reinterpret_cast<rdh::DevNode4 *>(base + o_nodes4);
  ds.prims = reinterpret_cast<rdh::DevPrim *>(base + o_prims);
  ds.objects = reinterpret_cast<rdh::DevObject *>(base + o_obj);
  ds.n_nodes = (uint32_t)b.nodes.size();
  ds.n_nodes4 = (uint32_t)b.nodes4.size();
  ds.stack4 = b.stack4;
  ds.max_leaf = b.max_leaf;
  ds.n_prims = (uint32_t)b.prims.size();
  ds.r0 = b.r0;
  ds.depth = b.depth;
  float *acc = reinterpret_cast<float *>(base + o_acc);
  uchar4 *img = reinterpret_cast<uchar4 *>(base + o_img);
  rdh::DevJob job;
  rdh::fill_job(*scene, ds, acc, r->cfg.max_bounces, r->seed, &job);
  rdh::DevJob *d_job = reinterpret_cast<rdh::DevJob *>(base + o_job);
  const hipStream_t st = se.stream;
  struct Up {
    void *dst;
    const void *src;
    uint64_t bytes;
  } ups[] = {{ds.nodes, b.nodes.data(), sizeof(rdh::DevNode) * b.nodes.size()},
             {ds.nodes4, b.nodes4.data(), sizeof(rdh::DevNode4) * b.nodes4.size()},
             {ds.prims, b.prims.data(), sizeof(rdh::DevPrim) * b.prims.size()},
             {ds.objects, b.objects.data(), sizeof(rdh::DevObject) * b.objects.size()},
             {d_job, &job, sizeof(job)}};
  // every call starts from its own state: the accumulator is cleared by a
  // kernel and the control block by run_tickets; the rest is written before
  // it is read
  bool ok = true;
  for (const Up &u : ups)
    if (ok && u.bytes)
      ok = se.check(hipMemcpyAsync(u.dst, u.src, u.bytes, hipMemcpyHostToDevice, st), "upload scene", &err);
  ok = ok && se.check(rdh::launch_zero_acc(acc, 4 * npix, st), "zero accumulator", &err);
  // render_next_sample timers (cpu.rs:194-196, 212-218) around passes of
  // as many samples as the radiance slots hold, in sample order
  const std::vector<rdh::DevJob> host_jobs{job};
  const uint32_t total = r->cfg.max_sample_count;
  if (ok && total > 0) {
    r->prepare_timer.end_if_not_ended();
    r->render_timer.start_if_not_started();
    r->sample_timer.start();
    const uint64_t fit = std::min<uint64_t>(l_slots / npix, 0xffffffffull / npix);
    for (uint32_t done = 0; ok && done < total;) {
      rdh::DevTicket t{};
      t.job = 0;
      t.sample_base = done;
      t.spp = (uint32_t)std::min<uint64_t>(total - done, fit);
      t.paths = (uint32_t)(t.spp * npix);
      ok = se.run_tickets(d_job, {t}, host_jobs, &err);
      done += t.spp;
    }
  }
  const int grid = (int)std::min<uint64_t>((npix + 255) / 256, 8192);
  ok = ok &&
       se.check(rdh::launch_tonemap(d_job, (uint32_t)npix, (float)(ok ? total : 0), img, eb.counters,
                                    false, grid, st),
                "tonemap", &err) &&
       se.check(hipMemcpyAsync(rgba, img, 4 * npix, hipMemcpyDeviceToHost, st), "read image", &err);
  std::string sync_err;
  const bool synced = se.check(hipStreamSynchronize(st), "render", &sync_err);
  if (ok && !synced) err = sync_err;
  if (!ok || !synced) return device_fail(r, err);
  r->sample_count = total;
  if (total > 0) {
    r->render_timer.end();
    r->frame_timer.end();
    r->sample_timer.end_multiple(total);
  }
  return RD_OK;
}

rd_status rd_profiler(const rd_renderer *r, rd_timings *out) {
  if (!r || !out) return RD_ERR_INVALID_ARGUMENT;
  out->frame_ns = r->frame_timer.ns;
  out->sample_ns = r->sample_timer.ns;
  out->prepare_ns = r->prepare_timer.ns;
  out->render_ns = r->render_timer.ns;
  return RD_OK;
}

uint32_t rd_sample_count(const rd_renderer *r) { return r ? r->sample_count : 0; }
uint32_t rd_max_sample_count(const rd_renderer *r) { return r ? r->cfg.max_sample_count : 0; }
uint32_t rd_max_bounces(const rd_renderer *r) { return r ? r->cfg.max_bounces : 0; }
void rd_set_max_sample_count(rd_renderer *r, uint32_t count) {
  if (r) r->cfg.max_sample_count = count;
}
void rd_set_max_bounces(rd_renderer *r, uint32_t bounces) {
  if (r) r->cfg.max_bounces = bounces;
}

const char *rd_last_error(const rd_renderer *r) {
  if (r) return r->last_error.c_str();
  static thread_local std::string copy;
  std::lock_guard<std::mutex> lk(rdh::g_err_mu);
  copy = rdh::g_err;
  return copy.c_str();
}

rd_status rd_set_samples_per_pass(rd_renderer *r, uint32_t samples) {
  if (!r) return RD_ERR_INVALID_ARGUMENT;
  if (samples == 0) return r->fail(RD_ERR_INVALID_ARGUMENT, "samples per pass must be >= 1");
  r->samples_per_pass = samples;
  return RD_OK;
}

void rd_set_seed(rd_renderer *r, uint64_t seed) {
  if (r) r->seed = seed;
}

rd_status rd_read_accumulator(rd_renderer *r, float *out, size_t cap) {
  if (!r || !out) return RD_ERR_INVALID_ARGUMENT;
  if (!r->has_frame) return r->fail(RD_ERR_INVALID_ARGUMENT, "no frame");
  const uint64_t n = 4ull * r->W * r->H;
  if (cap < n) return r->fail(RD_ERR_BUFFER_TOO_SMALL, "accumulator buffer too small");
  std::string err;
  if (!r->eng.check(hipMemcpyAsync(out, r->acc, 4 * n, hipMemcpyDeviceToHost, r->eng.stream),
                    "read accumulator", &err) ||
      !r->eng.check(hipStreamSynchronize(r->eng.stream), "read accumulator", &err))
    return device_fail(r, err);
  return RD_OK;
}

rd_status rd_set_option(rd_renderer *r, const char *key, int64_t v) {
  if (!r || !key) return RD_ERR_INVALID_ARGUMENT;
  const std::string k(key);
  rdh::EngineOptions &o = r->eng.opt;
  if (k == "block") {
    if (v != 64 && v != 128 && v != 256) return r->fail(RD_ERR_INVALID_ARGUMENT, "block must be 64, 128 or 256");
    o.block = (int)v;
  } else if (k == "queue") {
    if (v < 1024 || v > (1ll << 28)) return r->fail(RD_ERR_INVALID_ARGUMENT, "queue must be in [1024, 2^28]");
    o.queue = (uint32_t)v;
  } else if (k == "leaf") {
    if (v < 1 || v > rdh::kMaxLeafSize) return r->fail(RD_ERR_INVALID_ARGUMENT, "leaf must be in 1..8");
    o.leaf = (int)v;
  } else if (k == "grid") {
    if (v < 0 || v > (1 << 20)) return r->fail(RD_ERR_INVALID_ARGUMENT, "grid out of range");
    o.grid = (int)v;
  } else if (k == "traversal") {
    if (v != 0 && v != 1) return r->fail(RD_ERR_INVALID_ARGUMENT, "traversal must be 0 or 1");
    o.brute_force = (int)v;
  } else if (k == "refill") {
    if (v < 0 || v > 64) return r->fail(RD_ERR_INVALID_ARGUMENT, "refill must be in 0..64");
    o.refill = (int)v;
  } else if (k == "bvh4") {
    if (v != 0 && v != 1) return r->fail(RD_ERR_INVALID_ARGUMENT, "bvh4 must be 0 or 1");
    o.bvh4 = (int)v;
  } else if (k == "pipes") {
    if (v < 0 || v > 4) return r->fail(RD_ERR_INVALID_ARGUMENT, "pipes must be in 0..4 (0 = automatic)");
    o.pipes = (int)v;
  } else if (k == "hit_reuse") {
    if (v != 0 && v != 1) return r->fail(RD_ERR_INVALID_ARGUMENT, "hit_reuse must be 0 or 1");
    o.hit_reuse = (int)v;
  } else if (k == "shade_blocks") {
    if (v < 0 || v > 64) return r->fail(RD_ERR_INVALID_ARGUMENT, "shade_blocks must be in 0..64");
    o.shade_blocks = (int)v;
  } else if (k == "lds_bvh") {
    if (v < 0 || v > (1 << 17)) return r->fail(RD_ERR_INVALID_ARGUMENT, "lds_bvh must be in 0..131072");
    o.lds_bvh = (uint32_t)v;
  } else if (k == "lds_objects") {
    if (v < 0 || v > (1 << 16)) return r->fail(RD_ERR_INVALID_ARGUMENT, "lds_objects must be in 0..65536");
    o.lds_objects = (uint32_t)v;
  } else if (k == "count") {
    o.count = v != 0;
  } else if (k == "kernel_timing") {
    o.kernel_timing = v != 0;
  } else if (k == "l_capacity") {
    if (v < (1 << 16) || v > (1ll << 30)) return r->fail(RD_ERR_INVALID_ARGUMENT, "l_capacity out of range");
    o.l_capacity = (uint64_t)v;
  } else if (k == "sync_every") {
    if (v < 1 || v > 1024) return r->fail(RD_ERR_INVALID_ARGUMENT, "sync_every out of range");
    o.sync_every = (int)v;
  } else if (k == "samples_per_pass") {
    return rd_set_samples_per_pass(r, (uint32_t)std::max<int64_t>(0, std::min<int64_t>(v, 0xffffffffll)));
  } else if (k == "lane_chunk") {
    if (v < 0 || (v > 0 && v < 64) || v > 65536 || (v & (v - 1)) != 0)
      return r->fail(RD_ERR_INVALID_ARGUMENT, "lane_chunk must be 0 or a power of two in 64..65536");
    o.lane_chunk_log2 = v ? (uint32_t)__builtin_ctzll((unsigned long long)v) : 0u;
  } else if (k == "semantics") {
    if (v != 0 && v != 1) return r->fail(RD_ERR_INVALID_ARGUMENT, "semantics must be 0 (CPU renderer) or 1 (Vulkan renderer)");
    o.vk = (int)v;
  } else {
    return r->fail(RD_ERR_INVALID_ARGUMENT, "unknown option " + k);
  }
  return RD_OK;
}

rd_status rd_get_option(const rd_renderer *r, const char *key, int64_t *v) {
  if (!r || !key || !v) return RD_ERR_INVALID_ARGUMENT;
  return option_value(r, key, v) ? RD_OK : RD_ERR_INVALID_ARGUMENT;
}

rd_status rd_get_counters(const rd_renderer *r, rd_counters *out) {
  if (!r || !out) return RD_ERR_INVALID_ARGUMENT;
  rdh::DevCounters c;
  if (hipMemcpy(&c, r->eng.counters, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
    return RD_ERR_DEVICE;
  static_assert(sizeof(rd_counters) == sizeof(rdh::DevCounters), "counter layout");
  std::memcpy(out, &c, sizeof(c));
  return RD_OK;
}

rd_status rd_reset_counters(rd_renderer *r) {
  if (!r) return RD_ERR_INVALID_ARGUMENT;
  if (hipMemsetAsync(r->eng.counters, 0, sizeof(rdh::DevCounters), r->eng.stream) != hipSuccess ||
      hipStreamSynchronize(r->eng.stream) != hipSuccess)
    return RD_ERR_DEVICE;
  for (int i = 0; i < 5; ++i) {
    r->eng.kernel_ms[i] = 0.0;
    r->eng.kernel_launches[i] = 0;
  }
  r->eng.iterations = 0;
  return RD_OK;
}

rd_status rd_get_kernel_times(const rd_renderer *r, rd_kernel_times *out) {
  if (!r || !out) return RD_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 5; ++i) {
    out->total_ms[i] = r->eng.kernel_ms[i];
    out->launches[i] = r->eng.kernel_launches[i];
  }
  return RD_OK;
}

}  // extern "C"

// ------------------------------------------------------------ debug hooks --
namespace {

struct DebugCtx {
  rdh::BuiltScene built;
  rdh::DevNode *nodes = nullptr;
  rdh::DevPrim *prims = nullptr;
  rdh::DevObject *objects = nullptr;
  rdh::DevJob *job = nullptr;
  std::vector<void *> extra;  // every device allocation above, freed once
  ~DebugCtx() {
    for (void *p : extra) (void)hipFree(p);
  }
  template <class T> T *dev(const T *host, size_t n) {
    T *d = nullptr;
    if (hipMalloc(&d, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
    extra.push_back(d);
    if (host && n && hipMemcpy(d, host, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
  }
};

rd_status debug_setup(const rd_scene *scene, int leaf, DebugCtx *c) {
  std::string err;
  if (!scene) { rdh::set_global_error("scene is NULL"); return RD_ERR_INVALID_ARGUMENT; }
  if (!rdh::build_scene(*scene, leaf, &c->built, &err)) {
    rdh::set_global_error(err);
    return RD_ERR_INVALID_ARGUMENT;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    rdh::set_global_error("no HIP device available");
    return RD_ERR_DEVICE;
  }
  auto &b = c->built;
  if (b.nodes.size()) c->nodes = c->dev(b.nodes.data(), b.nodes.size());
  if (b.prims.size()) c->prims = c->dev(b.prims.data(), b.prims.size());
  if (b.objects.size()) c->objects = c->dev(b.objects.data(), b.objects.size());
  rdh::DeviceScene ds;
  ds.nodes = c->nodes;
  ds.prims = c->prims;
  ds.objects = c->objects;
  ds.n_nodes = (uint32_t)b.nodes.size();
  ds.n_prims = (uint32_t)b.prims.size();
  ds.r0 = b.r0;
  rdh::DevJob j;
  rdh::fill_job(*scene, ds, nullptr, 0, 0, &j);
  c->job = nullptr;
  c->job = c->dev(&j, 1);
  if (!c->job) { rdh::set_global_error("device allocation failed"); return RD_ERR_OUT_OF_MEMORY; }
  return RD_OK;
}

rd_status finish(hipError_t e) {
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    rdh::set_global_error(hipGetErrorString(e));
    return RD_ERR_DEVICE;
  }
  return RD_OK;
}

}  // namespace

extern "C" {

rd_status rd_debug_trace(const rd_scene *scene, int leaf, int brute, const float *rays, uint32_t n,
                         int32_t *ids, float *ts, float *normals, uint8_t *front) {
  if (!rays || !ids || !ts || !normals || !front) return RD_ERR_INVALID_ARGUMENT;
  DebugCtx c;
  rd_status st = debug_setup(scene, leaf, &c);
  if (st != RD_OK) return st;
  const float *d_rays = c.dev(rays, 6ull * n);
  int32_t *d_ids = c.dev<int32_t>(nullptr, n);
  float *d_ts = c.dev<float>(nullptr, n);
  float *d_n = c.dev<float>(nullptr, 3ull * n);
  uint8_t *d_f = c.dev<uint8_t>(nullptr, n);
  if (!d_rays || !d_ids || !d_ts || !d_n || !d_f) return RD_ERR_OUT_OF_MEMORY;
  hipError_t e = rdh::launch_debug_trace(c.nodes, c.prims, (uint32_t)c.built.prims.size(),
                                         (uint32_t)c.built.nodes.size(), c.built.r0, c.objects,
                                         d_rays, n, brute, d_ids, d_ts, d_n, d_f, nullptr);
  if ((st = finish(e)) != RD_OK) return st;
  if (n) {
    (void)hipMemcpy(ids, d_ids, 4ull * n, hipMemcpyDeviceToHost);
    (void)hipMemcpy(ts, d_ts, 4ull * n, hipMemcpyDeviceToHost);
    (void)hipMemcpy(normals, d_n, 12ull * n, hipMemcpyDeviceToHost);
    (void)hipMemcpy(front, d_f, n, hipMemcpyDeviceToHost);
  }
  return finish(hipSuccess);
}

rd_status rd_debug_extend(const rd_scene *scene, int leaf, const rd_extend_variant *v,
                          const float *rays, uint32_t n, int32_t *ids, float *ts, float *normals,
                          uint8_t *front) {
  if (!v || !rays || !ids || !ts || !normals || !front) return RD_ERR_INVALID_ARGUMENT;
  if ((v->block != 64 && v->block != 128 && v->block != 256) || v->jobs < 1 || v->jobs > 64 ||
      v->refill < 0 || v->refill > 64 || v->grid < 0 || v->grid > (1 << 16) || n > (1u << 26) ||
      v->lane_chunk < 0 || (v->lane_chunk > 0 && v->lane_chunk < 64) || v->lane_chunk > 65536 ||
      (v->lane_chunk & (v->lane_chunk - 1)) != 0) {
    rdh::set_global_error("rd_debug_extend: variant out of range");
    return RD_ERR_INVALID_ARGUMENT;
  }
  DebugCtx c;
  rd_status st = debug_setup(scene, leaf, &c);
  if (st != RD_OK) return st;
  const rdh::BuiltScene &b = c.built;
  const bool single = v->jobs == 1, lanes = v->refill > 0;
  const uint32_t n4 = (uint32_t)b.nodes4.size();
  auto unsupported = [](const char *why) {
    rdh::set_global_error(std::string("rd_debug_extend: ") + why);
    return RD_ERR_UNSUPPORTED;
  };
  // the forms the engine selects (Engine::run_tickets, Streamer::args)
  if (v->bvh4 && (!lanes || v->wide || n4 == 0 || b.stack4 + 3 > (uint32_t)rdh::kStackDepth))
    return unsupported("no 4-wide walk for this scene or variant");
  if (!v->wide && b.nodes.size() > rdh::kNarrowStackNodes)
    return unsupported("more than 65535 nodes need the 32-bit stack");
  uint32_t lds_bytes = 0, rows = 0;
  if (v->lds) {
    const uint64_t nb = v->bvh4 ? sizeof(rdh::DevNode4) * (uint64_t)n4
                                : sizeof(rdh::DevNode) * (uint64_t)b.nodes.size();
    const uint64_t bytes = nb + sizeof(rdh::DevPrim) * (uint64_t)b.prims.size();
    rows = v->bvh4 ? b.stack4 + 3 : b.depth + 1;
    if (!single || !lanes || v->wide || b.nodes.empty() || bytes > rdh::EngineOptions().lds_bvh ||
        rows > (uint32_t)rdh::kStackDepth)
      return unsupported("the tree is not traversed from LDS for this scene or variant");
    lds_bytes = (uint32_t)bytes;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess)
    return finish(hipGetLastError());
  // one device copy of the scene per job, so each lane follows its own job's pointers
  std::vector<rdh::DevJob> jobs((size_t)v->jobs);
  for (int k = 0; k < v->jobs; ++k) {
    rdh::DeviceScene ds;
    ds.nodes = k == 0 ? c.nodes : (b.nodes.empty() ? nullptr : c.dev(b.nodes.data(), b.nodes.size()));
    ds.prims = k == 0 ? c.prims : (b.prims.empty() ? nullptr : c.dev(b.prims.data(), b.prims.size()));
    ds.objects = k == 0 ? c.objects : (b.objects.empty() ? nullptr : c.dev(b.objects.data(), b.objects.size()));
    ds.nodes4 = n4 ? c.dev(b.nodes4.data(), n4) : nullptr;
    ds.n_nodes = (uint32_t)b.nodes.size();
    ds.n_nodes4 = n4;
    ds.stack4 = b.stack4;
    ds.n_prims = (uint32_t)b.prims.size();
    ds.r0 = b.r0;
    ds.depth = b.depth;
    if ((!b.nodes.empty() && !ds.nodes) || (!b.prims.empty() && !ds.prims) || (n4 && !ds.nodes4))
      return RD_ERR_OUT_OF_MEMORY;
    rdh::fill_job(*scene, ds, nullptr, 1, 0, &jobs[k]);
  }
  std::vector<rdh::DevTicket> tickets((size_t)v->jobs);
  for (int k = 0; k < v->jobs; ++k) tickets[k].job = (uint32_t)k;
  // The queue: segment s holds rays [s * seg_cap, (s + 1) * seg_cap), so the
  // consumer order of the iteration is the ray order and slot i holds ray i.
  const uint32_t seg_cap = std::max<uint32_t>(1, (n + rdh::kSegments - 1) / rdh::kSegments);
  const uint64_t slots = (uint64_t)rdh::kSegments * seg_cap;
  std::vector<rdh::QRay> qray(slots, rdh::QRay{});
  std::vector<rdh::QDir> qdir(slots, rdh::QDir{});
  std::vector<rdh::QWord> qword(slots, rdh::QWord{});
  for (uint32_t i = 0; i < n; ++i) {
    const float *r6 = rays + 6ull * i;
    qray[i] = rdh::QRay{r6[0], r6[1], r6[2], r6[3]};
    qdir[i] = rdh::QDir{r6[4], r6[5]};
    qword[i] = rdh::QWord{(single ? 0u : i % (uint32_t)v->jobs) << 8, 0u};
  }
  rdh::QueueCtl ctl_h;
  std::memset(&ctl_h, 0, sizeof(ctl_h));
  for (int s = 0; s < rdh::kSegments; ++s) {
    const uint64_t lo = (uint64_t)s * seg_cap;
    ctl_h.count[0][s * rdh::kSegStride] = lo >= n ? 0u : (uint32_t)std::min<uint64_t>(seg_cap, n - lo);
  }
  ctl_h.n_tickets = (uint32_t)v->jobs;
  rdh::PathQueue q[2];
  for (int p = 0; p < 2; ++p) {
    q[p].ray = c.dev<rdh::QRay>(p == 0 ? qray.data() : nullptr, slots);
    q[p].dir = c.dev<rdh::QDir>(p == 0 ? qdir.data() : nullptr, slots);
    q[p].word = c.dev<rdh::QWord>(p == 0 ? qword.data() : nullptr, slots);
    q[p].att = c.dev<rdh::QAtt>(nullptr, slots);
    q[p].light = c.dev<rdh::QLight>(nullptr, slots);
    if (!q[p].ray || !q[p].dir || !q[p].word || !q[p].att || !q[p].light) return RD_ERR_OUT_OF_MEMORY;
  }
  rdh::IterArgs a{};
  a.q[0] = q[0];
  a.q[1] = q[1];
  a.hit_t = c.dev<float>(nullptr, slots);
  a.hit_id = c.dev<uint32_t>(nullptr, slots);
  a.ctl = c.dev(&ctl_h, 1);
  a.jobs = c.dev(jobs.data(), jobs.size());
  a.tickets = c.dev(tickets.data(), tickets.size());
  a.counters = c.dev<rdh::DevCounters>(nullptr, 1);
  float *d_rays = c.dev(rays, 6ull * std::max<uint32_t>(n, 1));
  float *d_n = c.dev<float>(nullptr, 3ull * std::max<uint32_t>(n, 1));
  uint8_t *d_f = c.dev<uint8_t>(nullptr, std::max<uint32_t>(n, 1));
  if (!a.hit_t || !a.hit_id || !a.ctl || !a.jobs || !a.tickets || !a.counters || !d_rays || !d_n || !d_f)
    return RD_ERR_OUT_OF_MEMORY;
  a.capacity = (uint32_t)slots;
  a.seg_cap = seg_cap;
  a.brute_force = 0;
  a.single_job = single ? 1 : 0;
  a.track_done = 0;
  a.wide_stack = v->wide ? 1 : 0;
  a.refill = v->refill;
  a.chunk_log2 = v->lane_chunk ? (uint32_t)__builtin_ctz((unsigned)v->lane_chunk) : 0u;
  a.lds_bvh_bytes = lds_bytes;
  a.lds_obj_bytes = 0;
  a.stack_rows = rows;
  a.bvh4 = v->bvh4 ? 1 : 0;
  a.leaf1 = single && v->bvh4 && b.max_leaf == 1 ? 1 : 0;
  a.slots = (uint32_t)slots;
  a.ticket_cap = (uint32_t)v->jobs;
  a.job_cap = (uint32_t)v->jobs;
  rdh::LaunchShape sh;
  sh.block = v->block;
  if (v->grid > 0) {
    sh.grid = v->grid;
  } else {
    const uint32_t dyn = rdh::extend_dynamic_lds(a, sh.block);
    sh.grid = cus * std::max(1, rdh::occupancy_blocks(1, sh.block, lanes, dyn ? lds_bytes : 0, dyn,
                                                      v->bvh4 != 0, single));
  }
  if (n && (st = finish(rdh::launch_extend(a, 0, sh, false, nullptr))) != RD_OK) return st;
  if (n && (st = finish(rdh::launch_debug_hits(c.objects, d_rays, a.hit_id, a.hit_t, n, d_n, d_f,
                                               nullptr))) != RD_OK)
    return st;
  std::vector<uint32_t> id_h(n);
  if (n) {
    (void)hipMemcpy(id_h.data(), a.hit_id, 4ull * n, hipMemcpyDeviceToHost);
    (void)hipMemcpy(ts, a.hit_t, 4ull * n, hipMemcpyDeviceToHost);
    (void)hipMemcpy(normals, d_n, 12ull * n, hipMemcpyDeviceToHost);
    (void)hipMemcpy(front, d_f, n, hipMemcpyDeviceToHost);
  }
  for (uint32_t i = 0; i < n; ++i) {
    ids[i] = id_h[i] == rdh::kNoHit ? -1 : (int32_t)id_h[i];
    if (ids[i] < 0) ts[i] = 0.0f;
  }
  return finish(hipSuccess);
}

rd_status rd_debug_scatter(const rd_scene *scene, const float *in, const uint32_t *bits,
                           uint32_t n, float *out) {
  if (!in || !bits || !out) return RD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t id;
    std::memcpy(&id, in + 16ull * i + 7, 4);
    if (!scene || id >= scene->object_count) {
      rdh::set_global_error("rd_debug_scatter: object id out of range");
      return RD_ERR_INVALID_ARGUMENT;
    }
  }
  DebugCtx c;
  rd_status st = debug_setup(scene, 4, &c);
  if (st != RD_OK) return st;
  const float *d_in = c.dev(in, 16ull * n);
  const uint32_t *d_bits = c.dev(bits, (uint64_t)rdh::kDrawWords * n);
  float *d_out = c.dev<float>(nullptr, 16ull * n);
  if (!d_in || !d_bits || !d_out) return RD_ERR_OUT_OF_MEMORY;
  if ((st = finish(rdh::launch_debug_scatter(c.objects, d_in, d_bits, n, d_out, nullptr))) != RD_OK)
    return st;
  if (n) (void)hipMemcpy(out, d_out, 64ull * n, hipMemcpyDeviceToHost);
  return finish(hipSuccess);
}

rd_status rd_debug_camera_rays(const rd_scene *scene, const uint32_t *pixels, uint32_t n,
                               float *out) {
  if (!pixels || !out || !scene) return RD_ERR_INVALID_ARGUMENT;
  const uint64_t np = (uint64_t)scene->camera.resolution_x * scene->camera.resolution_y;
  for (uint32_t i = 0; i < n; ++i)
    if (pixels[i] >= np) return RD_ERR_INVALID_ARGUMENT;
  DebugCtx c;
  rd_status st = debug_setup(scene, 4, &c);
  if (st != RD_OK) return st;
  const uint32_t *d_px = c.dev(pixels, n);
  float *d_out = c.dev<float>(nullptr, 6ull * n);
  if (!d_px || !d_out) return RD_ERR_OUT_OF_MEMORY;
  if ((st = finish(rdh::launch_debug_camera(c.job, d_px, n, d_out, nullptr))) != RD_OK) return st;
  if (n) (void)hipMemcpy(out, d_out, 24ull * n, hipMemcpyDeviceToHost);
  return finish(hipSuccess);
}

rd_status rd_debug_world(const rd_scene *scene, const float *dirs, uint32_t n, float *out) {
  if (!dirs || !out || !scene) return RD_ERR_INVALID_ARGUMENT;
  DebugCtx c;
  rd_status st = debug_setup(scene, 4, &c);
  if (st != RD_OK) return st;
  const float *d_in = c.dev(dirs, 3ull * n);
  float *d_out = c.dev<float>(nullptr, 3ull * n);
  if (!d_in || !d_out) return RD_ERR_OUT_OF_MEMORY;
  if ((st = finish(rdh::launch_debug_world(c.job, d_in, n, d_out, nullptr))) != RD_OK) return st;
  if (n) (void)hipMemcpy(out, d_out, 12ull * n, hipMemcpyDeviceToHost);
  return finish(hipSuccess);
}

rd_status rd_debug_philox(const uint32_t *in, uint32_t n, uint32_t *out) {
  if (!in || !out) return RD_ERR_INVALID_ARGUMENT;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) {
    (void)hipGetLastError();
    rdh::set_global_error("no HIP device available");
    return RD_ERR_DEVICE;
  }
  DebugCtx c;
  const uint32_t *d_in = c.dev(in, 6ull * n);
  uint32_t *d_out = c.dev<uint32_t>(nullptr, 4ull * n);
  if (!d_in || !d_out) return RD_ERR_OUT_OF_MEMORY;
  rd_status st = finish(rdh::launch_debug_philox(d_in, n, d_out, nullptr));
  if (st != RD_OK) return st;
  if (n) (void)hipMemcpy(out, d_out, 16ull * n, hipMemcpyDeviceToHost);
  return finish(hipSuccess);
}

rd_status rd_debug_copy_calibration(uint64_t bytes, uint32_t reps, float *ms) {
  if (bytes == 0 || bytes % 16 || bytes > (1ull << 35) || reps == 0 || reps > 1000)
    return RD_ERR_INVALID_ARGUMENT;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) {
    (void)hipGetLastError();
    rdh::set_global_error("no HIP device available");
    return RD_ERR_DEVICE;
  }
  DebugCtx c;
  uint8_t *src = c.dev<uint8_t>(nullptr, bytes), *dst = c.dev<uint8_t>(nullptr, bytes);
  if (!src || !dst) return RD_ERR_OUT_OF_MEMORY;
  if (hipMemset(src, 1, bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return finish(hipGetLastError());
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return finish(hipGetLastError());
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return finish(hipGetLastError());
  }
  hipError_t err = hipSuccess;
  for (uint32_t r = 0; r < reps && err == hipSuccess; ++r) {
    (void)hipEventRecord(e0, nullptr);
    err = rdh::launch_copy_calibration(src, dst, bytes, nullptr);
    (void)hipEventRecord(e1, nullptr);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float t = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
    if (ms) ms[r] = t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return finish(err);
}

rd_status rd_debug_host_bvh_info(const rd_scene *scene, int leaf, uint32_t *n_nodes,
                                 uint32_t *depth, char *err, size_t cap) {
  if (!scene || !n_nodes || !depth) return RD_ERR_INVALID_ARGUMENT;
  rdh::BuiltScene b;
  std::string e;
  bool ok = rdh::build_scene(*scene, leaf, &b, &e) && rdh::check_bvh(b, &e);
  *n_nodes = (uint32_t)b.nodes.size();
  *depth = b.depth;
  if (err && cap) {
    std::snprintf(err, cap, "%s", e.c_str());
  }
  return ok ? RD_OK : RD_ERR_INVALID_ARGUMENT;
}

rd_status rd_debug_host_scatter(const rd_scene *scene, const float *in, const uint32_t *bits,
                                uint32_t n, float *out) {
  if (!scene || !in || !bits || !out) return RD_ERR_INVALID_ARGUMENT;
  rdh::BuiltScene b;
  std::string e;
  if (!rdh::build_scene(*scene, 4, &b, &e)) {
    rdh::set_global_error(e);
    return RD_ERR_INVALID_ARGUMENT;
  }
  for (uint32_t i = 0; i < n; ++i) {
    const float *q = in + 16ull * i;
    uint32_t id;
    std::memcpy(&id, q + 7, 4);
    if (id >= scene->object_count) return RD_ERR_INVALID_ARGUMENT;
    const rdh::F3 o = rdh::f3(q[0], q[1], q[2]), d = rdh::f3(q[3], q[4], q[5]);
    const rdh::DevObject &ob = b.objects[id];
    const rdh::HitRec h = rdh::closest_hit(o, d, q[6], ob);
    rdh::FixedDraws dr{bits + (uint64_t)rdh::kDrawWords * i};
    const rdh::ScatterOut so = rdh::scatter(d, h, ob, dr, rdh::f3(q[8], q[9], q[10]),
                                            rdh::f3(q[11], q[12], q[13]));
    float *w = out + 16ull * i;
    const float v[16] = {so.o.x, so.o.y, so.o.z, so.d.x, so.d.y, so.d.z, so.atten.x, so.atten.y,
                         so.atten.z, so.light.x, so.light.y, so.light.z,
                         so.transmission ? 1.0f : 0.0f, h.n.x, h.n.y, h.n.z};
    std::memcpy(w, v, sizeof(v));
  }
  return RD_OK;
}

rd_status rd_debug_host_camera_rays(const rd_scene *scene, const uint32_t *pixels, uint32_t n,
                                    float *out) {
  if (!scene || !pixels || !out) return RD_ERR_INVALID_ARGUMENT;
  rdh::DeviceScene ds;
  rdh::DevJob j;
  rdh::fill_job(*scene, ds, nullptr, 0, 0, &j);
  const uint64_t np = (uint64_t)j.width * j.height;
  for (uint32_t i = 0; i < n; ++i) {
    if (pixels[i] >= np) return RD_ERR_INVALID_ARGUMENT;
    rdh::F3 o, d;
    rdh::camera_ray(j, pixels[i] % j.width, pixels[i] / j.width, &o, &d);
    const float v[6] = {o.x, o.y, o.z, d.x, d.y, d.z};
    std::memcpy(out + 6ull * i, v, sizeof(v));
  }
  return RD_OK;
}

rd_status rd_debug_host_trace(const rd_scene *scene, int leaf, const float *rays, uint32_t n,
                              int32_t *ids, float *ts) {
  if (!scene || !rays || !ids || !ts) return RD_ERR_INVALID_ARGUMENT;
  rdh::BuiltScene b;
  std::string e;
  if (!rdh::build_scene(*scene, leaf, &b, &e)) {
    rdh::set_global_error(e);
    return RD_ERR_INVALID_ARGUMENT;
  }
  for (uint32_t i = 0; i < n; ++i) ids[i] = rdh::host_trace(b, rays + 6ull * i, rays + 6ull * i + 3, ts + i);
  return RD_OK;
}

rd_status rd_debug_cache_classes(const uint64_t *bytes, uint32_t n, uint64_t *out) {
  if (n && (!bytes || !out)) return RD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n; ++i) out[i] = rdh::DeviceCache::size_class((size_t)bytes[i]);
  return RD_OK;
}

rd_status rd_debug_ring_ops(uint64_t capacity, const int64_t *ops, uint32_t n, int64_t *out) {
  if ((n && (!ops || !out)) || capacity == 0) return RD_ERR_INVALID_ARGUMENT;
  rdh::RadianceRing ring(capacity);
  uint32_t tickets = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (ops[i] >= 0) {
      uint64_t off = 0;
      out[i] = ring.alloc((uint64_t)ops[i], tickets, &off) ? (int64_t)off : -1;
      ++tickets;
    } else {
      ring.release((uint32_t)(-ops[i] - 1));
      out[i] = 0;
    }
  }
  return RD_OK;
}

rd_status rd_debug_host_bvh4_info(const rd_scene *scene, int leaf, uint32_t *n_nodes4,
                                  uint32_t *stack4, char *err, size_t cap) {
  if (!scene || !n_nodes4 || !stack4) return RD_ERR_INVALID_ARGUMENT;
  rdh::BuiltScene b;
  std::string e;
  bool ok = rdh::build_scene(*scene, leaf, &b, &e) && rdh::check_bvh4(b, &e);
  *n_nodes4 = (uint32_t)b.nodes4.size();
  *stack4 = b.stack4;
  if (err && cap) std::snprintf(err, cap, "%s", e.c_str());
  return ok ? RD_OK : RD_ERR_INVALID_ARGUMENT;
}

rd_status rd_debug_host_trace4(const rd_scene *scene, int leaf, const float *rays, uint32_t n,
                               int32_t *ids, float *ts) {
  if (!scene || !rays || !ids || !ts) return RD_ERR_INVALID_ARGUMENT;
  rdh::BuiltScene b;
  std::string e;
  if (!rdh::build_scene(*scene, leaf, &b, &e)) {
    rdh::set_global_error(e);
    return RD_ERR_INVALID_ARGUMENT;
  }
  for (uint32_t i = 0; i < n; ++i)
    ids[i] = rdh::host_trace4(b, rays + 6ull * i, rays + 6ull * i + 3, ts + i);
  return RD_OK;
}

}  // extern "C"
