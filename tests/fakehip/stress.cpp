// Host-code stress driver on the fake device (tests/fakehip): the library's
// engine, batch renderer, solo lane, autotuner and render server run
// unchanged against fake_runtime.cpp and fake_kernels.cpp, under
// AddressSanitizer + UndefinedBehaviorSanitizer or ThreadSanitizer, through
// the workloads of the benchmark:
//   plain     the per-scene renderer: frames over 1..4 pipes, several passes
//             per frame, progressive render_sample
//   batch     the 64-scene batch (seeds 1000-1063) with the default pipes and
//             solo lane, then with a 2-entry ticket table and a small
//             radiance ring (forced rewinds and wrap-around)
//   autotune  the 48-shape autotune on 8 of those scenes, then the batch
//   server    Poisson arrivals (20/s for 60 s, seed 7, scene seeds 2000+i,
//             8-spp passes), time compressed by --speed, from a submitting
//             thread while a waiting thread collects the results
//   bench     plain, batch + autotune, 64 sequential plain renders and a
//             server run in one process, the sequence of the benchmark run
//             whose GPU fault prompted this harness
// Every job's final image and accumulator carry the marks the fake kernels
// write (samples, seed), checked here; every violation the fake device
// records fails the run.  Test infrastructure only.
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/raydar_hip.h"
#include "../../include/raydar_hip_batch.h"
#include "fake_device.hpp"

namespace {

using Clock = std::chrono::steady_clock;
int g_fail = 0;

void fail(const std::string &m) {
  if (g_fail++ < 30) std::fprintf(stderr, "FAIL: %s\n", m.c_str());
}

// ---------------------------------------------------- scene recipes -------
// raydar_amd/scenegen.py: SplitMix64 draws, clustered scenes.
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform(double a = 0.0, double b = 1.0) {
    return a + (b - a) * ((double)(next() >> 40) * std::ldexp(1.0, -24));
  }
  int choice(int n) { return (int)(uniform() * n); }
  double log_uniform(double a, double b) { return std::exp(uniform(std::log(a), std::log(b))); }
  double normal() {
    const double u1 = 1.0 - uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

struct Scene {
  rd_scene s{};
  std::vector<rd_object> objs;
};

std::unique_ptr<Scene> clustered(uint64_t seed, int n, uint32_t w, uint32_t h) {
  auto sc = std::make_unique<Scene>();
  SplitMix64 r{seed};
  double centres[16][3];
  for (auto &c : centres) {
    c[0] = r.uniform(-30, 30);
    c[1] = r.uniform(0, 20);
    c[2] = r.uniform(-30, 30);
  }
  const double lo[3] = {-30, 0, -30}, hi[3] = {30, 20, 30};
  for (int i = 0; i < n; ++i) {
    rd_object o{};
    o.kind = i % 4 == 3 ? RD_CUBE : RD_SPHERE;
    o.size = (float)r.log_uniform(0.05, 1.0);
    const double *c = centres[r.choice(16)];
    for (int a = 0; a < 3; ++a)
      o.center[a] = (float)std::min(std::max(c[a] + 3.0 * r.normal(), lo[a]), hi[a]);
    const double u = r.uniform();
    const int cls = u < 0.4 ? 0 : u < 0.6 ? 1 : u < 0.8 ? 2 : 3;
    o.ior = 1.5f;
    if (cls == 0) {
      for (float &v : o.albedo) v = (float)r.uniform(0.1, 0.9);
      o.roughness = 1.0f;
    } else if (cls == 2) {
      for (float &v : o.albedo) v = (float)r.uniform(0.5, 1.0);
      o.roughness = (float)std::sqrt(r.uniform(0.0, 0.3));
      o.metallic = 1.0f;
    } else {
      o.albedo[0] = o.albedo[1] = o.albedo[2] = 1.0f;
      o.transmission = cls == 3 ? 1.0f : 0.0f;
    }
    sc->objs.push_back(o);
  }
  rd_camera &cam = sc->s.camera;
  const float pos[3] = {0, 25, 75}, tgt[3] = {0, 10, 0};
  std::memcpy(cam.position, pos, sizeof(pos));
  std::memcpy(cam.target, tgt, sizeof(tgt));
  cam.up[1] = 1.0f;
  cam.resolution_x = w;
  cam.resolution_y = h;
  cam.projection = RD_PROJECTION_PERSPECTIVE;
  cam.projection_param = 40.0f;
  cam.near_clip = 0.01f;
  cam.far_clip = 1000.0f;
  if (rd_camera_update_matrices(&cam) != RD_OK) fail("camera matrices");
  sc->s.world.kind = RD_WORLD_SKY_COLOR;
  const float top[3] = {0.5f, 0.7f, 1.0f}, bottom[3] = {1, 1, 1};
  std::memcpy(sc->s.world.top_color, top, sizeof(top));
  std::memcpy(sc->s.world.bottom_color, bottom, sizeof(bottom));
  sc->s.object_count = (uint32_t)sc->objs.size();
  sc->s.objects = sc->objs.data();
  return sc;
}

struct JobSpec {
  std::unique_ptr<Scene> scene;
  uint32_t spp, bounces;
  uint64_t seed;
  double t = 0;  // arrival time (serving)
  uint64_t paths() const {
    return (uint64_t)scene->s.camera.resolution_x * scene->s.camera.resolution_y * spp;
  }
};

std::vector<JobSpec> batch_jobs(uint64_t first = 1000, int count = 64) {
  const uint32_t res[5][2] = {{128, 128}, {256, 256}, {512, 512}, {1024, 768}, {1920, 1080}};
  const uint32_t spps[3] = {8, 32, 128}, bounces[3] = {4, 8, 12};
  std::vector<JobSpec> out;
  for (int i = 0; i < count; ++i) {
    const uint64_t seed = first + i;
    SplitMix64 r{seed ^ 0xB47C4};
    const int n = (int)std::lround(r.log_uniform(16, 2048));
    const auto *wh = res[r.choice(5)];
    const uint32_t spp = spps[r.choice(3)], nb = bounces[r.choice(3)];
    out.push_back(JobSpec{clustered(seed, n, wh[0], wh[1]), spp, nb, seed});
  }
  return out;
}

std::vector<JobSpec> serve_arrivals(double rate, double duration, uint64_t seed = 7,
                                    uint64_t first = 2000) {
  const uint32_t res[4][2] = {{256, 256}, {512, 512}, {1024, 768}, {1920, 1080}};
  const uint32_t spps[3] = {32, 64, 128};
  SplitMix64 r{seed};
  std::vector<JobSpec> out;
  double t = 0;
  for (int i = 0;; ++i) {
    t += -std::log(1.0 - r.uniform()) / rate;
    if (t >= duration) break;
    const uint64_t s = first + i;
    SplitMix64 sr{s ^ 0x5E4E};
    const int n = (int)std::lround(sr.log_uniform(16, 2048));
    const auto *wh = res[sr.choice(4)];
    const uint32_t spp = spps[sr.choice(3)];
    JobSpec j{clustered(s, n, wh[0], wh[1]), spp, 12, s};
    j.t = t;
    out.push_back(std::move(j));
  }
  return out;
}

// ------------------------------------------------------ outputs ----------
struct Out {
  std::vector<uint8_t> rgba;
  std::vector<float> acc;
  std::vector<uint8_t> progress;
};

rd_job_desc desc(const JobSpec &j, Out *o, bool acc, bool progress, uint32_t pass = 0) {
  const uint64_t npix = (uint64_t)j.scene->s.camera.resolution_x * j.scene->s.camera.resolution_y;
  o->rgba.assign(4 * npix, 0);
  if (acc) o->acc.assign(4 * npix, -1.0f);
  if (progress) o->progress.assign(4 * npix, 0);
  rd_job_desc d{};
  d.scene = &j.scene->s;
  d.samples = j.spp;
  d.max_bounces = j.bounces;
  d.seed = j.seed;
  d.pass_samples = pass;
  d.rgba = o->rgba.data();
  d.accumulator = acc ? o->acc.data() : nullptr;
  d.progress_rgba = progress ? o->progress.data() : nullptr;
  return d;
}

void check_out(const JobSpec &j, const Out &o, const std::string &what) {
  const uint8_t want[4] = {(uint8_t)j.spp, (uint8_t)j.seed, (uint8_t)(j.seed >> 8), 0xA5};
  if (o.rgba.size() < 4 || std::memcmp(o.rgba.data(), want, 4) != 0) {
    char b[200];
    std::snprintf(b, sizeof(b), "%s seed %llu: image mark %u %u %u %u, want %u %u %u %u",
                  what.c_str(), (unsigned long long)j.seed, o.rgba[0], o.rgba[1], o.rgba[2],
                  o.rgba[3], want[0], want[1], want[2], want[3]);
    fail(b);
  }
  if (!o.acc.empty() && (o.acc[0] != (float)j.spp || o.acc[1] != (float)(j.seed & 0xffff)))
    fail(what + " seed " + std::to_string(j.seed) + ": accumulator mark " +
         std::to_string(o.acc[0]) + " " + std::to_string(o.acc[1]));
  if (!o.progress.empty() && std::memcmp(o.progress.data(), want, 4) != 0)
    fail(what + " seed " + std::to_string(j.seed) + ": last progress image is not the final one");
}

bool device_clean(const char *phase) {
  const auto e = fakehip::errors();
  if (e.empty()) return true;
  std::fprintf(stderr, "fake device violations after %s: %zu\n", phase, e.size());
  for (size_t i = 0; i < e.size() && i < 20; ++i) std::fprintf(stderr, "  %s\n", e[i].c_str());
  g_fail += (int)e.size();
  return false;
}

// --------------------------------------------------------- scenarios -----
void plain(int frames) {
  auto jobs = batch_jobs(1000, 8);
  for (int pipes : {0, 1, 2, 3, 4}) {
    rd_config cfg{0, 12};
    rd_renderer *r = nullptr;
    if (rd_create(&cfg, &r) != RD_OK) return fail("rd_create");
    rd_set_option(r, "pipes", pipes);
    rd_set_option(r, "queue", 1 << 20);
    if (pipes == 3) rd_set_option(r, "l_capacity", 1 << 21);  // several passes per frame
    for (int f = 0; f < frames; ++f) {
      const JobSpec &j = jobs[(size_t)(f + pipes) % jobs.size()];
      rd_set_max_sample_count(r, j.spp);
      rd_set_max_bounces(r, j.bounces);
      rd_set_seed(r, j.seed);
      Out o;
      o.rgba.assign(4ull * j.scene->s.camera.resolution_x * j.scene->s.camera.resolution_y, 0);
      if (rd_render_frame(r, &j.scene->s, o.rgba.data(), o.rgba.size()) != RD_OK)
        fail(std::string("render_frame: ") + rd_last_error(r));
      o.acc.assign(o.rgba.size(), 0.0f);
      if (rd_read_accumulator(r, o.acc.data(), o.acc.size()) != RD_OK) fail("read_accumulator");
      check_out(j, o, "plain pipes " + std::to_string(pipes));
    }
    // the editor's loop: one pass of 3 samples per call
    const JobSpec &j = jobs[0];
    rd_set_max_sample_count(r, 10);
    rd_set_seed(r, j.seed);
    rd_set_samples_per_pass(r, 3);
    if (rd_new_frame(r, &j.scene->s) != RD_OK) fail("new_frame");
    std::vector<uint8_t> img(4ull * j.scene->s.camera.resolution_x * j.scene->s.camera.resolution_y);
    int produced = 1, calls = 0;
    while (produced) {
      if (rd_render_sample(r, nullptr, img.data(), img.size(), &produced) != RD_OK) {
        fail(std::string("render_sample: ") + rd_last_error(r));
        break;
      }
      calls += produced;
    }
    if (calls != 4 || img[0] != 10) fail("progressive passes");
    rd_destroy(r);
  }
}

void run_batch(rd_batch *b, std::vector<JobSpec> &jobs, const std::string &what, bool progress) {
  std::printf("  %s\n", what.c_str());
  std::fflush(stdout);
  std::vector<Out> outs(jobs.size());
  std::vector<rd_job_desc> d;
  for (size_t i = 0; i < jobs.size(); ++i)
    d.push_back(desc(jobs[i], &outs[i], true, progress && i % 5 == 0, progress ? 8 : 0));
  if (rd_batch_render(b, d.data(), (uint32_t)d.size()) != RD_OK)
    return fail(what + ": " + rd_batch_last_error(b));
  for (size_t i = 0; i < jobs.size(); ++i) check_out(jobs[i], outs[i], what);
}

void batch() {
  auto jobs = batch_jobs();
  {
    rd_batch *b = nullptr;
    if (rd_batch_create(&b) != RD_OK) return fail("batch create");
    run_batch(b, jobs, "batch default", false);
    run_batch(b, jobs, "batch default again", true);
    rd_batch_destroy(b);
  }
  struct Variant {
    const char *name;
    std::vector<std::pair<const char *, int64_t>> opts;
  } variants[] = {
      {"pipes1", {{"pipes", 1}}},
      {"pipes4 no graph", {{"pipes", 4}, {"graph", 0}}},
      {"ticket_cap 2", {{"ticket_cap", 2}, {"solo_paths", 0}}},
      {"small ring", {{"l_capacity", 1 << 22}, {"queue", 1 << 16}, {"solo_paths", 0}}},
      {"solo all", {{"solo_paths", 1}}},
      {"binary trees", {{"bvh4", 0}, {"block", 128}}},
  };
  for (const Variant &v : variants) {
    rd_batch *b = nullptr;
    if (rd_batch_create(&b) != RD_OK) return fail("batch create");
    for (const auto &kv : v.opts)
      if (rd_batch_set_option(b, kv.first, kv.second) != RD_OK)
        fail(std::string(v.name) + ": option " + kv.first + ": " + rd_batch_last_error(b));
    run_batch(b, jobs, std::string("batch ") + v.name, true);
    rd_batch_destroy(b);
  }
  // a job whose single sample outgrows the radiance ring is refused, not left waiting
  rd_batch *b = nullptr;
  if (rd_batch_create(&b) != RD_OK) return fail("batch create");
  rd_batch_set_option(b, "l_capacity", 1 << 20);
  rd_batch_set_option(b, "solo_paths", 0);
  Out o;
  rd_job_desc d = desc(jobs[0], &o, false, false);
  auto big = clustered(5, 20, 1920, 1080);
  d.scene = &big->s;
  if (rd_batch_render(b, &d, 1) != RD_ERR_INVALID_ARGUMENT) fail("oversize job not refused");
  rd_batch_destroy(b);
}

void autotune() {
  auto jobs = batch_jobs();
  rd_batch *b = nullptr;
  if (rd_batch_create(&b) != RD_OK) return fail("batch create");
  std::vector<Out> outs(8);
  std::vector<rd_job_desc> d;
  for (int i = 0; i < 8; ++i) d.push_back(desc(jobs[i], &outs[i], false, false));
  rd_tune_result t{};
  if (rd_batch_autotune(b, d.data(), 8, &t) != RD_OK) fail(std::string("autotune: ") + rd_batch_last_error(b));
  if (t.configs_tried != 48) fail("autotune tried " + std::to_string(t.configs_tried));
  std::printf("autotune cached=%u block=%d leaf=%d queue=%u\n", t.cached, t.block, t.leaf, t.queue);
  // (with RAYDAR_HIP_TUNE_CACHE naming the file of an earlier run, both mixes
  // come from the file: the first call's cached flag above says which)
  // the same mix again on the same device: the kept result, nothing timed
  const uint64_t kernels = fakehip::stats().kernels;
  rd_tune_result t2{};
  if (rd_batch_autotune(b, d.data(), 8, &t2) != RD_OK) fail("second autotune");
  if (!t2.cached || t2.block != t.block || t2.leaf != t.leaf || t2.queue != t.queue)
    fail("second autotune of the same mix was not served from the cache");
  if (fakehip::stats().kernels != kernels) fail("cached autotune launched kernels");
  // another mix is tuned afresh (unless an earlier run's file holds it)
  rd_tune_result t3{};
  if (rd_batch_autotune(b, d.data(), 7, &t3) != RD_OK || t3.cached != t.cached)
    fail("a different mix: cached " + std::to_string(t3.cached));
  std::printf("autotune other mix cached=%u\n", t3.cached);
  run_batch(b, jobs, "batch after autotune", false);
  rd_batch_destroy(b);
}

void server(double seconds, double speed, uint64_t budget) {
  auto arr = serve_arrivals(20.0, seconds);
  rd_server_config cfg{budget, 8, 0, 0, 0};
  rd_server *s = nullptr;
  if (rd_server_create(&cfg, &s) != RD_OK) return fail(std::string("server create: ") + rd_last_error(nullptr));
  std::vector<Out> outs(arr.size());
  std::vector<uint64_t> ids(arr.size(), 0);
  std::mutex mu;
  std::condition_variable cv;
  size_t submitted = 0;
  std::thread waiter([&] {
    for (size_t k = 0; k < arr.size(); ++k) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return submitted > k; });
      }
      if (!ids[k]) continue;
      rd_job_result r{};
      if (rd_server_wait(s, ids[k], -1, &r) != RD_OK || !r.done || r.status != RD_OK) {
        fail("server job " + std::to_string(k) + " status " + std::to_string(r.status));
        continue;
      }
      check_out(arr[k], outs[k], "server");
      outs[k] = Out();  // release the host image, as the benchmark's client does
    }
  });
  const auto t0 = Clock::now();
  for (size_t k = 0; k < arr.size(); ++k) {
    const auto due = t0 + std::chrono::duration_cast<Clock::duration>(
                              std::chrono::duration<double>(arr[k].t / speed));
    std::this_thread::sleep_until(due);
    rd_job_desc d = desc(arr[k], &outs[k], false, k % 7 == 0);
    uint64_t id = 0;
    if (rd_server_submit(s, &d, &id) != RD_OK) fail(std::string("submit: ") + rd_server_last_error(s));
    {
      std::lock_guard<std::mutex> lk(mu);
      ids[k] = id;
      submitted = k + 1;
    }
    cv.notify_all();
  }
  waiter.join();
  rd_server_stats st{};
  rd_server_stats_get(s, &st);
  rd_server_destroy(s);
  std::printf("server: %zu requests, %llu done, max resident %u, %.1f s wall\n", arr.size(),
              (unsigned long long)st.jobs_done, st.max_resident,
              std::chrono::duration<double>(Clock::now() - t0).count());
  if (st.jobs_done != arr.size()) fail("server finished " + std::to_string(st.jobs_done) + " jobs");
}

// rd_render_frame_noalloc: every device buffer from the static array; no
// allocation or free on the fake device during a call; the same marks as
// rd_render_frame; a scene beyond the array is refused.
void noalloc() {
  auto jobs = batch_jobs(1000, 16);
  rd_config cfg{1, 12};
  rd_renderer *r = nullptr;
  if (rd_create(&cfg, &r) != RD_OK) return fail("rd_create");
  for (const JobSpec &j : jobs) {
    rd_set_max_sample_count(r, j.spp);
    rd_set_max_bounces(r, j.bounces);
    rd_set_seed(r, j.seed);
    Out o;
    o.rgba.assign(4ull * j.scene->s.camera.resolution_x * j.scene->s.camera.resolution_y, 0);
    const auto before = fakehip::stats();
    if (rd_render_frame_noalloc(r, &j.scene->s, o.rgba.data(), o.rgba.size()) != RD_OK) {
      fail(std::string("render_frame_noalloc: ") + rd_last_error(r));
      continue;
    }
    const auto after = fakehip::stats();
    if (after.mallocs != before.mallocs || after.frees != before.frees)
      fail("render_frame_noalloc allocated or freed device memory");
    if (rd_sample_count(r) != j.spp) fail("render_frame_noalloc sample count");
    check_out(j, o, "noalloc");
  }
  auto huge = clustered(9, 50, 4096, 4096);
  std::vector<uint8_t> img(4ull * 4096 * 4096);
  if (rd_render_frame_noalloc(r, &huge->s, img.data(), img.size()) != RD_ERR_INVALID_ARGUMENT)
    fail("render_frame_noalloc accepted a scene beyond its static array");
  rd_destroy(r);
}

void sequential_plain() {
  auto jobs = batch_jobs();
  rd_config cfg{1, 12};
  rd_renderer *r = nullptr;
  if (rd_create(&cfg, &r) != RD_OK) return fail("rd_create");
  for (const JobSpec &j : jobs) {
    rd_set_max_sample_count(r, j.spp);
    rd_set_max_bounces(r, j.bounces);
    rd_set_seed(r, j.seed);
    Out o;
    o.rgba.assign(4ull * j.scene->s.camera.resolution_x * j.scene->s.camera.resolution_y, 0);
    if (rd_render_frame(r, &j.scene->s, o.rgba.data(), o.rgba.size()) != RD_OK)
      fail(std::string("sequential render_frame: ") + rd_last_error(r));
    check_out(j, o, "sequential plain");
  }
  rd_destroy(r);
}

}  // namespace

int main(int argc, char **argv) {
  std::string scenario = "all";
  double seconds = 60.0, speed = 20.0;
  int delay = 40;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--seconds" && i + 1 < argc) seconds = std::atof(argv[++i]);
    else if (a == "--speed" && i + 1 < argc) speed = std::atof(argv[++i]);
    else if (a == "--delay-us" && i + 1 < argc) delay = std::atoi(argv[++i]);
    else scenario = a;
  }
  fakehip::set_max_delay_us(delay);
  const auto t0 = Clock::now();
  auto phase = [&](const char *name, auto fn) {
    const auto p0 = Clock::now();
    fn();
    std::printf("%s done in %.1f s\n", name, std::chrono::duration<double>(Clock::now() - p0).count());
    std::fflush(stdout);
    device_clean(name);
  };
  if (scenario == "plain" || scenario == "all") phase("plain", [] { plain(3); });
  if (scenario == "noalloc" || scenario == "all") phase("noalloc", [] { noalloc(); });
  if (scenario == "batch" || scenario == "all") phase("batch", [] { batch(); });
  if (scenario == "autotune" || scenario == "all") phase("autotune", [] { autotune(); });
  if (scenario == "server" || scenario == "all")
    phase("server", [&] { server(seconds, speed, 200ull << 30); });
  if (scenario == "server_budget")  // admission waits: room for a few jobs only
    phase("server (small budget)", [&] { server(seconds, speed, 12ull << 30); });
  if (scenario == "bench") {
    phase("plain", [] { plain(2); });
    phase("autotune + batch", [] { autotune(); });
    phase("sequential plain", [] { sequential_plain(); });
    phase("server", [&] { server(seconds, speed, 200ull << 30); });
  }
  const auto st = fakehip::stats();
  std::printf("fake device: %llu ops, %llu kernels, %llu copies, %llu graph launches, %llu mallocs, "
              "%llu frees, %llu streams; %.1f s\n",
              (unsigned long long)st.ops, (unsigned long long)st.kernels, (unsigned long long)st.copies,
              (unsigned long long)st.graph_launches, (unsigned long long)st.mallocs,
              (unsigned long long)st.frees, (unsigned long long)st.streams,
              std::chrono::duration<double>(Clock::now() - t0).count());
  if (g_fail) {
    std::printf("stress FAILED: %d problems\n", g_fail);
    return 1;
  }
  std::printf("stress ok\n");
  return 0;
}
