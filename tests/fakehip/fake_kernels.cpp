// Fake kernels for the host-code harness: the launchers of kernels.hpp, run
// on the fake device (fake_runtime.cpp) in stream order.  They do not trace
// rays: they keep the wavefront's bookkeeping in aggregate (each queue as
// runs of paths of one ticket at one bounce), finish half of every run per
// bounce and all of it at the job's bounce limit, and write exactly the
// device state the host code reads back (queue counters, cursors, per-ticket
// finished counts).  On the way they check what the real kernels would
// touch:
//   - every buffer an iteration uses (queues, hits, control block, tickets,
//     job table, radiance arena) is a live allocation of the stated size
//   - the queue counters match the paths in flight; work indices fall inside
//     their tickets; tickets lie in the window the host published
//   - a job's scene buffers (nodes, primitives, objects) and accumulator are
//     live, were uploaded, and were not rewritten after the job's record was
//     (a buffer reused while the job is in flight); the job's record was not
//     rewritten after a ticket that refers to it was uploaded (a reused slot)
//   - a ticket's radiance slots are inside the arena and no other ticket
//     writes them between the first path's write and their accumulation
//   - passes are accumulated in sample order into a zeroed accumulator, and
//     a tonemap sees exactly the samples accumulated so far
// Accumulate writes (samples so far, seed & 0xffff) into the first two
// floats of the accumulator, and tonemap writes (spp, seed lo, seed hi,
// 0xA5) into the first pixel, so the driver can check every job's results.
// Test infrastructure only.
#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../raydar_amd/csrc/kernels.hpp"
#include "fake_device.hpp"

namespace rdh {

namespace {

struct Run {
  uint32_t ticket, bounce;
  uint64_t count;
};
struct LRange {
  uint64_t off, len;
  bool complete;
  uint32_t job, sample_base;
};
typedef std::tuple<const void *, uint32_t, uint64_t, uint32_t, uint32_t> TicketKey;

std::mutex g_k;  // fake-kernel state
std::map<const void *, std::vector<Run>> g_queue;      // by PathQueue::word
std::map<TicketKey, uint64_t> g_finished;             // paths finished per issued ticket
std::map<const void *, std::vector<LRange>> g_lmap;   // radiance ranges in use, by arena
std::map<const void *, uint64_t> g_acc;               // samples accumulated, by accumulator
std::once_flag g_hooked;

void err(const char *kernel, const std::string &m) { fakehip::error(std::string(kernel) + ": " + m); }

bool need(const void *p, uint64_t bytes, const char *kernel, const char *what) {
  std::string why;
  if (fakehip::live(p, bytes, &why)) return true;
  err(kernel, std::string(what) + ": " + why);
  return false;
}

void hook() {
  std::call_once(g_hooked, [] { fakehip::on_free([](const void *base, size_t size) {
    std::lock_guard<std::mutex> g(g_k);
    const char *b = (const char *)base, *e = b + size;
    auto in = [&](const void *p) { return (const char *)p >= b && (const char *)p < e; };
    for (auto it = g_queue.begin(); it != g_queue.end();)
      it = in(it->first) ? g_queue.erase(it) : std::next(it);
    for (auto it = g_lmap.begin(); it != g_lmap.end();) {
      if (in(it->first)) {
        for (const LRange &r : it->second)
          err("hipFree", "radiance arena freed with slots of job slot " + std::to_string(r.job) +
                             " not yet accumulated");
        it = g_lmap.erase(it);
      } else {
        ++it;
      }
    }
    for (auto it = g_acc.begin(); it != g_acc.end();) it = in(it->first) ? g_acc.erase(it) : std::next(it);
    for (auto it = g_finished.begin(); it != g_finished.end();)
      it = in(std::get<0>(it->first)) ? g_finished.erase(it) : std::next(it);
  }); });
}

struct Plan {
  uint32_t n_old = 0;
  uint64_t n_new = 0, cursor = 0;
};

Plan plan(const IterArgs &a, int p) {
  Plan pl;
  for (int s = 0; s < kSegments; ++s) {
    const uint32_t c = a.ctl->count[p][s * kSegStride];
    pl.n_old += c < a.seg_cap ? c : a.seg_cap;
  }
  pl.cursor = a.ctl->cursor[p];
  const unsigned long long avail = a.ctl->total - pl.cursor;
  const unsigned long long room = (unsigned long long)(a.capacity - pl.n_old);
  pl.n_new = avail < room ? avail : room;
  return pl;
}

const PathQueue &queue_of(const IterArgs &a, int p) { return p ? a.q[1] : a.q[0]; }

// Table sizes as allocated (from the device registry, not from what the host
// passes in IterArgs).
uint64_t ticket_cap(const IterArgs &a) { return fakehip::bytes_from(a.tickets) / sizeof(DevTicket); }
uint64_t job_cap(const IterArgs &a) { return fakehip::bytes_from(a.jobs) / sizeof(DevJob); }
uint64_t l_cap(const IterArgs &a) { return a.L ? fakehip::bytes_from(a.L) / 16 : 0; }

bool check_args(const IterArgs &a, const LaunchShape &sh, const char *k) {
  bool ok = need(a.ctl, sizeof(QueueCtl), k, "control block");
  if (sh.grid <= 0 || (sh.block != 64 && sh.block != 128 && sh.block != 256)) {
    err(k, "launch shape " + std::to_string(sh.grid) + " x " + std::to_string(sh.block));
    ok = false;
  }
  // every queue array must hold the kSegments * seg_cap slots the kernels index
  const uint64_t slots = (uint64_t)kSegments * a.seg_cap;
  if (a.capacity > slots) {
    err(k, "queue capacity does not fit the segments");
    ok = false;
  }
  for (int p = 0; p < 2; ++p) {
    const PathQueue &q = a.q[p];
    ok = need(q.ray, sizeof(QRay) * slots, k, "queue array") && ok;
    ok = need(q.dir, sizeof(QDir) * slots, k, "queue array") && ok;
    ok = need(q.word, sizeof(QWord) * slots, k, "queue array") && ok;
    ok = need(q.att, sizeof(QAtt) * slots, k, "queue array") && ok;
    ok = need(q.light, sizeof(QLight) * slots, k, "queue array") && ok;
  }
  ok = need(a.hit_t, 4ull * slots, k, "hit distances") && ok;
  ok = need(a.hit_id, 4ull * slots, k, "hit ids") && ok;
  ok = need(a.tickets, sizeof(DevTicket), k, "ticket table") && ok;
  ok = need(a.jobs, sizeof(DevJob), k, "job table") && ok;
  if (a.L) ok = need(a.L, 16, k, "radiance arena") && ok;
  if (a.track_done) ok = need(a.ticket_done, 8ull * ticket_cap(a), k, "ticket progress") && ok;
  if (a.first_hit_mode != 0) {
    // camera-ray hits: one single-job scene's pixels, whose camera paths the
    // first iteration must all generate (the queue holds a whole sample)
    if (!a.single_job || a.vk || a.first_hit_mode > 2) {
      err(k, "camera-ray hits outside a single-job CPU-semantics pass");
      ok = false;
    } else if (need(a.jobs, sizeof(DevJob), k, "job table")) {
      const uint64_t npix = (uint64_t)a.jobs[0].width * a.jobs[0].height;
      ok = need(a.first_hit, 8ull * npix, k, "camera-ray hits") && ok;
      if (a.capacity < npix) {
        err(k, "camera-ray hits with a queue smaller than one sample");
        ok = false;
      }
    }
  }
  return ok;
}

// What the kernels read through ticket `ti`.
bool check_ticket(const IterArgs &a, uint32_t ti, const char *k) {
  if (ti >= ticket_cap(a)) {
    err(k, "ticket " + std::to_string(ti) + " beyond the table");
    return false;
  }
  const DevTicket &T = a.tickets[ti];
  const uint32_t js = a.single_job ? 0u : T.job;
  if (js >= job_cap(a)) {
    err(k, "job slot " + std::to_string(js) + " beyond the table");
    return false;
  }
  const DevJob &J = a.jobs[js];
  const uint64_t e_ticket = fakehip::last_write(&T, sizeof(DevTicket));
  const uint64_t e_job = fakehip::last_write(&J, sizeof(DevJob));
  bool ok = true;
  if (!e_ticket) err(k, "ticket " + std::to_string(ti) + " was never uploaded"), ok = false;
  if (!e_job) err(k, "job slot " + std::to_string(js) + " was never uploaded"), ok = false;
  if (e_job > e_ticket)
    err(k, "job slot " + std::to_string(js) + " rewritten after ticket " + std::to_string(ti) +
               " that uses it was uploaded (slot reused while its work was in flight)"),
        ok = false;
  struct Buf {
    const void *p;
    uint64_t bytes;
    const char *what;
  } bufs[] = {{J.nodes, 64ull * J.n_nodes, "BVH nodes"},
              {J.nodes4, 128ull * J.n_nodes4, "4-wide BVH nodes"},
              {J.prims, 32ull * J.n_prims, "primitives"},
              {J.objects, 64ull * J.n_prims, "objects"}};
  for (const Buf &b : bufs) {
    if (!b.bytes) continue;
    if (!need(b.p, b.bytes, k, b.what)) {
      ok = false;
      continue;
    }
    const uint64_t e = fakehip::last_write(b.p, b.bytes);
    if (!e) err(k, std::string(b.what) + " of job slot " + std::to_string(js) + " never uploaded"), ok = false;
    if (e > e_job)
      err(k, std::string(b.what) + " of job slot " + std::to_string(js) +
                 " rewritten after the job's record (buffer reused while the job is in flight)"),
          ok = false;
  }
  ok = need(J.acc, 16ull * J.width * J.height, k, "accumulator") && ok;
  if (a.L && T.l_offset + T.paths > l_cap(a))
    err(k, "ticket " + std::to_string(ti) + " radiance slots beyond the arena"), ok = false;
  if (J.max_bounces == 0 || J.max_bounces > 255) err(k, "job bounce limit out of range"), ok = false;
  return ok;
}

// Largest ticket in [lo, n) with work_begin <= w, as the kernels search.
uint32_t find_ticket(const DevTicket *t, uint32_t lo, uint32_t n, uint64_t w) {
  uint32_t hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (t[mid].work_begin <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

void fake_generate(const IterArgs a, int p, const LaunchShape sh) {
  fakehip::count_kernel();
  std::lock_guard<std::mutex> g(g_k);
  if (!check_args(a, sh, "generate")) return;
  const Plan pl = plan(a, p);
  std::vector<Run> &runs = g_queue[queue_of(a, p).word];
  uint64_t inflight = 0;
  for (const Run &r : runs) inflight += r.count;
  if (inflight != pl.n_old)
    err("generate", "queue counters say " + std::to_string(pl.n_old) + " paths, " +
                        std::to_string(inflight) + " are in flight");
  if (!pl.n_new) return;
  const uint32_t lo = a.ctl->ticket_lo, n = a.ctl->n_tickets;
  if (n > ticket_cap(a) || lo >= n) {
    err("generate", "ticket window [" + std::to_string(lo) + ", " + std::to_string(n) + ") invalid");
    return;
  }
  uint64_t w = pl.cursor, left = pl.n_new;
  uint32_t ti = a.single_job ? 0u : find_ticket(a.tickets, lo, n, w);
  while (left) {
    if (ti >= n) {
      err("generate", "work index " + std::to_string(w) + " beyond the last ticket");
      return;
    }
    const DevTicket &T = a.tickets[ti];
    if (w < T.work_begin || w >= T.work_begin + T.paths) {
      err("generate", "work index " + std::to_string(w) + " outside its ticket " + std::to_string(ti));
      return;
    }
    check_ticket(a, ti, "generate");
    const uint64_t take = std::min<uint64_t>(left, T.work_begin + T.paths - w);
    runs.push_back(Run{ti, 0, take});
    w += take;
    left -= take;
    ++ti;
  }
}

void fake_extend(const IterArgs a, int p, const LaunchShape sh) {
  fakehip::count_kernel();
  std::lock_guard<std::mutex> g(g_k);
  if (!check_args(a, sh, "extend")) return;
  const Plan pl = plan(a, p);
  const std::vector<Run> &runs = g_queue[queue_of(a, p).word];
  uint64_t inflight = 0;
  std::set<uint32_t> seen;
  for (const Run &r : runs) {
    inflight += r.count;
    if (seen.insert(r.ticket).second) check_ticket(a, r.ticket, "extend");
  }
  if (inflight != pl.n_old + pl.n_new)
    err("extend", "iteration covers " + std::to_string(pl.n_old + pl.n_new) + " paths, " +
                      std::to_string(inflight) + " are queued");
  // queue[1-p] is refilled by shade; its previous contents were consumed
  g_queue[queue_of(a, 1 - p).word].clear();
  for (int s = 0; s < kSegments; ++s) a.ctl->count[1 - p][s * kSegStride] = 0;
  a.ctl->cursor[1 - p] = pl.cursor + pl.n_new;
}

void finish_paths(const IterArgs &a, uint32_t ti, uint64_t fin) {
  const DevTicket &T = a.tickets[ti];
  const TicketKey key{a.tickets, ti, T.work_begin, T.job, T.sample_base};
  uint64_t &done = g_finished[key];
  if (a.L) {
    std::vector<LRange> &ranges = g_lmap[a.L];
    if (done == 0) {
      for (const LRange &r : ranges)
        if (r.off < T.l_offset + T.paths && T.l_offset < r.off + r.len)
          err("shade", "ticket " + std::to_string(ti) + " (job slot " + std::to_string(T.job) +
                           ") writes radiance slots [" + std::to_string(T.l_offset) + ", " +
                           std::to_string(T.l_offset + T.paths) + ") that job slot " +
                           std::to_string(r.job) + " sample " + std::to_string(r.sample_base) +
                           " has not accumulated yet");
      ranges.push_back(LRange{T.l_offset, T.paths, false, T.job, T.sample_base});
    }
    done += fin;
    if (done > T.paths) err("shade", "ticket " + std::to_string(ti) + " finished more paths than it has");
    if (done == T.paths) {
      for (LRange &r : ranges)
        if (r.off == T.l_offset && r.len == T.paths && !r.complete) {
          r.complete = true;
          break;
        }
    }
  } else {
    done += fin;
  }
  if (done >= T.paths) g_finished.erase(key);
  if (a.track_done) a.ticket_done[ti] += fin;
}

void fake_shade(const IterArgs a, int p, const LaunchShape sh) {
  fakehip::count_kernel();
  std::lock_guard<std::mutex> g(g_k);
  if (!check_args(a, sh, "shade")) return;
  // LDS-resident objects: only for one job, and exactly that job's objects
  // (the kernel copies lds_obj_bytes from its object array)
  if (a.lds_obj_bytes &&
      (!a.single_job || a.lds_obj_bytes != sizeof(DevObject) * (uint64_t)a.jobs[0].n_prims ||
       a.lds_obj_bytes > 65536))
    err("shade", "LDS object copy of " + std::to_string(a.lds_obj_bytes) + " bytes does not match the job");
  const Plan pl = plan(a, p);
  std::vector<Run> &runs = g_queue[queue_of(a, p).word];
  std::vector<Run> &next = g_queue[queue_of(a, 1 - p).word];
  uint64_t inflight = 0, surv_total = 0;
  std::set<uint32_t> seen;
  for (const Run &r : runs) {
    inflight += r.count;
    if (seen.insert(r.ticket).second) check_ticket(a, r.ticket, "shade");
  }
  if (inflight != pl.n_old + pl.n_new)
    err("shade", "iteration covers " + std::to_string(pl.n_old + pl.n_new) + " paths, " +
                     std::to_string(inflight) + " are queued");
  for (const Run &r : runs) {
    if (r.ticket >= ticket_cap(a)) continue;  // reported by check_ticket
    const DevTicket &T = a.tickets[r.ticket];
    const DevJob &J = a.jobs[a.single_job ? 0u : T.job];
    const uint64_t fin = r.bounce + 1 >= J.max_bounces ? r.count : (r.count + 1) / 2;
    const uint64_t surv = r.count - fin;
    if (fin) finish_paths(a, r.ticket, fin);
    if (surv) {
      if (!next.empty() && next.back().ticket == r.ticket && next.back().bounce == r.bounce + 1)
        next.back().count += surv;
      else
        next.push_back(Run{r.ticket, r.bounce + 1, surv});
      surv_total += surv;
    }
  }
  runs.clear();
  // survivors spread over the segments as the workgroups append them
  const uint64_t per = (surv_total + kSegments - 1) / kSegments;
  uint64_t left = surv_total;
  for (int s = 0; s < kSegments; ++s) {
    const uint64_t c = std::min(left, per);
    if (c > a.seg_cap) a.ctl->overflow = 1;
    a.ctl->count[1 - p][s * kSegStride] += (uint32_t)c;
    left -= c;
  }
}

void fake_accumulate(const DevJob *job, const DevTicket T, const float4 *L) {
  fakehip::count_kernel();
  std::lock_guard<std::mutex> g(g_k);
  if (!need(job, sizeof(DevJob), "accumulate", "job record")) return;
  const DevJob &J = *job;
  const uint64_t npix = (uint64_t)J.width * J.height;
  if (!need(J.acc, 16 * npix, "accumulate", "accumulator")) return;
  if (L) {
    const uint64_t lo = T.l_offset, hi = T.l_offset + (uint64_t)T.spp * npix;
    if (!need(L + lo, 16 * (hi - lo), "accumulate", "radiance slots")) return;
    std::vector<LRange> &ranges = g_lmap[L];
    uint64_t covered = 0;
    bool ok = true;
    for (auto it = ranges.begin(); it != ranges.end();) {
      if (it->off < hi && lo < it->off + it->len) {
        if (it->off < lo || it->off + it->len > hi || !it->complete) {
          err("accumulate", "radiance slots [" + std::to_string(lo) + ", " + std::to_string(hi) +
                                ") of job sample " + std::to_string(T.sample_base) +
                                " overlap a range still being written or of another pass");
          ok = false;
          ++it;
          continue;
        }
        covered += it->len;
        it = ranges.erase(it);
      } else {
        ++it;
      }
    }
    if (ok && covered != hi - lo)
      err("accumulate", "pass of sample " + std::to_string(T.sample_base) + ": " +
                            std::to_string(covered) + " of " + std::to_string(hi - lo) +
                            " radiance slots were written before the accumulation");
  }
  uint64_t &samples = g_acc[J.acc];
  if (samples != T.sample_base)
    err("accumulate", "pass of samples [" + std::to_string(T.sample_base) + ", " +
                          std::to_string(T.sample_base + T.spp) + ") accumulated after " +
                          std::to_string(samples) + " samples (out of order)");
  samples = (uint64_t)T.sample_base + T.spp;
  J.acc[0] = (float)samples;
  J.acc[1] = (float)(J.seed & 0xffff);
}

}  // namespace

uint32_t extend_dynamic_lds(const IterArgs &a, int block) {
  if (a.refill <= 0 || a.lds_bvh_bytes == 0 || !a.single_job || a.wide_stack) return 0;
  return ((a.stack_rows * (uint32_t)block * 2u + 15u) & ~15u) + a.lds_bvh_bytes;
}

uint32_t shade_dynamic_lds(const IterArgs &a) { return a.single_job ? a.lds_obj_bytes : 0u; }

int occupancy_blocks(int which, int block, bool, uint32_t, uint32_t, bool, bool single) {
  const int waves = which == 1 ? 8 : which == 2 ? (single ? 7 : 5) : 8;
  return std::max(1, waves * 4 / (block / 64));
}

hipError_t launch_generate(const IterArgs &a, int parity, const LaunchShape &s, bool, hipStream_t st) {
  hook();
  fakehip::enqueue(st, [a, parity, s] { fake_generate(a, parity, s); });
  return hipSuccess;
}
hipError_t launch_extend(const IterArgs &a, int parity, const LaunchShape &s, bool, hipStream_t st) {
  hook();
  fakehip::enqueue(st, [a, parity, s] { fake_extend(a, parity, s); });
  return hipSuccess;
}
hipError_t launch_shade(const IterArgs &a, int parity, const LaunchShape &s, bool, hipStream_t st) {
  hook();
  fakehip::enqueue(st, [a, parity, s] { fake_shade(a, parity, s); });
  return hipSuccess;
}

hipError_t launch_accumulate(const DevJob *job, const DevTicket &t, const float4 *L, DevCounters *,
                             bool, int grid, hipStream_t st) {
  hook();
  if (grid <= 0) return hipErrorInvalidValue;
  const DevTicket tt = t;
  fakehip::enqueue(st, [job, tt, L] { fake_accumulate(job, tt, L); });
  return hipSuccess;
}

hipError_t launch_tonemap(const DevJob *job, uint32_t n, float spp, uchar4 *out, DevCounters *, bool,
                          int grid, hipStream_t st, bool) {
  hook();
  if (grid <= 0) return hipErrorInvalidValue;
  fakehip::enqueue(st, [job, n, spp, out] {
    fakehip::count_kernel();
    std::lock_guard<std::mutex> g(g_k);
    if (!need(job, sizeof(DevJob), "tonemap", "job record")) return;
    const DevJob &J = *job;
    if (!need(J.acc, 16ull * n, "tonemap", "accumulator") || !need(out, 4ull * n, "tonemap", "image"))
      return;
    const uint64_t samples = g_acc[J.acc];
    if ((uint64_t)spp != samples)
      err("tonemap", "image of " + std::to_string((uint64_t)spp) + " samples with " +
                         std::to_string(samples) + " accumulated");
    out[0] = make_uchar4((unsigned char)(uint64_t)spp, (unsigned char)J.seed,
                         (unsigned char)(J.seed >> 8), 0xA5);
  });
  return hipSuccess;
}

hipError_t launch_set_ctl(QueueCtl *ctl, unsigned long long total, uint32_t n, hipStream_t st) {
  hook();
  fakehip::enqueue(st, [ctl, total, n] {
    std::lock_guard<std::mutex> g(g_k);
    if (!need(ctl, sizeof(QueueCtl), "set_ctl", "control block")) return;
    for (int s = 0; s < kSegments; ++s) ctl->count[0][s * kSegStride] = ctl->count[1][s * kSegStride] = 0;
    ctl->overflow = ctl->violation = 0;
    ctl->cursor[0] = ctl->cursor[1] = 0;
    ctl->total = total;
    ctl->n_tickets = n;
    ctl->ticket_lo = 0;
  });
  return hipSuccess;
}

hipError_t launch_extend_work(QueueCtl *ctl, unsigned long long total, uint32_t n, uint32_t lo,
                              hipStream_t st) {
  hook();
  fakehip::enqueue(st, [ctl, total, n, lo] {
    std::lock_guard<std::mutex> g(g_k);
    if (!need(ctl, sizeof(QueueCtl), "extend_work", "control block")) return;
    ctl->total = total;
    ctl->n_tickets = n;
    ctl->ticket_lo = lo;
  });
  return hipSuccess;
}

hipError_t launch_zero_acc(float *acc, uint64_t n, hipStream_t st) {
  hook();
  fakehip::enqueue(st, [acc, n] {
    std::lock_guard<std::mutex> g(g_k);
    if (!need(acc, 4 * n, "zero", "accumulator")) return;
    g_acc[acc] = 0;
    std::memset(acc, 0, 4 * std::min<uint64_t>(n, 16));
  });
  return hipSuccess;
}

// parity hooks: not part of the engine under test
hipError_t launch_debug_trace(const DevNode *, const DevPrim *, uint32_t, uint32_t, float,
                              const DevObject *, const float *, uint32_t, int, int32_t *, float *,
                              float *, uint8_t *, hipStream_t) {
  return hipErrorInvalidValue;
}
hipError_t launch_debug_hits(const DevObject *, const float *, const uint32_t *, const float *,
                             uint32_t, float *, uint8_t *, hipStream_t) {
  return hipErrorInvalidValue;
}
hipError_t launch_debug_scatter(const DevObject *, const float *, const uint32_t *, uint32_t, float *,
                                hipStream_t) {
  return hipErrorInvalidValue;
}
hipError_t launch_debug_camera(const DevJob *, const uint32_t *, uint32_t, float *, hipStream_t) {
  return hipErrorInvalidValue;
}
hipError_t launch_debug_world(const DevJob *, const float *, uint32_t, float *, hipStream_t) {
  return hipErrorInvalidValue;
}
hipError_t launch_copy_calibration(const void *, void *, uint64_t, hipStream_t) {
  return hipErrorInvalidValue;
}
hipError_t launch_debug_philox(const uint32_t *, uint32_t, uint32_t *, hipStream_t) {
  return hipErrorInvalidValue;
}

}  // namespace rdh
