// Wavefront path-tracing kernels for CDNA4 (gfx950).  See kernels.hpp for the
// iteration structure and pathmath.hpp for the per-ray arithmetic.
#include "kernels.hpp"
#include "pathmath.hpp"

namespace rdh {

namespace {

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ uint32_t lane_id() { return lanes_below(~0ull); }

// Work of one iteration: survivors [0, base) plus n_new fresh paths.
struct IterPlan {
  uint32_t base, n_new, n;
  unsigned long long cursor;
};
__device__ __forceinline__ IterPlan plan_of(const IterArgs &a, int p) {
  IterPlan pl;
  pl.base = a.ctl->count[p];
  pl.cursor = a.ctl->cursor[p];
  const unsigned long long avail = a.ctl->total - pl.cursor;
  const unsigned long long room = (unsigned long long)(a.capacity - pl.base);
  pl.n_new = (uint32_t)(avail < room ? avail : room);
  pl.n = pl.base + pl.n_new;
  return pl;
}

// Largest ticket index whose work_begin <= w (tickets sorted by work_begin).
__device__ __forceinline__ uint32_t find_ticket(const DevTicket *t, uint32_t n,
                                                unsigned long long w) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (t[mid].work_begin <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void load_prim(const DevPrim *prims, uint32_t j, DevPrim *p) {
  const float4 *s = reinterpret_cast<const float4 *>(prims + j);
  const float4 a = s[0], b = s[1];
  p->a[0] = a.x; p->a[1] = a.y; p->a[2] = a.z; p->a[3] = a.w;
  p->b[0] = b.x; p->b[1] = b.y; p->b[2] = b.z; p->b[3] = b.w;
}

struct TraceCounts {
  uint32_t box = 0, sphere = 0, cube = 0;
};

template <bool COUNT>
__device__ __forceinline__ void test_leaf(const RayPre &r, const DevPrim *prims, int32_t code,
                                          float *bt, uint32_t *bid, TraceCounts &tc) {
  const uint32_t first = leaf_first(code), end = first + leaf_count(code);
  for (uint32_t j = first; j < end; ++j) {
    DevPrim p;
    load_prim(prims, j, &p);
    if (COUNT) {
      if (fbits(p.b[3]) & kTagCube) tc.cube++;
      else tc.sphere++;
    }
    test_prim(r, p, bt, bid);
  }
}

// Closest hit over one scene.  STACK points at this thread's column of the
// block's LDS stack (entry k at stack[k * STRIDE]).
template <int STRIDE, bool COUNT>
__device__ __forceinline__ void trace(const DevNode *__restrict__ nodes,
                                      const DevPrim *__restrict__ prims, uint32_t n_prims,
                                      uint32_t n_nodes, float r0, int brute, const RayPre &r,
                                      int32_t *stack, float *bt, uint32_t *bid,
                                      TraceCounts &tc) {
  *bid = kNoHit;
  *bt = 0.0f;
  if (n_prims == 0) return;
  const bool finite = isfinite(r.ox) && isfinite(r.oy) && isfinite(r.oz) && isfinite(r.dx) &&
                      isfinite(r.dy) && isfinite(r.dz);
  if (brute || !finite || n_nodes == 0) {
    // the reference's linear scan (cpu.rs:345-349); also the exact answer
    // for non-finite rays, whose box tests would be meaningless
    for (uint32_t j = 0; j < n_prims; ++j) {
      DevPrim p;
      load_prim(prims, j, &p);
      if (COUNT) {
        if (fbits(p.b[3]) & kTagCube) tc.cube++;
        else tc.sphere++;
      }
      test_prim(r, p, bt, bid);
    }
    return;
  }
  const SlabRay s = slab_ray(r, ray_tolerance(r, r0));
  int sp = 0;
  int32_t node = 0;
  for (;;) {
    const float4 *np = reinterpret_cast<const float4 *>(nodes + node);
    const float4 bx = np[0], by = np[1], bz = np[2];
    const int4 ch = reinterpret_cast<const int4 *>(np)[3];
    const float tcull = cull_limit(*bt, *bid);
    float t0, t1;
    bool h0 = slab_hit(s, bx.x, bx.z, by.x, by.z, bz.x, bz.z, tcull, &t0);
    bool h1 = slab_hit(s, bx.y, bx.w, by.y, by.w, bz.y, bz.w, tcull, &t1);
    if (COUNT) tc.box += 2;
    if (h0 && is_leaf(ch.x)) {
      test_leaf<COUNT>(r, prims, ch.x, bt, bid, tc);
      h0 = false;
    }
    if (h1 && is_leaf(ch.y)) {
      test_leaf<COUNT>(r, prims, ch.y, bt, bid, tc);
      h1 = false;
    }
    if (h0 && h1) {
      const bool first0 = t0 <= t1;
      stack[sp * STRIDE] = first0 ? ch.y : ch.x;
      ++sp;
      node = first0 ? ch.x : ch.y;
    } else if (h0) {
      node = ch.x;
    } else if (h1) {
      node = ch.y;
    } else {
      if (sp == 0) break;
      --sp;
      node = stack[sp * STRIDE];
    }
  }
}

// ------------------------------------------------------------ generate ----
template <bool COUNT>
__global__ void k_generate(IterArgs a, int p) {
  const IterPlan pl = plan_of(a, p);
  const uint32_t stride = gridDim.x * blockDim.x;
  const PathQueue q = a.q[p];
  uint32_t made = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < pl.n_new; i += stride) {
    const unsigned long long w = pl.cursor + i;
    const uint32_t ti = find_ticket(a.tickets, a.n_tickets, w);
    const DevTicket &T = a.tickets[ti];
    const DevJob &J = a.jobs[T.job];
    const uint32_t pidx = (uint32_t)(w - T.work_begin);
    const uint32_t npix = J.width * J.height;
    const uint32_t pixel = pidx % npix;
    F3 o, d;
    camera_ray(J, pixel % J.width, pixel / J.width, &o, &d);
    const uint32_t k = pl.base + i;
    q.ox[k] = o.x; q.oy[k] = o.y; q.oz[k] = o.z;
    q.dx[k] = d.x; q.dy[k] = d.y; q.dz[k] = d.z;
    q.ax[k] = 1.0f; q.ay[k] = 1.0f; q.az[k] = 1.0f;
    q.lx[k] = 0.0f; q.ly[k] = 0.0f; q.lz[k] = 0.0f;
    q.tb[k] = ti << 8;
    q.pidx[k] = pidx;
    ++made;
  }
  if (COUNT && made) atomicAdd(&a.counters->v[C_CAMERA], (unsigned long long)made);
}

// -------------------------------------------------------------- extend ----
template <int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_extend(IterArgs a, int p) {
  __shared__ int32_t stack[kStackDepth * BLOCK];
  const IterPlan pl = plan_of(a, p);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.ctl->count[1 - p] = 0;
    a.ctl->cursor[1 - p] = pl.cursor + pl.n_new;
  }
  const uint32_t stride = gridDim.x * BLOCK;
  const PathQueue q = a.q[p];
  TraceCounts tc;
  uint32_t rays = 0;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < pl.n; i += stride) {
    const F3 o = f3(q.ox[i], q.oy[i], q.oz[i]);
    const F3 d = f3(q.dx[i], q.dy[i], q.dz[i]);
    const DevJob &J = a.jobs[a.tickets[q.tb[i] >> 8].job];
    const RayPre r = ray_pre(o, d);
    float bt;
    uint32_t bid;
    trace<BLOCK, COUNT>(J.nodes, J.prims, J.n_prims, J.n_nodes, J.r0_tol, a.brute_force, r,
                        stack + threadIdx.x, &bt, &bid, tc);
    a.hit_t[i] = bt;
    a.hit_id[i] = bid;
    ++rays;
  }
  if (COUNT) {
    if (rays) atomicAdd(&a.counters->v[C_SEGMENTS], (unsigned long long)rays);
    if (tc.box) atomicAdd(&a.counters->v[C_BOX], (unsigned long long)tc.box);
    if (tc.sphere) atomicAdd(&a.counters->v[C_SPHERE], (unsigned long long)tc.sphere);
    if (tc.cube) atomicAdd(&a.counters->v[C_CUBE], (unsigned long long)tc.cube);
  }
}

// --------------------------------------------------------------- shade ----
template <bool COUNT>
__global__ void k_shade(IterArgs a, int p) {
  const IterPlan pl = plan_of(a, p);
  const PathQueue q = a.q[p], nq = a.q[1 - p];
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  uint32_t c_sph = 0, c_cube = 0, c_glass = 0, c_miss = 0;
  for (uint32_t base = wave * 64u; base < pl.n; base += n_waves * 64u) {
    const uint32_t i = base + lane;
    bool alive = false, finished = false;
    uint32_t ti = 0, pidx = 0, bounce = 0;
    F3 no = f3(0, 0, 0), nd = no, na = no, nl = no;
    uint64_t l_slot = 0;
    if (i < pl.n) {
      const F3 o = f3(q.ox[i], q.oy[i], q.oz[i]);
      const F3 d = f3(q.dx[i], q.dy[i], q.dz[i]);
      const F3 at = f3(q.ax[i], q.ay[i], q.az[i]);
      const F3 li = f3(q.lx[i], q.ly[i], q.lz[i]);
      const uint32_t tb = q.tb[i];
      ti = tb >> 8;
      bounce = tb & 0xffu;
      pidx = q.pidx[i];
      const DevTicket &T = a.tickets[ti];
      const DevJob &J = a.jobs[T.job];
      l_slot = T.l_offset + pidx;
      const uint32_t id = a.hit_id[i];
      if (id != kNoHit) {
        const DevObject ob = J.objects[id];
        const HitRec h = closest_hit(o, d, a.hit_t[i], ob);
        const uint32_t npix = J.width * J.height;
        BounceDraws dr;
        dr.init(J.seed, pidx % npix, T.sample_base + pidx / npix, bounce);
        const ScatterOut so = scatter(d, h, ob, dr, at, li);
        if (COUNT) {
          if (fbits(ob.m2[2]) == RD_SPHERE) c_sph++;
          else c_cube++;
          if (so.transmission) c_glass++;
        }
        ++bounce;
        nl = so.light;
        if (bounce < J.max_bounces) {
          alive = true;
          no = so.o; nd = so.d; na = so.atten;
        } else {
          finished = true;  // bounce budget spent: no sky term (cpu.rs:256)
        }
      } else {
        nl = add(li, mul(world_sample(J, d), at));  // cpu.rs:334-337
        finished = true;
        if (COUNT) c_miss++;
      }
    }
    // wave64 stream compaction of the survivors into q[1-p]
    const uint64_t am = __ballot(alive);
    if (am) {
      uint32_t wbase = 0;
      if (lane == 0) wbase = atomicAdd(&a.ctl->count[1 - p], (uint32_t)__popcll(am));
      wbase = __builtin_amdgcn_readfirstlane(wbase);
      if (alive) {
        const uint32_t k = wbase + lanes_below(am);
        nq.ox[k] = no.x; nq.oy[k] = no.y; nq.oz[k] = no.z;
        nq.dx[k] = nd.x; nq.dy[k] = nd.y; nq.dz[k] = nd.z;
        nq.ax[k] = na.x; nq.ay[k] = na.y; nq.az[k] = na.z;
        nq.lx[k] = nl.x; nq.ly[k] = nl.y; nq.lz[k] = nl.z;
        nq.tb[k] = (ti << 8) | bounce;
        nq.pidx[k] = pidx;
      }
    }
    if (finished) a.L[l_slot] = make_float4(nl.x, nl.y, nl.z, 1.0f);
    // finished-path counts per ticket, one atomic per distinct ticket per wave
    uint64_t fm = __ballot(finished);
    while (fm) {
      const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)fm) - 1);
      const uint32_t lt = (uint32_t)__builtin_amdgcn_readlane((int)ti, (int)leader);
      const uint64_t same = __ballot(finished && ti == lt);
      if (lane == leader) atomicAdd(&a.ticket_done[lt], (unsigned long long)__popcll(same));
      fm &= ~same;
    }
  }
  if (COUNT) {
    if (c_sph) atomicAdd(&a.counters->v[C_SPHERE_HIT], (unsigned long long)c_sph);
    if (c_cube) atomicAdd(&a.counters->v[C_CUBE_HIT], (unsigned long long)c_cube);
    if (c_glass) atomicAdd(&a.counters->v[C_GLASS], (unsigned long long)c_glass);
    if (c_miss) atomicAdd(&a.counters->v[C_MISS], (unsigned long long)c_miss);
  }
}

// ---------------------------------------------------------- accumulate ----
// Adds the ticket's samples to each pixel in sample order (cpu.rs:204-209).
// L == nullptr adds zero radiance (max_bounces == 0: no segment is traced).
template <bool COUNT>
__global__ void k_accumulate(const DevJob *job, DevTicket T, const float4 *L,
                             DevCounters *counters) {
  const DevJob &J = *job;
  const uint32_t npix = J.width * J.height;
  float4 *acc = reinterpret_cast<float4 *>(J.acc);
  uint32_t done = 0;
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix;
       px += gridDim.x * blockDim.x) {
    float4 v = acc[px];
    for (uint32_t s = 0; s < T.spp; ++s) {
      float4 c = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
      if (L) c = L[T.l_offset + (uint64_t)s * npix + px];
      v.x = v.x + c.x;
      v.y = v.y + c.y;
      v.z = v.z + c.z;
      v.w = v.w + 1.0f;
    }
    acc[px] = v;
    done += T.spp;
  }
  if (COUNT && done) atomicAdd(&counters->v[C_ACCUM], (unsigned long long)done);
}

// ------------------------------------------------------------- tonemap ----
template <bool COUNT>
__global__ void k_tonemap(const DevJob *job, uint32_t n, float spp, uchar4 *out,
                          DevCounters *counters) {
  const float4 *acc = reinterpret_cast<const float4 *>(job->acc);
  uint32_t done = 0;
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < n;
       px += gridDim.x * blockDim.x) {
    const float4 v = acc[px];
    out[px] = make_uchar4(tonemap_channel(v.x, spp), tonemap_channel(v.y, spp),
                          tonemap_channel(v.z, spp), tonemap_channel(v.w, spp));
    ++done;
  }
  if (COUNT && done) atomicAdd(&counters->v[C_TONEMAP], (unsigned long long)done);
}

__global__ void k_set_ctl(QueueCtl *ctl, unsigned long long total) {
  ctl->count[0] = 0;
  ctl->count[1] = 0;
  ctl->cursor[0] = 0;
  ctl->cursor[1] = 0;
  ctl->total = total;
}

__global__ void k_zero(float *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// --------------------------------------------------------- parity hooks ---
__global__ void k_debug_trace(const DevNode *nodes, const DevPrim *prims, uint32_t n_prims,
                              uint32_t n_nodes, float r0, const DevObject *objects,
                              const float *rays, uint32_t n, int brute, int32_t *ids, float *ts,
                              float *normals, uint8_t *front) {
  __shared__ int32_t stack[kStackDepth * 64];
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  const float *q = rays + 6 * (uint64_t)i;
  const F3 o = f3(q[0], q[1], q[2]), d = f3(q[3], q[4], q[5]);
  const RayPre r = ray_pre(o, d);
  float bt;
  uint32_t bid;
  TraceCounts tc;
  trace<64, false>(nodes, prims, n_prims, n_nodes, r0, brute, r, stack + threadIdx.x, &bt, &bid,
                   tc);
  if (bid == kNoHit) {
    ids[i] = -1;
    ts[i] = 0.0f;
    normals[3 * i] = normals[3 * i + 1] = normals[3 * i + 2] = 0.0f;
    front[i] = 0;
  } else {
    ids[i] = (int32_t)bid;
    ts[i] = bt;
    const HitRec h = closest_hit(o, d, bt, objects[bid]);
    normals[3 * i] = h.n.x; normals[3 * i + 1] = h.n.y; normals[3 * i + 2] = h.n.z;
    front[i] = h.front ? 1 : 0;
  }
}

// in / out: 16 floats per hit, the layout of the oracle's or_scatter_f32.
__global__ void k_debug_scatter(const DevObject *objects, const float *in, const uint32_t *bits,
                                uint32_t n, float *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *q = in + 16 * (uint64_t)i;
  const F3 o = f3(q[0], q[1], q[2]), d = f3(q[3], q[4], q[5]);
  const uint32_t id = fbits(q[7]);
  const DevObject ob = objects[id];
  const HitRec h = closest_hit(o, d, q[6], ob);
  FixedDraws dr{bits + 12 * (uint64_t)i};
  const ScatterOut so = scatter(d, h, ob, dr, f3(q[8], q[9], q[10]), f3(q[11], q[12], q[13]));
  float *w = out + 16 * (uint64_t)i;
  w[0] = so.o.x; w[1] = so.o.y; w[2] = so.o.z;
  w[3] = so.d.x; w[4] = so.d.y; w[5] = so.d.z;
  w[6] = so.atten.x; w[7] = so.atten.y; w[8] = so.atten.z;
  w[9] = so.light.x; w[10] = so.light.y; w[11] = so.light.z;
  w[12] = so.transmission ? 1.0f : 0.0f;
  w[13] = h.n.x; w[14] = h.n.y; w[15] = h.n.z;
}

__global__ void k_debug_camera(const DevJob *job, const uint32_t *pixels, uint32_t n,
                               float *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t px = pixels[i];
  F3 o, d;
  camera_ray(*job, px % job->width, px / job->width, &o, &d);
  float *w = out + 6 * (uint64_t)i;
  w[0] = o.x; w[1] = o.y; w[2] = o.z; w[3] = d.x; w[4] = d.y; w[5] = d.z;
}

__global__ void k_debug_world(const DevJob *job, const float *dirs, uint32_t n, float *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const F3 c = world_sample(*job, f3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
  out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
}

__global__ void k_debug_philox(const uint32_t *in, uint32_t n, uint32_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t *c = in + 6 * (uint64_t)i;
  const U4 r = philox4x32_10(c[0], c[1], c[2], c[3], c[4], c[5]);
  for (int k = 0; k < 4; ++k) out[4 * (uint64_t)i + k] = r.v[k];
}

uint32_t blocks_for(uint64_t n, uint32_t block) { return (uint32_t)((n + block - 1) / block); }

}  // namespace

// ------------------------------------------------------------ launchers ---
int occupancy_blocks(int which, int block) {
  int n = 0;
  hipError_t e = hipSuccess;
  if (which == 0) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_generate<false>, block, 0);
  } else if (which == 2) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_shade<false>, block, 0);
  } else if (block == 64) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_extend<64, false>, 64, 0);
  } else if (block == 128) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_extend<128, false>, 128, 0);
  } else {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_extend<256, false>, 256, 0);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 1;
  }
  return n;
}

hipError_t launch_generate(const IterArgs &a, int parity, const LaunchShape &s, bool count,
                           hipStream_t st) {
  if (count) hipLaunchKernelGGL(k_generate<true>, dim3(s.grid), dim3(s.block), 0, st, a, parity);
  else hipLaunchKernelGGL(k_generate<false>, dim3(s.grid), dim3(s.block), 0, st, a, parity);
  return hipGetLastError();
}

hipError_t launch_extend(const IterArgs &a, int parity, const LaunchShape &s, bool count,
                         hipStream_t st) {
#define RDH_EXTEND(B)                                                                       \
  case B:                                                                                   \
    if (count) hipLaunchKernelGGL((k_extend<B, true>), dim3(s.grid), dim3(B), 0, st, a, parity); \
    else hipLaunchKernelGGL((k_extend<B, false>), dim3(s.grid), dim3(B), 0, st, a, parity); \
    break;
  switch (s.block) {
    RDH_EXTEND(64)
    RDH_EXTEND(128)
    RDH_EXTEND(256)
    default:
      return hipErrorInvalidValue;
  }
#undef RDH_EXTEND
  return hipGetLastError();
}

hipError_t launch_shade(const IterArgs &a, int parity, const LaunchShape &s, bool count,
                        hipStream_t st) {
  if (count) hipLaunchKernelGGL(k_shade<true>, dim3(s.grid), dim3(s.block), 0, st, a, parity);
  else hipLaunchKernelGGL(k_shade<false>, dim3(s.grid), dim3(s.block), 0, st, a, parity);
  return hipGetLastError();
}

hipError_t launch_accumulate(const DevJob *job, const DevTicket &t, const float4 *L,
                             DevCounters *counters, bool count, int grid, hipStream_t st) {
  if (count) hipLaunchKernelGGL(k_accumulate<true>, dim3(grid), dim3(256), 0, st, job, t, L, counters);
  else hipLaunchKernelGGL(k_accumulate<false>, dim3(grid), dim3(256), 0, st, job, t, L, counters);
  return hipGetLastError();
}

hipError_t launch_tonemap(const DevJob *job, uint32_t n_pixels, float spp, uchar4 *out,
                          DevCounters *counters, bool count, int grid, hipStream_t st) {
  if (count) hipLaunchKernelGGL(k_tonemap<true>, dim3(grid), dim3(256), 0, st, job, n_pixels, spp, out, counters);
  else hipLaunchKernelGGL(k_tonemap<false>, dim3(grid), dim3(256), 0, st, job, n_pixels, spp, out, counters);
  return hipGetLastError();
}

hipError_t launch_set_ctl(QueueCtl *ctl, unsigned long long total, hipStream_t st) {
  hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(1), 0, st, ctl, total);
  return hipGetLastError();
}

hipError_t launch_zero_acc(float *acc, uint64_t n_floats, hipStream_t st) {
  const uint32_t grid = (uint32_t)std::min<uint64_t>(blocks_for(n_floats, 256), 8192);
  hipLaunchKernelGGL(k_zero, dim3(grid > 0 ? grid : 1), dim3(256), 0, st, acc, n_floats);
  return hipGetLastError();
}

hipError_t launch_debug_trace(const DevNode *nodes, const DevPrim *prims, uint32_t n_prims,
                              uint32_t n_nodes, float r0, const DevObject *objects,
                              const float *rays, uint32_t n, int brute, int32_t *ids, float *ts,
                              float *normals, uint8_t *front, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_trace, dim3(blocks_for(n, 64)), dim3(64), 0, st, nodes, prims,
                     n_prims, n_nodes, r0, objects, rays, n, brute, ids, ts, normals, front);
  return hipGetLastError();
}

hipError_t launch_debug_scatter(const DevObject *objects, const float *in, const uint32_t *bits,
                                uint32_t n, float *out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, st, objects, in,
                     bits, n, out);
  return hipGetLastError();
}

hipError_t launch_debug_camera(const DevJob *job, const uint32_t *pixels, uint32_t n, float *out,
                               hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_camera, dim3(blocks_for(n, 256)), dim3(256), 0, st, job, pixels, n,
                     out);
  return hipGetLastError();
}

hipError_t launch_debug_world(const DevJob *job, const float *dirs, uint32_t n, float *out,
                              hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_world, dim3(blocks_for(n, 256)), dim3(256), 0, st, job, dirs, n, out);
  return hipGetLastError();
}

hipError_t launch_debug_philox(const uint32_t *in, uint32_t n, uint32_t *out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_philox, dim3(blocks_for(n, 256)), dim3(256), 0, st, in, n, out);
  return hipGetLastError();
}

}  // namespace rdh
