// Wavefront path-tracing kernels for CDNA4 (gfx950).  See kernels.hpp for the
// iteration structure and pathmath.hpp for the per-ray arithmetic.
#include "kernels.hpp"
#include "pathmath.hpp"

// Target resident waves per SIMD for the latency-bound traversal kernel; the
// compiler caps its registers to fit (without spilling, checked at build).
#ifndef RDH_EXTEND_WAVES
#define RDH_EXTEND_WAVES 8
#endif

namespace rdh {

namespace {

// Bounds-checked builds (-DRDH_CHECKED) pass every computed index through
// chk(): an index outside its buffer sets bit `bit` of QueueCtl::violation
// (the host then fails the render) and is replaced by 0, so the access
// itself stays inside the buffer.  Production builds compile chk() away.
enum CheckBit {
  K_SLOT = 0, K_TICKET, K_JOB, K_NODE, K_PRIM, K_STACK, K_OBJECT, K_RADIANCE
};
template <class I>
__device__ __forceinline__ I chk(I i, I n, CheckBit bit, QueueCtl *ctl) {
#ifdef RDH_CHECKED
  if (i >= n) {
    atomicOr(&ctl->violation, 1u << bit);
    return I(0);
  }
#endif
  (void)n; (void)bit; (void)ctl;
  return i;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ uint32_t lane_id() { return lanes_below(~0ull); }

// Work of one iteration over queue[p]: segment s holds cnt[s] survivors
// followed by add[s] fresh paths (n_new in all, at most the nominal capacity
// in total).  Every workgroup of generate, extend and shade derives the same
// plan from the same counters, which no kernel of the iteration modifies.
struct SegPlan {
  uint32_t cnt[kSegments];
  uint32_t add[kSegments];
  uint32_t pre[kSegments + 1];      // prefix of cnt + add: consumers' linear order
  uint32_t pre_new[kSegments + 1];  // prefix of add: generate's linear order
  uint32_t n_new, n;
  unsigned long long cursor;
};

// survivors_only: the consumers' order covers only the survivors (cnt[s] of
// each segment, n_old in all), for an extend whose fresh paths already have
// their hits (first_hit_mode 2); n_new and the cursor are unchanged.
__device__ __forceinline__ void make_plan(const IterArgs &a, int p, SegPlan &sp,
                                          bool survivors_only = false) {
  if (threadIdx.x < kSegments) {
    // a counter above seg_cap only follows an overflow (already flagged for
    // the host); clamping keeps every slot index of this iteration in bounds
    const uint32_t c = a.ctl->count[p][threadIdx.x * kSegStride];
    sp.cnt[threadIdx.x] = c < a.seg_cap ? c : a.seg_cap;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t n_old = 0;
    for (int s = 0; s < kSegments; ++s) n_old += sp.cnt[s];
    const unsigned long long cursor = a.ctl->cursor[p];
    const unsigned long long avail = a.ctl->total - cursor;
    const unsigned long long room = (unsigned long long)(a.capacity - n_old);
    const uint32_t n_new = (uint32_t)(avail < room ? avail : room);
    uint32_t left = n_new;
    sp.pre[0] = 0;
    sp.pre_new[0] = 0;
    for (int s = 0; s < kSegments; ++s) {
      const uint32_t free_s = a.seg_cap - sp.cnt[s];
      const uint32_t add = left < free_s ? left : free_s;
      left -= add;
      sp.add[s] = add;
      sp.pre_new[s + 1] = sp.pre_new[s] + add;
      sp.pre[s + 1] = sp.pre[s] + sp.cnt[s] + (survivors_only ? 0u : add);
    }
    sp.n_new = n_new;
    sp.n = sp.pre[kSegments];
    sp.cursor = cursor;
  }
  __syncthreads();
}

// Largest s with pre[s] <= i (pre non-decreasing, pre[0] = 0 <= i).
__device__ __forceinline__ uint32_t segment_of(const uint32_t *pre, uint32_t i) {
  uint32_t lo = 0, hi = kSegments;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Queue slot of the i-th path of the iteration (consumers' order).  `s` is
// the caller's running segment: a lane's index only grows, so its segment
// only moves forward and each step costs one LDS read.
__device__ __forceinline__ uint32_t slot_of(const SegPlan &sp, uint32_t seg_cap, uint32_t i,
                                            uint32_t &s) {
  while (sp.pre[s + 1] <= i) ++s;
  return s * seg_cap + (i - sp.pre[s]);
}

// Finished-path counts per ticket, gathered in LDS by one shade workgroup and
// flushed to HBM once per launch (instead of one same-address atomic per wave).
constexpr uint32_t kDoneEntries = 32;
struct DoneTable {
  uint32_t n;                       // entries claimed (may exceed kDoneEntries)
  uint32_t ticket[kDoneEntries];
  uint32_t count[kDoneEntries];
};

// A slot's ticket is written once, after the slot is claimed, so any match
// found by a scan is genuine; two waves racing to insert the same ticket
// merely claim two slots.  With the table full, the count goes straight to
// HBM.
__device__ __forceinline__ void done_add(DoneTable &d, uint32_t t, uint32_t c,
                                         unsigned long long *global) {
  volatile uint32_t *vt = d.ticket;
  const uint32_t n = min(*(volatile uint32_t *)&d.n, kDoneEntries);
  for (uint32_t k = 0; k < n; ++k)
    if (vt[k] == t) {
      atomicAdd(&d.count[k], c);
      return;
    }
  const uint32_t k = atomicAdd(&d.n, 1u);
  if (k < kDoneEntries) {
    vt[k] = t;
    atomicAdd(&d.count[k], c);
  } else {
    atomicAdd(&global[t], (unsigned long long)c);
  }
}

// Queue of parity p.  Indexing a.q[p] with the runtime parity makes the
// compiler fetch the pointer table through a vector address and keep all 28
// pointers in VGPRs; a uniform select keeps them in scalar registers.
__device__ __forceinline__ PathQueue queue_of(const IterArgs &a, int p) {
  return p ? a.q[1] : a.q[0];
}

// A record of a table that no kernel writes during a launch, read through
// the constant address space: at a wave-uniform address that is a scalar
// load.  The pointer is rebuilt from its integer value, which the compiler
// cannot fold back to a global pointer (a plain cast it folds, and the loads
// stay vector loads).  The host pass of the compiler only checks the body.
template <class T>
__device__ __forceinline__ T load_constant(const T *p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(4))) T *)(uintptr_t)p;
#else
  return *p;
#endif
}

// Largest ticket index in [lo, n) whose work_begin <= w (tickets sorted by
// work_begin; the host keeps lo at or below the ticket holding the cursor).
__device__ __forceinline__ uint32_t find_ticket(const DevTicket *t, uint32_t lo, uint32_t n,
                                                unsigned long long w) {
  uint32_t hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (t[mid].work_begin <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}
// The same search at a wave-uniform work index: every probe is a scalar load.
__device__ __forceinline__ uint32_t find_ticket_uniform(const DevTicket *t, uint32_t lo, uint32_t n,
                                                        unsigned long long w) {
  uint32_t hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (load_constant(&t[mid].work_begin) <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Scene buffers are reached through pointers read from the job table, which
// the compiler cannot place in an address space; these loads name the global
// space explicitly so they become global_load (not flat_load, which also
// waits on the LDS counter the traversal stack uses).
typedef float VecF4 __attribute__((ext_vector_type(4)));
typedef float VecF2 __attribute__((ext_vector_type(2)));
typedef int VecI4 __attribute__((ext_vector_type(4)));
typedef unsigned int VecU4 __attribute__((ext_vector_type(4)));
typedef unsigned int VecU2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const VecF4 GlobalF4;
typedef __attribute__((address_space(1))) const VecF2 GlobalF2;
typedef __attribute__((address_space(1))) const VecI4 GlobalI4;
typedef __attribute__((address_space(1))) const VecU4 GlobalU4;
typedef __attribute__((address_space(1))) const VecU2 GlobalU2;
typedef __attribute__((address_space(1))) VecF4 GlobalF4Out;
typedef __attribute__((address_space(1))) VecF2 GlobalF2Out;
typedef __attribute__((address_space(1))) VecU4 GlobalU4Out;
typedef __attribute__((address_space(1))) VecU2 GlobalU2Out;

// A queued path's ray (QRay, QDir), and with WORD its two words.
template <bool WORD>
__device__ __forceinline__ void load_ray(const PathQueue &q, uint32_t k, F3 *o, F3 *d,
                                         uint32_t *tb, uint32_t *pw = nullptr) {
  const VecF4 a = *(GlobalF4 *)(q.ray + k);
  const VecF2 b = *(GlobalF2 *)(q.dir + k);
  *o = f3(a.x, a.y, a.z);
  *d = f3(a.w, b.x, b.y);
  if (WORD) {
    const VecU2 w = *(GlobalU2 *)(q.word + k);
    *tb = w.x;
    if (pw) *pw = w.y;
  }
}
// A path's ray and words, stored into slot k.
__device__ __forceinline__ void store_ray(const PathQueue &q, uint32_t k, F3 o, F3 d, uint32_t tb,
                                          uint32_t pidx) {
  *(GlobalF4Out *)(q.ray + k) = VecF4{o.x, o.y, o.z, d.x};
  *(GlobalF2Out *)(q.dir + k) = VecF2{d.y, d.z};
  *(GlobalU2Out *)(q.word + k) = VecU2{tb, pidx};
}

__device__ __forceinline__ void load_prim(const DevPrim *prims, uint32_t j, DevPrim *p) {
  GlobalF4 *s = (GlobalF4 *)(prims + j);
  const VecF4 a = s[0], b = s[1];
  p->a[0] = a.x; p->a[1] = a.y; p->a[2] = a.z; p->a[3] = a.w;
  p->b[0] = b.x; p->b[1] = b.y; p->b[2] = b.z; p->b[3] = b.w;
}

__device__ __forceinline__ DevObject object_from(VecF4 g, VecF4 m0, VecF4 m1, VecF4 m2) {
  DevObject o;
  o.geo[0] = g.x; o.geo[1] = g.y; o.geo[2] = g.z; o.geo[3] = g.w;
  o.m0[0] = m0.x; o.m0[1] = m0.y; o.m0[2] = m0.z; o.m0[3] = m0.w;
  o.m1[0] = m1.x; o.m1[1] = m1.y; o.m1[2] = m1.z; o.m1[3] = m1.w;
  o.m2[0] = m2.x; o.m2[1] = m2.y; o.m2[2] = m2.z; o.m2[3] = m2.w;
  return o;
}
__device__ __forceinline__ DevObject load_object(const DevObject *objects, uint32_t id) {
  GlobalF4 *s = (GlobalF4 *)(objects + id);
  return object_from(s[0], s[1], s[2], s[3]);
}

struct TraceCounts {
  uint32_t box = 0, sphere = 0, cube = 0;
  // wave-level events of the lanes kernel, each counted by one lane
  uint32_t wave_steps = 0, lane_steps = 0, sort_steps = 0, prim_trips = 0, cube_trips = 0;
  uint32_t refills = 0;
};

// True in exactly one lane of the lanes currently executing (the lowest).
__device__ __forceinline__ bool wave_leader() { return lanes_below(__ballot(true)) == 0; }

template <bool COUNT>
__device__ __forceinline__ void count_trip(TraceCounts &tc, const DevPrim &p) {
  if (COUNT) {
    const bool cube = (fbits(p.b[3]) & kTagCube) != 0;
    const uint64_t cubes = __ballot(cube);
    if (wave_leader()) {
      tc.prim_trips++;
      tc.cube_trips += cubes != 0;
    }
  }
}

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
// block's LDS stack (entry k at stack[k * STRIDE]); StackT is uint16_t when
// every BVH has at most 65535 nodes (half the LDS, more resident waves).
template <int STRIDE, bool COUNT, class StackT>
__device__ __forceinline__ void trace(const DevNode *__restrict__ nodes,
                                      const DevPrim *__restrict__ prims, uint32_t n_prims,
                                      uint32_t n_nodes, float r0, int brute, const RayPre &r,
                                      StackT *stack, float *bt, uint32_t *bid, TraceCounts &tc) {
  *bid = kNoHit;
  *bt = 0.0f;
  if (n_prims == 0) return;
  if (brute || !walkable(r) || n_nodes == 0) {
    // the reference's linear scan (cpu.rs:345-349); also the exact answer
    // for rays the box walk cannot represent (walkable())
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
      stack[sp * STRIDE] = (StackT)(first0 ? ch.y : ch.x);
      ++sp;
      node = first0 ? ch.x : ch.y;
    } else if (h0) {
      node = ch.x;
    } else if (h1) {
      node = ch.y;
    } else {
      if (sp == 0) break;
      --sp;
      node = (int32_t)stack[sp * STRIDE];
    }
  }
}

// Job and ticket of a path.  In single-job launches (the per-scene renderer)
// both are the first table entries, so every lane reads the same address and
// the loads become scalar.
template <bool SINGLE>
__device__ __forceinline__ uint32_t ticket_of(uint32_t tb) { return SINGLE ? 0u : (tb >> 8); }

// ------------------------------------------------------------ generate ----
// Pixel of the r-th camera path of a sample: 8x8 tiles, row by row inside a
// tile and tile by tile along each band of 8 rows (the columns right of the
// last whole tile row-major inside the band; a last band of fewer than 8 rows
// row-major), a permutation of [0, W*H).  A wave's 64 consecutive paths then
// start from one 8x8 tile instead of 64 pixels of one row: their rays and,
// after a bounce, their origins stay close together, which the closest-hit
// kernel's node fetches and traversal order profit from.  A path's draws,
// radiance slot and accumulation are keyed by its pixel, so the order changes
// no result.
__device__ __forceinline__ uint32_t tile_order_pixel(uint32_t r, uint32_t W, uint32_t H) {
  const uint32_t band_px = 8u * W;
  const uint32_t band = r / band_px;
  if ((band + 1u) * 8u > H) return r;
  const uint32_t q = r - band * band_px;
  const uint32_t full = W & ~7u;
  uint32_t x, y;
  if (q < full * 8u) {
    x = (q >> 6) * 8u + (q & 7u);
    y = band * 8u + ((q >> 3) & 7u);
  } else {
    const uint32_t e = q - full * 8u, rw = W - full;
    x = full + e % rw;
    y = band * 8u + e / rw;
  }
  return y * W + x;
}

// VK (Vulkan semantics, one sample per pass): the pinhole camera of
// raytrace.rgen:83-89; the pass of sample 0 also seeds each pixel's PCG
// state with x + y * 1000 (raytrace.rgen:50).
template <bool COUNT, bool SINGLE, bool VK = false>
__global__ void k_generate(IterArgs a, int p) {
  __shared__ SegPlan pl;
  // One job (SINGLE): its camera and its one ticket are read here, before
  // any global store, as scalar loads kept in scalar registers.  Read inside
  // the loop, after the first stores, they would be vector loads and a
  // memory round trip per iteration; a camera path needs no other read.
  DevJob job0;
  DevTicket ticket0;
  if (SINGLE) {
    job0 = a.jobs[0];
    ticket0 = a.tickets[0];
  }
  make_plan(a, p, pl);
  const uint32_t stride = gridDim.x * blockDim.x;
  const PathQueue q = queue_of(a, p);
  const uint32_t t_lo = a.ctl->ticket_lo, t_n = a.ctl->n_tickets;
  uint32_t made = 0, reused = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < pl.n_new; i += stride) {
    const unsigned long long w = pl.cursor + i;
    const uint32_t s = segment_of(pl.pre_new, i);
    const uint32_t k = chk(s * a.seg_cap + pl.cnt[s] + (i - pl.pre_new[s]), a.slots, K_SLOT, a.ctl);
    uint32_t ti = 0, pidx = 0, pixel = 0, pre = 0;
    F3 o, d;
    auto camera = [&](const DevJob &J, const DevTicket &T) {
      const uint32_t j = (uint32_t)(w - T.work_begin);
      const uint32_t npix = J.width * J.height;
      const uint32_t smp = j / npix;
      pixel = tile_order_pixel(j - smp * npix, J.width, J.height);
      pidx = smp * npix + pixel;
      if (VK) {
        const uint32_t x = pixel % J.width, y = pixel / J.width;
        camera_ray_vk(J, x, y, &o, &d);
        if (T.sample_base == 0) J.vk_rng[chk(pixel, npix, K_RADIANCE, a.ctl)] = x + y * 1000u;
      } else {
        camera_ray(J, pixel % J.width, pixel / J.width, &o, &d);
      }
    };
    if (SINGLE) {
      camera(job0, ticket0);
      if (a.first_hit_mode == 2) {
        // the pixel's camera-ray hit, recorded by this pass's first shade:
        // extend leaves this fresh path alone
        const HitWord h = a.first_hit[chk(pixel, job0.width * job0.height, K_RADIANCE, a.ctl)];
        a.hit_t[k] = bitsf(h.t1 - 1u);
        a.hit_id[k] = h.id;
        if (COUNT) ++reused;
      }
    } else {
      // the wave's lanes hold consecutive work indices: one search for the
      // first active lane's index (wave-uniform: scalar loads) serves nearly
      // all; lanes past that ticket's end search again on their own
      const unsigned long long w0 =
          ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(w >> 32)) << 32) |
          __builtin_amdgcn_readfirstlane((uint32_t)w);
      ti = find_ticket_uniform(a.tickets, t_lo, t_n, w0);
      const DevTicket t0 = load_constant(a.tickets + ti);
      if (w >= t0.work_begin + t0.paths) ti = find_ticket(a.tickets, ti, t_n, w);
      ti = chk(ti, a.ticket_cap, K_TICKET, a.ctl);
      // one round per distinct ticket among the lanes (nearly always one),
      // with the ticket and job records read as scalar loads
      uint64_t todo = __ballot(true);
      while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)ti, leader);
        // membership from the ballot's bit: under ti == t the compiler
        // would read the tables at ti again
        const uint64_t same = __ballot(ti == t);
        todo &= ~same;
        if ((same >> lane_id()) & 1ull) {
          const DevTicket T = load_constant(a.tickets + t);
          const DevJob J = load_constant(a.jobs + chk(T.job, a.job_cap, K_JOB, a.ctl));
          camera(J, T);
          if (J.first_hit) {
            // the pixel's camera-ray hit once a shade has recorded it (zero
            // until then): the path is flagged, and extend leaves it alone
            const HitWord h = J.first_hit[chk(pixel, J.width * J.height, K_RADIANCE, a.ctl)];
            if (h.t1 != 0u) {
              a.hit_t[k] = bitsf(h.t1 - 1u);
              a.hit_id[k] = h.id;
              pre = kPreHitFlag;
              if (COUNT) ++reused;
            }
          }
        }
      }
    }
    // attenuation and light of a camera path are (1,1,1) and 0 (cpu.rs:253-254):
    // shade uses the constants at bounce 0, so they are not stored
    store_ray(q, k, o, d, ti << 8, pidx | pre);
    ++made;
  }
  if (COUNT && made) atomicAdd(&a.counters->v[C_CAMERA], (unsigned long long)made);
  if (COUNT && reused) atomicAdd(&a.counters->v[C_REUSED], (unsigned long long)reused);
}

// -------------------------------------------------------------- extend ----
template <int BLOCK, bool COUNT, bool SINGLE, class StackT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(RDH_EXTEND_WAVES, 8)))
void k_extend(IterArgs a, int p) {
  __shared__ StackT stack[kStackDepth * BLOCK];
  __shared__ SegPlan pl;
  make_plan(a, p, pl, a.first_hit_mode == 2);
  // queue[1-p] is empty until shade fills it; nothing else reads it now
  if (blockIdx.x == 0 && threadIdx.x < kSegments) a.ctl->count[1 - p][threadIdx.x * kSegStride] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl->cursor[1 - p] = pl.cursor + pl.n_new;
  const uint32_t stride = gridDim.x * BLOCK;
  const PathQueue q = queue_of(a, p);
  TraceCounts tc;
  uint32_t rays = 0, seg = 0;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < pl.n; i += stride) {
    const uint32_t k = slot_of(pl, a.seg_cap, i, seg);
    F3 o, d;
    uint32_t tbk = 0, pwk = 0;
    load_ray<!SINGLE>(q, k, &o, &d, &tbk, &pwk);
    if (!SINGLE && (pwk & kPreHitFlag)) continue;  // generate wrote its hit
    const DevJob &J = a.jobs[SINGLE ? 0u : a.tickets[tbk >> 8].job];
    const RayPre r = ray_pre(o, d);
    float bt;
    uint32_t bid;
    trace<BLOCK, COUNT, StackT>(J.nodes, J.prims, J.n_prims, J.n_nodes, J.r0_tol, a.brute_force,
                                r, stack + threadIdx.x, &bt, &bid, tc);
    a.hit_t[k] = bt;
    a.hit_id[k] = bid;
    ++rays;
  }
  if (COUNT) {
    if (rays) atomicAdd(&a.counters->v[C_SEGMENTS], (unsigned long long)rays);
    if (tc.box) atomicAdd(&a.counters->v[C_BOX], (unsigned long long)tc.box);
    if (tc.sphere) atomicAdd(&a.counters->v[C_SPHERE], (unsigned long long)tc.sphere);
    if (tc.cube) atomicAdd(&a.counters->v[C_CUBE], (unsigned long long)tc.cube);
  }
}

// Comparator of the 4-wide traversal's sorting network: (ka, ca) gets the
// smaller key (a strictly smaller kb swaps, as the host traversal does).
__device__ __forceinline__ void order_pair(float &ka, int32_t &ca, float &kb, int32_t &cb) {
  const bool sw = kb < ka;
  const float k = sw ? kb : ka;
  const int32_t c = sw ? cb : ca;
  kb = sw ? ka : kb;
  cb = sw ? ca : cb;
  ka = k;
  ca = c;
}

// Persistent-lane closest hit.  Rays differ widely in traversal length, so
// a wave that traces 64 rays to completion keeps most lanes idle while its
// longest ray finishes.  Here each wave owns a contiguous range of queue
// indices (no shared counter: handing out chunks through one device-wide
// atomic measured twice as slow on short rays),
// and a lane that finishes takes the next index of its wave's range as soon
// as `refill` or more lanes of the wave are idle.  Each step of the loop
// visits one BVH node per active lane.  Indices only grow per lane, so
// slot_of's running segment stays valid.
//
// LDS = true (single job, small scene): the workgroup first copies the BVH
// nodes and primitives into LDS and traverses them there; the stack then
// has only as many rows as the tree has levels, both in dynamic LDS.
//
// W4 = true (single job): the 4-wide tree (DevJob::nodes4).  A step tests
// four child boxes, sorts the entered inner children by entry distance with
// five comparators, writes the three farther candidates above the stack top
// (kept as many as were entered, minus one) and reads the top entry; the
// stack needs DevJob::stack4 + 3 rows.
// Waves per SIMD the lanes kernel is compiled for: the multi-job 4-wide walk
// keeps each lane's job pointers in registers and needs 7 (72 VGPRs) to
// avoid spilling; every other variant fits 8.
constexpr int lanes_min_waves(bool single, bool w4) { return !single && w4 ? 7 : RDH_EXTEND_WAVES; }

// LEAF1 (single 4-wide jobs whose leaves hold one primitive each, IterArgs::
// leaf1): a hit leaf child contributes exactly one primitive, so the leaf
// counts are the hit flags and each trip's primitive is its child's first.
//
// VK (Vulkan semantics): hits count only inside [t_min, t_max]
// (test_prim_vk), and the cull starts at t_max instead of infinity.
template <int BLOCK, bool COUNT, bool SINGLE, class StackT, bool LDS, bool W4, bool LEAF1 = false,
          bool VK = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(lanes_min_waves(SINGLE, W4), 8)))
void k_extend_lanes(IterArgs a, int p) {
  const float cull0 = VK ? cull_limit(kVkTMax, 0u) : INFINITY;
  __shared__ StackT static_stack[LDS ? 1 : kStackDepth * BLOCK];
  extern __shared__ __align__(16) unsigned char dynamic_lds[];
  __shared__ SegPlan pl;
  StackT *stack = static_stack;
  VecF4 *lds_nodes = nullptr, *lds_prims = nullptr;
  if (LDS) {
    const DevJob &J = a.jobs[0];
    const uint32_t stack_bytes = (a.stack_rows * BLOCK * (uint32_t)sizeof(StackT) + 15u) & ~15u;
    stack = reinterpret_cast<StackT *>(dynamic_lds);
    lds_nodes = reinterpret_cast<VecF4 *>(dynamic_lds + stack_bytes);
    const uint32_t node_vecs = W4 ? 8 * J.n_nodes4 : 4 * J.n_nodes;
    lds_prims = lds_nodes + node_vecs;
    GlobalF4 *gn = W4 ? (GlobalF4 *)J.nodes4 : (GlobalF4 *)J.nodes;
    GlobalF4 *gp = (GlobalF4 *)J.prims;
    for (uint32_t i = threadIdx.x; i < node_vecs; i += BLOCK) lds_nodes[i] = gn[i];
    for (uint32_t i = threadIdx.x; i < 2 * J.n_prims; i += BLOCK) lds_prims[i] = gp[i];
  }
  make_plan(a, p, pl, a.first_hit_mode == 2);  // its barriers also publish the LDS copy
  // queue[1-p] is empty until shade fills it; nothing else reads it now
  if (blockIdx.x == 0 && threadIdx.x < kSegments) a.ctl->count[1 - p][threadIdx.x * kSegStride] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl->cursor[1 - p] = pl.cursor + pl.n_new;
  const PathQueue q = queue_of(a, p);
  const uint32_t n = pl.n;
  const uint32_t waves = gridDim.x * (BLOCK / 64);
  const uint32_t wave = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
  // A wave's share of the queue: one contiguous range (chunk_log2 = 0), or
  // the queue's chunks of 2^chunk_log2 >= 64 consecutive indices dealt
  // round-robin (wave w takes chunks w, w + waves, ...).  The queue keeps
  // neighbouring pixels together, so a contiguous share can be all sky or all
  // clutter and its wave ends early or late; dealt chunks keep neighbours
  // together only within a chunk and give every wave a sample of the whole
  // queue.  c_next and c_end count the wave's own indices (scalar registers:
  // both are wave-uniform); a refill of at most 64 lanes spans at most two
  // chunks, whose queue offsets are computed once per refill.
  const uint32_t cl = a.chunk_log2;
  const uint32_t cmask = (1u << cl) - 1u;
  uint32_t c_next, c_end;
  if (cl) {
    const uint32_t nc = n >> cl, rem = n & cmask;
    const uint32_t full = nc > wave ? (nc - wave - 1u) / waves + 1u : 0u;
    c_next = 0;
    c_end = (full << cl) + (nc % waves == wave ? rem : 0u);
  } else {
    c_next = (uint32_t)((uint64_t)n * wave / waves);
    c_end = (uint32_t)((uint64_t)n * (wave + 1) / waves);
  }
  c_next = __builtin_amdgcn_readfirstlane(c_next);
  c_end = __builtin_amdgcn_readfirstlane(c_end);
  // queue index of the wave's own index c (cl > 0)
  auto queue_index = [&](uint32_t c) { return ((((c >> cl) * waves + wave) << cl) | (c & cmask)); };
  bool more = c_next < c_end;  // wave-uniform: the wave still has rays to hand out
  StackT *stk = stack + threadIdx.x;
  const uint32_t refill = (uint32_t)a.refill;

  const DevJob &J0 = a.jobs[0];
  const DevNode *nodes = J0.nodes;
  const DevNode4 *nodes4 = J0.nodes4;
  const DevPrim *prims = J0.prims;
  // job-0 fields read once (a single-job launch needs no per-ray job loads)
  const uint32_t j0_prims = J0.n_prims, j0_nodes = J0.n_nodes;
  const float j0_r0 = J0.r0_tol;
#ifdef RDH_CHECKED
  uint32_t cur_nodes = 0, cur_prims = 0;
  const uint32_t rows = LDS ? a.stack_rows : (uint32_t)kStackDepth;
  const uint32_t j0_nodes4 = J0.n_nodes4;
#endif
  bool active = false;
  uint32_t k = 0, seg = 0, bid = kNoHit;
  int32_t node = 0;
  int sp = 0;
  float bt = 0.0f, tcull = INFINITY;
  RayPre r;
  SlabRay s;
  SlabNF nf;  // the 4-wide walk's near/far form of s
  TraceCounts tc;
  uint32_t rays = 0;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    if (more && (__popcll(idle) >= refill || idle == ~0ull)) {
      // hand the wave's next indices to its idle lanes, in lane order
      if (COUNT && lane_id() == 0) tc.refills++;
      const uint32_t take = min((uint32_t)__popcll(idle), c_end - c_next);
      const uint32_t rank = lanes_below(idle);
      const bool got = !active && rank < take;
      uint32_t mine = c_next + rank;
      if (cl) {
        // lanes below `left` stay in c_next's chunk, the others start the next
        const uint32_t left = cmask + 1u - (c_next & cmask);
        const uint32_t b0 = queue_index(c_next), b1 = queue_index(c_next + left) - left;
        mine = (rank < left ? b0 : b1) + rank;
      }
      c_next = __builtin_amdgcn_readfirstlane(c_next + take);
      more = c_next < c_end;
      // a refilled lane's start: the exact linear scan, or a walk from the root
      auto start = [&](F3 o, F3 d, uint32_t n_prims, uint32_t n_nodes, float r0) {
        r = ray_pre(o, d);
        bt = 0.0f;
        bid = kNoHit;
        if (n_prims == 0 || a.brute_force || !walkable(r) || n_nodes == 0) {
          // the reference's linear scan (cpu.rs:345-349), also the exact
          // answer for rays the box walk cannot represent (walkable()); done at once
          for (uint32_t j = 0; j < n_prims; ++j) {
            DevPrim pr;
            load_prim(prims, j, &pr);
            if (COUNT) {
              if (fbits(pr.b[3]) & kTagCube) tc.cube++;
              else tc.sphere++;
            }
            if (VK) test_prim_vk(r, pr, &bt, &bid);
            else test_prim(r, pr, &bt, &bid);
          }
          a.hit_t[k] = bid == kNoHit ? 0.0f : bt;
          a.hit_id[k] = bid;
          ++rays;
        } else {
          s = slab_ray(r, ray_tolerance(r, r0));
          if (W4) nf = slab_nf(s);
          node = 0;
          sp = 0;
          tcull = cull0;  // cull_limit of "no hit yet"
          active = true;
        }
      };
      if (SINGLE) {
        // (the start written out again: through the lambda the compiler
        // allocates this kernel's registers differently)
        if (got) {
          k = chk(slot_of(pl, a.seg_cap, mine, seg), a.slots, K_SLOT, a.ctl);
          F3 o, d;
          uint32_t tbk = 0;
          load_ray<false>(q, k, &o, &d, &tbk);
#ifdef RDH_CHECKED
          cur_nodes = W4 ? j0_nodes4 : j0_nodes;
          cur_prims = j0_prims;
#endif
          r = ray_pre(o, d);
          bt = 0.0f;
          bid = kNoHit;
          if (j0_prims == 0 || a.brute_force || !walkable(r) || j0_nodes == 0) {
            for (uint32_t j = 0; j < j0_prims; ++j) {
              DevPrim pr;
              load_prim(prims, j, &pr);
              if (COUNT) {
                if (fbits(pr.b[3]) & kTagCube) tc.cube++;
                else tc.sphere++;
              }
              if (VK) test_prim_vk(r, pr, &bt, &bid);
              else test_prim(r, pr, &bt, &bid);
            }
            a.hit_t[k] = bid == kNoHit ? 0.0f : bt;
            a.hit_id[k] = bid;
            ++rays;
          } else {
            s = slab_ray(r, ray_tolerance(r, j0_r0));
            if (W4) nf = slab_nf(s);
            node = 0;
            sp = 0;
            tcull = cull0;
            active = true;
          }
        }
      } else {
        // several jobs: one round per distinct ticket among the refilled
        // lanes (nearly always one), reading the ticket and job records at
        // a wave-uniform index as scalar loads; a refilled lane was idle, so
        // its tree pointers are free to take its new job's
        F3 o = f3(0.0f, 0.0f, 0.0f), d = o;
        uint32_t ti = 0, n_prims = 0, n_nodes = 0;
        float r0 = 0.0f;
        bool go = false;  // a refilled lane with a ray to trace
        if (got) {
          k = chk(slot_of(pl, a.seg_cap, mine, seg), a.slots, K_SLOT, a.ctl);
          uint32_t tbk = 0, pwk = 0;
          load_ray<true>(q, k, &o, &d, &tbk, &pwk);
          ti = chk(tbk >> 8, a.ticket_cap, K_TICKET, a.ctl);
          // a camera path whose hit generate wrote stays idle
          go = (pwk & kPreHitFlag) == 0u;
        }
        uint64_t todo = __ballot(go);
        while (todo) {
          const int leader = __ffsll((unsigned long long)todo) - 1;
          const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)ti, leader);
          // membership from the ballot's bit: under ti == t the compiler
          // would read the tables at ti again
          const uint64_t same = __ballot(go && ti == t);
          todo &= ~same;
          if ((same >> lane_id()) & 1ull) {
            const DevTicket T = load_constant(a.tickets + t);
            const DevJob J = load_constant(a.jobs + chk(T.job, a.job_cap, K_JOB, a.ctl));
            n_prims = J.n_prims;
            n_nodes = J.n_nodes;
            r0 = J.r0_tol;
            nodes = J.nodes;
            nodes4 = J.nodes4;
            prims = J.prims;
#ifdef RDH_CHECKED
            cur_nodes = W4 ? J.n_nodes4 : J.n_nodes;
            cur_prims = J.n_prims;
#endif
          }
        }
        if (go) start(o, d, n_prims, n_nodes, r0);
      }
    }
    if (__ballot(active) == 0) {
      if (!more) break;
      continue;
    }
    if (COUNT) {
      if (lane_id() == 0) tc.wave_steps++;
      tc.lane_steps += active;
    }
    if (active) {
      if constexpr (W4) {
#ifdef RDH_CHECKED
        node = (int32_t)chk((uint32_t)node, cur_nodes, K_NODE, a.ctl);
#endif
        // each axis's near and far planes of the four children, picked by
        // the ray's direction signs (slab_hit_nf)
        VecF4 xn, xf, yn, yf, zn, zf;
        VecI4 ch;
        if (LDS) {
          const unsigned char *np = reinterpret_cast<const unsigned char *>(lds_nodes + 8 * node);
          xn = *(const VecF4 *)(np + nf.ox); xf = *(const VecF4 *)((np - nf.ox) + 16);
          yn = *(const VecF4 *)(np + nf.oy); yf = *(const VecF4 *)((np - nf.oy) + 80);
          zn = *(const VecF4 *)(np + nf.oz); zf = *(const VecF4 *)((np - nf.oz) + 144);
          ch = *(const VecI4 *)(np + 96);
        } else {
          // 32-bit offsets from the tree's base keep the address arithmetic
          // (and, for one job, the base itself) out of 64-bit registers
          const __attribute__((address_space(1))) unsigned char *gb =
              (const __attribute__((address_space(1))) unsigned char *)nodes4;
          const uint32_t nb = (uint32_t)node * (uint32_t)sizeof(DevNode4);
          xn = *(GlobalF4 *)(gb + (nb + nf.ox)); xf = *(GlobalF4 *)(gb + ((nb - nf.ox) + 16u));
          yn = *(GlobalF4 *)(gb + (nb + nf.oy)); yf = *(GlobalF4 *)(gb + ((nb - nf.oy) + 80u));
          zn = *(GlobalF4 *)(gb + (nb + nf.oz)); zf = *(GlobalF4 *)(gb + ((nb - nf.oz) + 144u));
          ch = *(GlobalI4 *)(gb + (nb + 96u));
        }
        float t0, t1, t2, t3;
        const bool h0 = slab_hit_nf(nf, xn.x, xf.x, yn.x, yf.x, zn.x, zf.x, tcull, &t0);
        const bool h1 = slab_hit_nf(nf, xn.y, xf.y, yn.y, yf.y, zn.y, zf.y, tcull, &t1);
        const bool h2 = slab_hit_nf(nf, xn.z, xf.z, yn.z, yf.z, zn.z, zf.z, tcull, &t2);
        const bool h3 = slab_hit_nf(nf, xn.w, xf.w, yn.w, yf.w, zn.w, zf.w, tcull, &t3);
        if (COUNT) {  // empty slots (leaf code with no primitives) are not boxes of the tree
          const int32_t e = leaf_code(0, 0);
          tc.box += (ch.x != e) + (ch.y != e) + (ch.z != e) + (ch.w != e);
        }
        const bool l0 = h0 && is_leaf(ch.x), l1 = h1 && is_leaf(ch.y);
        const bool l2 = h2 && is_leaf(ch.z), l3 = h3 && is_leaf(ch.w);
        const uint32_t n0 = l0 ? (LEAF1 ? 1u : leaf_count(ch.x)) : 0u;
        const uint32_t p1 = n0 + (l1 ? (LEAF1 ? 1u : leaf_count(ch.y)) : 0u);
        const uint32_t p2 = p1 + (l2 ? (LEAF1 ? 1u : leaf_count(ch.z)) : 0u);
        const uint32_t n_leaf = p2 + (l3 ? (LEAF1 ? 1u : leaf_count(ch.w)) : 0u);
        // trip t tests f_k + t for the k-th hit leaf (one primitive each
        // with LEAF1: trip t is that leaf's first primitive)
        const uint32_t f0 = leaf_first(ch.x), f1 = leaf_first(ch.y) - (LEAF1 ? 0u : n0);
        const uint32_t f2 = leaf_first(ch.z) - (LEAF1 ? 0u : p1);
        const uint32_t f3 = leaf_first(ch.w) - (LEAF1 ? 0u : p2);
        // The next node and the stack are settled before the primitive tests
        // (they depend only on this node's boxes), so the children's keys
        // and ids are dead while the primitives are tested.
        // entered inner children in order of entry distance
        const bool i0 = h0 && !l0, i1 = h1 && !l1, i2 = h2 && !l2, i3 = h3 && !l3;
        const int cnt = (int)i0 + (int)i1 + (int)i2 + (int)i3;
        int32_t c0;
        if (__ballot(cnt > 1) == 0) {
          // no lane entered two inner children: the entered one (if any) is
          // the nearest and nothing is pushed
          c0 = i0 ? ch.x : i1 ? ch.y : i2 ? ch.z : ch.w;
        } else {
          if (COUNT && wave_leader()) tc.sort_steps++;
          float k0 = i0 ? t0 : INFINITY, k1 = i1 ? t1 : INFINITY;
          float k2 = i2 ? t2 : INFINITY, k3 = i3 ? t3 : INFINITY;
          int32_t c1 = ch.y, c2 = ch.z, c3 = ch.w;
          c0 = ch.x;
          order_pair(k0, c0, k1, c1);
          order_pair(k2, c2, k3, c3);
          order_pair(k0, c0, k2, c2);
          order_pair(k1, c1, k3, c3);
          order_pair(k1, c1, k2, c2);
#ifdef RDH_CHECKED
          sp = (int)chk((uint32_t)sp + 2u, rows, K_STACK, a.ctl) - 2;
          if (sp < 0) sp = 0;
#endif
          stk[sp * BLOCK] = (StackT)(cnt == 4 ? c3 : cnt == 3 ? c2 : c1);
          stk[(sp + 1) * BLOCK] = (StackT)(cnt == 4 ? c2 : c1);
          stk[(sp + 2) * BLOCK] = (StackT)c1;
        }
        const int top = sp > 0 ? sp - 1 : 0;
        const int32_t popped = (int32_t)stk[top * BLOCK];
        const bool done = cnt == 0 && sp == 0;
        if (cnt > 0) {
          node = c0;
          sp += cnt - 1;
        } else if (sp > 0) {
          node = popped;
          sp = top;
        }
        if (n_leaf) {
          for (uint32_t t = 0; t < n_leaf; ++t) {
            DevPrim pr;
            uint32_t j = LEAF1 ? (t < n0 ? f0 : t < p1 ? f1 : t < p2 ? f2 : f3)
                               : (t < n0 ? f0 + t : t < p1 ? f1 + t : t < p2 ? f2 + t : f3 + t);
#ifdef RDH_CHECKED
            j = chk(j, cur_prims, K_PRIM, a.ctl);
#endif
            if (LDS) {
              const VecF4 pa = lds_prims[2 * j], pb = lds_prims[2 * j + 1];
              pr.a[0] = pa.x; pr.a[1] = pa.y; pr.a[2] = pa.z; pr.a[3] = pa.w;
              pr.b[0] = pb.x; pr.b[1] = pb.y; pr.b[2] = pb.z; pr.b[3] = pb.w;
            } else {
              load_prim(prims, j, &pr);
            }
            count_trip<COUNT>(tc, pr);
            if (COUNT) {
              if (fbits(pr.b[3]) & kTagCube) tc.cube++;
              else tc.sphere++;
            }
            if (VK) test_prim_vk(r, pr, &bt, &bid);
            else test_prim(r, pr, &bt, &bid);
          }
          tcull = VK && bid == kNoHit ? cull0 : cull_limit(bt, bid);
        }
        if (done) {
          a.hit_t[k] = bid == kNoHit ? 0.0f : bt;
          a.hit_id[k] = bid;
          active = false;
          ++rays;
        }
      } else {
        // one node: both child boxes, then the primitives of the hit leaf
        // children in one loop (lanes with a leaf on either side run it
        // together; the closest hit does not depend on the test order)
  #ifdef RDH_CHECKED
        node = (int32_t)chk((uint32_t)node, cur_nodes, K_NODE, a.ctl);
  #endif
        VecF4 bx, by, bz;
        VecI4 ch;
        if (LDS) {
          const VecF4 *np = lds_nodes + 4 * node;
          bx = np[0]; by = np[1]; bz = np[2];
          ch = __builtin_bit_cast(VecI4, np[3]);
        } else {
          GlobalF4 *np = (GlobalF4 *)(nodes + node);
          bx = np[0]; by = np[1]; bz = np[2];
          ch = ((GlobalI4 *)np)[3];
        }
        float t0, t1;
        const bool h0 = slab_hit(s, bx.x, bx.z, by.x, by.z, bz.x, bz.z, tcull, &t0);
        const bool h1 = slab_hit(s, bx.y, bx.w, by.y, by.w, bz.y, bz.w, tcull, &t1);
        if (COUNT) tc.box += 2;
        const bool l0 = h0 && is_leaf(ch.x), l1 = h1 && is_leaf(ch.y);
        const uint32_t n0 = l0 ? leaf_count(ch.x) : 0u;
        const uint32_t n_leaf = n0 + (l1 ? leaf_count(ch.y) : 0u);
        if (n_leaf) {
          const uint32_t f0 = leaf_first(ch.x), f1 = leaf_first(ch.y) - n0;
          for (uint32_t t = 0; t < n_leaf; ++t) {
            DevPrim pr;
            uint32_t j = t < n0 ? f0 + t : f1 + t;
  #ifdef RDH_CHECKED
            j = chk(j, cur_prims, K_PRIM, a.ctl);
  #endif
            if (LDS) {
              const VecF4 pa = lds_prims[2 * j], pb = lds_prims[2 * j + 1];
              pr.a[0] = pa.x; pr.a[1] = pa.y; pr.a[2] = pa.z; pr.a[3] = pa.w;
              pr.b[0] = pb.x; pr.b[1] = pb.y; pr.b[2] = pb.z; pr.b[3] = pb.w;
            } else {
              load_prim(prims, j, &pr);
            }
            count_trip<COUNT>(tc, pr);
            if (COUNT) {
              if (fbits(pr.b[3]) & kTagCube) tc.cube++;
              else tc.sphere++;
            }
            if (VK) test_prim_vk(r, pr, &bt, &bid);
            else test_prim(r, pr, &bt, &bid);
          }
          // changes only when a primitive is tested
          tcull = VK && bid == kNoHit ? cull0 : cull_limit(bt, bid);
        }
        // next node without branches: the far child is written above the
        // stack top every step (kept only when both children are entered;
        // slot sp < depth is in bounds) and the top entry is read every step
        const bool i0 = h0 && !l0, i1 = h1 && !l1;
        const bool near0 = i0 && (!i1 || t0 <= t1);
        const int32_t near_c = near0 ? ch.x : ch.y;
  #ifdef RDH_CHECKED
        sp = (int)chk((uint32_t)sp, rows, K_STACK, a.ctl);
  #endif
        stk[sp * BLOCK] = (StackT)(near0 ? ch.y : ch.x);
        const int top = sp > 0 ? sp - 1 : 0;
        const int32_t popped = (int32_t)stk[top * BLOCK];
        if (i0 || i1) {
          node = near_c;
          sp += (i0 && i1) ? 1 : 0;
        } else if (sp > 0) {
          node = popped;
          sp = top;
        } else {
          a.hit_t[k] = bid == kNoHit ? 0.0f : bt;
          a.hit_id[k] = bid;
          active = false;
          ++rays;
        }
      }
    }
  }
  if (COUNT) {
    if (rays) atomicAdd(&a.counters->v[C_SEGMENTS], (unsigned long long)rays);
    if (tc.box) atomicAdd(&a.counters->v[C_BOX], (unsigned long long)tc.box);
    if (tc.sphere) atomicAdd(&a.counters->v[C_SPHERE], (unsigned long long)tc.sphere);
    if (tc.cube) atomicAdd(&a.counters->v[C_CUBE], (unsigned long long)tc.cube);
    const uint32_t w[6] = {tc.wave_steps, tc.lane_steps, tc.sort_steps, tc.prim_trips,
                           tc.cube_trips, tc.refills};
    for (int i = 0; i < 6; ++i)
      if (w[i]) atomicAdd(&a.counters->v[C_WAVE_STEPS + i], (unsigned long long)w[i]);
  }
}

// --------------------------------------------------------------- shade ----
// Survivors are compacted per workgroup: a ballot + mbcnt gives each lane its
// slot inside its wave, an LDS scan over the block's waves gives each wave its
// offset, and one atomic per workgroup iteration reserves the block's range
// in its own segment (blockIdx % kSegments) of the next queue.  The host
// sizes segments so that the blocks sharing one can never overfill it; the
// overflow flag is a guard, not a mode of operation.
// Single-job shade is held to at least 7 waves/SIMD (it needs 63 VGPRs and
// runs at 8).  The multi-job variant (92 VGPRs, 5 waves) would spill at 6,
// so it keeps the compiler's choice.
constexpr int shade_min_waves(bool single) { return single ? 7 : 1; }

// VK (Vulkan semantics, single job, one sample per pass): the scatter and
// sky of raytrace.rgen, drawing from the pixel's PCG state, which the path
// (its pixel's only one in the pass) reads and writes back at every hit.
template <int BLOCK, bool COUNT, bool SINGLE, bool VK = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(shade_min_waves(SINGLE), 8)))
void k_shade(IterArgs a, int p) {
  constexpr int kWaves = BLOCK / 64;
  __shared__ uint32_t wave_off[kWaves];
  __shared__ uint32_t block_base;
  __shared__ SegPlan pl;
  __shared__ DoneTable done;
  if (threadIdx.x == 0) done.n = 0;
  if (threadIdx.x < kDoneEntries) {
    done.ticket[threadIdx.x] = 0xffffffffu;
    done.count[threadIdx.x] = 0;
  }
  // A small scene's objects are copied into LDS (lds_obj_bytes, launch
  // constant): the object read follows the hit id, and from LDS it does not
  // add a second global-memory latency to every path.
  extern __shared__ __align__(16) unsigned char dynamic_lds[];
  // One job (SINGLE): the job and its one ticket are read here, before the
  // kernel's first global store, as scalar loads whose values stay in scalar
  // registers.  Read through the pointers inside the loop, after stores,
  // each field would be a vector load and a memory round trip in every
  // path's chain.
  DevJob job0;
  DevTicket ticket0;
  if (SINGLE) {
    job0 = a.jobs[0];
    ticket0 = a.tickets[0];
  }
  const bool lds_obj = SINGLE && a.lds_obj_bytes != 0;
  VecF4 *lds_objects = reinterpret_cast<VecF4 *>(dynamic_lds);
  if (lds_obj) {
    GlobalF4 *g = (GlobalF4 *)a.jobs[0].objects;
    for (uint32_t v = threadIdx.x; v < a.lds_obj_bytes / 16u; v += BLOCK) lds_objects[v] = g[v];
  }
  make_plan(a, p, pl);  // its barriers also publish the table's reset and the LDS copy
  const PathQueue q = queue_of(a, p), nq = queue_of(a, 1 - p);
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x >> 6;
  const uint32_t out_seg = blockIdx.x % kSegments;
  uint32_t *out_count = &a.ctl->count[1 - p][out_seg * kSegStride];
  uint32_t c_sph = 0, c_cube = 0, c_glass = 0, c_miss = 0, seg = 0;
  for (uint32_t base = blockIdx.x * BLOCK; base < pl.n; base += gridDim.x * BLOCK) {
    const uint32_t i = base + threadIdx.x;
    const bool has = i < pl.n;
    bool alive = false, finished = false, pre = false;
    uint32_t ti = 0, pidx = 0, bounce = 0, id = kNoHit;
    float hit_t = 0.0f;
    F3 o = f3(0, 0, 0), d = o, at = o, li = o;
    F3 no = o, nd = o, na = o, nl = o;
    uint64_t l_slot = 0;
    if (has) {
      const uint32_t k = chk(slot_of(pl, a.seg_cap, i, seg), a.slots, K_SLOT, a.ctl);
      // the ray, its words and its hit in one batch of reads; attenuation
      // and light only where the bounce and lit flag say they exist (read
      // whatever the flags say, they cost more in traffic than the round
      // trip they save)
      const VecF4 ra = *(GlobalF4 *)(q.ray + k);
      const VecF2 rd = *(GlobalF2 *)(q.dir + k);
      const VecU2 wd = *(GlobalU2 *)(q.word + k);
      id = a.hit_id[k];
      hit_t = a.hit_t[k];
      o = f3(ra.x, ra.y, ra.z);
      d = f3(ra.w, rd.x, rd.y);
      const uint32_t tb = wd.x;
      ti = chk(ticket_of<SINGLE>(tb), a.ticket_cap, K_TICKET, a.ctl);
      bounce = tb & 0xffu;
      const uint32_t pw = wd.y;
      pidx = pw & kPathIndexMask;
      pre = (pw & kPreHitFlag) != 0u;
      at = f3(1.0f, 1.0f, 1.0f);
      if (bounce != 0) {
        const VecF4 aw = *(GlobalF4 *)(q.att + k);
        at = f3(aw.x, aw.y, aw.z);
        // light stays exactly (+0, +0, +0) until an emitter or the sky adds
        // to it; only lit paths carry it in the queue (x beside the
        // attenuation, y and z in QLight)
        if (pw & kLitFlag) {
          const VecF2 l2 = *(GlobalF2 *)(q.light + k);
          li = f3(aw.w, l2.x, l2.y);
        }
      }
    }
    // the path's bounce with its ticket T and job J
    auto shade_path = [&](const DevJob &J, const DevTicket &T) {
      l_slot = chk((unsigned long long)(T.l_offset + pidx), a.l_cap, K_RADIANCE, a.ctl);
      if (bounce == 0 && (SINGLE ? a.first_hit_mode == 1 : (J.first_hit != nullptr && !pre))) {
        // a camera path's hit is its pixel's for every sample (distance bits
        // + 1, so that zero means not recorded)
        const uint32_t npix = J.width * J.height;
        HitWord *fh = SINGLE ? a.first_hit : J.first_hit;
        fh[chk(pidx % npix, npix, K_RADIANCE, a.ctl)] = HitWord{fbits(hit_t) + 1u, id};
      }
      if (id != kNoHit) {
        const uint32_t oid = chk(id, J.n_prims, K_OBJECT, a.ctl);
        const DevObject ob = lds_obj ? object_from(lds_objects[4 * oid], lds_objects[4 * oid + 1],
                                                   lds_objects[4 * oid + 2], lds_objects[4 * oid + 3])
                                     : load_object(J.objects, oid);
        const HitRec h = closest_hit(o, d, hit_t, ob);
        const uint32_t npix = J.width * J.height;
        ScatterOut so;
        if (VK) {
          uint32_t *rng = &J.vk_rng[chk(pidx % npix, npix, K_RADIANCE, a.ctl)];
          uint32_t st = *rng;
          so = scatter_vk(d, h, ob, st, at, li);
          *rng = st;
        } else {
          BounceDraws dr;
          dr.init(J.seed, pidx % npix, T.sample_base + pidx / npix, bounce);
          so = scatter(d, h, ob, dr, at, li);
        }
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
        // cpu.rs:334-337 (raytrace.rgen:178)
        nl = add(li, mul(VK ? world_sample_vk(J, d) : world_sample(J, d), at));
        finished = true;
        if (COUNT) c_miss++;
      }
    };
    if (SINGLE) {
      if (has) shade_path(job0, ticket0);
    } else {
      // A wave's paths nearly always share one ticket (the queue keeps a
      // ticket's paths together): one round per distinct ticket, with the
      // ticket and its job read at a wave-uniform index from the constant
      // address space (neither changes during a launch), i.e. as scalar
      // loads that the scalar cache serves, instead of two dependent vector
      // loads and their job-field loads in every path's chain.
      uint64_t todo = __ballot(has);
      while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)ti, leader);
        // membership is taken from the ballot's bit, not from ti == t: under
        // that condition the compiler would read the tables at ti again
        const uint64_t same = __ballot(has && ti == t);
        todo &= ~same;
        if ((same >> lane) & 1ull) {
          const DevTicket T = load_constant(a.tickets + t);
          const DevJob J = load_constant(a.jobs + chk(T.job, a.job_cap, K_JOB, a.ctl));
          shade_path(J, T);
        }
      }
    }
    const uint64_t am = __ballot(alive);
    if (lane == 0) wave_off[wid] = (uint32_t)__popcll(am);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = wave_off[w];
        wave_off[w] = tot;
        tot += c;
      }
      uint32_t b0 = tot ? atomicAdd(out_count, tot) : 0u;
      if (b0 + tot > a.seg_cap) {
        a.ctl->overflow = 1u;  // never written past the segment: the host fails the render
        b0 = 0xffffffffu;
      }
      block_base = b0;
    }
    __syncthreads();
    if (alive && block_base != 0xffffffffu) {
      const uint32_t k = chk(out_seg * a.seg_cap + block_base + wave_off[wid] + lanes_below(am), a.slots,
                             K_SLOT, a.ctl);
      const bool lit = (fbits(nl.x) | fbits(nl.y) | fbits(nl.z)) != 0u;
      store_ray(nq, k, no, nd, (ti << 8) | bounce, pidx | (lit ? kLitFlag : 0u));
      *(GlobalF4Out *)(nq.att + k) = VecF4{na.x, na.y, na.z, nl.x};
      if (lit) *(GlobalF2Out *)(nq.light + k) = VecF2{nl.y, nl.z};
    }
    if (finished) a.L[l_slot] = make_float4(nl.x, nl.y, nl.z, 1.0f);
    if (a.track_done) {
      // finished paths per ticket: each wave folds its lanes per distinct
      // ticket into the block's LDS table, flushed to HBM once per launch
      uint64_t fm = __ballot(finished);
      while (fm) {
        const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)fm) - 1);
        const uint32_t lt = (uint32_t)__builtin_amdgcn_readlane((int)ti, (int)leader);
        const uint64_t same = __ballot(finished && ti == lt);
        if (lane == leader) done_add(done, lt, (uint32_t)__popcll(same), a.ticket_done);
        fm &= ~same;
      }
    }
    __syncthreads();  // wave_off / block_base are rewritten next iteration
  }
  // the loop ended with a barrier: every entry of the table is final
  if (a.track_done && threadIdx.x < done.n && threadIdx.x < kDoneEntries && done.count[threadIdx.x])
    atomicAdd(&a.ticket_done[chk(done.ticket[threadIdx.x], a.ticket_cap, K_TICKET, a.ctl)],
              (unsigned long long)done.count[threadIdx.x]);
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
// VK: the Vulkan renderer's UNORM store (rounds) instead of print_frame_buffer.
template <bool COUNT, bool VK = false>
__global__ void k_tonemap(const DevJob *job, uint32_t n, float spp, uchar4 *out,
                          DevCounters *counters) {
  const float4 *acc = reinterpret_cast<const float4 *>(job->acc);
  uint32_t done = 0;
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < n;
       px += gridDim.x * blockDim.x) {
    const float4 v = acc[px];
    if (VK)
      out[px] = make_uchar4(tonemap_channel_vk(v.x, spp), tonemap_channel_vk(v.y, spp),
                            tonemap_channel_vk(v.z, spp), tonemap_channel_vk(v.w, spp));
    else
      out[px] = make_uchar4(tonemap_channel(v.x, spp), tonemap_channel(v.y, spp),
                            tonemap_channel(v.z, spp), tonemap_channel(v.w, spp));
    ++done;
  }
  if (COUNT && done) atomicAdd(&counters->v[C_TONEMAP], (unsigned long long)done);
}

__global__ void k_set_ctl(QueueCtl *ctl, unsigned long long total, uint32_t n_tickets) {
  for (int s = 0; s < kSegments; ++s) {
    ctl->count[0][s * kSegStride] = 0;
    ctl->count[1][s * kSegStride] = 0;
  }
  ctl->overflow = 0;
  ctl->violation = 0;
  ctl->cursor[0] = 0;
  ctl->cursor[1] = 0;
  ctl->total = total;
  ctl->n_tickets = n_tickets;
  ctl->ticket_lo = 0;
}

// Appends work between iterations (the streaming engine): never touches the
// ping-pong fields the kernels own.
__global__ void k_extend_work(QueueCtl *ctl, unsigned long long total, uint32_t n_tickets,
                              uint32_t ticket_lo) {
  ctl->total = total;
  ctl->n_tickets = n_tickets;
  ctl->ticket_lo = ticket_lo;
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
  trace<64, false, int32_t>(nodes, prims, n_prims, n_nodes, r0, brute, r, stack + threadIdx.x,
                            &bt, &bid, tc);
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

__global__ void k_debug_hits(const DevObject *objects, const float *rays, const uint32_t *ids,
                             const float *ts, uint32_t n, float *normals, uint8_t *front) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (ids[i] == kNoHit) {
    normals[3 * i] = normals[3 * i + 1] = normals[3 * i + 2] = 0.0f;
    front[i] = 0;
    return;
  }
  const float *q = rays + 6 * (uint64_t)i;
  const HitRec h = closest_hit(f3(q[0], q[1], q[2]), f3(q[3], q[4], q[5]), ts[i], objects[ids[i]]);
  normals[3 * i] = h.n.x; normals[3 * i + 1] = h.n.y; normals[3 * i + 2] = h.n.z;
  front[i] = h.front ? 1 : 0;
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
  FixedDraws dr{bits + kDrawWords * (uint64_t)i};
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

// Plain 16-byte-per-lane copy: the reference point for HBM traffic counters.
__global__ void k_copy_calibration(const uint4 *src, uint4 *dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

uint32_t blocks_for(uint64_t n, uint32_t block) { return (uint32_t)((n + block - 1) / block); }

}  // namespace

// ------------------------------------------------------------ launchers ---
// Kernel instance for a launch: block size, counting build, single-job
// specialisation and stack width are template parameters.
template <bool COUNT, bool SINGLE>
static const void *extend_fn(int block, bool wide, bool lanes, bool lds, bool w4, bool leaf1) {
#define RDH_EXT(B)                                                                           \
  if (block == B) {                                                                          \
    if (lanes && !wide && w4) {                                                              \
      if (SINGLE && leaf1)                                                                   \
        return lds ? (const void *)k_extend_lanes<B, COUNT, true, uint16_t, true, true, true> \
                   : (const void *)k_extend_lanes<B, COUNT, true, uint16_t, false, true, true>; \
      if (lds && SINGLE) return (const void *)k_extend_lanes<B, COUNT, true, uint16_t, true, true>; \
      return (const void *)k_extend_lanes<B, COUNT, SINGLE, uint16_t, false, true>;          \
    }                                                                                        \
    if (lanes && lds && SINGLE && !wide)                                                     \
      return (const void *)k_extend_lanes<B, COUNT, true, uint16_t, true, false>;            \
    if (lanes)                                                                               \
      return wide ? (const void *)k_extend_lanes<B, COUNT, SINGLE, int32_t, false, false>    \
                  : (const void *)k_extend_lanes<B, COUNT, SINGLE, uint16_t, false, false>;  \
    return wide ? (const void *)k_extend<B, COUNT, SINGLE, int32_t>              \
                : (const void *)k_extend<B, COUNT, SINGLE, uint16_t>;            \
  }
  RDH_EXT(64)
  RDH_EXT(128)
  RDH_EXT(256)
#undef RDH_EXT
  return nullptr;
}
template <bool COUNT, bool SINGLE>
static const void *shade_fn(int block) {
  if (block == 64) return (const void *)k_shade<64, COUNT, SINGLE>;
  if (block == 128) return (const void *)k_shade<128, COUNT, SINGLE>;
  if (block == 256) return (const void *)k_shade<256, COUNT, SINGLE>;
  return nullptr;
}
// The Vulkan-semantics kernels: single job, 256 threads, persistent lanes,
// no counting build.
static const void *vk_kernel_fn(int which, bool wide, bool lds, bool w4) {
  if (which == 0) return (const void *)k_generate<false, true, true>;
  if (which == 2) return (const void *)k_shade<256, false, true, true>;
  if (w4 && !wide)
    return lds ? (const void *)k_extend_lanes<256, false, true, uint16_t, true, true, false, true>
               : (const void *)k_extend_lanes<256, false, true, uint16_t, false, true, false, true>;
  if (lds && !wide)
    return (const void *)k_extend_lanes<256, false, true, uint16_t, true, false, false, true>;
  return wide ? (const void *)k_extend_lanes<256, false, true, int32_t, false, false, false, true>
              : (const void *)k_extend_lanes<256, false, true, uint16_t, false, false, false, true>;
}
static const void *kernel_fn(int which, int block, bool count, bool single, bool wide,
                             bool lanes = false, bool lds = false, bool w4 = false,
                             bool leaf1 = false) {
  if (which == 0) {
    if (count) return single ? (const void *)k_generate<true, true> : (const void *)k_generate<true, false>;
    return single ? (const void *)k_generate<false, true> : (const void *)k_generate<false, false>;
  }
  if (which == 1) {
    if (count)
      return single ? extend_fn<true, true>(block, wide, lanes, lds, w4, leaf1)
                    : extend_fn<true, false>(block, wide, lanes, lds, w4, leaf1);
    return single ? extend_fn<false, true>(block, wide, lanes, lds, w4, leaf1)
                  : extend_fn<false, false>(block, wide, lanes, lds, w4, leaf1);
  }
  if (count) return single ? shade_fn<true, true>(block) : shade_fn<true, false>(block);
  return single ? shade_fn<false, true>(block) : shade_fn<false, false>(block);
}

int occupancy_blocks(int which, int block, bool lanes, uint32_t lds_bvh_bytes,
                     uint32_t dynamic_lds, bool w4, bool single) {
  const void *fn = kernel_fn(which, block, false, single, false, lanes, lds_bvh_bytes > 0, w4);
  int n = 0;
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, block, dynamic_lds) != hipSuccess) {
    (void)hipGetLastError();
    return 1;
  }
  return n;
}

uint32_t extend_dynamic_lds(const IterArgs &a, int block) {
  if (a.refill <= 0 || a.lds_bvh_bytes == 0 || !a.single_job || a.wide_stack) return 0;
  return ((a.stack_rows * (uint32_t)block * 2u + 15u) & ~15u) + a.lds_bvh_bytes;
}

uint32_t shade_dynamic_lds(const IterArgs &a) { return a.single_job ? a.lds_obj_bytes : 0u; }

static hipError_t launch_stage(int which, const IterArgs &a, int parity, const LaunchShape &s,
                               bool count, hipStream_t st) {
  const uint32_t dyn = which == 1 ? extend_dynamic_lds(a, s.block) : which == 2 ? shade_dynamic_lds(a) : 0;
  const void *fn;
  if (a.vk) {
    // the host admits Vulkan semantics only for these launches
    if (count || s.block != 256 || !a.single_job || a.refill <= 0) return hipErrorInvalidValue;
    fn = vk_kernel_fn(which, a.wide_stack != 0, dyn > 0, a.bvh4 != 0);
  } else {
    fn = kernel_fn(which, s.block, count, a.single_job != 0, a.wide_stack != 0, a.refill > 0,
                   dyn > 0, a.bvh4 != 0, a.leaf1 != 0);
  }
  if (!fn) return hipErrorInvalidValue;
  IterArgs args = a;
  int par = parity;
  void *kargs[] = {&args, &par};
  return hipLaunchKernel(fn, dim3(s.grid), dim3(s.block), kargs, dyn, st);
}

hipError_t launch_generate(const IterArgs &a, int parity, const LaunchShape &s, bool count,
                           hipStream_t st) {
  return launch_stage(0, a, parity, s, count, st);
}

hipError_t launch_extend(const IterArgs &a, int parity, const LaunchShape &s, bool count,
                         hipStream_t st) {
  return launch_stage(1, a, parity, s, count, st);
}

hipError_t launch_shade(const IterArgs &a, int parity, const LaunchShape &s, bool count,
                        hipStream_t st) {
  return launch_stage(2, a, parity, s, count, st);
}

hipError_t launch_accumulate(const DevJob *job, const DevTicket &t, const float4 *L,
                             DevCounters *counters, bool count, int grid, hipStream_t st) {
  if (count) hipLaunchKernelGGL(k_accumulate<true>, dim3(grid), dim3(256), 0, st, job, t, L, counters);
  else hipLaunchKernelGGL(k_accumulate<false>, dim3(grid), dim3(256), 0, st, job, t, L, counters);
  return hipGetLastError();
}

hipError_t launch_tonemap(const DevJob *job, uint32_t n_pixels, float spp, uchar4 *out,
                          DevCounters *counters, bool count, int grid, hipStream_t st, bool vk) {
  if (vk) hipLaunchKernelGGL((k_tonemap<false, true>), dim3(grid), dim3(256), 0, st, job, n_pixels, spp, out, counters);
  else if (count) hipLaunchKernelGGL(k_tonemap<true>, dim3(grid), dim3(256), 0, st, job, n_pixels, spp, out, counters);
  else hipLaunchKernelGGL(k_tonemap<false>, dim3(grid), dim3(256), 0, st, job, n_pixels, spp, out, counters);
  return hipGetLastError();
}

hipError_t launch_set_ctl(QueueCtl *ctl, unsigned long long total, uint32_t n_tickets,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(1), 0, st, ctl, total, n_tickets);
  return hipGetLastError();
}

hipError_t launch_extend_work(QueueCtl *ctl, unsigned long long total, uint32_t n_tickets,
                              uint32_t ticket_lo, hipStream_t st) {
  hipLaunchKernelGGL(k_extend_work, dim3(1), dim3(1), 0, st, ctl, total, n_tickets, ticket_lo);
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

hipError_t launch_debug_hits(const DevObject *objects, const float *rays, const uint32_t *ids,
                             const float *ts, uint32_t n, float *normals, uint8_t *front,
                             hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_hits, dim3(blocks_for(n, 256)), dim3(256), 0, st, objects, rays, ids,
                     ts, n, normals, front);
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

hipError_t launch_copy_calibration(const void *src, void *dst, uint64_t bytes, hipStream_t st) {
  const uint64_t n = bytes / 16;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_calibration, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                     static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
  return hipGetLastError();
}

hipError_t launch_debug_philox(const uint32_t *in, uint32_t n, uint32_t *out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_philox, dim3(blocks_for(n, 256)), dim3(256), 0, st, in, n, out);
  return hipGetLastError();
}

}  // namespace rdh
