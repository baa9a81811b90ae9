// Device memory layout of the wavefront path tracer (host + device view).
//
// Everything a kernel touches lives in HBM in one of these forms:
//   DevNode    64 B  binary BVH node holding BOTH children's boxes, so one
//                    node fetch (4 x 16 B) feeds two slab tests
//   DevPrim    32 B  primitive in BVH leaf order, pre-digested for the exact
//                    hit tests of src/renderer/cpu.rs:34-98
//   DevObject  64 B  geometry + material by ORIGINAL object index (the index
//                    the closest-hit tie rule and the shading use)
//   DevJob           one scene being rendered: camera, world, scene buffers,
//                    accumulator, RNG seed, bounce limit
//   DevTicket        one pass of one job: `paths` consecutive path indices
//                    s*W*H + p, s in [0, spp); radiance slots in the L arena
//   Path queue       five arrays of packed path fields, 56 B per path in flight
#pragma once

#include <stdint.h>

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#define __device__
#endif
#endif

namespace rdh {

constexpr int kMaxLeafSize = 8;
constexpr int kStackDepth = 32;          // >= max BVH depth the builder emits
constexpr uint32_t kNarrowStackNodes = 65535;  // 16-bit LDS stack entries suffice
constexpr uint32_t kNoHit = 0x7fffffffu; // best-id sentinel: "no hit yet"
// Path queue word `pidx`: the path's index in its ticket in bits 0..30 (a
// pass holds at most 2^31 paths: l_capacity <= 2^31), and in bit 31 whether
// the path's light is not (+0, +0, +0); the light words of a queue slot are
// read only when it is set (y and z are also written only then).
constexpr uint32_t kLitFlag = 0x80000000u;
// a camera path whose hit generate already wrote (several jobs: its job's
// per-pixel hit table); path indices stay below 2^30 (radiance arenas hold
// at most 2^30 slots)
constexpr uint32_t kPreHitFlag = 0x40000000u;
constexpr uint32_t kPathIndexMask = 0x3fffffffu;
constexpr int kDrawWords = 8;            // random words per bounce (two Philox blocks)

// Leaf child code: bit 31 set, first prim in bits 4..30, count in bits 0..3.
__host__ __device__ inline bool is_leaf(int32_t c) { return c < 0; }
__host__ __device__ inline int32_t leaf_code(uint32_t first, uint32_t count) {
  return (int32_t)(0x80000000u | (first << 4) | count);
}
__host__ __device__ inline uint32_t leaf_first(int32_t c) { return ((uint32_t)c & 0x7fffffffu) >> 4; }
__host__ __device__ inline uint32_t leaf_count(int32_t c) { return (uint32_t)c & 15u; }

struct alignas(16) DevNode {
  float lo_x[2], hi_x[2];  // child 0 / child 1 x slab  (lo0, lo1, hi0, hi1)
  float lo_y[2], hi_y[2];
  float lo_z[2], hi_z[2];
  int32_t child[2];        // >= 0: internal node index; < 0: leaf code
  uint32_t pad[2];
};
static_assert(sizeof(DevNode) == 64, "DevNode must be 64 bytes");

// 4-wide node (the binary tree collapsed): the boxes of up to four children,
// axis-major with child k in lane k, and their codes.  An unused child is an
// empty leaf, leaf_code(0, 0): hitting its box tests nothing.
struct alignas(16) DevNode4 {
  float lo_x[4], hi_x[4];
  float lo_y[4], hi_y[4];
  float lo_z[4], hi_z[4];
  int32_t child[4];
  uint32_t pad[4];
};
static_assert(sizeof(DevNode4) == 128, "DevNode4 must be 128 bytes");

// Primitive tag in DevPrim::b.w: original object id | kind << 31.
struct alignas(16) DevPrim {
  // sphere: a = (cx, cy, cz, dot(C,C)), b = (r*r, 0, 0, tag)
  // cube:   a = (min x, y, z, 0),       b = (max x, y, z, tag)
  float a[4];
  float b[4];
};
static_assert(sizeof(DevPrim) == 32, "DevPrim must be 32 bytes");

constexpr uint32_t kTagCube = 0x80000000u;

struct alignas(16) DevObject {
  float geo[4];   // center xyz, size (radius or side length)
  float m0[4];    // albedo xyz, roughness^2 (cpu.rs:262-263)
  float m1[4];    // emission_color * emission_strength xyz (cpu.rs:332-333), metallic
  float m2[4];    // transmission, ior, kind (as uint bits), 0
};
static_assert(sizeof(DevObject) == 64, "DevObject must be 64 bytes");

// A pixel's camera-ray hit: distance bits + 1 (0: not recorded yet) and id.
struct alignas(8) HitWord {
  uint32_t t1;
  uint32_t id;
};

struct DevJob {
  const DevNode *nodes;
  const DevPrim *prims;
  const DevObject *objects;
  float *acc;               // W*H*4 RGBA32F accumulator (cpu.rs:113)
  uint32_t n_prims;
  uint32_t n_nodes;
  uint32_t width, height;
  uint32_t max_bounces;
  uint32_t world_kind;
  uint64_t seed;
  float inv_proj[16];       // column-major, as stored in the scene (camera.rs:29)
  float inv_view[16];       // column-major (camera.rs:28)
  float position[4];
  float top[4], bottom[4], solid[4];
  float r0_tol;             // |o|_1 above which a ray needs the dynamic box margin
  uint32_t depth;           // BVH levels (a traversal stack needs fewer entries)
  uint32_t n_nodes4;        // the same tree as 4-wide nodes
  uint32_t stack4;          // stack entries a 4-wide traversal can hold
  uint32_t max_leaf;        // primitives in the largest leaf (1: one per leaf)
  const DevNode4 *nodes4;
  uint32_t *vk_rng;         // Vulkan semantics: each pixel's PCG state between passes
  HitWord *first_hit;       // several jobs: each pixel's camera-ray hit, all zero until
                            // the job's first camera paths are shaded
};

struct DevTicket {
  uint64_t work_begin;      // first global work index of this ticket
  uint64_t l_offset;        // first radiance slot in the L arena
  uint32_t job;
  uint32_t sample_base;     // global sample index of local sample 0
  uint32_t spp;             // samples in this pass
  uint32_t paths;           // spp * W * H
};

// Path queue (one of two ping-pong buffers): five arrays indexed by slot,
// each holding a path's fields in the order the kernels use them, so that a
// path moves in 16- and 8-byte loads and stores (a wave's accesses to
// consecutive slots stay contiguous), the closest-hit kernel reads exactly
// its ray (QRay + QDir), and a kernel needs five base addresses per queue.
struct alignas(16) QRay {
  float ox, oy, oz, dx;     // origin, direction x
};
struct alignas(8) QDir {
  float dy, dz;             // direction y, z
};
struct alignas(8) QWord {
  uint32_t tb;              // ticket << 8 | bounce
  uint32_t pidx;            // path index inside its ticket (s_local * W*H + pixel) | lit flag
};
struct alignas(16) QAtt {
  float ax, ay, az;         // attenuation (cpu.rs:254)
  float lx;                 // light x (cpu.rs:253); valid only with the lit flag
};
struct alignas(8) QLight {
  float ly, lz;             // light y, z; read and written only with the lit flag
};
struct PathQueue {
  QRay *ray;
  QDir *dir;
  QWord *word;
  QAtt *att;
  QLight *light;
};

// Per-iteration control block in device memory (ping-pong by parity).
// Each path queue is split into kSegments segments of seg_cap slots.  Shade
// workgroup b appends its survivors to segment b % kSegments, so the
// reservation atomics spread over kSegments counters on separate cache lines
// instead of serialising on one address.  Consumers see the queue as the
// concatenation of the segments' filled prefixes.
constexpr int kSegments = 32;
constexpr int kSegStride = 32;  // uint32 stride between counters: 128 bytes

// The host only changes n_tickets, ticket_lo and total, between iterations.
struct QueueCtl {
  uint32_t n_tickets;       // tickets visible to generate
  uint32_t ticket_lo;       // no path before this ticket is still to be generated
  uint32_t overflow;        // set if a segment bound was ever exceeded (internal error)
  uint32_t violation;       // bounds-checked builds: bit b = check b failed (internal error)
  unsigned long long cursor[2];  // next global work index to generate
  unsigned long long total;      // end of the work list
  unsigned long long pad1;
  uint32_t pad2[24];             // counters start on their own cache line
  uint32_t count[2][kSegments * kSegStride];  // paths per segment of queue[p]
};

// Algorithmic event counters (collected only by the counting build).
struct DevCounters {
  unsigned long long v[18];
};
enum CounterIndex {
  C_CAMERA = 0, C_SEGMENTS, C_BOX, C_SPHERE, C_CUBE, C_SPHERE_HIT, C_CUBE_HIT,
  C_GLASS, C_MISS, C_ACCUM, C_TONEMAP,
  C_WAVE_STEPS, C_LANE_STEPS, C_SORT_STEPS, C_PRIM_TRIPS, C_CUBE_TRIPS, C_REFILLS, C_REUSED
};

}  // namespace rdh
