// Wavefront kernels of the HIP path tracer and their host-side launchers.
//
// One iteration of parity p (queues q[p] -> q[1-p]) is three launches:
//   generate  appends new camera paths (cpu.rs:234-251) after the survivors
//             in q[p], taking global work indices from the cursor
//   extend    closest hit for every path in q[p] (cpu.rs:344-394) through the
//             BVH with an LDS stack, or by linear scan
//   shade     scatter or sky (cpu.rs:256-338); survivors are compacted into
//             q[1-p] with wave64 ballot + mbcnt and one atomic per wave;
//             finished paths write their radiance to the L arena
// The launches read the queue sizes from device memory (QueueCtl), so the
// host never waits between them and an iteration can be captured in a graph.
// A pass is finished when the cursor reached the end of the work list and the
// queue is empty; accumulate then adds each pixel's samples in sample order
// (cpu.rs:204-209) and tonemap writes RGBA8 (cpu.rs:221-230).
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace rdh {

struct IterArgs {
  PathQueue q[2];
  float *hit_t;
  uint32_t *hit_id;
  QueueCtl *ctl;
  const DevJob *jobs;
  const DevTicket *tickets;         // the ticket count lives in QueueCtl
  uint32_t capacity;                // paths in flight, all segments together
  uint32_t seg_cap;                 // slots per queue segment
  float4 *L;                        // radiance arena (rgb, 1)
  unsigned long long *ticket_done;  // finished paths per ticket
  DevCounters *counters;            // used by the counting build only
  int brute_force;                  // 1: linear scan instead of the BVH
  int single_job;                   // 1: every path belongs to job 0 / ticket 0
  int track_done;                   // 1: count finished paths per ticket
  int wide_stack;                   // 1: some BVH has more than 65535 nodes
  int refill;                       // extend: 0 = one ray per lane per pass over the
                                    // queue; k > 0 = persistent lanes, a wave takes new
                                    // rays once k or more of its lanes are idle
  uint32_t lds_bvh_bytes;           // > 0: single job whose nodes + primitives (this many
                                    // bytes) each extend workgroup copies into LDS
  uint32_t stack_rows;              // traversal stack entries per lane with lds_bvh_bytes
  uint32_t lds_obj_bytes;           // > 0: single job whose objects (this many bytes)
                                    // each shade workgroup copies into LDS
  int bvh4;                         // single job: traverse DevJob::nodes4
  int leaf1;                        // single 4-wide job whose leaves hold one primitive each
  uint32_t chunk_log2;               // extend lanes: 0 = each wave takes one contiguous
                                    // range of the queue; c > 0 = chunks of 2^c indices
                                    // dealt round-robin to the waves
  int vk;                           // 1: the Vulkan renderer's semantics (single job,
                                    // block 256, one sample per pass, no counting)
  // Camera-ray hits reused (single job): a pixel's camera ray is the same for
  // every sample (cpu.rs:234-251 takes no draw), so its closest hit is too.
  HitWord *first_hit;               // per pixel
  int first_hit_mode;               // 0 off; 1 shade records each bounce-0 path's hit;
                                    // 2 generate gives fresh paths their pixel's hit
                                    // and extend traces only the survivors
  // Buffer sizes, read only by bounds-checked builds (RDH_CHECKED):
  uint32_t slots;                   // entries of each queue array and hit array
  uint32_t ticket_cap;              // ticket table entries
  uint32_t job_cap;                 // job table entries
  unsigned long long l_cap;         // radiance arena entries
};

// Dynamic LDS of the extend launch for these arguments (0: static only).
uint32_t extend_dynamic_lds(const IterArgs &a, int block);
// Dynamic LDS of the shade launch (the job's objects, or 0).
uint32_t shade_dynamic_lds(const IterArgs &a);

struct LaunchShape {
  int block = 256;   // 64, 128 or 256
  int grid = 0;      // workgroups per launch
};

hipError_t launch_generate(const IterArgs &a, int parity, const LaunchShape &s, bool count,
                           hipStream_t st);
hipError_t launch_extend(const IterArgs &a, int parity, const LaunchShape &s, bool count,
                         hipStream_t st);
hipError_t launch_shade(const IterArgs &a, int parity, const LaunchShape &s, bool count,
                        hipStream_t st);
hipError_t launch_accumulate(const DevJob *job, const DevTicket &t, const float4 *L,
                             DevCounters *counters, bool count, int grid, hipStream_t st);
// vk: the Vulkan renderer's rounding UNORM store instead of print_frame_buffer
hipError_t launch_tonemap(const DevJob *job, uint32_t n_pixels, float spp, uchar4 *out,
                          DevCounters *counters, bool count, int grid, hipStream_t st,
                          bool vk = false);
// Workgroups of kernel `which` (0 generate, 1 extend, 2 shade) that fit on one CU.
int occupancy_blocks(int which, int block, bool lanes = false, uint32_t lds_bvh_bytes = 0,
                     uint32_t dynamic_lds = 0, bool w4 = false, bool single = true);

hipError_t launch_set_ctl(QueueCtl *ctl, unsigned long long total, uint32_t n_tickets,
                          hipStream_t st);
hipError_t launch_extend_work(QueueCtl *ctl, unsigned long long total, uint32_t n_tickets,
                              uint32_t ticket_lo, hipStream_t st);
hipError_t launch_zero_acc(float *acc, uint64_t n_floats, hipStream_t st);

// Parity-test entry kernels (single rays / single hits).
hipError_t launch_debug_trace(const DevNode *nodes, const DevPrim *prims, uint32_t n_prims,
                              uint32_t n_nodes, float r0, const DevObject *objects,
                              const float *rays, uint32_t n, int brute, int32_t *ids, float *ts,
                              float *normals, uint8_t *front, hipStream_t st);
// closest_hit normal and front flag (the shading kernels' prologue) of n
// (ray, t, id) triples; id == kNoHit gives zeros.
hipError_t launch_debug_hits(const DevObject *objects, const float *rays, const uint32_t *ids,
                             const float *ts, uint32_t n, float *normals, uint8_t *front,
                             hipStream_t st);
hipError_t launch_debug_scatter(const DevObject *objects, const float *in, const uint32_t *bits,
                                uint32_t n, float *out, hipStream_t st);
hipError_t launch_debug_camera(const DevJob *job, const uint32_t *pixels, uint32_t n, float *out,
                               hipStream_t st);
hipError_t launch_debug_world(const DevJob *job, const float *dirs, uint32_t n, float *out,
                              hipStream_t st);
// dst[0, bytes) = src[0, bytes); bytes a multiple of 16, at most 2^35.
hipError_t launch_copy_calibration(const void *src, void *dst, uint64_t bytes, hipStream_t st);
hipError_t launch_debug_philox(const uint32_t *in, uint32_t n, uint32_t *out, hipStream_t st);

}  // namespace rdh
