/*
 * raydar_hip_debug.h — per-stage parity hooks of the HIP backend.
 *
 * These entry points have no counterpart in the reference: raydar's renderer
 * exposes whole frames only (src/renderer/mod.rs:25-35).  They run single
 * stages of the GPU hot path on caller-supplied inputs so the test-suite can
 * compare each stage with the CPU oracle on identical numbers:
 *   camera rays      cpu.rs:234-251
 *   closest hit      cpu.rs:344-394 (BVH or linear scan)
 *   scatter          cpu.rs:262-333 with explicit draw words
 *   world sample     world.rs:17-34
 *   Philox-4x32-10   the counter RNG that replaces thread_rng
 * Device hooks need a GPU; the rd_debug_host_* hooks run the BVH builder and
 * its traversal rule on the host (no GPU) for the builder's CPU tests.
 */
#ifndef RAYDAR_HIP_DEBUG_H
#define RAYDAR_HIP_DEBUG_H

#include "raydar_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Closest hit of n rays (n x 6 floats: origin, direction) on the GPU.
 * ids: object index or -1; ts: hit distance (0 on miss); normals: n x 3
 * closest_hit normal after the front-face flip; front: front-face flag.
 * brute != 0 scans every object instead of traversing the BVH. */
rd_status rd_debug_trace(const rd_scene *scene, int leaf_size, int brute, const float *rays,
                         uint32_t n, int32_t *ids, float *ts, float *normals, uint8_t *front);

/* Closest hit of n rays through the closest-hit kernel the wavefront loop
 * launches (k_extend_lanes / k_extend, cpu.rs:344-394), not a separate
 * debug walk: the rays are placed in a path queue and one extend launch of
 * the chosen form runs over it.  Outputs as rd_debug_trace; normals and the
 * front flag come from the shading kernels' closest_hit on the device.
 *   bvh4    1 = the 4-wide tree, 0 = the binary tree
 *   lds     1 = the tree copied into LDS (single job, small scenes)
 *   wide    1 = 32-bit traversal stack, 0 = 16-bit
 *   jobs    1 = single-job launch; k in 2..64 = multi-job launch over a job
 *           table of k entries, each with its own device copy of the scene;
 *           ray i belongs to job i % k (through its ticket)
 *   refill  0 = one ray per lane (k_extend); 1..64 = persistent lanes that
 *           take new rays once that many lanes of the wave are idle
 *   block   64, 128 or 256; grid = workgroups (0 = the occupancy grid)
 * RD_ERR_UNSUPPORTED: the renderer never launches this form for this scene
 * (a 4-wide stack bound above the kernel's, a tree too large for LDS, ...). */
typedef struct rd_extend_variant {
  int32_t bvh4, lds, wide, jobs, refill, block, grid;
  int32_t lane_chunk; /* 0, or the queue chunk (64..65536) the waves are dealt (render option "lane_chunk") */
} rd_extend_variant;
rd_status rd_debug_extend(const rd_scene *scene, int leaf_size, const rd_extend_variant *variant,
                          const float *rays, uint32_t n, int32_t *ids, float *ts, float *normals,
                          uint8_t *front);

/* Scatter of n hits on the GPU.  in: n x 16 floats = ray origin(3),
 * direction(3), t, object id (uint32 bits), attenuation(3), light(3), 2 unused;
 * bits: n x 8 draw words, one bounce's two Philox blocks in slot order
 * (0 transmission test, 1 Fresnel or metallic test, 2 roughness test,
 * 4-6 the unit-sphere offset of the direction taken).  out: n x 16 floats
 * = new origin(3), new direction(3), attenuation(3), light(3), transmission
 * flag, normal(3). */
rd_status rd_debug_scatter(const rd_scene *scene, const float *in, const uint32_t *bits,
                           uint32_t n, float *out);

/* Camera rays (n x 6) for the given pixel indices (y * W + x). */
rd_status rd_debug_camera_rays(const rd_scene *scene, const uint32_t *pixels, uint32_t n,
                               float *out);

/* World colour (n x 3) for n directions. */
rd_status rd_debug_world(const rd_scene *scene, const float *dirs, uint32_t n, float *out);

/* Philox-4x32-10 on the GPU: in n x 6 words (counter[4], key[2]), out n x 4. */
rd_status rd_debug_philox(const uint32_t *in, uint32_t n, uint32_t *out);

/* A plain device copy of `bytes` (multiple of 16, <= 32 GiB) between two
 * fresh buffers, `reps` launches of k_copy_calibration, each timed with
 * hipEvents into ms[reps] (may be NULL).  It reads and writes exactly
 * `bytes` of HBM per launch: the reference point that converts hardware
 * traffic counters (FETCH_SIZE / WRITE_SIZE) into bytes. */
rd_status rd_debug_copy_calibration(uint64_t bytes, uint32_t reps, float *ms);

/* Host: build the BVH and report node count, depth and whether its
 * structural invariants hold (message in err when they do not). */
rd_status rd_debug_host_bvh_info(const rd_scene *scene, int leaf_size, uint32_t *n_nodes,
                                 uint32_t *depth, char *err, size_t err_capacity);

/* Host: closest hit through the BVH with the kernels' culling rule. */
rd_status rd_debug_host_trace(const rd_scene *scene, int leaf_size, const float *rays,
                              uint32_t n, int32_t *ids, float *ts);

/* Host: the streaming engine's radiance-ring allocator.  ops[i] >= 0 reserves
 * that many slots for the next ticket (tickets are numbered 0, 1, ... in
 * reservation order, refused ones included) and writes the offset, or -1 if
 * it does not fit now; ops[i] < 0 releases ticket -ops[i]-1 and writes 0. */
rd_status rd_debug_ring_ops(uint64_t capacity, const int64_t *ops, uint32_t n, int64_t *out);

/* Host: the size class the reused-buffer cache allocates for a request of
 * `bytes` (buffers are reused by class): out[i] for bytes[i]. */
rd_status rd_debug_cache_classes(const uint64_t *bytes, uint32_t n, uint64_t *out);

/* Host: the same tree collapsed to 4-wide nodes.  bvh4_info reports its node
 * count and the stack depth its traversal needs (the kernel keeps
 * stack4 + 3 rows) and checks that every primitive sits in exactly one leaf;
 * trace4 walks it step for step as the extend kernel does and returns id -2
 * for a ray whose stack would outgrow that bound. */
rd_status rd_debug_host_bvh4_info(const rd_scene *scene, int leaf_size, uint32_t *n_nodes4,
                                  uint32_t *stack4, char *err, size_t err_capacity);
rd_status rd_debug_host_trace4(const rd_scene *scene, int leaf_size, const float *rays,
                               uint32_t n, int32_t *ids, float *ts);

/* Host builds of the kernels' per-ray functions (the same source compiled
 * for the CPU), with the layouts of rd_debug_scatter / rd_debug_camera_rays,
 * so the CPU test-suite can check their logic against the oracle. */
rd_status rd_debug_host_scatter(const rd_scene *scene, const float *in, const uint32_t *bits,
                                uint32_t n, float *out);
rd_status rd_debug_host_camera_rays(const rd_scene *scene, const uint32_t *pixels, uint32_t n,
                                    float *out);

#ifdef __cplusplus
}
#endif

#endif /* RAYDAR_HIP_DEBUG_H */
