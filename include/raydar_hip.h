/*
 * raydar_hip.h — C ABI of the MI355X (gfx950) path-tracing backend for raydar.
 *
 * This is the drop-in boundary.  Each entry point below replaces one method of
 * raydar's `trait Renderer` (src/renderer/mod.rs:25-35) or one step of its
 * bootstrap (src/cli/mod.rs:31-57), so a thin FFI crate can implement the
 * trait on top of it exactly as src/renderer/cpu.rs does natively.
 *
 * Conventions
 *   - Every call is synchronous with respect to the caller's buffers: when a
 *     call returns, output buffers are filled and input scenes may be freed.
 *   - The library owns all device memory; callers own every host buffer.
 *   - `rd_scene` is a flattened, caller-owned copy of raydar's serde `Scene`
 *     (src/scene/mod.rs:13-18); it is read only during the call.
 *   - Errors: every fallible call returns rd_status; the message of the last
 *     failure on a renderer is available from rd_last_error().  The trait has
 *     no error channel, so the Rust shim panics on a non-OK status inside trait
 *     methods, as the Vulkan backend does (src/renderer/vulkan.rs:116-522).
 *   - One caller per renderer handle at a time (the trait takes &mut self).
 */
#ifndef RAYDAR_HIP_H
#define RAYDAR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rd_status {
  RD_OK = 0,
  RD_ERR_INVALID_ARGUMENT = 1, /* null pointer, zero-size image, bad enum     */
  RD_ERR_UNSUPPORTED = 2,      /* World::Transparent (src/scene/world.rs:31)  */
  RD_ERR_OUT_OF_MEMORY = 3,    /* device or host allocation failed            */
  RD_ERR_DEVICE = 4,           /* a HIP runtime call failed                   */
  RD_ERR_BUFFER_TOO_SMALL = 5, /* caller buffer smaller than W*H*4 bytes      */
  RD_ERR_PARSE = 6,            /* malformed .rscn                             */
  RD_ERR_IO = 7,               /* file could not be opened / written          */
  RD_ERR_BUDGET = 8,           /* render server: job exceeds the HBM budget   */
  RD_ERR_NOT_FOUND = 9         /* render server: unknown job id               */
} rd_status;

/* Geometry tag of src/scene/objects.rs:6-10. */
typedef enum rd_geometry_kind { RD_SPHERE = 0, RD_CUBE = 1 } rd_geometry_kind;

/* World variants of src/scene/world.rs:6-14. */
typedef enum rd_world_kind {
  RD_WORLD_SKY_COLOR = 0,
  RD_WORLD_SOLID_COLOR = 1,
  RD_WORLD_TRANSPARENT = 2
} rd_world_kind;

/* Projection variants of src/scene/camera.rs:7-11. */
typedef enum rd_projection_kind {
  RD_PROJECTION_PERSPECTIVE = 0,
  RD_PROJECTION_ORTHOGRAPHIC = 1
} rd_projection_kind;

/* One `Object` (src/scene/objects.rs:13-17) with its `Material`
 * (src/scene/material.rs:4-13).  `size` is Sphere::radius or
 * Cube::side_length (objects.rs:41-50).  64 bytes. */
typedef struct rd_object {
  uint32_t kind; /* rd_geometry_kind */
  float center[3];
  float size;
  float albedo[3];
  float roughness;
  float metallic;
  float emission_color[3];
  float emission_strength;
  float transmission;
  float ior;
} rd_object;

/* All serialised fields of `Camera` (src/scene/camera.rs:13-30).  Matrices are
 * column-major exactly as serialised ({x,y,z,w} columns).  The renderer reads
 * only position, resolution and the two inverse matrices, as cpu.rs:234-251
 * does; the rest is carried so a scene round-trips through the ABI. */
typedef struct rd_camera {
  float position[3];
  float target[3];
  float up[3];
  uint32_t resolution_x;
  uint32_t resolution_y;
  uint32_t projection; /* rd_projection_kind */
  float projection_param; /* fov in degrees, or orthographic size */
  float near_clip;
  float far_clip;
  float view_matrix[16];
  float proj_matrix[16];
  float inverse_view_matrix[16];
  float inverse_proj_matrix[16];
} rd_camera;

/* `World` (src/scene/world.rs:6-14). */
typedef struct rd_world {
  uint32_t kind; /* rd_world_kind */
  float top_color[3];
  float bottom_color[3];
  float solid_color[3];
} rd_world;

/* `Scene` (src/scene/mod.rs:13-18). */
typedef struct rd_scene {
  rd_camera camera;
  rd_world world;
  uint32_t object_count;
  const rd_object *objects;
} rd_scene;

/* `RendererConfig` (src/renderer/mod.rs:11-23). */
typedef struct rd_config {
  uint32_t max_sample_count;
  uint32_t max_bounces;
} rd_config;

/* `Profiler` (src/renderer/timing.rs:9-19): durations in nanoseconds, -1 when
 * the timer has no duration (Option::None, timing.rs:124-126). */
typedef struct rd_timings {
  int64_t frame_ns;
  int64_t sample_ns;
  int64_t prepare_ns;
  int64_t render_ns;
} rd_timings;

typedef struct rd_renderer rd_renderer;

/* RendererConfig::default(): 1024 samples, 12 bounces (mod.rs:16-23). */
rd_status rd_config_default(rd_config *out);

/* CpuRenderer::new (cpu.rs:186-191) / VulkanRenderer::new (vulkan.rs:458):
 * selects HIP device 0, creates the library's stream.  Fails cleanly (no
 * panic) so CommonArgs::initialize (cli/mod.rs:31) can return Err. */
rd_status rd_create(const rd_config *config, rd_renderer **out);
void rd_destroy(rd_renderer *r);

/* Renderer::new_frame (mod.rs:28; cpu.rs:135-140): snapshot the scene into
 * device memory (BVH build + upload), zero the accumulator, sample_count = 0,
 * start the frame and prepare timers. */
rd_status rd_new_frame(rd_renderer *r, const rd_scene *scene);

/* Renderer::render_sample (mod.rs:27; cpu.rs:142-158): if sample_count has
 * reached max_sample_count sets *produced = 0 (None) and returns RD_OK.
 * Otherwise renders one pass of `samples_per_pass` samples (1 by default, as
 * the CPU renderer; see rd_set_samples_per_pass), tonemaps the running
 * average into `rgba` (W*H*4 bytes, row 0 = top) and sets *produced = 1.
 * `scene` may be NULL after rd_new_frame; if it is given and its resolution
 * differs from the snapshot, the scene is re-snapshotted and the accumulator
 * reallocated (cpu.rs:400-417). */
rd_status rd_render_sample(rd_renderer *r, const rd_scene *scene,
                           uint8_t *rgba, size_t rgba_capacity, int *produced);

/* Renderer::render_frame (mod.rs:26; cpu.rs:119-133): new_frame, then all
 * max_sample_count samples, then the tonemap into `rgba`. */
rd_status rd_render_frame(rd_renderer *r, const rd_scene *scene,
                          uint8_t *rgba, size_t rgba_capacity);

/* rd_render_frame without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same semantics, images, sample count and profiler
 * timers as rd_render_frame, except that the renderer's resident frame
 * (rd_read_accumulator, later rd_render_sample passes) is not touched.
 * Every device buffer of the call (BVH, primitives, accumulator, image,
 * path queues, radiance slots, control block) is carved from one static
 * device array of 192 MiB per GPU, set up by rd_create and held until the
 * process ends; the call itself makes no hipMalloc, hipFree or other
 * allocating call and uses one stream, with passes of as many samples as
 * the leftover radiance slots hold.  Limit: 36 bytes per pixel (accumulator,
 * image, one sample of radiance slots) plus the BVH (about 352 bytes per
 * object) must fit beside 60 MiB of path queues: 1920x1080 with up to about
 * 180,000 objects, or up to about 3.8 M pixels without objects; larger
 * scenes get RD_ERR_INVALID_ARGUMENT with the byte count.  The array keeps the last
 * call's contents and is shared by every renderer on the GPU: calls from one
 * process are serialised by a lock, so calls that would overlap on one GPU
 * wait for each other. */
rd_status rd_render_frame_noalloc(rd_renderer *r, const rd_scene *scene, uint8_t *rgba,
                                  size_t rgba_capacity);

/* Renderer::profiler (mod.rs:29). */
rd_status rd_profiler(const rd_renderer *r, rd_timings *out);

/* Renderer getters / setters (mod.rs:30-34). */
uint32_t rd_sample_count(const rd_renderer *r);
uint32_t rd_max_sample_count(const rd_renderer *r);
uint32_t rd_max_bounces(const rd_renderer *r);
void rd_set_max_sample_count(rd_renderer *r, uint32_t count);
void rd_set_max_bounces(rd_renderer *r, uint32_t bounces);

/* Message of the last failed call on `r` (or of the last failed rd_create /
 * scene call when r is NULL).  Never NULL; empty when there was no error. */
const char *rd_last_error(const rd_renderer *r);

/* ---- Extensions with no counterpart in the trait ------------------------ */

/* Samples added per render_sample call (default 1 = CpuRenderer semantics;
 * the render server uses 8).  0 is rejected. */
rd_status rd_set_samples_per_pass(rd_renderer *r, uint32_t samples);

/* Seed of the counter-based RNG.  An image depends only on (seed, pixel,
 * sample index, bounce, draw slot), never on launch shape or job mix. */
void rd_set_seed(rd_renderer *r, uint64_t seed);

/* new_frame for the scene already on the device: zero the accumulator and the
 * sample count and restart the frame timers, without rebuilding or
 * re-uploading the scene (the editor's re-render of an unchanged scene,
 * raydar_editor.rs:95-99, and the render server's resident scenes). */
rd_status rd_restart_frame(rd_renderer *r);

/* Read the running FP32 accumulator (W*H*4 floats: summed r,g,b and the
 * sample count in alpha, as the CpuRenderer frame_buffer, cpu.rs:113,204-209). */
rd_status rd_read_accumulator(rd_renderer *r, float *out, size_t float_capacity);

/* Tuning knobs, normally chosen by the autotuner.  Keys:
 *   "block"       workgroup size of the wavefront kernels: 64, 128 or 256
 *   "queue"       path-queue capacity (paths in flight per pipe), >= 1024,
 *                 default 2^24
 *   "leaf"        BVH leaf size 1..8, default 1 (takes effect at the next new_frame)
 *   "grid"        workgroups per persistent launch (0 = automatic)
 *   "traversal"   0 = BVH, 1 = brute-force linear scan (cpu.rs:345-349)
 *   "refill"      closest-hit lanes: 0 = each lane traces its rays to the end
 *                 in launch order; k in 1..64 = a wave hands new rays to its
 *                 idle lanes once k or more are idle (default 48)
 *   "lane_chunk"  closest-hit lanes: 0 = each wave takes one contiguous range of
 *                 the path queue; c = 64..65536 (a power of two) = the queue's
 *                 chunks of c paths are dealt round-robin to the waves
 *                 (default 64)
 *   "lds_bvh"     scenes whose BVH nodes + primitives take at most this many
 *                 bytes are traversed from LDS (default 14336; 0 = never)
 *   "lds_objects" scenes whose objects (64 B each) take at most this many
 *                 bytes are shaded from a copy in LDS (default 16384; 0 = never)
 *   "bvh4"        1 = traverse the tree collapsed to 4-wide nodes (default),
 *                 0 = the binary tree
 *   "pipes"       a pass of k >= 2 samples is divided by samples over up to
 *                 min(pipes, k) independent streams whose kernels run
 *                 concurrently, each taking at least min(queue/2, 2^22) paths
 *                 (1..4; 1 = one stream; 0 = automatic, the default: 2 for
 *                 scenes traversed from LDS, else 3); read-only
 *                 "pipes_used" reports the last pass's count
 *   "shade_blocks" shade workgroups per CU (0 = 7 waves per SIMD for one
 *                  scene, as many as fit for several)
 *   "hit_reuse"    1 (default): a pixel's camera ray, the same for every
 *                  sample, is traced once per pass (or frame) and its hit
 *                  reused by the other samples; 0: traced for every sample
 *                  (same images)
 *   "count"       1 = collect algorithmic event counters in the next passes
 *   "semantics"   0 = the CPU renderer's (default); 1 = the Vulkan renderer's
 *                 (src/renderer/vulkan.rs, shaders/vulkan/raytrace.rgen):
 *                 pixel-centre pinhole camera from look_at_rh(position,
 *                 target, -up), one PCG stream per pixel seeded x + y*1000
 *                 through all samples (the seed is not used), unnormalised
 *                 random_in_unit_sphere, hits only in [1e-3, 1e4], 1e-3
 *                 offsets, GLSL reflect/refract/mix, a rounding UNORM image,
 *                 and rd_render_sample rendering every remaining sample at
 *                 once (vulkan.rs:109-184).  Takes effect at the next
 *                 new_frame; needs block 256, refill > 0 and no counting;
 *                 not available in rd_render_frame_noalloc or the batch and
 *                 server handles. */
rd_status rd_set_option(rd_renderer *r, const char *key, int64_t value);
rd_status rd_get_option(const rd_renderer *r, const char *key, int64_t *value);

/* Algorithmic event counters collected while option "count" was 1. */
typedef struct rd_counters {
  uint64_t camera_rays;    /* paths started                               */
  uint64_t segments;       /* closest-hit queries (rays traced), including
                              misses                                       */
  uint64_t box_tests;      /* BVH child-box slab tests                     */
  uint64_t sphere_tests;   /* hit_sphere evaluations (cpu.rs:34-62)        */
  uint64_t cube_tests;     /* hit_cube evaluations (cpu.rs:64-98)          */
  uint64_t sphere_hits;    /* closest hits on spheres                      */
  uint64_t cube_hits;      /* closest hits on cubes                        */
  uint64_t glass_scatters; /* scatters that took the transmission branch   */
  uint64_t misses;         /* sky samples (world.rs:17-34)                 */
  uint64_t accumulates;    /* per-path accumulations (cpu.rs:204-209)      */
  uint64_t tonemapped_px;  /* pixels written by the tonemap                */
  /* SIMD efficiency of the closest-hit kernel (wave = 64 lanes):           */
  uint64_t wave_steps;     /* traversal iterations of waves with a ray     */
  uint64_t lane_steps;     /* lanes holding a ray, summed over those       */
  uint64_t sort_steps;     /* wave iterations that ordered two or more
                              entered inner children                       */
  uint64_t prim_trips;     /* wave iterations of the primitive loop        */
  uint64_t cube_trips;     /* those in which a lane tested a cube          */
  uint64_t refills;        /* times a wave handed new rays to idle lanes   */
  /* camera paths that took their pixel's closest hit, recorded by an earlier
     path of the same pass, instead of a query (a pixel's camera ray is the
     same for every sample); sphere_hits + cube_hits + misses equals
     segments + reused_hits                                                  */
  uint64_t reused_hits;
} rd_counters;
rd_status rd_get_counters(const rd_renderer *r, rd_counters *out);
rd_status rd_reset_counters(rd_renderer *r);

/* Per-kernel device time (hipEvent pairs on the library's stream) of the
 * passes rendered while option "count" was 0 and option "kernel_timing" was
 * 1.  Index: 0 generate, 1 extend, 2 shade, 3 accumulate, 4 tonemap. */
typedef struct rd_kernel_times {
  double total_ms[5];
  uint64_t launches[5];
} rd_kernel_times;
rd_status rd_get_kernel_times(const rd_renderer *r, rd_kernel_times *out);

/* Camera::update_matrices (src/scene/camera.rs:210-231): recomputes
 * view_matrix = look_at_lh(position, target, up), proj_matrix =
 * perspective(projection_param degrees, aspect, near, far) or
 * ortho(-size*aspect, size*aspect, -size, size, near, far), and both inverses,
 * in FP32 in cgmath 0.18's operation order (aspect = resolution_x /
 * resolution_y).  The editor's camera edits (orbit, pan, zoom, setters,
 * camera.rs:62-208) all end in this call.  RD_ERR_INVALID_ARGUMENT, with the
 * camera unchanged, for a zero height, an unknown projection, perspective
 * parameters cgmath asserts against (fov outside (0, 180) degrees, near <= 0,
 * far <= near) or a singular matrix: the reference panics on each. */
rd_status rd_camera_update_matrices(rd_camera *camera);

/* The camera the Vulkan renderer hands its ray-generation shader
 * (VulkanRenderer::new_frame, src/renderer/vulkan.rs:251-272): the inverse of
 * look_at_rh(position, target, -up), column-major, FP32 in cgmath 0.18's
 * order.  The "semantics" = 1 render option uses it with the scene's stored
 * inverse projection.  RD_ERR_INVALID_ARGUMENT for a singular view (the
 * reference unwraps the inverse). */
rd_status rd_camera_vulkan_inverse_view(const rd_camera *camera, float *out_inverse_view);

/* ---- .rscn I/O (serde layout of the src/scene modules) ----------------- */

/* Parse a .rscn file (cli/mod.rs:32-37).  Numbers are parsed as f64 and
 * rounded to f32, as serde_json does.  Free with rd_scene_free. */
rd_status rd_scene_load_rscn(const char *path, rd_scene **out);
rd_status rd_scene_parse_rscn(const char *json, size_t len, rd_scene **out);
/* Write a scene as pretty JSON in the serde layout (raydar_editor.rs:76-93). */
rd_status rd_scene_save_rscn(const rd_scene *scene, const char *path);
void rd_scene_free(rd_scene *scene);

#ifdef __cplusplus
}
#endif

#endif /* RAYDAR_HIP_H */
