/*
 * raydar_hip_batch.h — batched rendering and the render server.
 *
 * Neither has a counterpart in the reference: raydar renders one scene per
 * Renderer (src/renderer/mod.rs:25-35).  Both sit next to the per-scene C ABI
 * (raydar_hip.h) and share its kernels: every job's passes become tickets in
 * one work list, and a single wavefront of persistent generate / extend /
 * shade launches (replayed from a captured hipGraph) renders paths of all jobs
 * mixed in the same queues.  A job's image depends only on its scene, seed,
 * sample count and bounce limit — never on what else is in flight — because
 * draws are keyed by (seed, pixel, sample, bounce, slot) and each pixel's
 * samples are accumulated in sample order, pass by pass.
 */
#ifndef RAYDAR_HIP_BATCH_H
#define RAYDAR_HIP_BATCH_H

#include "raydar_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One render job.  The scene is copied (and its BVH built) during the submit
 * call.  Output buffers, when given, must stay valid until the job is
 * reported done; they receive the final image (W*H*4 bytes, as
 * print_frame_buffer, cpu.rs:221-230) and the FP32 accumulator (W*H*4 floats).
 * progress_rgba, when given, is refreshed after every pass (the editor's
 * render_sample loop, raydar_editor.rs:95-115). */
typedef struct rd_job_desc {
  const rd_scene *scene;
  uint32_t samples;       /* samples per pixel, >= 1                        */
  uint32_t max_bounces;   /* >= 1                                           */
  uint64_t seed;
  uint32_t pass_samples;  /* samples per pass; 0 = the whole job in one pass */
  uint8_t *rgba;
  float *accumulator;
  uint8_t *progress_rgba;
} rd_job_desc;

/* ---- batched renderer --------------------------------------------------- */
typedef struct rd_batch rd_batch;

typedef struct rd_batch_stats {
  int64_t elapsed_ns;     /* submit of the first job to the last job's image  */
  uint64_t paths;         /* camera samples rendered                           */
  uint64_t iterations;    /* generate/extend/shade rounds                      */
  uint64_t graph_launches;
  uint64_t pumps;         /* host synchronisations                             */
  int64_t build_ns;       /* host BVH build time, summed over the jobs         */
  uint32_t host_threads;  /* threads that built the jobs' scenes               */
} rd_batch_stats;

rd_status rd_batch_create(rd_batch **out);
void rd_batch_destroy(rd_batch *b);
/* Keys: "block" (64/128/256), "queue" (paths in flight), "leaf" (1..8, default 1),
 * "grid" (workgroups, 0 = auto), "refill" (0..64, as rd_set_option),
 * "lane_chunk" (0 or 64..65536, as rd_set_option, default 64),
 * "bvh4" (1 = 4-wide trees while every resident job's walk fits the
 * kernel's stack, the default; 0 = binary trees),
 * "graph" (1 = replay a captured hipGraph),
 * "l_capacity" (radiance slots, 2^16..2^30), "iters_per_pump" (even, >= 2),
 * "ticket_cap" (passes in the work list, default 2^20, fixed once rendering
 * started: the list starts over whenever the engine has drained, and at the
 * cap new passes wait for that), read-only "rewinds" (first pipe's count),
 * "pipes" (1..4, default 3; settable until the first render): independent
 * engines, each with its own stream, queues and radiance ring; every job is
 * given whole to the one with the fewest pending paths and the host pumps
 * them in turn, so their kernels overlap on the GPU,
 * "solo_paths" (default 64 Mi = 67108864; 0 = off): a job of at least this
 * many paths (W * H * samples) without a progress image renders on the solo
 * lane, a per-scene engine (the rd_render_frame code path: single-job
 * kernels, small trees in LDS, samples split over pipes) driven by a host
 * thread of its own, one job after another, while the pipes take the other
 * jobs; read-only "solo_jobs" (jobs the lane has rendered, cumulative).
 * Images do not depend on these options. */
rd_status rd_batch_set_option(rd_batch *b, const char *key, int64_t value);
rd_status rd_batch_get_option(const rd_batch *b, const char *key, int64_t *value);
/* Render all jobs to completion in one mixed wavefront (and the solo lane). */
rd_status rd_batch_render(rd_batch *b, const rd_job_desc *jobs, uint32_t n_jobs);
rd_status rd_batch_stats_get(const rd_batch *b, rd_batch_stats *out);
const char *rd_batch_last_error(const rd_batch *b);

/* Autotuner: times the sample jobs over block {64,128,256} x queue capacity
 * {2^20, 2^22, 2^23, 2^24} x BVH leaf size {1,2,4,8}, keeps the fastest
 * setting on the batch and returns it; default_ms is the time at the
 * library's defaults (block 256, leaf 1, queue 2^24).
 *
 * Results are kept per device: the key is the GPU (name, architecture, CU
 * count, memory size, PCI location), the batch's pipe count and solo
 * threshold, and the shape of the sample mix (per job: object count,
 * resolution, samples, bounce limit).  Tuning the same mix again on the same
 * GPU returns the kept result at once (cached = 1) and applies it.  The
 * cache lives in the process; with RAYDAR_HIP_TUNE_CACHE=<file> set it is
 * also read from and appended to that file (one line per result), so later
 * processes on the same machine start tuned. */
typedef struct rd_tune_result {
  int32_t block;
  int32_t leaf;
  uint32_t queue;
  double best_ms;
  double default_ms;
  uint32_t configs_tried;
  uint32_t cached;        /* 1: taken from the per-device cache, nothing timed */
} rd_tune_result;
rd_status rd_batch_autotune(rd_batch *b, const rd_job_desc *jobs, uint32_t n_jobs,
                            rd_tune_result *out);

/* ---- render server ------------------------------------------------------ */
typedef struct rd_server rd_server;

typedef struct rd_server_config {
  uint64_t hbm_budget_bytes;  /* device memory for admitted jobs + engine     */
  uint32_t pass_samples;      /* default samples per pass (8)                 */
  int32_t block;              /* 0 = default                                  */
  uint32_t queue;             /* 0 = default                                  */
  int32_t leaf;               /* 0 = default                                  */
} rd_server_config;

typedef struct rd_job_result {
  int32_t status;             /* rd_status of the job                         */
  int32_t done;               /* 1 when the final image has been written      */
  uint32_t samples_done;
  uint32_t passes_done;
  int64_t queued_ns;          /* submit -> admission (HBM budget wait)        */
  int64_t latency_ns;         /* submit -> final image in the caller's buffer */
  uint64_t device_bytes;      /* HBM held while resident                      */
} rd_job_result;

typedef struct rd_server_stats {
  uint64_t jobs_submitted, jobs_done, jobs_rejected;
  uint64_t paths;             /* camera samples rendered                      */
  uint64_t iterations;
  int64_t busy_ns;            /* time the worker had work                     */
  uint64_t hbm_in_use, hbm_peak;  /* admitted jobs' device memory             */
  uint64_t engine_bytes;          /* queues, radiance ring, tables (fixed)    */
  uint32_t jobs_resident, jobs_queued, max_resident;
} rd_server_stats;

/* The server runs two pipes (see rd_batch_set_option); hbm_budget_bytes
 * covers both engines' buffers and every admitted job. */
rd_status rd_server_create(const rd_server_config *cfg, rd_server **out);
/* Stops accepting work, finishes every submitted job, frees the device. */
void rd_server_destroy(rd_server *s);
/* Thread-safe.  RD_ERR_BUDGET when the job alone can never fit the budget. */
rd_status rd_server_submit(rd_server *s, const rd_job_desc *job, uint64_t *job_id);
/* Non-blocking status of a job. */
rd_status rd_server_poll(rd_server *s, uint64_t job_id, rd_job_result *out);
/* Blocks until the job is done or timeout_ms passes (-1 = no timeout);
 * returns RD_OK with out->done = 0 on timeout. */
rd_status rd_server_wait(rd_server *s, uint64_t job_id, int64_t timeout_ms, rd_job_result *out);
rd_status rd_server_stats_get(rd_server *s, rd_server_stats *out);
const char *rd_server_last_error(const rd_server *s);

#ifdef __cplusplus
}
#endif

#endif /* RAYDAR_HIP_BATCH_H */
