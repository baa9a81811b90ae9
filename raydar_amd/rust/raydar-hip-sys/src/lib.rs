//! Raw bindings to libraydar_hip: the C ABI declared in include/raydar_hip.h
//! (the per-scene renderer that implements raydar's `trait Renderer`,
//! src/renderer/mod.rs:25-35) and include/raydar_hip_batch.h (batching and
//! the render server).  Layouts are `#[repr(C)]` copies of the header's
//! structs; every function is `unsafe` to call.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int};

/// `enum rd_status` (an `int` in the C ABI).
pub type rd_status = c_int;
pub const RD_OK: rd_status = 0;
pub const RD_ERR_INVALID_ARGUMENT: rd_status = 1;
pub const RD_ERR_UNSUPPORTED: rd_status = 2;
pub const RD_ERR_OUT_OF_MEMORY: rd_status = 3;
pub const RD_ERR_DEVICE: rd_status = 4;
pub const RD_ERR_BUFFER_TOO_SMALL: rd_status = 5;
pub const RD_ERR_PARSE: rd_status = 6;
pub const RD_ERR_IO: rd_status = 7;
pub const RD_ERR_BUDGET: rd_status = 8;
pub const RD_ERR_NOT_FOUND: rd_status = 9;

pub const RD_SPHERE: u32 = 0;
pub const RD_CUBE: u32 = 1;
pub const RD_WORLD_SKY_COLOR: u32 = 0;
pub const RD_WORLD_SOLID_COLOR: u32 = 1;
pub const RD_WORLD_TRANSPARENT: u32 = 2;
pub const RD_PROJECTION_PERSPECTIVE: u32 = 0;
pub const RD_PROJECTION_ORTHOGRAPHIC: u32 = 1;

/// `Object` + `Material` (src/scene/objects.rs:13-17, material.rs:4-13).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rd_object {
    pub kind: u32,
    pub center: [f32; 3],
    pub size: f32,
    pub albedo: [f32; 3],
    pub roughness: f32,
    pub metallic: f32,
    pub emission_color: [f32; 3],
    pub emission_strength: f32,
    pub transmission: f32,
    pub ior: f32,
}

/// Every serialised field of `Camera` (src/scene/camera.rs:13-30);
/// matrices column-major.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rd_camera {
    pub position: [f32; 3],
    pub target: [f32; 3],
    pub up: [f32; 3],
    pub resolution_x: u32,
    pub resolution_y: u32,
    pub projection: u32,
    pub projection_param: f32,
    pub near_clip: f32,
    pub far_clip: f32,
    pub view_matrix: [f32; 16],
    pub proj_matrix: [f32; 16],
    pub inverse_view_matrix: [f32; 16],
    pub inverse_proj_matrix: [f32; 16],
}

/// `World` (src/scene/world.rs:6-14).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rd_world {
    pub kind: u32,
    pub top_color: [f32; 3],
    pub bottom_color: [f32; 3],
    pub solid_color: [f32; 3],
}

/// `Scene` (src/scene/mod.rs:13-18); `objects` is read only during a call.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rd_scene {
    pub camera: rd_camera,
    pub world: rd_world,
    pub object_count: u32,
    pub objects: *const rd_object,
}

/// `RendererConfig` (src/renderer/mod.rs:11-23).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rd_config {
    pub max_sample_count: u32,
    pub max_bounces: u32,
}

/// `Profiler` durations in ns, -1 for `None` (src/renderer/timing.rs:9-19).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rd_timings {
    pub frame_ns: i64,
    pub sample_ns: i64,
    pub prepare_ns: i64,
    pub render_ns: i64,
}

/// Opaque renderer handle.
#[repr(C)]
pub struct rd_renderer {
    _private: [u8; 0],
}

/// Opaque batch renderer / render server handles (raydar_hip_batch.h).
#[repr(C)]
pub struct rd_batch {
    _private: [u8; 0],
}
#[repr(C)]
pub struct rd_server {
    _private: [u8; 0],
}

/// One render job of a batch or of the server (raydar_hip_batch.h).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rd_job_desc {
    pub scene: *const rd_scene,
    pub samples: u32,
    pub max_bounces: u32,
    pub seed: u64,
    pub pass_samples: u32,
    pub rgba: *mut u8,
    pub accumulator: *mut f32,
    pub progress_rgba: *mut u8,
}

extern "C" {
    pub fn rd_config_default(out: *mut rd_config) -> rd_status;
    pub fn rd_create(config: *const rd_config, out: *mut *mut rd_renderer) -> rd_status;
    pub fn rd_destroy(r: *mut rd_renderer);
    pub fn rd_new_frame(r: *mut rd_renderer, scene: *const rd_scene) -> rd_status;
    pub fn rd_render_sample(
        r: *mut rd_renderer,
        scene: *const rd_scene,
        rgba: *mut u8,
        rgba_capacity: usize,
        produced: *mut c_int,
    ) -> rd_status;
    pub fn rd_render_frame(
        r: *mut rd_renderer,
        scene: *const rd_scene,
        rgba: *mut u8,
        rgba_capacity: usize,
    ) -> rd_status;
    pub fn rd_profiler(r: *const rd_renderer, out: *mut rd_timings) -> rd_status;
    pub fn rd_sample_count(r: *const rd_renderer) -> u32;
    pub fn rd_max_sample_count(r: *const rd_renderer) -> u32;
    pub fn rd_max_bounces(r: *const rd_renderer) -> u32;
    pub fn rd_set_max_sample_count(r: *mut rd_renderer, count: u32);
    pub fn rd_set_max_bounces(r: *mut rd_renderer, bounces: u32);
    pub fn rd_last_error(r: *const rd_renderer) -> *const c_char;

    pub fn rd_set_samples_per_pass(r: *mut rd_renderer, samples: u32) -> rd_status;
    pub fn rd_set_seed(r: *mut rd_renderer, seed: u64);
    pub fn rd_restart_frame(r: *mut rd_renderer) -> rd_status;
    pub fn rd_read_accumulator(r: *mut rd_renderer, out: *mut f32, float_capacity: usize) -> rd_status;
    pub fn rd_set_option(r: *mut rd_renderer, key: *const c_char, value: i64) -> rd_status;
    pub fn rd_get_option(r: *const rd_renderer, key: *const c_char, value: *mut i64) -> rd_status;

    pub fn rd_camera_update_matrices(camera: *mut rd_camera) -> rd_status;
    pub fn rd_camera_vulkan_inverse_view(camera: *const rd_camera, out_inverse_view: *mut f32) -> rd_status;

    pub fn rd_scene_load_rscn(path: *const c_char, out: *mut *mut rd_scene) -> rd_status;
    pub fn rd_scene_parse_rscn(json: *const c_char, len: usize, out: *mut *mut rd_scene) -> rd_status;
    pub fn rd_scene_save_rscn(scene: *const rd_scene, path: *const c_char) -> rd_status;
    pub fn rd_scene_free(scene: *mut rd_scene);

    pub fn rd_batch_create(out: *mut *mut rd_batch) -> rd_status;
    pub fn rd_batch_destroy(b: *mut rd_batch);
    pub fn rd_batch_set_option(b: *mut rd_batch, key: *const c_char, value: i64) -> rd_status;
    pub fn rd_batch_render(b: *mut rd_batch, jobs: *const rd_job_desc, n_jobs: u32) -> rd_status;
    pub fn rd_batch_last_error(b: *const rd_batch) -> *const c_char;
}

#[cfg(test)]
mod tests {
    use super::*;
    use std::mem::size_of;

    #[test]
    fn layouts_match_the_header() {
        assert_eq!(size_of::<rd_object>(), 64);
        assert_eq!(size_of::<rd_world>(), 40);
        assert_eq!(size_of::<rd_config>(), 8);
        assert_eq!(size_of::<rd_timings>(), 32);
        assert_eq!(size_of::<rd_camera>(), 4 * (9 + 6 + 64));
    }
}
