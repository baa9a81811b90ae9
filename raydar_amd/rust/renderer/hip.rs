//! HIP backend for raydar: `trait Renderer` (src/renderer/mod.rs:25-35) over
//! the raydar-hip-sys C ABI, rendering on one MI355X (gfx950).
//!
//! Install as src/renderer/hip.rs, next to cpu.rs and vulkan.rs, with
//! `pub mod hip;` in src/renderer/mod.rs (INTEGRATION.md lists the other
//! edits: the --hip flag and its arm in src/cli/mod.rs, the renderer name in
//! src/main.rs, the dependency in Cargo.toml).
//!
//! Semantics follow CpuRenderer (src/renderer/cpu.rs), whose images the
//! library reproduces: render_sample adds one sample per call, render_frame
//! renders every sample, and the four profiler timers are driven as
//! cpu.rs:135-137 and 193-219 drive them.  render_frame runs all samples as
//! one pass, so its sample timer holds their average (Timer::end_multiple),
//! as the Vulkan renderer's does (vulkan.rs:175-177).  The trait has no error
//! channel: a failed call panics with the library's message, as the Vulkan
//! renderer's unwraps do; construction returns Err instead.
use std::ffi::CStr;
use std::ptr;

use image::RgbaImage;
use raydar_hip_sys as sys;

use super::{timing::Profiler, Renderer, RendererConfig};
use crate::scene::{camera::Projection, objects::Geometry, world::World, Scene};

pub struct HipRenderer {
    handle: *mut sys::rd_renderer,
    profiler: Profiler,
    /// Flattened objects of the scene of the current call (the C ABI reads
    /// them only during the call).
    objects: Vec<sys::rd_object>,
}

fn last_error(handle: *const sys::rd_renderer) -> String {
    // rd_last_error never returns NULL
    unsafe { CStr::from_ptr(sys::rd_last_error(handle)) }
        .to_string_lossy()
        .into_owned()
}

fn matrix(m: cgmath::Matrix4<f32>) -> [f32; 16] {
    // cgmath's Matrix4 fields x, y, z, w are its columns; rd_camera is column-major
    [
        m.x.x, m.x.y, m.x.z, m.x.w, m.y.x, m.y.y, m.y.z, m.y.w, m.z.x, m.z.y, m.z.z, m.z.w, m.w.x,
        m.w.y, m.w.z, m.w.w,
    ]
}

/// The scene as the C ABI's flattened, caller-owned copy.
fn flatten(scene: &Scene, objects: &mut Vec<sys::rd_object>) -> sys::rd_scene {
    objects.clear();
    for object in &scene.objects {
        let (kind, center, size) = match &object.geometry {
            Geometry::Sphere(s) => (sys::RD_SPHERE, s.center, s.radius),
            Geometry::Cube(c) => (sys::RD_CUBE, c.center, c.side_length),
        };
        let m = &object.material;
        objects.push(sys::rd_object {
            kind,
            center: [center.x, center.y, center.z],
            size,
            albedo: [m.albedo.x, m.albedo.y, m.albedo.z],
            roughness: m.roughness,
            metallic: m.metallic,
            emission_color: [m.emission_color.x, m.emission_color.y, m.emission_color.z],
            emission_strength: m.emission_strength,
            transmission: m.transmission,
            ior: m.ior,
        });
    }
    let cam = &scene.camera;
    let (projection, projection_param) = match cam.projection() {
        Projection::Perspective { fov } => (sys::RD_PROJECTION_PERSPECTIVE, fov.0),
        Projection::Orthographic { size } => (sys::RD_PROJECTION_ORTHOGRAPHIC, size),
    };
    let (p, t, up) = (cam.position(), cam.target(), cam.up());
    let camera = sys::rd_camera {
        position: [p.x, p.y, p.z],
        target: [t.x, t.y, t.z],
        up: [up.x, up.y, up.z],
        resolution_x: cam.resolution_x(),
        resolution_y: cam.resolution_y(),
        projection,
        projection_param,
        near_clip: cam.near_clip(),
        far_clip: cam.far_clip(),
        view_matrix: matrix(cam.view_matrix()),
        proj_matrix: matrix(cam.proj_matrix()),
        inverse_view_matrix: matrix(cam.inverse_view_matrix()),
        inverse_proj_matrix: matrix(cam.inverse_proj_matrix()),
    };
    let mut world = sys::rd_world::default();
    match &scene.world {
        World::SkyColor {
            top_color,
            bottom_color,
        } => {
            world.kind = sys::RD_WORLD_SKY_COLOR;
            world.top_color = [top_color.x, top_color.y, top_color.z];
            world.bottom_color = [bottom_color.x, bottom_color.y, bottom_color.z];
        }
        World::SolidColor(c) => {
            world.kind = sys::RD_WORLD_SOLID_COLOR;
            world.solid_color = [c.x, c.y, c.z];
        }
        // the library answers RD_ERR_UNSUPPORTED, where World::sample has todo!()
        World::Transparent => world.kind = sys::RD_WORLD_TRANSPARENT,
    }
    sys::rd_scene {
        camera,
        world,
        object_count: objects.len() as u32,
        objects: objects.as_ptr(),
    }
}

impl HipRenderer {
    /// Selects HIP device 0 and creates the library's stream (CpuRenderer::new,
    /// cpu.rs:186-191).  Err when no usable GPU is present.
    pub fn new(config: RendererConfig) -> Result<Self, String> {
        let cfg = sys::rd_config {
            max_sample_count: config.max_sample_count,
            max_bounces: config.max_bounces,
        };
        let mut handle: *mut sys::rd_renderer = ptr::null_mut();
        let status = unsafe { sys::rd_create(&cfg, &mut handle) };
        if status != sys::RD_OK {
            return Err(format!("HIP renderer: {}", last_error(ptr::null())));
        }
        Ok(Self {
            handle,
            profiler: Profiler::default(),
            objects: Vec::new(),
        })
    }

    /// A render option of the library (rd_set_option), e.g. ("semantics", 1)
    /// for the Vulkan backend's images; it applies from the next new_frame.
    pub fn set_option(&mut self, key: &str, value: i64) -> Result<(), String> {
        let k = std::ffi::CString::new(key).map_err(|e| e.to_string())?;
        let status = unsafe { sys::rd_set_option(self.handle, k.as_ptr(), value) };
        if status != sys::RD_OK {
            return Err(last_error(self.handle));
        }
        Ok(())
    }

    fn check(&self, status: sys::rd_status, call: &str) {
        if status != sys::RD_OK {
            panic!("HIP renderer: {} failed ({}): {}", call, status, last_error(self.handle));
        }
    }

    fn blank_image(scene: &Scene) -> RgbaImage {
        RgbaImage::new(scene.camera.resolution_x(), scene.camera.resolution_y())
    }

    /// One rd_render_sample call of `samples` samples into `image`.
    fn render_pass(&mut self, scene: &Scene, samples: u32, image: &mut RgbaImage) {
        let s = flatten(scene, &mut self.objects);
        let status = unsafe { sys::rd_set_samples_per_pass(self.handle, samples) };
        self.check(status, "rd_set_samples_per_pass");
        let mut produced: std::os::raw::c_int = 0;
        let (data, len) = (image.as_mut_ptr(), image.len());
        let status = unsafe { sys::rd_render_sample(self.handle, &s, data, len, &mut produced) };
        self.check(status, "rd_render_sample");
    }
}

impl Drop for HipRenderer {
    fn drop(&mut self) {
        unsafe { sys::rd_destroy(self.handle) }
    }
}

impl Renderer for HipRenderer {
    fn render_frame(&mut self, scene: &Scene) -> RgbaImage {
        // cpu.rs:119-133
        self.new_frame(scene);
        let mut image = Self::blank_image(scene);
        let remaining = self.max_sample_count() - self.sample_count();
        if remaining > 0 {
            // render_next_sample's timers (cpu.rs:194-196, 212-218) around one
            // pass of every remaining sample
            self.profiler.prepare_timer.end_if_not_ended();
            self.profiler.render_timer.start_if_not_started();
            self.profiler.sample_timer.start();
            self.render_pass(scene, remaining, &mut image);
            self.profiler.render_timer.end();
            self.profiler.frame_timer.end();
            self.profiler.sample_timer.end_multiple(remaining);
        }
        image
    }

    fn render_sample(&mut self, scene: &Scene) -> Option<RgbaImage> {
        // cpu.rs:142-158: one sample per call, None once all are rendered
        if self.sample_count() >= self.max_sample_count() {
            return None;
        }
        let mut image = Self::blank_image(scene);
        self.profiler.prepare_timer.end_if_not_ended();
        self.profiler.render_timer.start_if_not_started();
        self.profiler.sample_timer.start();
        self.render_pass(scene, 1, &mut image);
        if self.sample_count() == self.max_sample_count() {
            self.profiler.render_timer.end();
            self.profiler.frame_timer.end();
        }
        self.profiler.sample_timer.end();
        Some(image)
    }

    fn new_frame(&mut self, scene: &Scene) {
        // cpu.rs:135-140: the snapshot (BVH build and upload) is the prepare step
        self.profiler.frame_timer.start();
        self.profiler.prepare_timer.start();
        let s = flatten(scene, &mut self.objects);
        let status = unsafe { sys::rd_new_frame(self.handle, &s) };
        self.check(status, "rd_new_frame");
    }

    fn profiler(&self) -> &Profiler {
        &self.profiler
    }

    fn sample_count(&self) -> u32 {
        unsafe { sys::rd_sample_count(self.handle) }
    }

    fn max_sample_count(&self) -> u32 {
        unsafe { sys::rd_max_sample_count(self.handle) }
    }

    fn max_bounces(&self) -> u32 {
        unsafe { sys::rd_max_bounces(self.handle) }
    }

    fn set_max_sample_count(&mut self, count: u32) {
        unsafe { sys::rd_set_max_sample_count(self.handle, count) }
    }

    fn set_max_bounces(&mut self, bounces: u32) {
        unsafe { sys::rd_set_max_bounces(self.handle, bounces) }
    }
}
