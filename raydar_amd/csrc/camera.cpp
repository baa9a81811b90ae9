// Camera::update_matrices (src/scene/camera.rs:210-231) in FP32, for callers
// that edit a camera through the C ABI (the editor's orbit / pan / zoom and
// setters all end in update_matrices, camera.rs:62-208) and for the CLI's
// default scene.
//
// The reference computes with cgmath 0.18 on f32 (not vendored here).  Its
// operations are restated in their order, every one an IEEE f32 operation
// (compiled with -ffp-contract=off: Rust never fuses a*b+c):
//   dot(a, b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   normalize(v)   = v * (1 / sqrt(dot(v, v)))
//   cross(a, b)    = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   look_at_lh(e, c, up) = look_to_rh(e, -(c - e), up): f = normalize(dir),
//       s = normalize(cross(f, up)), u = cross(s, f); columns (s.x, u.x, -f.x, 0),
//       (s.y, u.y, -f.y, 0), (s.z, u.z, -f.z, 0), (-dot(e,s), -dot(e,u), dot(e,f), 1)
//   perspective(Deg fovy, aspect, n, f): rad = fovy * f32(pi/180),
//       cot = 1 / tan(rad / 2); c0r0 = cot / aspect, c1r1 = cot,
//       c2r2 = (f + n) / (n - f), c2r3 = -1, c3r2 = ((2 * f) * n) / (n - f)
//   ortho(l, r, b, t, n, f): c0r0 = 2 / (r - l), c1r1 = 2 / (t - b),
//       c2r2 = -2 / (f - n), c3r0 = -(r + l) / (r - l), c3r1 = -(t + b) / (t - b),
//       c3r2 = -(f + n) / (f - n), c3r3 = 1
//   invert(M): det by cofactors of row 0, None when |det| <= f32::EPSILON
//       (ulps_eq! against zero); entry (col i, row j) of the inverse =
//       det3(transpose(M) without column i and row j) * (-1)^(i+j) * (1 / det)
// The tests pin the result to the matrices stored in the reference's scene
// files (scenes/default.rscn, scenes/benchmark.rscn), which cgmath wrote.
#include <cmath>
#include <cstring>

#include "../../include/raydar_hip.h"

namespace {

struct V3 {
  float x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 normalize(V3 v) { return scale(v, 1.0f / std::sqrt(dot(v, v))); }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 load(const float *p) { return {p[0], p[1], p[2]}; }

// m[c * 4 + r]: column c, row r (cgmath's m[c][r])
void look_to_rh(V3 eye, V3 dir, V3 up, float *m) {
  const V3 f = normalize(dir);
  const V3 s = normalize(cross(f, up));
  const V3 u = cross(s, f);
  const float v[16] = {s.x, u.x, -f.x, 0.0f, s.y, u.y, -f.y, 0.0f,
                       s.z, u.z, -f.z, 0.0f, -dot(eye, s), -dot(eye, u), dot(eye, f), 1.0f};
  std::memcpy(m, v, sizeof(v));
}
void look_at_lh(V3 eye, V3 center, V3 up, float *m) { look_to_rh(eye, neg(sub(center, eye)), up, m); }
void look_at_rh(V3 eye, V3 center, V3 up, float *m) { look_to_rh(eye, sub(center, eye), up, m); }

void perspective(float fovy_deg, float aspect, float n, float f, float *m) {
  const float rad = fovy_deg * (float)(M_PI / 180.0);
  const float cot = 1.0f / std::tan(rad / 2.0f);
  std::memset(m, 0, 16 * sizeof(float));
  m[0] = cot / aspect;
  m[5] = cot;
  m[10] = (f + n) / (n - f);
  m[11] = -1.0f;
  m[14] = (2.0f * f * n) / (n - f);
}

void ortho(float l, float r, float b, float t, float n, float f, float *m) {
  std::memset(m, 0, 16 * sizeof(float));
  m[0] = 2.0f / (r - l);
  m[5] = 2.0f / (t - b);
  m[10] = -2.0f / (f - n);
  m[12] = -(r + l) / (r - l);
  m[13] = -(t + b) / (t - b);
  m[14] = -(f + n) / (f - n);
  m[15] = 1.0f;
}

// Matrix3::determinant with a[c][r] = a[c * 3 + r]
float det3(const float *a) {
  return a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
         a[6] * (a[1] * a[5] - a[4] * a[2]);
}

// the 3x3 matrix of columns `cols` (three of four) of x, each without row `skip`
void minor3(const float *x, const int cols[3], int skip, float *out) {
  for (int c = 0; c < 3; ++c) {
    int k = 0;
    for (int r = 0; r < 4; ++r)
      if (r != skip) out[c * 3 + k++] = x[cols[c] * 4 + r];
  }
}

float det4(const float *m) {
  // Matrix4::determinant: row-0 cofactor expansion; minor k drops column k
  // and row 0, its entries in the (column, row) layout of Matrix3::new
  float mk[4][9];
  for (int k = 0; k < 4; ++k) {
    int cols[3], n = 0;
    for (int c = 0; c < 4; ++c)
      if (c != k) cols[n++] = c;
    // Matrix3::new(m[c0][1], m[c1][1], m[c2][1], m[c0][2], ...) fills column 0
    // with row 1 of the three columns, column 1 with row 2, column 2 with row 3
    for (int r = 1; r < 4; ++r)
      for (int j = 0; j < 3; ++j) mk[k][(r - 1) * 3 + j] = m[cols[j] * 4 + r];
  }
  return m[0] * det3(mk[0]) - m[4] * det3(mk[1]) + m[8] * det3(mk[2]) - m[12] * det3(mk[3]);
}

bool invert(const float *m, float *out) {
  const float det = det4(m);
  if (std::fabs(det) <= 1.1920929e-7f) return false;  // ulps_eq!(det, 0): f32::EPSILON
  const float inv_det = 1.0f / det;
  float t[16];  // transpose
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) t[c * 4 + r] = m[r * 4 + c];
  for (int i = 0; i < 4; ++i) {
    int cols[3], n = 0;
    for (int c = 0; c < 4; ++c)
      if (c != i) cols[n++] = c;
    for (int j = 0; j < 4; ++j) {
      float a[9];
      minor3(t, cols, j, a);  // Matrix3::from_cols of t's columns, truncate_n(j)
      const float sign = ((i + j) & 1) ? -1.0f : 1.0f;
      out[i * 4 + j] = det3(a) * sign * inv_det;
    }
  }
  return true;
}

}  // namespace

extern "C" rd_status rd_camera_update_matrices(rd_camera *c) {
  if (!c || c->resolution_y == 0) return RD_ERR_INVALID_ARGUMENT;
  float view[16], proj[16], iv[16], ip[16];
  look_at_lh(load(c->position), load(c->target), load(c->up), view);
  const float aspect = (float)c->resolution_x / (float)c->resolution_y;
  if (c->projection == RD_PROJECTION_PERSPECTIVE) {
    // PerspectiveFov's asserts (cgmath 0.18): 0 < fovy < pi, aspect > 0,
    // near > 0, far > near
    const float fov = c->projection_param;
    if (!(fov > 0.0f && fov < 180.0f && aspect > 0.0f && c->near_clip > 0.0f &&
          c->far_clip > c->near_clip))
      return RD_ERR_INVALID_ARGUMENT;
    perspective(fov, aspect, c->near_clip, c->far_clip, proj);
  } else if (c->projection == RD_PROJECTION_ORTHOGRAPHIC) {
    const float s = c->projection_param;
    ortho(-s * aspect, s * aspect, -s, s, c->near_clip, c->far_clip, proj);
  } else {
    return RD_ERR_INVALID_ARGUMENT;
  }
  // the reference unwraps both inverses (camera.rs:228-230): a singular
  // matrix panics there, and is an error here, with the camera unchanged
  if (!invert(view, iv) || !invert(proj, ip)) return RD_ERR_INVALID_ARGUMENT;
  std::memcpy(c->view_matrix, view, sizeof(view));
  std::memcpy(c->proj_matrix, proj, sizeof(proj));
  std::memcpy(c->inverse_view_matrix, iv, sizeof(iv));
  std::memcpy(c->inverse_proj_matrix, ip, sizeof(ip));
  return RD_OK;
}

// The Vulkan renderer's camera (vulkan.rs:251-272): view = look_at_rh(position,
// target, up * -1) ("left-handed to right-handed and flip Y"), handed to the
// ray-generation shader inverted.  cgmath's `up * -1.0` negates exactly.
extern "C" rd_status rd_camera_vulkan_inverse_view(const rd_camera *c, float *out_inverse_view) {
  if (!c || !out_inverse_view) return RD_ERR_INVALID_ARGUMENT;
  float view[16], iv[16];
  look_at_rh(load(c->position), load(c->target), neg(load(c->up)), view);
  if (!invert(view, iv)) return RD_ERR_INVALID_ARGUMENT;  // vulkan.rs:271 unwraps
  std::memcpy(out_inverse_view, iv, sizeof(iv));
  return RD_OK;
}
