// Oracle: CPU restatement of raydar's CpuRenderer path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// or the timed CPU baseline; the product path (raydar_amd/) never calls it.
//
// The reference is Rust (bvpav/raydar) and cannot be built here (no cargo or
// rustc, crates not vendored), so there is no oracle/_ref build.  This file
// restates, op for op in the order written:
//   src/renderer/cpu.rs:16-98     Ray, hit_sphere, hit_cube
//   src/renderer/cpu.rs:193-230   render_next_sample, print_frame_buffer
//   src/renderer/cpu.rs:233-394   per_pixel, trace_ray, closest_hit
//   src/utils/mod.rs:6-55         reflect, refract, can_refract, random_in_unit_sphere
//   src/scene/world.rs:16-34      World::sample
// Third-party arithmetic restated (not vendored, parity unpinned where the
// pinning checks in tests/ do not reach): cgmath 0.18.0 dot = (x*x'+y*y')+z*z',
// normalize(v) = v * (1/|v|), Matrix4*Vector4 = sum of columns in order,
// lerp(a,b,t) = a + (b-a)*t; Rust powi(y,5) = y*((y*y)*(y*y)); Rust f32::min/max
// = IEEE minNum/maxNum (fmin/fmax); ordered-float 4.6.0: NaN sorts above all.
//
// The reference draws from rand 0.8.5 thread_rng (OS-seeded ChaCha12, not
// reproducible).  Here every draw is an explicit uniform from a counter-based
// Philox-4x32-10 stream (Salmon et al., SC'11) keyed by (seed) and counted by
// (pixel, sample, bounce, slot), so the HIP kernels can consume identical
// numbers.  Per bounce: slot 0 transmission test (cpu.rs:279), 1 Fresnel
// test (cpu.rs:295) or metallic test (cpu.rs:305), 2 roughness test
// (cpu.rs:308), 4-6 the random_in_unit_sphere of the direction taken
// (diffuse cpu.rs:266-267, specular cpu.rs:276 or refracted cpu.rs:300:
// exclusive branches, so they share one triple).
//
// Build: -O2 -ffp-contract=off, no fast-math (Rust never contracts a*b+c).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/raydar_hip.h"

namespace {

// ---------------------------------------------------------------- Philox ---
// Philox-4x32 with 10 rounds, written from Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11, section 3.2.
const uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
const uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

void philox4x32_10(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
    uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// The 8 draw slots of one bounce: two Philox blocks (slot s is lane s%4 of
// block s/4).  Counter = (pixel, sample, bounce*4 + block, 0); key = seed.
constexpr int kDrawWords = 8;
void bounce_bits(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce,
                 uint32_t bits[kDrawWords]) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (uint32_t b = 0; b < 2; ++b) {
    const uint32_t ctr[4] = {pixel, sample, bounce * 4u + b, 0u};
    philox4x32_10(ctr, key, bits + 4 * b);
  }
}

// U[0,1) with 24 bits: (x >> 8) * 2^-24, exact in f32 (stands in for
// rand::random::<f32>(), cpu.rs:280,296,306,309).
inline float unit_f32(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// ---------------------------------------------------------------- vectors --
template <class T> struct V3 { T x, y, z; };
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator-(V3<T> a) { return mk(-a.x, -a.y, -a.z); }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <class T> inline V3<T> mul_elem(V3<T> a, V3<T> b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> inline V3<T> div_elem(V3<T> a, V3<T> b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T mag2(V3<T> a) { return dot(a, a); }
template <class T> inline T mag(V3<T> a) { return std::sqrt(mag2(a)); }
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / mag(a)); }
template <class T> inline T rmin(T a, T b) { return std::fmin(a, b); }
template <class T> inline T rmax(T a, T b) { return std::fmax(a, b); }
template <class T> inline T rsignum(T a) { return std::isnan(a) ? a : std::copysign(T(1), a); }
template <class T> inline V3<T> load3(const float *p) { return mk((T)p[0], (T)p[1], (T)p[2]); }

// utils/mod.rs:14-16
template <class T> inline V3<T> reflect(V3<T> v, V3<T> n) { return v - (n * dot(v, n)) * T(2); }
// utils/mod.rs:25-35 (the release-mode assert!s are not restated)
template <class T> inline V3<T> refract(V3<T> v, V3<T> n, T ior_ratio) {
  T cos_theta = rmin(dot(v, -n), T(1));
  V3<T> perp = (v + n * cos_theta) * ior_ratio;
  T s = std::sqrt(std::fabs(T(1) - mag2(perp)));
  V3<T> par = n * (-s);
  return perp + par;
}
// utils/mod.rs:37-44
template <class T> inline bool can_refract(V3<T> v, V3<T> n, T ior_ratio) {
  T cos_theta = rmin(dot(v, -n), T(1));
  T sin_theta = std::sqrt(T(1) - cos_theta * cos_theta);
  return ior_ratio * sin_theta <= T(1);
}
// utils/mod.rs:47-55: normalize(U[-1,1]^3), each component 2u-1 (exact in f32).
template <class T> inline V3<T> random_in_unit_sphere(const uint32_t *bits) {
  V3<T> v = mk((T)(unit_f32(bits[0]) * 2.0f - 1.0f), (T)(unit_f32(bits[1]) * 2.0f - 1.0f),
               (T)(unit_f32(bits[2]) * 2.0f - 1.0f));
  return normalize(v);
}

template <class T> struct Ray { V3<T> o, d; };

// cpu.rs:34-62
template <class T> inline bool hit_sphere(const Ray<T> &r, V3<T> c, T radius, T *t_out) {
  T a = dot(r.d, r.d);
  T k = dot(r.o, r.d) - dot(r.d, c);
  T cc = dot(r.o, r.o) - T(2) * dot(r.o, c) + dot(c, c) - radius * radius;
  T disc = k * k - a * cc;
  if (disc < T(0)) return false;
  T s = std::sqrt(disc);
  T t1 = (-k - s) / a;
  T t2 = (-k + s) / a;
  if (t1 >= T(0)) { *t_out = t1; return true; }
  if (t2 >= T(0)) { *t_out = t2; return true; }
  return false;
}

// cpu.rs:64-98
template <class T> inline bool hit_cube(const Ray<T> &r, V3<T> c, T side, T *t_out) {
  V3<T> half = mk(side * T(0.5), side * T(0.5), side * T(0.5));
  V3<T> mn = c - half, mx = c + half;
  V3<T> t1 = div_elem(mn - r.o, r.d);
  V3<T> t2 = div_elem(mx - r.o, r.d);
  T tmin = rmax(rmax(rmin(t1.x, t2.x), rmin(t1.y, t2.y)), rmin(t1.z, t2.z));
  T tmax = rmin(rmin(rmax(t1.x, t2.x), rmax(t1.y, t2.y)), rmax(t1.z, t2.z));
  if (tmax < T(0)) return false;
  if (tmin > tmax) return false;
  *t_out = (tmin < T(0)) ? tmax : tmin;
  return true;
}

// ordered-float key order: all numbers by value (-0 == +0), NaN above all.
template <class T> inline bool key_less(T a, T b) {
  if (std::isnan(a)) return false;
  if (std::isnan(b)) return true;
  return a < b;
}

// cpu.rs:344-352 trace_ray: linear scan, min_by_key keeps the FIRST minimum.
template <class T> inline int trace(const rd_scene &s, const Ray<T> &r, T *t_best) {
  int best = -1;
  T bt = T(0);
  for (uint32_t i = 0; i < s.object_count; ++i) {
    const rd_object &o = s.objects[i];
    T t;
    bool h = (o.kind == RD_SPHERE) ? hit_sphere(r, load3<T>(o.center), (T)o.size, &t)
                                   : hit_cube(r, load3<T>(o.center), (T)o.size, &t);
    if (h && (best < 0 || key_less(t, bt))) { best = (int)i; bt = t; }
  }
  *t_best = bt;
  return best;
}

template <class T> struct Hit { V3<T> p, n; bool front; };

// cpu.rs:354-394 closest_hit
template <class T> inline Hit<T> closest_hit(const Ray<T> &r, T t, const rd_object &o) {
  Hit<T> h;
  h.p = r.o + r.d * t;
  V3<T> c = load3<T>(o.center);
  V3<T> n;
  if (o.kind == RD_SPHERE) {
    n = normalize(h.p - c);
  } else {
    V3<T> l = h.p - c;
    T half = (T)o.size / T(2);
    T xd = std::fabs(std::fabs(l.x) - half);
    T yd = std::fabs(std::fabs(l.y) - half);
    T zd = std::fabs(std::fabs(l.z) - half);
    if (xd < yd && xd < zd) n = mk(rsignum(l.x), T(0), T(0));
    else if (yd < zd) n = mk(T(0), rsignum(l.y), T(0));
    else n = mk(T(0), T(0), rsignum(l.z));
  }
  h.front = dot(n, r.d) <= T(0);
  if (!h.front) n = -n;
  h.n = n;
  return h;
}

// world.rs:17-34 (Transparent is rejected before rendering starts).
template <class T> inline V3<T> world_sample(const rd_world &w, V3<T> d) {
  if (w.kind == RD_WORLD_SOLID_COLOR) return load3<T>(w.solid_color);
  V3<T> up = mk(T(0), T(1), T(0));
  T cs = dot(d, up) / (mag(d) * mag(up));
  V3<T> bottom = load3<T>(w.bottom_color), top = load3<T>(w.top_color);
  return bottom + (top - bottom) * ((cs + T(1)) * T(0.5));
}

// Result of one scatter (cpu.rs:262-333).
template <class T> struct Scatter { Ray<T> next; V3<T> atten, light; bool transmission; };

// cpu.rs:262-333 with explicit slot uniforms (bits[0..7]).
// ORACLE_MUTANT_* macros plant a known mistake for the test that proves the
// pins in tests/test_oracle.py detect it; the shipped oracle defines none.
template <class T>
inline Scatter<T> scatter(const Ray<T> &ray, const Hit<T> &h, const rd_object &o,
                          const uint32_t *bits, V3<T> atten, V3<T> light) {
#ifdef ORACLE_MUTANT_PLAIN_ROUGHNESS
  T roughness = (T)o.roughness;
#else
  T roughness = (T)o.roughness * (T)o.roughness;
#endif
  T metallic = (T)o.metallic;
  T transmission = (T)o.transmission;
  V3<T> diffuse = h.n + random_in_unit_sphere<T>(bits + 4);
  if (dot(diffuse, h.n) < T(0)) diffuse = -diffuse;
  V3<T> perfect = reflect(ray.d, h.n);
  V3<T> offset = random_in_unit_sphere<T>(bits + 4) * roughness;
  V3<T> specular = normalize(perfect + offset);
  bool trans = (T)unit_f32(bits[0]) < transmission;
  V3<T> dir;
  if (trans) {
    T ior = (T)o.ior;
    if (h.front) ior = T(1) / ior;
    V3<T> rd = normalize(ray.d);
    T cos_theta = rmin(dot(rd, -h.n), T(1));
    T q = (ior - T(1)) / (ior + T(1));
    T r0 = q * q;
    T y = T(1) - cos_theta;
    T y2 = y * y;
    T y5 = y * (y2 * y2);
    T refl = r0 + (T(1) - r0) * y5;
    if (refl < (T)unit_f32(bits[1]) && can_refract(rd, h.n, ior)) {
      V3<T> refr = refract(rd, h.n, ior);
      V3<T> off2 = random_in_unit_sphere<T>(bits + 4) * roughness;
      dir = normalize(refr + off2);
    } else {
      dir = specular;
    }
  } else if ((T)unit_f32(bits[1]) < metallic) {
    dir = specular;
  } else if ((T)unit_f32(bits[2]) < roughness) {
    dir = diffuse;
  } else {
    dir = specular;
  }
  V3<T> off = trans ? dir : h.n;
  Scatter<T> out;
  out.next.o = h.p + off * (T)0.0001f;
  out.next.d = dir;
#ifndef ORACLE_MUTANT_NO_DEGENERATE_RULE
  if (mag2(out.next.d) < (T)1e-10f) out.next.d = h.n;
#endif
  out.atten = mul_elem(atten, load3<T>(o.albedo));
  V3<T> emission = load3<T>(o.emission_color) * (T)o.emission_strength;
  out.light = light + emission;
  out.transmission = trans;
  return out;
}

// cpu.rs:234-251 camera ray for pixel (x, y).
template <class T> inline Ray<T> camera_ray(const rd_scene &s, uint32_t x, uint32_t y) {
  const rd_camera &cam = s.camera;
  T u = (T)x / (T)cam.resolution_x;
  T v = T(1) - (T)y / (T)cam.resolution_y;
  T q[4] = {u * T(2) - T(1), v * T(2) - T(1), T(-1), T(-1)};
  const float *ip = cam.inverse_proj_matrix;
  T p[4];
  for (int r = 0; r < 4; ++r)
    p[r] = (T)ip[0 + r] * q[0] + (T)ip[4 + r] * q[1] + (T)ip[8 + r] * q[2] + (T)ip[12 + r] * q[3];
  T w = p[3];
  for (int r = 0; r < 4; ++r) p[r] = p[r] / w;
  const float *iv = cam.inverse_view_matrix;
  T wv[3];
  for (int r = 0; r < 3; ++r)
    wv[r] = (T)iv[0 + r] * p[0] + (T)iv[4 + r] * p[1] + (T)iv[8 + r] * p[2] + (T)iv[12 + r] * p[3];
  Ray<T> ray;
  ray.o = load3<T>(cam.position);
  ray.d = -normalize(mk(wv[0], wv[1], wv[2]));
  return ray;
}

// cpu.rs:233-342 per_pixel for one (pixel, sample).
template <class T>
inline V3<T> per_pixel(const rd_scene &s, uint32_t x, uint32_t y, uint32_t sample,
                       uint32_t max_bounces, uint64_t seed) {
  Ray<T> ray = camera_ray<T>(s, x, y);
  V3<T> light = mk(T(0), T(0), T(0));
  V3<T> atten = mk(T(1), T(1), T(1));
  const uint32_t pixel = y * s.camera.resolution_x + x;
  for (uint32_t b = 0; b < max_bounces; ++b) {
    T t;
    int id = trace(s, ray, &t);
    if (id >= 0) {
      const rd_object &o = s.objects[id];
      Hit<T> h = closest_hit(ray, t, o);
      uint32_t bits[kDrawWords];
      bounce_bits(seed, pixel, sample, b, bits);
      Scatter<T> sc = scatter(ray, h, o, bits, atten, light);
      ray = sc.next;
      atten = sc.atten;
      light = sc.light;
    } else {
      light = light + mul_elem(world_sample(s.world, ray.d), atten);
      break;
    }
  }
  return light;
}

template <class T>
void render(const rd_scene *s, uint32_t spp, uint32_t max_bounces, uint64_t seed,
            uint32_t sample_base, const uint32_t *pixels, uint32_t n_pixels, T *acc,
            int threads) {
  const uint32_t W = s->camera.resolution_x;
  const int64_t n = pixels ? (int64_t)n_pixels : (int64_t)W * s->camera.resolution_y;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
  for (int64_t i = 0; i < n; ++i) {
    uint32_t p = pixels ? pixels[i] : (uint32_t)i;
    uint32_t x = p % W, y = p / W;
    T *a = acc + 4 * i;
    for (uint32_t k = 0; k < spp; ++k) {
      V3<T> c = per_pixel<T>(*s, x, y, sample_base + k, max_bounces, seed);
      a[0] = a[0] + c.x;
      a[1] = a[1] + c.y;
      a[2] = a[2] + c.z;
      a[3] = a[3] + T(1);
    }
  }
}

// ------------------------------------------ Vulkan-renderer semantics -------
// The optional mode that follows the reference's Vulkan backend instead of
// its CPU renderer (shaders/vulkan/raytrace.rgen, raytrace.rint,
// raytrace_sphere.rchit, raytrace_triangle.rchit; src/renderer/vulkan.rs).
// GLSL leaves the rounding of its built-ins to the implementation; this
// restatement fixes every one as IEEE f32 in the written order:
//   mat4 * vec4  = ((m[0]*v.x + m[1]*v.y) + m[2]*v.z) + m[3]*v.w per row
//   normalize(v) = v * (1 / sqrt(dot(v, v))), length(v) = sqrt(dot(v, v))
//   mix(x, y, a) = x*(1 - a) + y*a, reflect(I, N) = I - (2*dot(N, I))*N
//   refract(I, N, eta): k = 1 - (eta*eta)*(1 - d*d), d = dot(N, I);
//       k < 0 -> 0, else eta*I - (eta*d + sqrt(k))*N
//   pow(x, 2) = x*x, pow(y, 5) = y*((y*y)*(y*y))
// The hardware traversal is restated as: every object is intersected, a
// reported t counts if 1e-3 <= t <= 1e4 (the rgen's t_min / t_max), the
// least such t wins, equal t go to the lower object index.  A cube is the
// 12-triangle box of vulkan.rs:722-893, taken as its closest surface
// crossing at or after t_min (the slab test of the CPU renderer), with the
// face normal of the crossed face (the axis rule of the CPU renderer).
const float kVkTMin = 1e-3f, kVkTMax = 1e4f;

// raytrace.rgen:52-63: the output is a permutation of the OLD state; the
// state then steps by the LCG.  rand() = float(word) / float(UINT_MAX).
inline uint32_t vk_pcg(uint32_t &state) {
  const uint32_t s = state;
  state = state * 747796405u + 2891336453u;
  const uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (word >> 22u) ^ word;
}
inline float vk_unit(uint32_t word) { return (float)word / 4294967295.0f; }

// A draw source: the PCG stream of one pixel, or a caller's list of uniforms.
struct VkDraws {
  uint32_t *state = nullptr;
  const float *list = nullptr;
  uint32_t used = 0;
  float rand() {
    ++used;
    return list ? list[used - 1] : vk_unit(vk_pcg(*state));
  }
  // raytrace.rgen:65-71: rand_range(-1, 1) per component, not normalised
  V3<float> rius() {
    const float x = -1.0f + (1.0f - -1.0f) * rand();
    const float y = -1.0f + (1.0f - -1.0f) * rand();
    const float z = -1.0f + (1.0f - -1.0f) * rand();
    return mk(x, y, z);
  }
};

// raytrace.rint:9-28 with the [t_min, t_max] acceptance of the report.
inline bool vk_hit_sphere(const Ray<float> &r, V3<float> c, float radius, float *t_out) {
  const float a = dot(r.d, r.d);
  const float k = dot(r.o, r.d) - dot(r.d, c);
  const float cc = dot(r.o, r.o) - 2.0f * dot(r.o, c) + dot(c, c) - radius * radius;
  const float disc = k * k - a * cc;
  if (!(disc >= 0.0f)) return false;
  const float s = std::sqrt(disc);
  const float t1 = (-k - s) / a;
  const float t2 = (-k + s) / a;
  float t;
  if (t1 >= 0.0f) t = t1;        // reported; a t1 below t_min is refused,
  else if (t2 >= 0.0f) t = t2;   // and t2 is then not reported at all
  else return false;
  if (!(t >= kVkTMin && t <= kVkTMax)) return false;
  *t_out = t;
  return true;
}

// The cube's closest surface crossing in [t_min, t_max].
inline bool vk_hit_cube(const Ray<float> &r, V3<float> c, float side, float *t_out) {
  V3<float> half = mk(side * 0.5f, side * 0.5f, side * 0.5f);
  V3<float> mn = c - half, mx = c + half;
  V3<float> t1 = div_elem(mn - r.o, r.d);
  V3<float> t2 = div_elem(mx - r.o, r.d);
  const float tmin = rmax(rmax(rmin(t1.x, t2.x), rmin(t1.y, t2.y)), rmin(t1.z, t2.z));
  const float tmax = rmin(rmin(rmax(t1.x, t2.x), rmax(t1.y, t2.y)), rmax(t1.z, t2.z));
  if (!(tmax >= kVkTMin)) return false;
  if (tmin > tmax) return false;
  const float t = tmin >= kVkTMin ? tmin : tmax;
  if (!(t <= kVkTMax)) return false;
  *t_out = t;
  return true;
}

inline int vk_trace(const rd_scene &s, const Ray<float> &r, float *t_best) {
  int best = -1;
  float bt = 0.0f;
  for (uint32_t i = 0; i < s.object_count; ++i) {
    const rd_object &o = s.objects[i];
    float t;
    const bool h = (o.kind == RD_SPHERE) ? vk_hit_sphere(r, load3<float>(o.center), o.size, &t)
                                         : vk_hit_cube(r, load3<float>(o.center), o.size, &t);
    if (h && (best < 0 || t < bt)) { best = (int)i; bt = t; }
  }
  *t_best = bt;
  return best;
}

// raytrace.rgen:83-89: pixel centre, inverse projection, normalised target
// turned by the inverse view (look_at_rh(position, target, -up) inverted,
// vulkan.rs:252-272; passed in as vk_inv_view).
inline void mat_vec(const float *m, const float v[4], float out[4]) {
  for (int r = 0; r < 4; ++r)
    out[r] = m[0 + r] * v[0] + m[4 + r] * v[1] + m[8 + r] * v[2] + m[12 + r] * v[3];
}
inline Ray<float> vk_camera_ray(const rd_scene &s, const float *vk_inv_view, uint32_t x,
                                uint32_t y) {
  const rd_camera &cam = s.camera;
  const float u = ((float)x + 0.5f) / (float)cam.resolution_x;
  const float v = ((float)y + 0.5f) / (float)cam.resolution_y;
  const float dx = u * 2.0f - 1.0f, dy = v * 2.0f - 1.0f;
  const float zero_w[4] = {0.0f, 0.0f, 0.0f, 1.0f};
  float o[4], tgt[4], dir[4];
  mat_vec(vk_inv_view, zero_w, o);
  const float q[4] = {dx, dy, 1.0f, 1.0f};
  mat_vec(cam.inverse_proj_matrix, q, tgt);
  const V3<float> n = normalize(mk(tgt[0], tgt[1], tgt[2]));
  const float nv[4] = {n.x, n.y, n.z, 0.0f};
  mat_vec(vk_inv_view, nv, dir);
  Ray<float> ray;
  ray.o = mk(o[0], o[1], o[2]);
  ray.d = mk(dir[0], dir[1], dir[2]);
  return ray;
}

// raytrace.rgen:43-48 (a solid world is both colours, vulkan.rs:288-299)
inline V3<float> vk_world(const rd_world &w, V3<float> d) {
  V3<float> top, bottom;
  if (w.kind == RD_WORLD_SOLID_COLOR) {
    top = bottom = load3<float>(w.solid_color);
  } else {
    top = load3<float>(w.top_color);
    bottom = load3<float>(w.bottom_color);
  }
  const V3<float> up = mk(0.0f, 1.0f, 0.0f);
  const float cs = dot(d, up) / (mag(d) * mag(up));
  const float a = (cs + 1.0f) * 0.5f;
  return bottom * (1.0f - a) + top * a;
}

// raytrace.rgen:106-172: one hit's scatter, drawing in the shader's order.
inline Scatter<float> vk_scatter(const Ray<float> &ray, const Hit<float> &h, const rd_object &o,
                                 VkDraws &dr, V3<float> atten, V3<float> light) {
  const float roughness = o.roughness * o.roughness;
  V3<float> diffuse = h.n + dr.rius();
  if (dot(diffuse, h.n) < 0.0f) diffuse = -diffuse;
  const float d2 = 2.0f * dot(h.n, ray.d);
  const V3<float> perfect = ray.d - h.n * d2;
  const V3<float> offset = dr.rius() * roughness;
  const V3<float> specular = normalize(perfect + offset);
  const bool trans = dr.rand() < o.transmission;
  V3<float> dir;
  if (trans) {
    float ior = o.ior;
    if (h.front) ior = 1.0f / ior;
    const V3<float> rd = normalize(ray.d);
    const float cos_theta = rmin(dot(rd, -h.n), 1.0f);
    const float q = (ior - 1.0f) / (ior + 1.0f);
    const float r0 = q * q;
    const float y = 1.0f - cos_theta;
    const float y2 = y * y;
    const float refl = r0 + (1.0f - r0) * (y * (y2 * y2));
    const float u = dr.rand();
    const float c2 = rmin(dot(rd, -h.n), 1.0f);
    const bool can = ior * std::sqrt(1.0f - c2 * c2) <= 1.0f;
    if (refl < u && can) {
      const float dn = dot(h.n, rd);
      const float k = 1.0f - (ior * ior) * (1.0f - dn * dn);
      V3<float> refr = mk(0.0f, 0.0f, 0.0f);
      if (!(k < 0.0f)) refr = rd * ior - h.n * (ior * dn + std::sqrt(k));
      const V3<float> off2 = dr.rius() * roughness;
      dir = normalize(refr + off2);
    } else {
      dir = specular;
    }
  } else if (dr.rand() < o.metallic) {
    dir = specular;
  } else if (dr.rand() < roughness) {
    dir = diffuse;
  } else {
    dir = specular;
  }
  const V3<float> off = trans ? dir : h.n;
  Scatter<float> out;
  out.next.o = h.p + off * 0.001f;
  out.next.d = dir;
  if (mag2(out.next.d) < 1e-10f) out.next.d = h.n;
  out.atten = mul_elem(atten, load3<float>(o.albedo));
  out.light = light + load3<float>(o.emission_color) * o.emission_strength;
  out.transmission = trans;
  return out;
}

// raytrace.rgen:79-183 for one pixel: every sample in order, one PCG stream
// seeded x + y*1000 (raytrace.rgen:50) running through all of them.
inline void vk_pixel(const rd_scene &s, const float *vk_inv_view, uint32_t x, uint32_t y,
                     uint32_t spp, uint32_t max_bounces, float acc[4]) {
  uint32_t state = x + y * 1000u;
  VkDraws dr;
  dr.state = &state;
  const Ray<float> cam = vk_camera_ray(s, vk_inv_view, x, y);
  for (uint32_t k = 0; k < spp; ++k) {
    Ray<float> ray = cam;
    V3<float> light = mk(0.0f, 0.0f, 0.0f), atten = mk(1.0f, 1.0f, 1.0f);
    for (uint32_t b = 0; b < max_bounces; ++b) {
      float t;
      const int id = vk_trace(s, ray, &t);
      if (id < 0) {
        light = light + mul_elem(vk_world(s.world, ray.d), atten);
        break;
      }
      const rd_object &o = s.objects[id];
      const Hit<float> h = closest_hit(ray, t, o);
      const Scatter<float> sc = vk_scatter(ray, h, o, dr, atten, light);
      ray = sc.next;
      atten = sc.atten;
      light = sc.light;
    }
    acc[0] = acc[0] + light.x;
    acc[1] = acc[1] + light.y;
    acc[2] = acc[2] + light.z;
    acc[3] = acc[3] + 1.0f;
  }
}

}  // namespace

extern "C" {

// ---- Vulkan-renderer semantics (raytrace.rgen); vk_inv_view is the inverse
// of look_at_rh(position, target, -up) (vulkan.rs:252-272), column-major.

// `n` words of the PCG stream starting from `state`.
void or_vk_pcg(uint32_t state, uint32_t n, uint32_t *out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = vk_pcg(state);
}

void or_vk_camera_rays(const rd_scene *s, const float *vk_inv_view, const uint32_t *pixels,
                       uint32_t n, float *out) {
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t x = pixels[i] % s->camera.resolution_x, y = pixels[i] / s->camera.resolution_x;
    const Ray<float> r = vk_camera_ray(*s, vk_inv_view, x, y);
    float *o = out + 6 * i;
    o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.d.x; o[4] = r.d.y; o[5] = r.d.z;
  }
}

// Closest hit in [t_min, t_max] for n rays (n x 6): id (-1 = miss), t, normal, front.
void or_vk_trace(const rd_scene *s, const float *rays, uint32_t n, int32_t *ids, float *ts,
                 float *normals, uint8_t *front) {
  for (uint32_t i = 0; i < n; ++i) {
    const float *q = rays + 6 * i;
    const Ray<float> r{mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5])};
    float t = 0.f;
    const int id = vk_trace(*s, r, &t);
    ids[i] = id;
    ts[i] = id >= 0 ? t : 0.f;
    V3<float> nn = mk(0.f, 0.f, 0.f);
    uint8_t f = 0;
    if (id >= 0) {
      const Hit<float> h = closest_hit(r, t, s->objects[id]);
      nn = h.n;
      f = h.front;
    }
    normals[3 * i] = nn.x; normals[3 * i + 1] = nn.y; normals[3 * i + 2] = nn.z;
    front[i] = f;
  }
}

// One scatter with caller-chosen uniforms: in = ray o(3) d(3), t, object id
// (float bits), atten(3), light(3); draws: up to 11 uniforms.  out = next
// o(3) d(3), atten(3), light(3), transmission flag, uniforms used.
void or_vk_scatter(const rd_scene *s, const float *in, const float *draws, float *out) {
  const Ray<float> r{mk(in[0], in[1], in[2]), mk(in[3], in[4], in[5])};
  int32_t id;
  std::memcpy(&id, &in[7], 4);
  const rd_object &o = s->objects[id];
  const Hit<float> h = closest_hit(r, in[6], o);
  VkDraws dr;
  dr.list = draws;
  const Scatter<float> sc = vk_scatter(r, h, o, dr, mk(in[8], in[9], in[10]),
                                       mk(in[11], in[12], in[13]));
  out[0] = sc.next.o.x; out[1] = sc.next.o.y; out[2] = sc.next.o.z;
  out[3] = sc.next.d.x; out[4] = sc.next.d.y; out[5] = sc.next.d.z;
  out[6] = sc.atten.x; out[7] = sc.atten.y; out[8] = sc.atten.z;
  out[9] = sc.light.x; out[10] = sc.light.y; out[11] = sc.light.z;
  out[12] = sc.transmission ? 1.f : 0.f;
  out[13] = (float)dr.used;
}

void or_vk_world(const rd_scene *s, const float *dirs, uint32_t n, float *out) {
  for (uint32_t i = 0; i < n; ++i) {
    const V3<float> c = vk_world(s->world, mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
  }
}

// The listed pixels (or all, pixels == NULL), all `spp` samples each, summed
// into acc (n x 4) in sample order.
void or_vk_render(const rd_scene *s, const float *vk_inv_view, uint32_t spp, uint32_t max_bounces,
                  const uint32_t *pixels, uint32_t n_pixels, float *acc, int threads) {
  const uint32_t W = s->camera.resolution_x;
  const int64_t n = pixels ? (int64_t)n_pixels : (int64_t)W * s->camera.resolution_y;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t p = pixels ? pixels[i] : (uint32_t)i;
    vk_pixel(*s, vk_inv_view, p % W, p / W, spp, max_bounces, acc + 4 * i);
  }
}

// imageStore of accumulated / spp into R8G8B8A8_UNORM (raytrace.rgen:186,
// vulkan.rs:196): clamp to [0, 1], scale by 255, round to nearest (ties to
// even); NaN -> 0.
void or_vk_tonemap(const float *acc, uint64_t n_pixels, uint32_t sample_count, uint8_t *rgba) {
  const float spp = (float)sample_count;
  for (uint64_t i = 0; i < 4 * n_pixels; ++i) {
    float v = acc[i] / spp;
    if (std::isnan(v)) { rgba[i] = 0; continue; }
    v = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
    rgba[i] = (uint8_t)std::nearbyint(v * 255.0f);
  }
}

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox4x32_10(ctr, key, out);
}

void or_bounce_bits(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce,
                    uint32_t bits[kDrawWords]) {
  bounce_bits(seed, pixel, sample, bounce, bits);
}

// Camera rays for n pixel indices: out = n x 6 (origin, direction).
void or_camera_rays_f32(const rd_scene *s, const uint32_t *pixels, uint32_t n, float *out) {
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t x = pixels[i] % s->camera.resolution_x, y = pixels[i] / s->camera.resolution_x;
    Ray<float> r = camera_ray<float>(*s, x, y);
    float *o = out + 6 * i;
    o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.d.x; o[4] = r.d.y; o[5] = r.d.z;
  }
}

// Closest hit for n rays (n x 6 floats): object id (-1 = miss), t, the
// closest_hit normal (cpu.rs:354-394) and front-face flag.
void or_trace_f32(const rd_scene *s, const float *rays, uint32_t n, int32_t *ids, float *ts,
                  float *normals, uint8_t *front) {
  for (uint32_t i = 0; i < n; ++i) {
    const float *q = rays + 6 * i;
    Ray<float> r{mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5])};
    float t = 0.f;
    int id = trace(*s, r, &t);
    ids[i] = id;
    ts[i] = id >= 0 ? t : 0.f;
    if (id >= 0) {
      Hit<float> h = closest_hit(r, t, s->objects[id]);
      normals[3 * i] = h.n.x; normals[3 * i + 1] = h.n.y; normals[3 * i + 2] = h.n.z;
      front[i] = h.front;
    } else {
      normals[3 * i] = normals[3 * i + 1] = normals[3 * i + 2] = 0.f;
      front[i] = 0;
    }
  }
}

// Scatter for n hits.  in: n x 16 floats = ray o(3) d(3), t, object id (as
// float bits), atten(3), light(3), pad(2); bits: n x 8.  out: n x 16 floats =
// next o(3) d(3), atten(3), light(3), transmission flag, hit p... (see tests).
void or_scatter_f32(const rd_scene *s, const float *in, const uint32_t *bits, uint32_t n,
                    float *out) {
  for (uint32_t i = 0; i < n; ++i) {
    const float *q = in + 16 * i;
    Ray<float> r{mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5])};
    float t = q[6];
    int32_t id;
    std::memcpy(&id, &q[7], 4);
    const rd_object &o = s->objects[id];
    Hit<float> h = closest_hit(r, t, o);
    Scatter<float> sc = scatter(r, h, o, bits + kDrawWords * i, mk(q[8], q[9], q[10]),
                                mk(q[11], q[12], q[13]));
    float *w = out + 16 * i;
    w[0] = sc.next.o.x; w[1] = sc.next.o.y; w[2] = sc.next.o.z;
    w[3] = sc.next.d.x; w[4] = sc.next.d.y; w[5] = sc.next.d.z;
    w[6] = sc.atten.x; w[7] = sc.atten.y; w[8] = sc.atten.z;
    w[9] = sc.light.x; w[10] = sc.light.y; w[11] = sc.light.z;
    w[12] = sc.transmission ? 1.f : 0.f;
    w[13] = h.n.x; w[14] = h.n.y; w[15] = h.n.z;
  }
}

// Sky / solid world colour for n directions (world.rs:17-34).
void or_world_f32(const rd_scene *s, const float *dirs, uint32_t n, float *out) {
  for (uint32_t i = 0; i < n; ++i) {
    V3<float> c = world_sample(s->world, mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
  }
}

// Accumulate `spp` samples (sample indices sample_base..) for the listed
// pixels (or all pixels when pixels == NULL) into acc (n x 4), in sample order
// exactly as render_next_sample (cpu.rs:193-219) adds them.
void or_render_f32(const rd_scene *s, uint32_t spp, uint32_t max_bounces, uint64_t seed,
                   uint32_t sample_base, const uint32_t *pixels, uint32_t n_pixels, float *acc,
                   int threads) {
  render<float>(s, spp, max_bounces, seed, sample_base, pixels, n_pixels, acc, threads);
}

// The same estimator in FP64 (the high-spp "truth").
void or_render_f64(const rd_scene *s, uint32_t spp, uint32_t max_bounces, uint64_t seed,
                   uint32_t sample_base, const uint32_t *pixels, uint32_t n_pixels, double *acc,
                   int threads) {
  render<double>(s, spp, max_bounces, seed, sample_base, pixels, n_pixels, acc, threads);
}

// print_frame_buffer (cpu.rs:221-230): u8((acc/spp).clamp(0,1)*255), Rust's
// saturating float->u8 cast (NaN -> 0).
void or_tonemap(const float *acc, uint64_t n_pixels, uint32_t sample_count, uint8_t *rgba) {
  const float spp = (float)sample_count;
  for (uint64_t i = 0; i < 4 * n_pixels; ++i) {
    float v = acc[i] / spp;
    if (std::isnan(v)) { rgba[i] = 0; continue; }
    v = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
    rgba[i] = (uint8_t)(v * 255.0f);
  }
}

}  // extern "C"
