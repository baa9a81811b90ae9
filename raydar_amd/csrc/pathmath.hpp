// Per-ray arithmetic of raydar's CPU renderer, restated for the GPU kernels.
//
// Every function performs the reference's IEEE-754 FP32 operations in the
// reference's order, so that the HIP kernels return the same bits as the CPU
// renderer for the same inputs.  That requires building with
// -ffp-contract=off (no a*b+c fusion: Rust never fuses), correctly rounded
// division and square root (HIP's default for float), and no fast-math.
//
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z        cgmath 0.18 InnerSpace::dot
//   normalize(v) = v * (1 / |v|)                     cgmath 0.18 normalize_to
//   Matrix4 * Vector4 = ((c0*v.x + c1*v.y) + c2*v.z) + c3*v.w
//
// Only the BVH box test (slab_interval) is free arithmetic: it is a
// conservative cull, never an answer, and may use FMA.
#pragma once

#include <math.h>
#include <stdint.h>

#include "../../include/raydar_hip.h"
#include "layout.hpp"

#if defined(__HIPCC__)
#define RD_HD __host__ __device__ __forceinline__
#else
#define RD_HD inline
#endif

namespace rdh {

struct F3 {
  float x, y, z;
};
RD_HD F3 f3(float x, float y, float z) { return F3{x, y, z}; }
RD_HD F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
RD_HD F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
RD_HD F3 neg(F3 a) { return f3(-a.x, -a.y, -a.z); }
RD_HD F3 scale(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
RD_HD F3 mul(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
RD_HD float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RD_HD F3 normalize(F3 v) { return scale(v, 1.0f / sqrtf(dot(v, v))); }
// Rust f32::signum: +-1 by the sign bit (so -0.0 -> -1.0), NaN stays NaN.
RD_HD float signum(float v) { return isnan(v) ? v : copysignf(1.0f, v); }
RD_HD uint32_t fbits(float v) { return __builtin_bit_cast(uint32_t, v); }
RD_HD float bitsf(uint32_t v) { return __builtin_bit_cast(float, v); }
// sqrtf, correctly rounded.  On the GPU, for x >= 2^-96 (no scaling needed),
// the hardware square root s (within 1 ulp) is corrected by the sign of the
// exact residuals x - s'*s of its neighbours s' = s -+ 1 ulp (fma is exact
// enough to decide which of the three is the rounded root); other x take
// sqrtf.
RD_HD float sqrt_rn(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (x >= 0x1p-96f) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = bitsf(fbits(s) - 1u), su = bitsf(fbits(s) + 1u);
    const float r = fmaf(-sd, s, x) <= 0.0f ? sd : s;
    return fmaf(-su, s, x) > 0.0f ? su : r;
  }
#endif
  return sqrtf(x);
}

// -------------------------------------------------------------- Philox ----
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, section 3.2).
struct U4 {
  uint32_t v[4];
};
RD_HD U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                       uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  U4 out;
  out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
  return out;
}

// Draw slots of one bounce.  Slot s is word s%4 of Philox block s/4 with
// counter (pixel, sample, bounce*4 + block, 0) and key (seed lo, seed hi).
//   block 0: 0 transmission test (cpu.rs:279), 1 Fresnel test (glass,
//            cpu.rs:295) or metallic test (cpu.rs:305), 2 roughness test
//            (cpu.rs:308), 3 unused
//   block 1: 4-6 the random_in_unit_sphere of the direction the bounce
//            takes: diffuse (cpu.rs:266-267), specular offset (cpu.rs:276)
//            or refraction offset (cpu.rs:300); 7 unused
// The reference draws both the diffuse and the specular offsets before it
// picks one; the branches are exclusive, so one triple serves whichever is
// taken and every hit costs two Philox blocks.  A block is generated the
// first time one of its slots is read (a draw depends only on its counter,
// so the order of generation changes no result).
struct BounceDraws {
  uint32_t pixel, sample, bounce, k0, k1;
  U4 b[2];
  uint32_t have;
  RD_HD void init(uint64_t seed, uint32_t px, uint32_t smp, uint32_t bnc) {
    pixel = px; sample = smp; bounce = bnc;
    k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
    have = 0;
  }
  RD_HD uint32_t slot(int s) {
    const int blk = s >> 2;
    if (!(have & (1u << blk))) {
      b[blk] = philox4x32_10(pixel, sample, bounce * 4u + (uint32_t)blk, 0u, k0, k1);
      have |= 1u << blk;
    }
    return b[blk].v[s & 3];
  }
};

// Explicit draws (the parity hooks feed the same words the oracle gets).
struct FixedDraws {
  const uint32_t *w;
  RD_HD uint32_t slot(int s) { return w[s]; }
};

// U[0,1) from the top 24 bits, exact in f32; U[-1,1) = 2u - 1, also exact.
RD_HD float unit_f32(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }
RD_HD float signed_unit_f32(uint32_t x) { return unit_f32(x) * 2.0f - 1.0f; }

// ------------------------------------------------------------- camera -----
// cpu.rs:199-202 uv, cpu.rs:234-251 ray through the inverse matrices.
RD_HD void camera_ray(const DevJob &j, uint32_t x, uint32_t y, F3 *o, F3 *d) {
  const float u = (float)x / (float)j.width;
  const float v = 1.0f - (float)y / (float)j.height;
  const float q0 = u * 2.0f - 1.0f, q1 = v * 2.0f - 1.0f, q2 = -1.0f, q3 = -1.0f;
  float p[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    p[r] = j.inv_proj[r] * q0 + j.inv_proj[4 + r] * q1 + j.inv_proj[8 + r] * q2 +
           j.inv_proj[12 + r] * q3;
  const float w = p[3];
#pragma unroll
  for (int r = 0; r < 4; ++r) p[r] = p[r] / w;
  float wv[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
    wv[r] = j.inv_view[r] * p[0] + j.inv_view[4 + r] * p[1] + j.inv_view[8 + r] * p[2] +
            j.inv_view[12 + r] * p[3];
  *o = f3(j.position[0], j.position[1], j.position[2]);
  *d = neg(normalize(f3(wv[0], wv[1], wv[2])));
}

// ------------------------------------------------------------ hit tests ---
// Per-ray invariants of hit_sphere hoisted out of the primitive loop; they
// are the same FP32 values the reference recomputes for every object.
struct RayPre {
  float ox, oy, oz, dx, dy, dz;
  float a;   // dot(d, d)
  float od;  // dot(o, d)
  float oo;  // dot(o, o)
};
RD_HD RayPre ray_pre(F3 o, F3 d) {
  RayPre r;
  r.ox = o.x; r.oy = o.y; r.oz = o.z; r.dx = d.x; r.dy = d.y; r.dz = d.z;
  r.a = dot(d, d);
  r.od = dot(o, d);
  r.oo = dot(o, o);
  return r;
}

// cpu.rs:34-62 hit_sphere.  p.a = (C, dot(C,C)), p.b.x = r*r.
RD_HD bool hit_sphere(const RayPre &r, const float *pa, const float *pb, float *t) {
  const float dC = r.dx * pa[0] + r.dy * pa[1] + r.dz * pa[2];
  const float oC = r.ox * pa[0] + r.oy * pa[1] + r.oz * pa[2];
  const float k = r.od - dC;
  const float c = r.oo - 2.0f * oC + pa[3] - pb[0];
  const float disc = k * k - r.a * c;
  if (disc < 0.0f) return false;
  const float s = sqrt_rn(disc);
  const float t1 = (-k - s) / r.a;
  if (t1 >= 0.0f) { *t = t1; return true; }
  const float t2 = (-k + s) / r.a;
  if (t2 >= 0.0f) { *t = t2; return true; }
  return false;
}

// cpu.rs:64-98 hit_cube from its six slab quotients (IEEE minNum/maxNum).
RD_HD bool cube_from_quotients(float t1x, float t2x, float t1y, float t2y, float t1z, float t2z,
                               float *t) {
  const float tmin = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
  const float tmax = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
  if (tmax < 0.0f) return false;
  if (tmin > tmax) return false;
  *t = (tmin < 0.0f) ? tmax : tmin;
  return true;
}

// cpu.rs:64-98 hit_cube (true division).  p.a = min, p.b = max.
RD_HD bool hit_cube(const RayPre &r, const float *pa, const float *pb, float *t) {
  return cube_from_quotients((pa[0] - r.ox) / r.dx, (pb[0] - r.ox) / r.dx, (pa[1] - r.oy) / r.dy,
                             (pb[1] - r.oy) / r.dy, (pa[2] - r.oz) / r.dz, (pb[2] - r.oz) / r.dz,
                             t);
}

// min_by_key(OrderedFloat(t)) over the objects in index order (cpu.rs:345-349):
// the smallest t wins, -0 == +0, NaN sorts above every number, and among
// equal keys the lowest object index wins.  Written without branches: a
// number beats NaN, a smaller number wins, and equal keys (two equal
// numbers, or two NaNs) go to the lower index.
RD_HD bool better_hit(float t, uint32_t id, float bt, uint32_t bid) {
  const bool tn = isnan(t), bn = isnan(bt);
  return (bid == kNoHit) | (bn & !tn) | (t < bt) | (((tn & bn) | (t == bt)) & (id < bid));
}

// Test one primitive and fold it into (best_t, best_id).
RD_HD void test_prim(const RayPre &r, const DevPrim &p, float *bt, uint32_t *bid) {
  const uint32_t tag = fbits(p.b[3]);
  float t;
  const bool h = (tag & kTagCube) ? hit_cube(r, p.a, p.b, &t) : hit_sphere(r, p.a, p.b, &t);
  const uint32_t id = tag & ~kTagCube;
  if (h && better_hit(t, id, *bt, *bid)) { *bt = t; *bid = id; }
}

// ----------------------------------------------------------- BVH cull -----
// Conservative ray/box interval.  A = (o + tol) * inv, B = (o - tol) * inv
// expand the box by `tol` on every side: entry = max over axes of
// min(lo*inv - A, hi*inv - B), exit = min of the max; clamped at t = 0.
struct SlabRay {
  float ix, iy, iz;     // 1 / d with |d_i| clamped away from 0 (finite)
  float ax, ay, az;     // (o + tol) * inv
  float bx, by, bz;     // (o - tol) * inv
};
// On the GPU the reciprocal is the hardware's (within 1 ulp instead of half
// an ulp): the box test is a cull, and the rounding margin every box carries
// (1e-6 L, the padding note of bvh.cpp) covers the extra 2^-24 relative error
// of each plane distance; the near/far plane order below holds for any
// positive scale.
RD_HD float safe_inv(float d) {
  const float m = fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d;
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(m);
#else
  return 1.0f / m;
#endif
}
// Whether the box walk answers exactly as the scan would: finite origin and
// direction, and no direction component that safe_inv would clamp unless it
// is an exact zero next to a nonzero component (an axis-parallel ray, whose
// slabs safe_inv scales correctly).  An all-zero direction or a tiny nonzero
// component makes the reference's divisions by d (cpu.rs:74-79) overflow
// differently from the clamped scale; such rays take the linear scan.
RD_HD bool tiny_nonzero(float d) { return d != 0.0f && fabsf(d) < 1e-30f; }
RD_HD bool walkable(const RayPre &r) {
  return isfinite(r.ox) && isfinite(r.oy) && isfinite(r.oz) && isfinite(r.dx) && isfinite(r.dy) &&
         isfinite(r.dz) && !(r.dx == 0.0f && r.dy == 0.0f && r.dz == 0.0f) &&
         !tiny_nonzero(r.dx) && !tiny_nonzero(r.dy) && !tiny_nonzero(r.dz);
}
RD_HD SlabRay slab_ray(const RayPre &r, float tol) {
  SlabRay s;
  s.ix = safe_inv(r.dx); s.iy = safe_inv(r.dy); s.iz = safe_inv(r.dz);
  // lo side uses (o + tol) so that lo - tol - o = lo - (o + tol)
  s.ax = (r.ox + tol) * s.ix; s.ay = (r.oy + tol) * s.iy; s.az = (r.oz + tol) * s.iz;
  s.bx = (r.ox - tol) * s.ix; s.by = (r.oy - tol) * s.iy; s.bz = (r.oz - tol) * s.iz;
  return s;
}
RD_HD bool slab_hit(const SlabRay &s, float lox, float hix, float loy, float hiy, float loz,
                    float hiz, float tcull, float *tenter) {
  const float x0 = fmaf(lox, s.ix, -s.ax), x1 = fmaf(hix, s.ix, -s.bx);
  const float y0 = fmaf(loy, s.iy, -s.ay), y1 = fmaf(hiy, s.iy, -s.by);
  const float z0 = fmaf(loz, s.iz, -s.az), z1 = fmaf(hiz, s.iz, -s.bz);
  const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
  const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
  *tenter = tn;
  return tn <= tf && tn <= tcull;
}

// The same cull with each axis's planes told apart by the direction's sign.
// For ix > 0, lo <= hi and a >= b (o + tol >= o - tol, scaled by ix) give
// lo*ix - a <= hi*ix - b exactly, and rounding keeps the order, so
// min(x0, x1) = fma(lo, ix, -a) and max(x0, x1) = fma(hi, ix, -b); for
// ix < 0 the roles swap.  slab_hit_nf therefore returns slab_hit's entry
// distance and verdict for every box with lo <= hi, without the per-axis
// min/max.  (An empty slot, lo = +inf and hi = -inf, misses here; slab_hit
// "enters" it as a leaf of zero primitives.  Either way nothing is tested.)
// The ray keeps, per axis, the byte offset of the near plane's float4 in a
// DevNode4 (lo_x at 0, hi_x at 16, lo_y at 32, ...; the far plane sits at
// 16 - ox, 80 - oy, 144 - oz) and the subtrahends in near/far order.
struct SlabNF {
  float ix, iy, iz;
  float nx, ny, nz;     // subtrahend of the near plane: a if ix > 0, else b
  float fx, fy, fz;     // of the far plane
  uint32_t ox, oy, oz;  // byte offset of the near plane in a DevNode4
};
RD_HD SlabNF slab_nf(const SlabRay &s) {
  SlabNF n;
  n.ix = s.ix; n.iy = s.iy; n.iz = s.iz;
  const bool sx = s.ix < 0.0f, sy = s.iy < 0.0f, sz = s.iz < 0.0f;
  n.nx = sx ? s.bx : s.ax; n.fx = sx ? s.ax : s.bx;
  n.ny = sy ? s.by : s.ay; n.fy = sy ? s.ay : s.by;
  n.nz = sz ? s.bz : s.az; n.fz = sz ? s.az : s.bz;
  n.ox = sx ? 16u : 0u;
  n.oy = sy ? 48u : 32u;
  n.oz = sz ? 80u : 64u;
  return n;
}
RD_HD bool slab_hit_nf(const SlabNF &s, float xn, float xf, float yn, float yf, float zn, float zf,
                       float tcull, float *tenter) {
  const float x0 = fmaf(xn, s.ix, -s.nx), x1 = fmaf(xf, s.ix, -s.fx);
  const float y0 = fmaf(yn, s.iy, -s.ny), y1 = fmaf(yf, s.iy, -s.fy);
  const float z0 = fmaf(zn, s.iz, -s.nz), z1 = fmaf(zf, s.iz, -s.fz);
  const float tn = fmaxf(fmaxf(x0, y0), fmaxf(z0, 0.0f));
  const float tf = fminf(fminf(x1, y1), fminf(z1, tcull));
  *tenter = tn;
  return tn <= tf;
}

// Per-ray dynamic margin: zero while |o|_1 stays inside the radius the static
// padding covers; beyond it, 2e-3 |o|_1 covers the sphere test's error growth.
RD_HD float ray_tolerance(const RayPre &r, float r0) {
  const float m = fabsf(r.ox) + fabsf(r.oy) + fabsf(r.oz);
  return m > r0 ? 2e-3f * m : 0.0f;
}

// Cull threshold from the current best hit (slightly loosened).
RD_HD float cull_limit(float bt, uint32_t bid) {
  if (bid == kNoHit || isnan(bt)) return INFINITY;
  return bt * 1.0001f + 1e-30f;
}

// -------------------------------------------------------- closest hit -----
// cpu.rs:354-394: position, outward normal, front-face flip.
struct HitRec {
  F3 p, n;
  bool front;
};
RD_HD HitRec closest_hit(F3 o, F3 d, float t, const DevObject &ob) {
  HitRec h;
  h.p = add(o, scale(d, t));
  const F3 c = f3(ob.geo[0], ob.geo[1], ob.geo[2]);
  F3 n;
  if (fbits(ob.m2[2]) == RD_SPHERE) {
    n = normalize(sub(h.p, c));
  } else {
    const F3 l = sub(h.p, c);
    const float half = ob.geo[3] / 2.0f;
    const float xd = fabsf(fabsf(l.x) - half);
    const float yd = fabsf(fabsf(l.y) - half);
    const float zd = fabsf(fabsf(l.z) - half);
    if (xd < yd && xd < zd) n = f3(signum(l.x), 0.0f, 0.0f);
    else if (yd < zd) n = f3(0.0f, signum(l.y), 0.0f);
    else n = f3(0.0f, 0.0f, signum(l.z));
  }
  h.front = dot(n, d) <= 0.0f;
  if (!h.front) n = neg(n);
  h.n = n;
  return h;
}

// ------------------------------------------------------------- scatter ----
// utils/mod.rs:14-16
RD_HD F3 reflect(F3 v, F3 n) { return sub(v, scale(scale(n, dot(v, n)), 2.0f)); }
// utils/mod.rs:25-35 (release-mode assert!s not reproduced)
RD_HD F3 refract(F3 v, F3 n, float eta) {
  const float cos_t = fminf(dot(v, neg(n)), 1.0f);
  const F3 perp = scale(add(v, scale(n, cos_t)), eta);
  const float s = sqrtf(fabsf(1.0f - dot(perp, perp)));
  return add(perp, scale(n, -s));
}
// utils/mod.rs:37-44
RD_HD bool can_refract(F3 v, F3 n, float eta) {
  const float cos_t = fminf(dot(v, neg(n)), 1.0f);
  const float sin_t = sqrtf(1.0f - cos_t * cos_t);
  return eta * sin_t <= 1.0f;
}
// utils/mod.rs:47-55: normalize(U[-1,1]^3)
template <class Draws> RD_HD F3 unit_sphere_draw(Draws &dr, int s0) {
  return normalize(f3(signed_unit_f32(dr.slot(s0)), signed_unit_f32(dr.slot(s0 + 1)),
                      signed_unit_f32(dr.slot(s0 + 2))));
}

struct ScatterOut {
  F3 o, d, atten, light;
  bool transmission;
};

// cpu.rs:262-333 for one hit.  The reference computes the diffuse and the
// specular direction before choosing; here the choice (which depends only on
// draws 6-8, the material and can_refract) comes first and only the chosen
// direction is computed, with the same operations.  Draws are keyed by slot,
// so the result is identical.
template <class Draws>
RD_HD ScatterOut scatter(F3 ray_d, const HitRec &h, const DevObject &ob, Draws &dr, F3 atten,
                         F3 light) {
  const float rough = ob.m0[3];  // roughness^2, precomputed with the same multiply
  const float metallic = ob.m1[3];
  const float transmission = ob.m2[0];
  const bool trans = unit_f32(dr.slot(0)) < transmission;
  enum { kSpecular, kDiffuse, kRefract } pick = kSpecular;
  F3 rd = f3(0.0f, 0.0f, 0.0f);
  float eta = 1.0f;
  if (trans) {
    float ior = ob.m2[1];
    if (h.front) ior = 1.0f / ior;
    rd = normalize(ray_d);
    const float cos_t = fminf(dot(rd, neg(h.n)), 1.0f);
    const float q = (ior - 1.0f) / (ior + 1.0f);
    const float r0 = q * q;
    const float y = 1.0f - cos_t;
    const float y2 = y * y;
    const float y5 = y * (y2 * y2);
    const float refl = r0 + (1.0f - r0) * y5;
    if (refl < unit_f32(dr.slot(1)) && can_refract(rd, h.n, ior)) pick = kRefract;
    eta = ior;
  } else if (unit_f32(dr.slot(1)) < metallic) {
    pick = kSpecular;
  } else if (unit_f32(dr.slot(2)) < rough) {
    pick = kDiffuse;
  }
  // Every branch offsets by the same unit-sphere draw, and the refracted and
  // the reflected direction are normalized alike: computing the draw once
  // and sharing the normalization keeps a wave whose lanes take different
  // branches from running those parts once per branch (same operations per
  // lane, so the same bits).
  const F3 u = unit_sphere_draw(dr, 4);
  F3 dir;
  if (pick == kDiffuse) {
    dir = add(h.n, u);
    if (dot(dir, h.n) < 0.0f) dir = neg(dir);
  } else {
    const F3 base = pick == kRefract ? refract(rd, h.n, eta) : reflect(ray_d, h.n);
    dir = normalize(add(base, scale(u, rough)));
  }
  ScatterOut out;
  const F3 off = trans ? dir : h.n;
  out.o = add(h.p, scale(off, 0.0001f));
  out.d = dir;
  if (dot(out.d, out.d) < 1e-10f) out.d = h.n;
  out.atten = mul(atten, f3(ob.m0[0], ob.m0[1], ob.m0[2]));
  out.light = add(light, f3(ob.m1[0], ob.m1[1], ob.m1[2]));
  out.transmission = trans;
  return out;
}

// --------------------------------------------------------------- world ----
// world.rs:17-34 (Transparent is rejected at new_frame).
RD_HD F3 world_sample(const DevJob &j, F3 d) {
  if (j.world_kind == RD_WORLD_SOLID_COLOR) return f3(j.solid[0], j.solid[1], j.solid[2]);
  const F3 up = f3(0.0f, 1.0f, 0.0f);
  const float cs = dot(d, up) / (sqrtf(dot(d, d)) * sqrtf(dot(up, up)));
  const F3 bottom = f3(j.bottom[0], j.bottom[1], j.bottom[2]);
  const F3 top = f3(j.top[0], j.top[1], j.top[2]);
  return add(bottom, scale(sub(top, bottom), (cs + 1.0f) * 0.5f));
}

// ------------------------------------------------------------- tonemap ----
// cpu.rs:224: ((acc / spp).clamp(0, 1) * 255) as u8; Rust's cast saturates
// and maps NaN to 0.
RD_HD uint8_t tonemap_channel(float acc, float spp) {
  float v = acc / spp;
  if (isnan(v)) return 0;
  v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  return (uint8_t)(v * 255.0f);
}

// ======================================= Vulkan-renderer semantics ========
// The optional mode (render option "semantics" = 1) that follows the
// reference's Vulkan backend: shaders/vulkan/raytrace.rgen, raytrace.rint,
// raytrace_sphere.rchit, raytrace_triangle.rchit and the camera, cube and
// output set-up of src/renderer/vulkan.rs.  GLSL leaves its built-ins'
// rounding to the implementation; here each is IEEE f32 in the written order
// (normalize(v) = v * (1/sqrt(dot(v,v))), mix(x,y,a) = x*(1-a) + y*a,
// pow(x,2) = x*x, pow(y,5) = y*((y*y)*(y*y))).  A cube is the box of its 12
// triangles (vulkan.rs:722-893), hit at its closest surface crossing inside
// the ray's [t_min, t_max], with the crossed face's normal.
constexpr float kVkTMin = 1e-3f;  // raytrace.rgen:103
constexpr float kVkTMax = 1e4f;   // raytrace.rgen:105

// raytrace.rgen:52-57: returns the RXS-M-XS output of the current state and
// steps the state by the LCG.
RD_HD uint32_t vk_pcg(uint32_t &state) {
  const uint32_t s = state;
  state = s * 747796405u + 2891336453u;
  const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}
// raytrace.rgen:59-67: float(word) / UINT_MAX (UINT_MAX rounds to 2^32 in
// f32, so U[0,1]); rand_range(-1, 1) = -1 + 2u.
RD_HD float vk_rand(uint32_t &state) { return (float)vk_pcg(state) / 4294967295.0f; }
RD_HD float vk_rand_signed(uint32_t &state) { return -1.0f + 2.0f * vk_rand(state); }
// raytrace.rgen:69-71, not normalised; x, y, z drawn in that order
RD_HD F3 vk_rius(uint32_t &state) {
  const float x = vk_rand_signed(state);
  const float y = vk_rand_signed(state);
  const float z = vk_rand_signed(state);
  return f3(x, y, z);
}

// raytrace.rint:16-27 and the report's [t_min, t_max] test: t1 is reported
// when t1 >= 0 (and refused below t_min, t2 then not tried), else t2.
RD_HD bool hit_sphere_vk(const RayPre &r, const float *pa, const float *pb, float *t) {
  const float dC = r.dx * pa[0] + r.dy * pa[1] + r.dz * pa[2];
  const float oC = r.ox * pa[0] + r.oy * pa[1] + r.oz * pa[2];
  const float k = r.od - dC;
  const float c = r.oo - 2.0f * oC + pa[3] - pb[0];
  const float disc = k * k - r.a * c;
  if (!(disc >= 0.0f)) return false;
  const float s = sqrt_rn(disc);
  float tt = (-k - s) / r.a;
  if (!(tt >= 0.0f)) {
    tt = (-k + s) / r.a;
    if (!(tt >= 0.0f)) return false;
  }
  *t = tt;
  return tt >= kVkTMin && tt <= kVkTMax;
}

// The box's first crossing at or after t_min: its entry, or its exit for a
// ray that enters before t_min.
RD_HD bool hit_cube_vk(const RayPre &r, const float *pa, const float *pb, float *t) {
  const float t1x = (pa[0] - r.ox) / r.dx, t2x = (pb[0] - r.ox) / r.dx;
  const float t1y = (pa[1] - r.oy) / r.dy, t2y = (pb[1] - r.oy) / r.dy;
  const float t1z = (pa[2] - r.oz) / r.dz, t2z = (pb[2] - r.oz) / r.dz;
  const float tmin = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
  const float tmax = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
  if (!(tmax >= kVkTMin) || tmin > tmax) return false;
  *t = tmin >= kVkTMin ? tmin : tmax;
  return *t <= kVkTMax;
}

// The hardware keeps the least accepted t; equal t go to the lower index.
// Accepted t are numbers, so better_hit reduces to that order.
RD_HD void test_prim_vk(const RayPre &r, const DevPrim &p, float *bt, uint32_t *bid) {
  const uint32_t tag = fbits(p.b[3]);
  float t;
  const bool h = (tag & kTagCube) ? hit_cube_vk(r, p.a, p.b, &t) : hit_sphere_vk(r, p.a, p.b, &t);
  const uint32_t id = tag & ~kTagCube;
  if (h && better_hit(t, id, *bt, *bid)) { *bt = t; *bid = id; }
}

// raytrace.rgen:83-89.  j.inv_view holds the inverse of look_at_rh(position,
// target, -up) (vulkan.rs:252-272), j.inv_proj the scene's inverse projection.
RD_HD void camera_ray_vk(const DevJob &j, uint32_t x, uint32_t y, F3 *o, F3 *d) {
  const float u = ((float)x + 0.5f) / (float)j.width;
  const float v = ((float)y + 0.5f) / (float)j.height;
  const float q0 = u * 2.0f - 1.0f, q1 = v * 2.0f - 1.0f;
  float org[3], tgt[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    org[r] = j.inv_view[r] * 0.0f + j.inv_view[4 + r] * 0.0f + j.inv_view[8 + r] * 0.0f +
             j.inv_view[12 + r] * 1.0f;
    tgt[r] = j.inv_proj[r] * q0 + j.inv_proj[4 + r] * q1 + j.inv_proj[8 + r] * 1.0f +
             j.inv_proj[12 + r] * 1.0f;
  }
  const F3 n = normalize(f3(tgt[0], tgt[1], tgt[2]));
  float dir[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
    dir[r] = j.inv_view[r] * n.x + j.inv_view[4 + r] * n.y + j.inv_view[8 + r] * n.z +
             j.inv_view[12 + r] * 0.0f;
  *o = f3(org[0], org[1], org[2]);
  *d = f3(dir[0], dir[1], dir[2]);
}

// raytrace.rgen:43-48; a solid world is both colours (vulkan.rs:288-299).
RD_HD F3 world_sample_vk(const DevJob &j, F3 d) {
  const bool solid = j.world_kind == RD_WORLD_SOLID_COLOR;
  const F3 top = solid ? f3(j.solid[0], j.solid[1], j.solid[2]) : f3(j.top[0], j.top[1], j.top[2]);
  const F3 bottom =
      solid ? f3(j.solid[0], j.solid[1], j.solid[2]) : f3(j.bottom[0], j.bottom[1], j.bottom[2]);
  const F3 up = f3(0.0f, 1.0f, 0.0f);
  const float cs = dot(d, up) / (sqrtf(dot(d, d)) * sqrtf(dot(up, up)));
  const float a = (cs + 1.0f) * 0.5f;
  return add(scale(bottom, 1.0f - a), scale(top, a));
}

// raytrace.rgen:106-172 for one hit, drawing from the pixel's PCG state in
// the shader's order: diffuse offset (3), specular offset (3), transmission
// test, then the Fresnel test and, when refracting, the refraction offset
// (3); or the metallic test and, for a non-metal, the roughness test.
RD_HD ScatterOut scatter_vk(F3 ray_d, const HitRec &h, const DevObject &ob, uint32_t &state,
                            F3 atten, F3 light) {
  const float rough = ob.m0[3];  // roughness * roughness (raytrace.rgen:118)
  F3 diffuse = add(h.n, vk_rius(state));
  if (dot(diffuse, h.n) < 0.0f) diffuse = neg(diffuse);
  const F3 perfect = sub(ray_d, scale(h.n, 2.0f * dot(h.n, ray_d)));  // GLSL reflect
  const F3 specular = normalize(add(perfect, scale(vk_rius(state), rough)));
  const bool trans = vk_rand(state) < ob.m2[0];
  F3 dir = specular;
  if (trans) {
    float ior = ob.m2[1];
    if (h.front) ior = 1.0f / ior;
    const F3 rd = normalize(ray_d);
    const float cos_t = fminf(dot(rd, neg(h.n)), 1.0f);
    const float q = (ior - 1.0f) / (ior + 1.0f);
    const float r0 = q * q;
    const float y = 1.0f - cos_t;
    const float y2 = y * y;
    const float refl = r0 + (1.0f - r0) * (y * (y2 * y2));
    const float u = vk_rand(state);
    if (refl < u && can_refract(rd, h.n, ior)) {
      // GLSL refract(I, N, eta)
      const float dn = dot(h.n, rd);
      const float k = 1.0f - (ior * ior) * (1.0f - dn * dn);
      F3 refr = f3(0.0f, 0.0f, 0.0f);
      if (!(k < 0.0f)) refr = sub(scale(rd, ior), scale(h.n, ior * dn + sqrtf(k)));
      dir = normalize(add(refr, scale(vk_rius(state), rough)));
    }
  } else if (!(vk_rand(state) < ob.m1[3])) {
    if (vk_rand(state) < rough) dir = diffuse;
  }
  ScatterOut out;
  const F3 off = trans ? dir : h.n;
  out.o = add(h.p, scale(off, 0.001f));
  out.d = dir;
  if (dot(out.d, out.d) < 1e-10f) out.d = h.n;
  out.atten = mul(atten, f3(ob.m0[0], ob.m0[1], ob.m0[2]));
  out.light = add(light, f3(ob.m1[0], ob.m1[1], ob.m1[2]));
  out.transmission = trans;
  return out;
}

// imageStore of accumulated / spp into R8G8B8A8_UNORM (raytrace.rgen:186,
// vulkan.rs:196): clamp, scale by 255, round to nearest even; NaN -> 0.
RD_HD uint8_t tonemap_channel_vk(float acc, float spp) {
  float v = acc / spp;
  if (isnan(v)) return 0;
  v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  return (uint8_t)rintf(v * 255.0f);
}

}  // namespace rdh
