// Host-side scene digestion and binned-SAH BVH build (see bvh.hpp).
//
// Padding.  hit_sphere (cpu.rs:34-62) forms c = |o|^2 - 2 o.C + |C|^2 - r^2 in
// FP32.  Its rounding error is at most ~17u L^2 with L = |o| + |C| + r and
// u = 2^-24 (k and the discriminant add less), so the FP32 test can report a
// hit for rays passing up to sqrt(r^2 + 17u L^2) - r outside the true sphere,
// and the hit point it reports lies within that distance of the sphere.  Such
// "phantom" hits are part of the reference's answer, so every primitive box is
// grown by twice that excess (2e-6 L^2 inside the square root) for every ray
// origin with |o|_1 <= r0, plus 1e-6 L for the box test's own rounding.  Rays
// starting beyond r0 (only possible after hitting a primitive larger than the
// rest of the scene, e.g. a ground sphere's far side) add a per-ray margin of
// 2e-3 |o|_1 at traversal time (ray_tolerance in pathmath.hpp).  The CPU tests
// compare this BVH against the linear scan on adversarial rays.
#include "bvh.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>

#include "pathmath.hpp"

namespace rdh {

namespace {

// Item boxes are the padded primitive boxes rounded outward to FP32, so every
// node box (a union of item boxes) is exact in FP32 and still conservative.
struct Item {
  float lo[3], hi[3], c[3];
  uint32_t id;
};

struct Box {
  float lo[3], hi[3];
  Box() {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::numeric_limits<float>::infinity();
      hi[a] = -std::numeric_limits<float>::infinity();
    }
  }
  void grow(const float *l, const float *h) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], l[a]);
      hi[a] = std::max(hi[a], h[a]);
    }
  }
  void grow(const Box &b) { grow(b.lo, b.hi); }
  double area() const {
    double e[3];
    for (int a = 0; a < 3; ++a) e[a] = std::max(0.0, (double)hi[a] - (double)lo[a]);
    return 2.0 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]);
  }
};

double l1(const float *v) { return std::fabs((double)v[0]) + std::fabs((double)v[1]) + std::fabs((double)v[2]); }

// sqrt(r^2 + 2e-6 L^2) - r, computed without cancellation.
double sphere_excess(double r, double L) {
  const double q = 2e-6 * L * L;
  return q / (std::sqrt(r * r + q) + r);
}

float round_down(double v) {
  float f = (float)v;
  if ((double)f > v) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}
float round_up(double v) {
  float f = (float)v;
  if ((double)f < v) f = std::nextafter(f, std::numeric_limits<float>::infinity());
  return f;
}

// Top-down binned SAH.  A node with at most leaf_size primitives becomes a
// leaf; larger ones split at the cheapest of kBins-1 planes on each axis,
// with a traversal cost of one primitive test per node visit.  All three
// axes are binned in one pass over the node's items.  Near the stack-depth
// limit the split falls back to the centroid median, which bounds the depth
// by ceil(log2(n / leaf_size)) more levels.
struct Builder {
  std::vector<Item> items;
  std::vector<DevNode> nodes;
  std::vector<uint32_t> order;
  int leaf_size = 4;
  uint32_t max_depth_seen = 0;
  uint32_t max_leaf = 0;
  double sah = 0.0;
  static constexpr int kBins = 16;
  static constexpr int kMaxDepth = kStackDepth - 2;

  static int ceil_log2(size_t n) {
    int k = 0;
    while (((size_t)1 << k) < n) ++k;
    return k;
  }

  // Returns the child code for items[b, e).
  int32_t build(size_t b, size_t e, uint32_t depth, const Box &bounds) {
    if (e - b <= (size_t)leaf_size) return make_leaf(b, e, depth);
    Box lb, rb;
    const size_t mid = split(b, e, depth, bounds, &lb, &rb);
    const uint32_t idx = (uint32_t)nodes.size();
    nodes.emplace_back();
    max_depth_seen = std::max(max_depth_seen, depth);
    sah += bounds.area();
    const int32_t c0 = build(b, mid, depth + 1, lb);
    const int32_t c1 = build(mid, e, depth + 1, rb);
    DevNode &nd = nodes[idx];
    std::memset(&nd, 0, sizeof(nd));
    set_box(nd, 0, lb);
    set_box(nd, 1, rb);
    nd.child[0] = c0;
    nd.child[1] = c1;
    return (int32_t)idx;
  }

  static void set_box(DevNode &nd, int k, const Box &b) {
    nd.lo_x[k] = b.lo[0]; nd.hi_x[k] = b.hi[0];
    nd.lo_y[k] = b.lo[1]; nd.hi_y[k] = b.hi[1];
    nd.lo_z[k] = b.lo[2]; nd.hi_z[k] = b.hi[2];
  }

  int32_t make_leaf(size_t b, size_t e, uint32_t depth) {
    const uint32_t first = (uint32_t)order.size();
    max_leaf = std::max(max_leaf, (uint32_t)(e - b));
    for (size_t i = b; i < e; ++i) order.push_back(items[i].id);
    max_depth_seen = std::max(max_depth_seen, depth);
    return leaf_code(first, (uint32_t)(e - b));
  }

  struct Binning {
    float lo[3], scale[3];
    int bin(const Item &it, int a) const {
      const int k = (int)((it.c[a] - lo[a]) * scale[a]);
      return std::min(std::max(k, 0), kBins - 1);
    }
  };

  // Splits items[b, e) in place; returns the first index of the right half
  // and the two halves' bounds.
  size_t split(size_t b, size_t e, uint32_t depth, const Box &bounds, Box *lb, Box *rb) {
    const size_t n = e - b;
    Box cb;
    for (size_t i = b; i < e; ++i) cb.grow(items[i].c, items[i].c);
    const int need = ceil_log2((n + leaf_size - 1) / leaf_size);
    size_t mid = b;
    if ((int)depth + need + 1 < kMaxDepth) {
      Binning bn;
      bool any = false;
      for (int a = 0; a < 3; ++a) {
        const float ext = cb.hi[a] - cb.lo[a];
        bn.lo[a] = cb.lo[a];
        // (an extent so small that the scale overflows bins nothing on its axis)
        const float scale = ext > 0.f ? (float)kBins * (1.0f - 1e-6f) / ext : 0.f;
        bn.scale[a] = std::isfinite(scale) ? scale : 0.f;
        any |= bn.scale[a] > 0.f;
      }
      int axis = -1, plane = 0;
      if (n == 2 && any) {
        // Two items: every plane that separates them gives the same two
        // children at the same cost, so the binned search below would pick
        // the first axis on which their bins differ, with the lower bin on
        // the left.  Same split, without filling and sweeping the bins.
        for (int a = 0; a < 3 && axis < 0; ++a) {
          if (bn.scale[a] == 0.f) continue;
          const int k0 = bn.bin(items[b], a), k1 = bn.bin(items[b + 1], a);
          if (k0 != k1) {
            axis = a;
            if (k0 > k1) std::swap(items[b], items[b + 1]);
          }
        }
        if (axis >= 0) {
          *lb = Box();
          *rb = Box();
          lb->grow(items[b].lo, items[b].hi);
          rb->grow(items[b + 1].lo, items[b + 1].hi);
          return b + 1;
        }
        any = false;  // no axis separates them: the median split below
      }
      Box bin_box[3][kBins];
      uint32_t bin_n[3][kBins] = {};
      if (any) {
        for (size_t i = b; i < e; ++i) {
          const Item &it = items[i];
          for (int a = 0; a < 3; ++a) {
            const int k = bn.bin(it, a);
            bin_n[a][k]++;
            bin_box[a][k].grow(it.lo, it.hi);
          }
        }
        const double inv_area = 1.0 / std::max(bounds.area(), 1e-30);
        double best = std::numeric_limits<double>::infinity();
        for (int a = 0; a < 3; ++a) {
          if (bn.scale[a] == 0.f) continue;
          double left_cost[kBins];
          Box acc;
          uint32_t cnt = 0;
          for (int k = 0; k < kBins - 1; ++k) {
            acc.grow(bin_box[a][k]);
            cnt += bin_n[a][k];
            left_cost[k] = cnt ? acc.area() * cnt : -1.0;
          }
          acc = Box();
          cnt = 0;
          for (int k = kBins - 1; k > 0; --k) {
            acc.grow(bin_box[a][k]);
            cnt += bin_n[a][k];
            if (cnt == 0 || left_cost[k - 1] < 0.0) continue;
            const double c = 1.0 + (left_cost[k - 1] + acc.area() * cnt) * inv_area;
            if (c < best) {
              best = c;
              axis = a;
              plane = k;
            }
          }
        }
      }
      if (axis >= 0) {
        auto it = std::partition(items.begin() + b, items.begin() + e,
                                 [&](const Item &x) { return bn.bin(x, axis) < plane; });
        mid = (size_t)(it - items.begin());
        *lb = Box();
        *rb = Box();
        for (int k = 0; k < plane; ++k) lb->grow(bin_box[axis][k]);
        for (int k = plane; k < kBins; ++k) rb->grow(bin_box[axis][k]);
      }
    }
    if (mid == b || mid == e) {
      // Median split on the widest centroid axis (also for coincident centroids).
      int a = 0;
      for (int k = 1; k < 3; ++k)
        if (cb.hi[k] - cb.lo[k] > cb.hi[a] - cb.lo[a]) a = k;
      mid = b + n / 2;
      std::nth_element(items.begin() + b, items.begin() + mid, items.begin() + e,
                       [&](const Item &x, const Item &y) {
                         return x.c[a] < y.c[a] || (x.c[a] == y.c[a] && x.id < y.id);
                       });
      *lb = Box();
      *rb = Box();
      for (size_t i = b; i < mid; ++i) lb->grow(items[i].lo, items[i].hi);
      for (size_t i = mid; i < e; ++i) rb->grow(items[i].lo, items[i].hi);
    }
    return mid;
  }
};

// Collapses the binary tree into 4-wide nodes: a node takes its two
// children and repeatedly replaces the inner child of largest surface area
// by that child's two children, up to four.  Children follow their parent.
// need = stack entries a traversal below the node can hold (an entered node
// defers all but one of its k entered inner children).
// Collapses the binary tree into 4-wide nodes, choosing which binary nodes
// to keep by dynamic programming so that the summed surface area of the
// 4-wide nodes (each visit tests four child boxes, and a node is visited
// with probability proportional to its area) is minimal.  F(x, k) is the
// least cost of representing binary subtree x in at most k slots of its
// parent: either x stays a node (one slot, its area plus its two children
// spread over four slots) or x dissolves and its children share the k slots.
// Leaves take one slot and cost nothing here (they are tested whatever the
// collapse).  The traversal's stack bound is computed on the result.
struct Collapser {
  const std::vector<DevNode> &n2;
  std::vector<DevNode4> out;
  uint32_t depth = 0;
  struct Child {
    int32_t code;
    float box[6];  // lo x, hi x, lo y, hi y, lo z, hi z
  };
  std::vector<std::array<double, 5>> F;   // F[x][k], k = 1..4
  std::vector<std::array<int8_t, 5>> how;  // 0: x stays a node; j: j slots to child 0
  std::vector<int8_t> split4;              // x as a node: slots for child 0

  explicit Collapser(const std::vector<DevNode> &nodes)
      : n2(nodes), F(nodes.size()), how(nodes.size()), split4(nodes.size(), 1) {}

  static double area(const Child &c) {
    const double ex = std::max(0.0, (double)c.box[1] - c.box[0]);
    const double ey = std::max(0.0, (double)c.box[3] - c.box[2]);
    const double ez = std::max(0.0, (double)c.box[5] - c.box[4]);
    return ex * ey + ey * ez + ez * ex;
  }
  Child child_of(int32_t node, int k) const {
    const DevNode &m = n2[node];
    return Child{m.child[k], {m.lo_x[k], m.hi_x[k], m.lo_y[k], m.hi_y[k], m.lo_z[k], m.hi_z[k]}};
  }
  double cost(int32_t c, int k) const { return is_leaf(c) ? 0.0 : F[c][k]; }

  // Fills F / how / split4 for the subtree of binary node b (area a).
  void solve(int32_t b, double a) {
    const Child c0 = child_of(b, 0), c1 = child_of(b, 1);
    if (!is_leaf(c0.code)) solve(c0.code, area(c0));
    if (!is_leaf(c1.code)) solve(c1.code, area(c1));
    double best = std::numeric_limits<double>::infinity();
    for (int j = 1; j <= 3; ++j) {
      const double c = cost(c0.code, j) + cost(c1.code, 4 - j);
      if (c < best) {
        best = c;
        split4[b] = (int8_t)j;
      }
    }
    const double as_node = a + best;
    F[b][1] = as_node;
    how[b][1] = 0;
    for (int k = 2; k <= 4; ++k) {
      double bk = as_node;
      int8_t h = 0;
      for (int j = 1; j < k; ++j) {
        const double c = cost(c0.code, j) + cost(c1.code, k - j);
        if (c < bk) {
          bk = c;
          h = (int8_t)j;
        }
      }
      F[b][k] = bk;
      how[b][k] = h;
    }
  }

  // The slots item c takes when given at most k of them.
  void emit(const Child &c, int k, std::vector<Child> *slots) const {
    if (is_leaf(c.code) || how[c.code][k] == 0) {
      slots->push_back(c);
      return;
    }
    const int j = how[c.code][k];
    emit(child_of(c.code, 0), j, slots);
    emit(child_of(c.code, 1), k - j, slots);
  }

  int32_t collapse(int32_t b, uint32_t level, uint32_t *need) {
    std::vector<Child> ch;
    ch.reserve(4);
    emit(child_of(b, 0), split4[b], &ch);
    emit(child_of(b, 1), 4 - split4[b], &ch);
    const int cnt = (int)ch.size();
    const uint32_t idx = (uint32_t)out.size();
    out.emplace_back();
    depth = std::max(depth, level);
    uint32_t inner = 0, deepest = 0;
    for (int i = 0; i < cnt; ++i)
      if (!is_leaf(ch[i].code)) {
        uint32_t nc = 0;
        ch[i].code = collapse(ch[i].code, level + 1, &nc);
        ++inner;
        deepest = std::max(deepest, nc);
      }
    *need = inner ? inner - 1 + deepest : 0;
    DevNode4 &nd = out[idx];
    std::memset(&nd, 0, sizeof(nd));
    for (int i = 0; i < 4; ++i) {
      if (i < cnt) {
        nd.lo_x[i] = ch[i].box[0]; nd.hi_x[i] = ch[i].box[1];
        nd.lo_y[i] = ch[i].box[2]; nd.hi_y[i] = ch[i].box[3];
        nd.lo_z[i] = ch[i].box[4]; nd.hi_z[i] = ch[i].box[5];
        nd.child[i] = ch[i].code;
      } else {
        nd.lo_x[i] = nd.lo_y[i] = nd.lo_z[i] = std::numeric_limits<float>::infinity();
        nd.hi_x[i] = nd.hi_y[i] = nd.hi_z[i] = -std::numeric_limits<float>::infinity();
        nd.child[i] = leaf_code(0, 0);
      }
    }
    return (int32_t)idx;
  }
};

bool finite3(const float *v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

}  // namespace

bool build_scene(const rd_scene &scene, int leaf_size, BuiltScene *out, std::string *err) {
  const uint32_t n = scene.object_count;
  if (n > 0 && scene.objects == nullptr) {
    *err = "scene.objects is NULL with object_count > 0";
    return false;
  }
  if (n >= (1u << 27)) {
    *err = "too many objects (limit 2^27)";
    return false;
  }
  if (leaf_size < 1 || leaf_size > kMaxLeafSize) {
    *err = "leaf size must be in 1..8";
    return false;
  }
  out->objects.resize(n);
  out->prims.clear();
  out->nodes.clear();
  // Scene radius for the padding rule (L1 norms bound L2 norms).
  double r_all = l1(scene.camera.position), r_small = r_all;
  for (uint32_t i = 0; i < n; ++i) {
    const rd_object &o = scene.objects[i];
    if (o.kind != RD_SPHERE && o.kind != RD_CUBE) {
      *err = "object " + std::to_string(i) + ": unknown geometry kind";
      return false;
    }
    if (!finite3(o.center) || !std::isfinite(o.size)) {
      *err = "object " + std::to_string(i) + ": non-finite geometry";
      return false;
    }
    r_all = std::max(r_all, l1(o.center) + 2.0 * std::fabs((double)o.size));
  }
  for (uint32_t i = 0; i < n; ++i) {
    const rd_object &o = scene.objects[i];
    const double ext = l1(o.center) + 2.0 * std::fabs((double)o.size);
    if (std::fabs((double)o.size) <= 0.25 * r_all) r_small = std::max(r_small, ext);
  }
  out->r0 = (float)r_small;

  Builder bld;
  bld.leaf_size = leaf_size;
  bld.items.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    const rd_object &o = scene.objects[i];
    DevObject &d = out->objects[i];
    std::memset(&d, 0, sizeof(d));
    d.geo[0] = o.center[0]; d.geo[1] = o.center[1]; d.geo[2] = o.center[2]; d.geo[3] = o.size;
    d.m0[0] = o.albedo[0]; d.m0[1] = o.albedo[1]; d.m0[2] = o.albedo[2];
    d.m0[3] = o.roughness * o.roughness;                         // cpu.rs:262-263
    d.m1[0] = o.emission_color[0] * o.emission_strength;         // cpu.rs:332-333
    d.m1[1] = o.emission_color[1] * o.emission_strength;
    d.m1[2] = o.emission_color[2] * o.emission_strength;
    d.m1[3] = o.metallic;
    d.m2[0] = o.transmission;
    d.m2[1] = o.ior;
    d.m2[2] = bitsf(o.kind);

    Item &it = bld.items[i];
    it.id = i;
    const double size = std::fabs((double)o.size);
    const double huge = size > 0.25 * r_all;
    const double L = (huge ? r_all : r_small) + l1(o.center) + size;
    double half, pad;
    if (o.kind == RD_SPHERE) {
      half = size;
      pad = 2.0 * sphere_excess(size, L);
    } else {
      half = 0.5 * size;
      pad = 2e-5 * L;
    }
    pad += 1e-6 * (r_all + l1(o.center) + size) + 1e-7;
    for (int a = 0; a < 3; ++a) {
      it.c[a] = o.center[a];
      it.lo[a] = round_down((double)o.center[a] - half - pad);
      it.hi[a] = round_up((double)o.center[a] + half + pad);
    }
  }
  if (n > 0) {
    Box root;
    for (auto &it : bld.items) root.grow(it.lo, it.hi);
    if (n <= (uint32_t)leaf_size) {
      // a root node whose second child is an empty leaf
      bld.nodes.emplace_back();
      DevNode &nd = bld.nodes[0];
      std::memset(&nd, 0, sizeof(nd));
      const int32_t leaf = bld.make_leaf(0, n, 1);
      Builder::set_box(nd, 0, root);
      Builder::set_box(nd, 1, root);
      nd.child[0] = leaf;
      nd.child[1] = leaf_code(0, 0);
      bld.max_depth_seen = 1;
    } else {
      bld.build(0, n, 1, root);
    }
  }
  out->nodes = std::move(bld.nodes);
  out->depth = bld.max_depth_seen;
  out->max_leaf = bld.max_leaf;
  out->nodes4.clear();
  out->depth4 = 0;
  out->stack4 = 0;
  if (!out->nodes.empty()) {
    Collapser c(out->nodes);
    c.out.reserve(out->nodes.size());
    c.solve(0, 0.0);
    uint32_t need = 0;
    c.collapse(0, 1, &need);
    out->nodes4 = std::move(c.out);
    out->depth4 = c.depth;
    out->stack4 = need;
  }
  out->sah_cost = bld.sah;
  out->prims.resize(n);
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t id = bld.order[k];
    const rd_object &o = scene.objects[id];
    DevPrim &p = out->prims[k];
    std::memset(&p, 0, sizeof(p));
    if (o.kind == RD_SPHERE) {
      p.a[0] = o.center[0]; p.a[1] = o.center[1]; p.a[2] = o.center[2];
      // dot(C, C) and r*r in the reference's FP32 order (cpu.rs:42-43)
      p.a[3] = o.center[0] * o.center[0] + o.center[1] * o.center[1] + o.center[2] * o.center[2];
      p.b[0] = o.size * o.size;
      p.b[3] = bitsf(id);
    } else {
      // cpu.rs:65-71: half = side * 0.5, min = C - half, max = C + half
      const float h = o.size * 0.5f;
      for (int a = 0; a < 3; ++a) {
        p.a[a] = o.center[a] - h;
        p.b[a] = o.center[a] + h;
      }
      p.b[3] = bitsf(id | kTagCube);
    }
  }
  return true;
}

bool check_bvh(const BuiltScene &b, std::string *err) {
  const size_t n = b.prims.size();
  if (n == 0) return true;
  if (b.depth + 1 > (uint32_t)kStackDepth) {
    *err = "BVH deeper than the traversal stack";
    return false;
  }
  std::vector<int> seen(n, 0);
  for (size_t i = 0; i < b.nodes.size(); ++i) {
    const DevNode &nd = b.nodes[i];
    for (int k = 0; k < 2; ++k) {
      const int32_t c = nd.child[k];
      if (is_leaf(c)) {
        const uint32_t f = leaf_first(c), cnt = leaf_count(c);
        if (f + cnt > n || cnt > (uint32_t)kMaxLeafSize) {
          *err = "leaf out of range";
          return false;
        }
        for (uint32_t j = f; j < f + cnt; ++j) seen[j]++;
      } else if (c <= (int32_t)i || (size_t)c >= b.nodes.size()) {
        *err = "child index not after its parent";
        return false;
      }
    }
  }
  for (size_t j = 0; j < n; ++j)
    if (seen[j] != 1) {
      *err = "primitive " + std::to_string(j) + " referenced " + std::to_string(seen[j]) + " times";
      return false;
    }
  return true;
}

int32_t host_trace(const BuiltScene &b, const float o[3], const float d[3], float *t_out) {
  const RayPre r = ray_pre(f3(o[0], o[1], o[2]), f3(d[0], d[1], d[2]));
  float bt = 0.f;
  uint32_t bid = kNoHit;
  if (!walkable(r) || b.nodes.empty()) {
    for (const DevPrim &p : b.prims) test_prim(r, p, &bt, &bid);
  } else {
    const SlabRay s = slab_ray(r, ray_tolerance(r, b.r0));
    int32_t stack[kStackDepth];
    int sp = 0;
    int32_t node = 0;
    for (;;) {
      const DevNode &nd = b.nodes[node];
      const float tc = cull_limit(bt, bid);
      float t0, t1;
      bool h0 = slab_hit(s, nd.lo_x[0], nd.hi_x[0], nd.lo_y[0], nd.hi_y[0], nd.lo_z[0], nd.hi_z[0], tc, &t0);
      bool h1 = slab_hit(s, nd.lo_x[1], nd.hi_x[1], nd.lo_y[1], nd.hi_y[1], nd.lo_z[1], nd.hi_z[1], tc, &t1);
      const int32_t c0 = nd.child[0], c1 = nd.child[1];
      if (h0 && is_leaf(c0)) {
        for (uint32_t j = leaf_first(c0), e = j + leaf_count(c0); j < e; ++j) test_prim(r, b.prims[j], &bt, &bid);
        h0 = false;
      }
      if (h1 && is_leaf(c1)) {
        for (uint32_t j = leaf_first(c1), e = j + leaf_count(c1); j < e; ++j) test_prim(r, b.prims[j], &bt, &bid);
        h1 = false;
      }
      if (h0 && h1) {
        const bool first0 = t0 <= t1;
        stack[sp++] = first0 ? c1 : c0;
        node = first0 ? c0 : c1;
      } else if (h0) {
        node = c0;
      } else if (h1) {
        node = c1;
      } else {
        if (sp == 0) break;
        node = stack[--sp];
      }
    }
  }
  *t_out = bid == kNoHit ? 0.f : bt;
  return bid == kNoHit ? -1 : (int32_t)bid;
}

}  // namespace rdh

namespace rdh {

bool check_bvh4(const BuiltScene &b, std::string *err) {
  const size_t n = b.prims.size();
  if (n == 0) return b.nodes4.empty() || true;
  std::vector<int> seen(n, 0);
  for (size_t i = 0; i < b.nodes4.size(); ++i) {
    const DevNode4 &nd = b.nodes4[i];
    for (int k = 0; k < 4; ++k) {
      const int32_t c = nd.child[k];
      if (is_leaf(c)) {
        const uint32_t f = leaf_first(c), cnt = leaf_count(c);
        if (f + cnt > n || cnt > (uint32_t)kMaxLeafSize) {
          *err = "4-wide leaf out of range";
          return false;
        }
        for (uint32_t j = f; j < f + cnt; ++j) seen[j]++;
      } else if (c <= (int32_t)i || (size_t)c >= b.nodes4.size()) {
        *err = "4-wide child index not after its parent";
        return false;
      }
    }
  }
  for (size_t j = 0; j < n; ++j)
    if (seen[j] != 1) {
      *err = "primitive " + std::to_string(j) + " referenced " + std::to_string(seen[j]) +
             " times in the 4-wide tree";
      return false;
    }
  return true;
}

// One step of the 4-wide traversal exactly as the extend kernel takes it:
// the entered inner children are sorted by entry distance with a
// five-comparator network, the three farther candidates are written above
// the stack top every step (kept as many as were entered, minus one) and the
// top entry is read every step.  Rows needed: stack4 + 3.
int32_t host_trace4(const BuiltScene &b, const float o[3], const float d[3], float *t_out) {
  const RayPre r = ray_pre(f3(o[0], o[1], o[2]), f3(d[0], d[1], d[2]));
  float bt = 0.f;
  uint32_t bid = kNoHit;
  if (!walkable(r) || b.nodes4.empty()) {
    for (const DevPrim &p : b.prims) test_prim(r, p, &bt, &bid);
  } else {
    const SlabRay s = slab_ray(r, ray_tolerance(r, b.r0));
    const uint32_t rows = b.stack4 + 3;
    std::vector<int32_t> stk(rows, 0);
    uint32_t sp = 0;
    int32_t node = 0;
    float tcull = INFINITY;
    for (;;) {
      const DevNode4 &nd = b.nodes4[node];
      float t[4];
      bool h[4], l[4], in[4];
      uint32_t cnt_leaf[4], n_leaf = 0;
      for (int c = 0; c < 4; ++c) {
        h[c] = slab_hit(s, nd.lo_x[c], nd.hi_x[c], nd.lo_y[c], nd.hi_y[c], nd.lo_z[c], nd.hi_z[c],
                        tcull, &t[c]);
        l[c] = h[c] && is_leaf(nd.child[c]);
        in[c] = h[c] && !l[c];
        cnt_leaf[c] = l[c] ? leaf_count(nd.child[c]) : 0u;
        n_leaf += cnt_leaf[c];
      }
      if (n_leaf) {
        for (int c = 0; c < 4; ++c)
          for (uint32_t j = 0; j < cnt_leaf[c]; ++j)
            test_prim(r, b.prims[leaf_first(nd.child[c]) + j], &bt, &bid);
        tcull = cull_limit(bt, bid);
      }
      float key[4];
      int32_t code[4];
      for (int c = 0; c < 4; ++c) {
        key[c] = in[c] ? t[c] : INFINITY;
        code[c] = nd.child[c];
      }
      auto cmpx = [&](int a, int c) {
        if (key[c] < key[a]) {
          std::swap(key[a], key[c]);
          std::swap(code[a], code[c]);
        }
      };
      cmpx(0, 1); cmpx(2, 3); cmpx(0, 2); cmpx(1, 3); cmpx(1, 2);
      const uint32_t cnt = (uint32_t)in[0] + in[1] + in[2] + in[3];
      if (sp + 2 >= rows) {
        *t_out = 0.f;
        return -2;  // the stack bound is wrong
      }
      stk[sp] = cnt == 4 ? code[3] : cnt == 3 ? code[2] : code[1];
      stk[sp + 1] = cnt == 4 ? code[2] : code[1];
      stk[sp + 2] = code[1];
      const uint32_t top = sp > 0 ? sp - 1 : 0;
      const int32_t popped = stk[top];
      if (cnt > 0) {
        node = code[0];
        sp += cnt - 1;
      } else if (sp > 0) {
        node = popped;
        sp = top;
      } else {
        break;
      }
    }
  }
  *t_out = bid == kNoHit ? 0.f : bt;
  return bid == kNoHit ? -1 : (int32_t)bid;
}

}  // namespace rdh
