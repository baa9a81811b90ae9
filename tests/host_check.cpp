// Host-side consistency and memory-safety driver, built with AddressSanitizer
// and UndefinedBehaviorSanitizer by test_host_sanitizers.py (no GPU).
//
// Covers the code that digests untrusted input on the host: the .rscn reader
// (valid, truncated and corrupted files), the BVH builder with its 4-wide
// collapse (random, clustered and degenerate scenes at every leaf size), the
// host traversals of both trees against a linear scan over every primitive
// (the reference's closest-hit rule, src/renderer/cpu.rs:344-352), and the
// radiance-ring allocator of the streaming engine.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../raydar_amd/csrc/bvh.hpp"
#include "../raydar_amd/csrc/pathmath.hpp"
#include "../raydar_amd/csrc/ring.hpp"

extern "C" rd_status rd_scene_parse_rscn(const char *json, size_t len, rd_scene **out);
extern "C" void rd_scene_free(rd_scene *scene);

namespace rdh {
void set_global_error(const std::string &) {}
}  // namespace rdh

using namespace rdh;

static int failures = 0;

static void fail(const std::string &what) {
  if (failures++ < 20) std::fprintf(stderr, "FAIL: %s\n", what.c_str());
}

// Closest hit over every primitive (the linear scan), same culling rule.
static int32_t scan(const BuiltScene &b, const float o[3], const float d[3], float *t) {
  const RayPre r = ray_pre(f3(o[0], o[1], o[2]), f3(d[0], d[1], d[2]));
  float bt = 0.f;
  uint32_t bid = kNoHit;
  for (const DevPrim &p : b.prims) test_prim(r, p, &bt, &bid);
  *t = bid == kNoHit ? 0.f : bt;
  return bid == kNoHit ? -1 : (int32_t)bid;
}

static void check_scene(const rd_scene &sc, const std::string &name, std::mt19937 &g, int rays) {
  std::uniform_real_distribution<float> u(0.f, 1.f);
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  for (uint32_t i = 0; i < sc.object_count; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], sc.objects[i].center[a]);
      hi[a] = std::max(hi[a], sc.objects[i].center[a]);
    }
  for (int leaf : {1, 2, 4, 8}) {
    BuiltScene b;
    std::string err;
    if (!build_scene(sc, leaf, &b, &err)) {
      fail(name + ": build failed: " + err);
      continue;
    }
    if (!check_bvh(b, &err)) fail(name + " leaf " + std::to_string(leaf) + ": " + err);
    if (!check_bvh4(b, &err)) fail(name + " leaf " + std::to_string(leaf) + " (4-wide): " + err);
    for (int k = 0; k < rays; ++k) {
      float o[3], d[3];
      for (int a = 0; a < 3; ++a) {
        const float span = std::max(hi[a] - lo[a], 1.0f);
        o[a] = lo[a] + span * (1.6f * u(g) - 0.3f);
        d[a] = 2.f * u(g) - 1.f;
      }
      if (k % 7 == 0) d[k % 3] = 0.f;                   // axis-parallel
      if (k % 11 == 0) d[0] = d[1] = 0.f;               // along one axis
      if (k % 13 == 0 && sc.object_count) {             // aimed at an object centre
        const rd_object &ob = sc.objects[(k / 13) % sc.object_count];
        for (int a = 0; a < 3; ++a) d[a] = ob.center[a] - o[a];
      }
      if (k % 17 == 0) d[0] = d[1] = d[2] = 0.f;        // no direction at all
      if (k % 19 == 0) d[k % 3] = 1e-33f;               // one tiny component
      if (k % 23 == 0)                                  // all tiny
        for (int a = 0; a < 3; ++a) d[a] = (2.f * u(g) - 1.f) * 1e-35f;
      float ts, t2, t4;
      const int32_t is = scan(b, o, d, &ts);
      const int32_t i2 = host_trace(b, o, d, &t2);
      const int32_t i4 = host_trace4(b, o, d, &t4);
      if (i2 != is || (is >= 0 && std::memcmp(&t2, &ts, 4) != 0))
        fail(name + " leaf " + std::to_string(leaf) + ": binary walk differs from the scan");
      if (i4 != is || (is >= 0 && std::memcmp(&t4, &ts, 4) != 0))
        fail(name + " leaf " + std::to_string(leaf) + ": 4-wide walk differs from the scan (" +
             std::to_string(i4) + " vs " + std::to_string(is) + ")");
    }
  }
}

static rd_scene make_scene(std::vector<rd_object> &objs) {
  rd_scene s;
  std::memset(&s, 0, sizeof(s));
  s.camera.resolution_x = 8;
  s.camera.resolution_y = 8;
  s.object_count = (uint32_t)objs.size();
  s.objects = objs.empty() ? nullptr : objs.data();
  return s;
}

static rd_object object(uint32_t kind, float x, float y, float z, float size) {
  rd_object o;
  std::memset(&o, 0, sizeof(o));
  o.kind = kind;
  o.center[0] = x;
  o.center[1] = y;
  o.center[2] = z;
  o.size = size;
  o.ior = 1.5f;
  return o;
}

static std::string read_file(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

int main(int argc, char **argv) {
  std::mt19937 g(12345);
  std::uniform_real_distribution<float> u(0.f, 1.f);

  // scene files: parse, then the BVH checks on what they hold
  for (int i = 1; i < argc; ++i) {
    const std::string text = read_file(argv[i]);
    if (text.empty()) {
      fail(std::string("cannot read ") + argv[i]);
      continue;
    }
    rd_scene *sc = nullptr;
    if (rd_scene_parse_rscn(text.data(), text.size(), &sc) != RD_OK || !sc) {
      fail(std::string("cannot parse ") + argv[i]);
      continue;
    }
    check_scene(*sc, argv[i], g, 300);
    rd_scene_free(sc);
    // truncated files: every prefix length on a stride, then single bytes
    // replaced by characters that break the syntax or the numbers
    const size_t stride = std::max<size_t>(1, text.size() / 400);
    for (size_t len = 0; len < text.size(); len += stride) {
      rd_scene *t = nullptr;
      if (rd_scene_parse_rscn(text.data(), len, &t) == RD_OK && t) rd_scene_free(t);
    }
    const char junk[] = "{}[],:\"-+.eE0x9 \\\n\x01\xff";
    for (int k = 0; k < 400; ++k) {
      std::string c = text;
      c[g() % c.size()] = junk[g() % (sizeof(junk) - 1)];
      rd_scene *t = nullptr;
      if (rd_scene_parse_rscn(c.data(), c.size(), &t) == RD_OK && t) rd_scene_free(t);
    }
  }

  // random, clustered and degenerate scenes
  for (int n : {0, 1, 2, 3, 17, 100, 400}) {
    std::vector<rd_object> objs;
    for (int i = 0; i < n; ++i)
      objs.push_back(object(g() % 2, 40.f * u(g) - 20.f, 40.f * u(g) - 20.f, 40.f * u(g) - 20.f,
                            0.01f + 3.f * u(g)));
    rd_scene s = make_scene(objs);
    check_scene(s, "random " + std::to_string(n), g, 400);
  }
  {
    std::vector<rd_object> objs;
    std::normal_distribution<float> nd(0.f, 1.5f);
    for (int c = 0; c < 8; ++c) {
      const float cx = 50.f * u(g), cy = 50.f * u(g), cz = 20.f * u(g);
      for (int i = 0; i < 60; ++i)
        objs.push_back(object(g() % 4 == 0, cx + nd(g), cy + nd(g), cz + nd(g), 0.05f + u(g)));
    }
    rd_scene s = make_scene(objs);
    check_scene(s, "clustered", g, 400);
  }
  {
    std::vector<rd_object> objs(200, object(0, 1.f, 2.f, 3.f, 0.5f));  // coincident
    objs.push_back(object(1, 0.f, 0.f, 0.f, 0.f));                     // zero-size cube
    objs.push_back(object(0, 1e4f, -1e4f, 5e3f, 1e3f));                // far and large
    rd_scene s = make_scene(objs);
    check_scene(s, "degenerate", g, 400);
  }
  {
    // centres 1e-39 apart: the binning scale overflows on that axis
    std::vector<rd_object> objs;
    for (int i = 0; i < 20; ++i) objs.push_back(object(i % 2, (float)i * 1e-39f, 0.f, 0.f, 0.5f));
    for (int i = 0; i < 20; ++i) objs.push_back(object(0, 3.f, (float)i * 1e-38f, 1.f, 0.25f));
    rd_scene s = make_scene(objs);
    check_scene(s, "denormal spacing", g, 400);
  }

  // radiance ring: live ranges never overlap and stay inside the capacity
  for (uint64_t cap : {7ull, 64ull, 1000ull}) {
    RadianceRing ring(cap);
    std::vector<std::pair<uint64_t, uint64_t>> live;  // (offset, size) by ticket
    std::vector<bool> alive;
    for (int k = 0; k < 5000; ++k) {
      const bool any = std::find(alive.begin(), alive.end(), true) != alive.end();
      if (any && u(g) < 0.45f) {
        uint32_t t;
        do t = g() % alive.size(); while (!alive[t]);
        alive[t] = false;
        ring.release(t);
      } else {
        const uint64_t n = 1 + g() % (cap / 2 + 1);
        uint64_t off = 0;
        const uint32_t t = (uint32_t)live.size();
        const bool ok = ring.alloc(n, t, &off);
        live.push_back({off, n});
        alive.push_back(ok);
        if (!ok) {
          if (!any && n <= cap) fail("ring refused a reservation while empty");
          continue;
        }
        if (off + n > cap) fail("ring range outside the capacity");
        for (uint32_t x = 0; x < t; ++x)
          if (alive[x] && !(off + n <= live[x].first || live[x].first + live[x].second <= off))
            fail("ring ranges overlap");
      }
    }
  }

  if (failures) {
    std::fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  std::printf("host checks ok\n");
  return 0;
}
