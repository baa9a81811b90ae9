// raydar_hip — headless renderer on the HIP backend, the counterpart of
// raydar's `raydar` binary (src/main.rs:10-24) with the flags of CommonArgs /
// RaydarArgs (src/cli/mod.rs:10-27, 60-69):
//
//   raydar_hip [--max-sample-count N] [--max-bounces N] [-o|--output PATH] [SCENE.rscn]
//
// Without a scene file it renders Scene::default() (src/scene/mod.rs:20-67).
// Prints the renderer info (main.rs:26-59) and the four profiler timers
// (main.rs:61-107), and writes the RGBA8 image as PNG.
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/raydar_hip.h"

namespace {

// 8-bit RGBA PNG (zlib stream in one IDAT chunk).
bool write_png(const char *path, const uint8_t *rgba, uint32_t w, uint32_t h) {
  std::vector<uint8_t> raw;
  raw.reserve((size_t)h * (4ull * w + 1));
  for (uint32_t y = 0; y < h; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), rgba + 4ull * w * y, rgba + 4ull * w * (y + 1));
  }
  uLongf zlen = compressBound(raw.size());
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), raw.size(), 6) != Z_OK) return false;
  z.resize(zlen);
  FILE *f = std::fopen(path, "wb");
  if (!f) return false;
  auto be32 = [](uint32_t v, uint8_t *p) {
    p[0] = v >> 24; p[1] = v >> 16; p[2] = v >> 8; p[3] = v;
  };
  auto chunk = [&](const char *type, const uint8_t *data, uint32_t len) {
    uint8_t hdr[8];
    be32(len, hdr);
    std::memcpy(hdr + 4, type, 4);
    std::fwrite(hdr, 1, 8, f);
    if (len) std::fwrite(data, 1, len, f);
    uLong crc = crc32(0, hdr + 4, 4);
    if (len) crc = crc32(crc, data, len);
    uint8_t c[4];
    be32((uint32_t)crc, c);
    std::fwrite(c, 1, 4, f);
  };
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::fwrite(sig, 1, 8, f);
  uint8_t ihdr[13];
  be32(w, ihdr);
  be32(h, ihdr + 4);
  ihdr[8] = 8;   // bit depth
  ihdr[9] = 6;   // RGBA
  ihdr[10] = ihdr[11] = ihdr[12] = 0;
  chunk("IHDR", ihdr, 13);
  chunk("IDAT", z.data(), (uint32_t)z.size());
  chunk("IEND", nullptr, 0);
  return std::fclose(f) == 0;
}

// Scene::default() (src/scene/mod.rs:20-67).  Camera::new ends in
// update_matrices (camera.rs:33-60), restated in FP32 by
// rd_camera_update_matrices: the matrices are the ones cgmath stores in
// scenes/default.rscn, bit for bit.
bool default_scene(rd_scene *s, std::vector<rd_object> *objs) {
  std::memset(s, 0, sizeof(*s));
  rd_camera &c = s->camera;
  const float pos[3] = {-3.09f, 0.03f, -1.16f};
  std::memcpy(c.position, pos, sizeof(pos));
  c.up[1] = 1.0f;
  c.resolution_x = 854;
  c.resolution_y = 480;
  c.projection = RD_PROJECTION_PERSPECTIVE;
  c.projection_param = 90.0f;
  c.near_clip = 0.01f;
  c.far_clip = 1000.0f;
  if (rd_camera_update_matrices(&c) != RD_OK) return false;
  s->world.kind = RD_WORLD_SKY_COLOR;
  const float top[3] = {0.53f, 0.8f, 0.92f}, bottom[3] = {1.0f, 1.0f, 1.0f};
  std::memcpy(s->world.top_color, top, sizeof(top));
  std::memcpy(s->world.bottom_color, bottom, sizeof(bottom));
  auto obj = [](uint32_t kind, float x, float y, float z, float size) {
    rd_object o;
    std::memset(&o, 0, sizeof(o));
    o.kind = kind;
    o.center[0] = x; o.center[1] = y; o.center[2] = z;
    o.size = size;
    o.albedo[0] = o.albedo[1] = o.albedo[2] = 0.8f;  // Material::default()
    o.roughness = 0.5f;
    o.ior = 1.5f;
    return o;
  };
  objs->clear();
  rd_object glass = obj(RD_SPHERE, 0.0f, 0.001f, 0.0f, 1.0f);
  glass.albedo[0] = glass.albedo[1] = glass.albedo[2] = 1.0f;
  glass.roughness = 0.2f;
  glass.transmission = 1.0f;
  rd_object floor = obj(RD_CUBE, 0.0f, -101.0f, 0.0f, 200.0f);
  floor.albedo[0] = floor.albedo[1] = 0.34f;
  floor.albedo[2] = 0.44f;
  rd_object lamp = obj(RD_CUBE, 7.0f, 3.0f, 0.0f, 1.8f);
  lamp.emission_color[0] = 0.8f; lamp.emission_color[1] = 0.5f; lamp.emission_color[2] = 0.2f;
  lamp.emission_strength = 30.0f;
  objs->push_back(glass);
  objs->push_back(floor);
  objs->push_back(lamp);
  s->object_count = (uint32_t)objs->size();
  s->objects = objs->data();
  return true;
}

int usage(const char *argv0) {
  std::fprintf(stderr,
               "usage: %s [--max-sample-count N] [--max-bounces N] [-o|--output PATH] "
               "[--write-scene PATH] [SCENE_FILE]\n"
               "  --write-scene PATH  write the scene that would be rendered as .rscn and exit\n",
               argv0);
  return 2;
}

bool parse_u32(const char *s, uint32_t *out) {
  char *end = nullptr;
  unsigned long v = std::strtoul(s, &end, 10);
  if (!s[0] || *end || v > 0xffffffffUL) return false;
  *out = (uint32_t)v;
  return true;
}

void print_ms(const char *label, int64_t ns, bool *ok, const char *missing) {
  if (ns < 0) {
    std::fprintf(stderr, "Error: %s\n", missing);
    *ok = false;
    return;
  }
  std::printf("%s %lldms\n", label, (long long)(ns / 1000000));
}

}  // namespace

int main(int argc, char **argv) {
  rd_config cfg;
  rd_config_default(&cfg);
  std::string output = "output.png";
  const char *scene_path = nullptr, *write_scene = nullptr;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if ((a == "--max-sample-count" || a == "--max-bounces" || a == "-o" || a == "--output" ||
         a == "--write-scene") &&
        i + 1 >= argc)
      return usage(argv[0]);
    if (a == "--write-scene") {
      write_scene = argv[++i];
    } else if (a == "--max-sample-count") {
      if (!parse_u32(argv[++i], &cfg.max_sample_count)) return usage(argv[0]);
    } else if (a == "--max-bounces") {
      if (!parse_u32(argv[++i], &cfg.max_bounces)) return usage(argv[0]);
    } else if (a == "-o" || a == "--output") {
      output = argv[++i];
    } else if (a == "-h" || a == "--help") {
      return usage(argv[0]), 0;
    } else if (!a.empty() && a[0] == '-') {
      return usage(argv[0]);
    } else if (!scene_path) {
      scene_path = argv[i];
    } else {
      return usage(argv[0]);
    }
  }
  rd_scene *loaded = nullptr;
  rd_scene def;
  std::vector<rd_object> def_objs;
  const rd_scene *scene = nullptr;
  if (scene_path) {
    if (rd_scene_load_rscn(scene_path, &loaded) != RD_OK) {
      std::fprintf(stderr, "Error: %s\n", rd_last_error(nullptr));
      return 1;
    }
    scene = loaded;
  } else {
    if (!default_scene(&def, &def_objs)) {
      std::fprintf(stderr, "Error: default camera matrices are singular\n");
      return 1;
    }
    scene = &def;
  }
  if (write_scene) {
    const rd_status st = rd_scene_save_rscn(scene, write_scene);
    if (st != RD_OK) std::fprintf(stderr, "Error: %s\n", rd_last_error(nullptr));
    rd_scene_free(loaded);
    return st == RD_OK ? 0 : 1;
  }
  rd_renderer *r = nullptr;
  if (rd_create(&cfg, &r) != RD_OK) {
    std::fprintf(stderr, "Error: %s\n", rd_last_error(nullptr));
    rd_scene_free(loaded);
    return 1;
  }
  std::printf("=== Raydar (HIP) ===\n");
  std::printf("Renderer: HIP (MI355X, gfx950)\n");
  std::printf("Max Samples: %u\n", rd_max_sample_count(r));
  std::printf("Max Bounces: %u\n", rd_max_bounces(r));
  std::printf("Resolution: %ux%u\n", scene->camera.resolution_x, scene->camera.resolution_y);
  std::printf("Objects: %u\n", scene->object_count);
  const uint32_t W = scene->camera.resolution_x, H = scene->camera.resolution_y;
  std::vector<uint8_t> img((size_t)W * H * 4);
  int rc = 0;
  if (rd_render_frame(r, scene, img.data(), img.size()) != RD_OK) {
    std::fprintf(stderr, "Error: %s\n", rd_last_error(r));
    rc = 1;
  } else if (!write_png(output.c_str(), img.data(), W, H)) {
    std::fprintf(stderr, "Error: Cannot save image\n");
    rc = 1;
  } else {
    rd_timings t;
    rd_profiler(r, &t);
    std::printf("\n=== Render Profiling Metrics ===\n");
    bool ok = true;
    print_ms("Scene Preparation:", t.prepare_ns, &ok, "Prepare timer not started");
    if (ok) print_ms("Render Time:", t.render_ns, &ok, "Render timer not started");
    if (ok) print_ms("Last Sample Time:", t.sample_ns, &ok, "Sample timer not started");
    if (ok) {
      std::printf("----------------------------------------\n");
      print_ms("Total Frame Time:", t.frame_ns, &ok, "Frame timer not started");
    }
    if (!ok) rc = 1;
  }
  rd_destroy(r);
  rd_scene_free(loaded);
  return rc;
}
