// .rscn reader / writer: the serde_json layout of raydar's Scene
// (src/scene/mod.rs:13-18, camera.rs:13-30, objects.rs:6-50, material.rs:4-13,
// world.rs:6-14), as loaded by CommonArgs::initialize (src/cli/mod.rs:32-37)
// and saved by the editor (src/bin/raydar_editor.rs:76-93).
//
// Numbers are parsed as f64 (strtod) and rounded to f32, as serde_json does
// for f32 fields.  The writer prints the shortest f32 text that round-trips.
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/raydar_hip.h"

namespace rdh {

void set_global_error(const std::string &msg);  // api.hip

namespace {

struct JVal {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  double num = 0.0;
  bool b = false;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal *get(const char *k) const {
    if (kind != Obj) return nullptr;
    for (auto &kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct Parser {
  const char *p, *end;
  std::string err;
  int depth = 0;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
  }
  bool fail(const char *m) {
    if (err.empty()) err = m;
    return false;
  }
  bool lit(const char *s) {
    size_t n = std::strlen(s);
    if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return false;
    p += n;
    return true;
  }
  bool string(std::string *out) {
    if (p >= end || *p != '"') return fail("expected string");
    ++p;
    while (p < end && *p != '"') {
      if (*p == '\\') {
        ++p;
        if (p >= end) return fail("bad escape");
        switch (*p) {
          case '"': out->push_back('"'); break;
          case '\\': out->push_back('\\'); break;
          case '/': out->push_back('/'); break;
          case 'b': out->push_back('\b'); break;
          case 'f': out->push_back('\f'); break;
          case 'n': out->push_back('\n'); break;
          case 'r': out->push_back('\r'); break;
          case 't': out->push_back('\t'); break;
          case 'u': {
            if (end - p < 5) return fail("bad \\u escape");
            unsigned v = 0;
            for (int k = 1; k <= 4; ++k) {
              char c = p[k];
              v <<= 4;
              if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
              else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
              else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
              else return fail("bad \\u escape");
            }
            if (v < 0x80) out->push_back((char)v);
            else out->push_back('?');
            p += 4;
            break;
          }
          default: return fail("bad escape");
        }
        ++p;
      } else {
        out->push_back(*p++);
      }
    }
    if (p >= end) return fail("unterminated string");
    ++p;
    return true;
  }
  bool value(JVal *v) {
    if (++depth > 64) return fail("nesting too deep");
    ws();
    if (p >= end) return fail("unexpected end of input");
    bool ok = true;
    if (*p == '{') {
      ++p;
      v->kind = JVal::Obj;
      ws();
      if (p < end && *p == '}') {
        ++p;
      } else {
        for (;;) {
          ws();
          std::string k;
          if (!string(&k)) { ok = false; break; }
          ws();
          if (p >= end || *p != ':') { ok = fail("expected ':'"); break; }
          ++p;
          v->obj.emplace_back(std::move(k), JVal());
          if (!value(&v->obj.back().second)) { ok = false; break; }
          ws();
          if (p < end && *p == ',') { ++p; continue; }
          if (p < end && *p == '}') { ++p; break; }
          ok = fail("expected ',' or '}'");
          break;
        }
      }
    } else if (*p == '[') {
      ++p;
      v->kind = JVal::Arr;
      ws();
      if (p < end && *p == ']') {
        ++p;
      } else {
        for (;;) {
          v->arr.emplace_back();
          if (!value(&v->arr.back())) { ok = false; break; }
          ws();
          if (p < end && *p == ',') { ++p; continue; }
          if (p < end && *p == ']') { ++p; break; }
          ok = fail("expected ',' or ']'");
          break;
        }
      }
    } else if (*p == '"') {
      v->kind = JVal::Str;
      ok = string(&v->str);
    } else if (lit("true")) {
      v->kind = JVal::Bool;
      v->b = true;
    } else if (lit("false")) {
      v->kind = JVal::Bool;
    } else if (lit("null")) {
      v->kind = JVal::Null;
    } else {
      // JSON number
      const char *s = p;
      if (p < end && *p == '-') ++p;
      while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' ||
                         *p == '+' || *p == '-'))
        ++p;
      if (p == s) return fail("unexpected character");
      std::string tok(s, p);
      char *e = nullptr;
      v->num = std::strtod(tok.c_str(), &e);
      if (e == nullptr || *e != '\0') return fail("malformed number");
      v->kind = JVal::Num;
    }
    --depth;
    return ok;
  }
};

struct Extract {
  std::string err;
  bool num(const JVal *v, const char *what, float *out) {
    if (!v || v->kind != JVal::Num) {
      if (err.empty()) err = std::string("missing or non-numeric field '") + what + "'";
      return false;
    }
    *out = (float)v->num;
    return true;
  }
  bool u32(const JVal *v, const char *what, uint32_t *out) {
    if (!v || v->kind != JVal::Num || v->num < 0 || v->num > 4294967295.0 ||
        v->num != std::floor(v->num)) {
      if (err.empty()) err = std::string("field '") + what + "' must be a u32";
      return false;
    }
    *out = (uint32_t)v->num;
    return true;
  }
  bool vec3(const JVal *v, const char *what, float *out) {
    if (!v || v->kind != JVal::Obj) {
      if (err.empty()) err = std::string("missing vector '") + what + "'";
      return false;
    }
    return num(v->get("x"), what, out) && num(v->get("y"), what, out + 1) &&
           num(v->get("z"), what, out + 2);
  }
  bool mat4(const JVal *v, const char *what, float *out) {
    if (!v || v->kind != JVal::Obj) {
      if (err.empty()) err = std::string("missing matrix '") + what + "'";
      return false;
    }
    const char *cols[4] = {"x", "y", "z", "w"};
    for (int c = 0; c < 4; ++c) {
      const JVal *col = v->get(cols[c]);
      if (!col || col->kind != JVal::Obj) {
        if (err.empty()) err = std::string("malformed matrix '") + what + "'";
        return false;
      }
      for (int r = 0; r < 4; ++r)
        if (!num(col->get(cols[r]), what, out + 4 * c + r)) return false;
    }
    return true;
  }
};

struct SceneBlob {
  rd_scene scene;
};

rd_status parse_scene(const char *json, size_t len, rd_scene **out) {
  Parser ps{json, json + len, std::string(), 0};
  JVal root;
  if (!ps.value(&root)) {
    set_global_error("cannot parse scene file: " + ps.err);
    return RD_ERR_PARSE;
  }
  ps.ws();
  if (ps.p != ps.end) {
    set_global_error("cannot parse scene file: trailing characters");
    return RD_ERR_PARSE;
  }
  Extract ex;
  rd_camera cam;
  std::memset(&cam, 0, sizeof(cam));
  rd_world world;
  std::memset(&world, 0, sizeof(world));
  const JVal *c = root.get("camera");
  if (!c || c->kind != JVal::Obj) {
    set_global_error("cannot parse scene file: missing 'camera'");
    return RD_ERR_PARSE;
  }
  bool ok = ex.vec3(c->get("position"), "position", cam.position) &&
            ex.vec3(c->get("target"), "target", cam.target) &&
            ex.vec3(c->get("up"), "up", cam.up) &&
            ex.u32(c->get("resolution_x"), "resolution_x", &cam.resolution_x) &&
            ex.u32(c->get("resolution_y"), "resolution_y", &cam.resolution_y) &&
            ex.num(c->get("near_clip"), "near_clip", &cam.near_clip) &&
            ex.num(c->get("far_clip"), "far_clip", &cam.far_clip) &&
            ex.mat4(c->get("view_matrix"), "view_matrix", cam.view_matrix) &&
            ex.mat4(c->get("proj_matrix"), "proj_matrix", cam.proj_matrix) &&
            ex.mat4(c->get("inverse_view_matrix"), "inverse_view_matrix",
                    cam.inverse_view_matrix) &&
            ex.mat4(c->get("inverse_proj_matrix"), "inverse_proj_matrix",
                    cam.inverse_proj_matrix);
  if (ok) {
    const JVal *pr = c->get("projection");
    const JVal *pp = pr ? pr->get("Perspective") : nullptr;
    const JVal *po = pr ? pr->get("Orthographic") : nullptr;
    if (pp) {
      cam.projection = RD_PROJECTION_PERSPECTIVE;
      ok = ex.num(pp->get("fov"), "fov", &cam.projection_param);
    } else if (po) {
      cam.projection = RD_PROJECTION_ORTHOGRAPHIC;
      ok = ex.num(po->get("size"), "size", &cam.projection_param);
    } else {
      ok = false;
      ex.err = "unknown camera projection";
    }
  }
  const JVal *w = ok ? root.get("world") : nullptr;
  if (ok) {
    if (w && w->kind == JVal::Str && w->str == "Transparent") {
      world.kind = RD_WORLD_TRANSPARENT;
    } else if (w && w->get("SkyColor")) {
      world.kind = RD_WORLD_SKY_COLOR;
      const JVal *s = w->get("SkyColor");
      ok = ex.vec3(s->get("top_color"), "top_color", world.top_color) &&
           ex.vec3(s->get("bottom_color"), "bottom_color", world.bottom_color);
    } else if (w && w->get("SolidColor")) {
      world.kind = RD_WORLD_SOLID_COLOR;
      ok = ex.vec3(w->get("SolidColor"), "SolidColor", world.solid_color);
    } else {
      ok = false;
      ex.err = "missing or unknown 'world'";
    }
  }
  const JVal *objs = ok ? root.get("objects") : nullptr;
  if (ok && (!objs || objs->kind != JVal::Arr)) {
    ok = false;
    ex.err = "missing 'objects' array";
  }
  if (!ok) {
    set_global_error("cannot parse scene file: " + ex.err);
    return RD_ERR_PARSE;
  }
  const size_t n = objs->arr.size();
  void *mem = std::calloc(1, sizeof(SceneBlob) + sizeof(rd_object) * (n ? n : 1));
  if (!mem) return RD_ERR_OUT_OF_MEMORY;
  SceneBlob *blob = static_cast<SceneBlob *>(mem);
  rd_object *arr = reinterpret_cast<rd_object *>(static_cast<char *>(mem) + sizeof(SceneBlob));
  for (size_t i = 0; i < n && ok; ++i) {
    const JVal &o = objs->arr[i];
    rd_object &d = arr[i];
    const JVal *g = o.get("geometry");
    const JVal *m = o.get("material");
    if (!g || !m) {
      ok = false;
      ex.err = "object without geometry or material";
      break;
    }
    if (const JVal *s = g->get("Sphere")) {
      d.kind = RD_SPHERE;
      ok = ex.vec3(s->get("center"), "center", d.center) &&
           ex.num(s->get("radius"), "radius", &d.size);
    } else if (const JVal *cb = g->get("Cube")) {
      d.kind = RD_CUBE;
      ok = ex.vec3(cb->get("center"), "center", d.center) &&
           ex.num(cb->get("side_length"), "side_length", &d.size);
    } else {
      ok = false;
      ex.err = "unknown geometry";
    }
    ok = ok && ex.vec3(m->get("albedo"), "albedo", d.albedo) &&
         ex.num(m->get("roughness"), "roughness", &d.roughness) &&
         ex.num(m->get("metallic"), "metallic", &d.metallic) &&
         ex.vec3(m->get("emission_color"), "emission_color", d.emission_color) &&
         ex.num(m->get("emission_strength"), "emission_strength", &d.emission_strength) &&
         ex.num(m->get("transmission"), "transmission", &d.transmission) &&
         ex.num(m->get("ior"), "ior", &d.ior);
  }
  if (!ok) {
    std::free(mem);
    set_global_error("cannot parse scene file: " + ex.err);
    return RD_ERR_PARSE;
  }
  blob->scene.camera = cam;
  blob->scene.world = world;
  blob->scene.object_count = (uint32_t)n;
  blob->scene.objects = arr;
  *out = &blob->scene;
  return RD_OK;
}

std::string fmt(float v) {
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof(buf), v);
  std::string s(buf, r.ptr);
  if (s.find_first_of(".eE") == std::string::npos && s.find("inf") == std::string::npos &&
      s.find("nan") == std::string::npos)
    s += ".0";
  return s;
}

struct Writer {
  std::string out;
  void ind(int n) { out.append(2 * n, ' '); }
  void vec(int n, const char *name, const float *v, bool comma) {
    ind(n);
    out += "\"" + std::string(name) + "\": {\n";
    const char *k[3] = {"x", "y", "z"};
    for (int i = 0; i < 3; ++i) {
      ind(n + 1);
      out += "\"" + std::string(k[i]) + "\": " + fmt(v[i]) + (i < 2 ? ",\n" : "\n");
    }
    ind(n);
    out += comma ? "},\n" : "}\n";
  }
  void mat(int n, const char *name, const float *m, bool comma) {
    ind(n);
    out += "\"" + std::string(name) + "\": {\n";
    const char *k[4] = {"x", "y", "z", "w"};
    for (int c = 0; c < 4; ++c) {
      ind(n + 1);
      out += "\"" + std::string(k[c]) + "\": {\n";
      for (int r = 0; r < 4; ++r) {
        ind(n + 2);
        out += "\"" + std::string(k[r]) + "\": " + fmt(m[4 * c + r]) + (r < 3 ? ",\n" : "\n");
      }
      ind(n + 1);
      out += c < 3 ? "},\n" : "}\n";
    }
    ind(n);
    out += comma ? "},\n" : "}\n";
  }
  void num(int n, const char *name, float v, bool comma) {
    ind(n);
    out += "\"" + std::string(name) + "\": " + fmt(v) + (comma ? ",\n" : "\n");
  }
};

}  // namespace

}  // namespace rdh

extern "C" {

rd_status rd_scene_parse_rscn(const char *json, size_t len, rd_scene **out) {
  if (!json || !out) {
    rdh::set_global_error("rd_scene_parse_rscn: NULL argument");
    return RD_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  return rdh::parse_scene(json, len, out);
}

rd_status rd_scene_load_rscn(const char *path, rd_scene **out) {
  if (!path || !out) {
    rdh::set_global_error("rd_scene_load_rscn: NULL argument");
    return RD_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    rdh::set_global_error(std::string("cannot open scene file ") + path);
    return RD_ERR_IO;
  }
  std::string data;
  char buf[65536];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) data.append(buf, n);
  std::fclose(f);
  return rdh::parse_scene(data.data(), data.size(), out);
}

void rd_scene_free(rd_scene *scene) {
  // rd_scene is the first member of the blob allocated by parse_scene
  std::free(scene);
}

rd_status rd_scene_save_rscn(const rd_scene *s, const char *path) {
  if (!s || !path || (s->object_count && !s->objects)) {
    rdh::set_global_error("rd_scene_save_rscn: NULL argument");
    return RD_ERR_INVALID_ARGUMENT;
  }
  rdh::Writer w;
  const rd_camera &c = s->camera;
  w.out += "{\n";
  w.ind(1);
  w.out += "\"camera\": {\n";
  w.vec(2, "position", c.position, true);
  w.vec(2, "target", c.target, true);
  w.vec(2, "up", c.up, true);
  w.ind(2);
  w.out += "\"resolution_x\": " + std::to_string(c.resolution_x) + ",\n";
  w.ind(2);
  w.out += "\"resolution_y\": " + std::to_string(c.resolution_y) + ",\n";
  w.ind(2);
  w.out += "\"projection\": {\n";
  w.ind(3);
  if (c.projection == RD_PROJECTION_ORTHOGRAPHIC) {
    w.out += "\"Orthographic\": {\n";
    w.num(4, "size", c.projection_param, false);
  } else {
    w.out += "\"Perspective\": {\n";
    w.num(4, "fov", c.projection_param, false);
  }
  w.ind(3);
  w.out += "}\n";
  w.ind(2);
  w.out += "},\n";
  w.num(2, "near_clip", c.near_clip, true);
  w.num(2, "far_clip", c.far_clip, true);
  w.mat(2, "view_matrix", c.view_matrix, true);
  w.mat(2, "proj_matrix", c.proj_matrix, true);
  w.mat(2, "inverse_view_matrix", c.inverse_view_matrix, true);
  w.mat(2, "inverse_proj_matrix", c.inverse_proj_matrix, false);
  w.ind(1);
  w.out += "},\n";
  w.ind(1);
  if (s->world.kind == RD_WORLD_TRANSPARENT) {
    w.out += "\"world\": \"Transparent\",\n";
  } else if (s->world.kind == RD_WORLD_SOLID_COLOR) {
    w.out += "\"world\": {\n";
    w.vec(2, "SolidColor", s->world.solid_color, false);
    w.ind(1);
    w.out += "},\n";
  } else {
    w.out += "\"world\": {\n";
    w.ind(2);
    w.out += "\"SkyColor\": {\n";
    w.vec(3, "top_color", s->world.top_color, true);
    w.vec(3, "bottom_color", s->world.bottom_color, false);
    w.ind(2);
    w.out += "}\n";
    w.ind(1);
    w.out += "},\n";
  }
  w.ind(1);
  w.out += s->object_count ? "\"objects\": [\n" : "\"objects\": []\n";
  for (uint32_t i = 0; i < s->object_count; ++i) {
    const rd_object &o = s->objects[i];
    w.ind(2);
    w.out += "{\n";
    w.ind(3);
    w.out += "\"geometry\": {\n";
    w.ind(4);
    w.out += o.kind == RD_CUBE ? "\"Cube\": {\n" : "\"Sphere\": {\n";
    w.vec(5, "center", o.center, true);
    w.num(5, o.kind == RD_CUBE ? "side_length" : "radius", o.size, false);
    w.ind(4);
    w.out += "}\n";
    w.ind(3);
    w.out += "},\n";
    w.ind(3);
    w.out += "\"material\": {\n";
    w.vec(4, "albedo", o.albedo, true);
    w.num(4, "roughness", o.roughness, true);
    w.num(4, "metallic", o.metallic, true);
    w.vec(4, "emission_color", o.emission_color, true);
    w.num(4, "emission_strength", o.emission_strength, true);
    w.num(4, "transmission", o.transmission, true);
    w.num(4, "ior", o.ior, false);
    w.ind(3);
    w.out += "}\n";
    w.ind(2);
    w.out += i + 1 < s->object_count ? "},\n" : "}\n";
  }
  if (s->object_count) {
    w.ind(1);
    w.out += "]\n";
  }
  w.out += "}";
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    rdh::set_global_error(std::string("cannot write scene file ") + path);
    return RD_ERR_IO;
  }
  const bool ok = std::fwrite(w.out.data(), 1, w.out.size(), f) == w.out.size();
  std::fclose(f);
  if (!ok) {
    rdh::set_global_error(std::string("cannot write scene file ") + path);
    return RD_ERR_IO;
  }
  return RD_OK;
}

}  // extern "C"
