"""raydar's scene model and its .rscn file format, on the Python side.

A `Scene` mirrors `struct Scene` (src/scene/mod.rs:13-18); the .rscn file is
the serde_json layout of that struct, read here with the json module (numbers
parsed as f64, then rounded to f32, as serde_json + f32 fields do,
src/cli/mod.rs:32-37).  `Scene.as_rd()` flattens it into the C ABI's
`rd_scene` (include/raydar_hip.h) without copying the object array.

Camera matrices are recomputed only when a scene is built from parameters
(`make_camera`), with cgmath 0.18's formulas as used by Camera::update_matrices
(src/scene/camera.rs:210-231); loaded scenes keep their stored matrices, as
the reference does (it never recomputes on load).
"""
import json
import math

import numpy as np

from . import abi

_F32 = np.float32


def _v3(d):
    return [float(_F32(d["x"])), float(_F32(d["y"])), float(_F32(d["z"]))]


def _m4(d):
    # column-major: {x: col0, y: col1, z: col2, w: col3}, each {x,y,z,w}
    out = []
    for c in "xyzw":
        col = d[c]
        out += [float(_F32(col[r])) for r in "xyzw"]
    return out


def _fmt(v):
    # shortest f32 representation, like serde_json's ryu output for f32
    s = str(_F32(v))
    if "e" not in s and "." not in s and "inf" not in s and "nan" not in s:
        s += ".0"
    return s


def _json_dump(obj, indent=0):
    pad = "  " * indent
    if isinstance(obj, dict):
        if not obj:
            return "{}"
        items = [f'{pad}  {json.dumps(k)}: {_json_dump(v, indent + 1)}' for k, v in obj.items()]
        return "{\n" + ",\n".join(items) + "\n" + pad + "}"
    if isinstance(obj, list):
        if not obj:
            return "[]"
        items = [f"{pad}  {_json_dump(v, indent + 1)}" for v in obj]
        return "[\n" + ",\n".join(items) + "\n" + pad + "]"
    if isinstance(obj, bool):
        return "true" if obj else "false"
    if isinstance(obj, (int, np.integer)):
        return str(int(obj))
    if isinstance(obj, (float, np.floating)):
        return _fmt(obj)
    return json.dumps(obj)


def _vec(v):
    return {"x": float(v[0]), "y": float(v[1]), "z": float(v[2])}


def _mat(m):
    return {c: {r: float(m[4 * i + j]) for j, r in enumerate("xyzw")} for i, c in enumerate("xyzw")}


class Scene:
    """Camera + world + objects (src/scene/mod.rs:13-18)."""

    def __init__(self, camera, world, objects):
        self.camera = camera  # dict with the rd_camera field names
        self.world = world  # dict: kind, top_color, bottom_color, solid_color
        self.objects = np.ascontiguousarray(objects, dtype=abi.OBJECT_DTYPE)

    # ------------------------------------------------------------ .rscn ----
    @classmethod
    def from_json(cls, text):
        d = json.loads(text)
        cam = d["camera"]
        proj = cam["projection"]
        if "Perspective" in proj:
            pk, pp = abi.RD_PROJECTION_PERSPECTIVE, float(_F32(proj["Perspective"]["fov"]))
        else:
            pk, pp = abi.RD_PROJECTION_ORTHOGRAPHIC, float(_F32(proj["Orthographic"]["size"]))
        camera = dict(
            position=_v3(cam["position"]), target=_v3(cam["target"]), up=_v3(cam["up"]),
            resolution_x=int(cam["resolution_x"]), resolution_y=int(cam["resolution_y"]),
            projection=pk, projection_param=pp,
            near_clip=float(_F32(cam["near_clip"])), far_clip=float(_F32(cam["far_clip"])),
            view_matrix=_m4(cam["view_matrix"]), proj_matrix=_m4(cam["proj_matrix"]),
            inverse_view_matrix=_m4(cam["inverse_view_matrix"]),
            inverse_proj_matrix=_m4(cam["inverse_proj_matrix"]),
        )
        w = d["world"]
        world = dict(kind=abi.RD_WORLD_TRANSPARENT, top_color=[0.0] * 3,
                     bottom_color=[0.0] * 3, solid_color=[0.0] * 3)
        if isinstance(w, dict) and "SkyColor" in w:
            world.update(kind=abi.RD_WORLD_SKY_COLOR, top_color=_v3(w["SkyColor"]["top_color"]),
                         bottom_color=_v3(w["SkyColor"]["bottom_color"]))
        elif isinstance(w, dict) and "SolidColor" in w:
            world.update(kind=abi.RD_WORLD_SOLID_COLOR, solid_color=_v3(w["SolidColor"]))
        elif w != "Transparent":
            raise ValueError(f"unknown world variant {w!r}")
        objs = np.zeros(len(d["objects"]), dtype=abi.OBJECT_DTYPE)
        for i, o in enumerate(d["objects"]):
            g = o["geometry"]
            if "Sphere" in g:
                objs[i]["kind"] = abi.RD_SPHERE
                objs[i]["center"] = _v3(g["Sphere"]["center"])
                objs[i]["size"] = g["Sphere"]["radius"]
            elif "Cube" in g:
                objs[i]["kind"] = abi.RD_CUBE
                objs[i]["center"] = _v3(g["Cube"]["center"])
                objs[i]["size"] = g["Cube"]["side_length"]
            else:
                raise ValueError(f"unknown geometry {list(g)}")
            m = o["material"]
            objs[i]["albedo"] = _v3(m["albedo"])
            objs[i]["emission_color"] = _v3(m["emission_color"])
            for k in ("roughness", "metallic", "emission_strength", "transmission", "ior"):
                objs[i][k] = m[k]
        return cls(camera, world, objs)

    @classmethod
    def load(cls, path):
        with open(path) as f:
            return cls.from_json(f.read())

    def to_json(self):
        c = self.camera
        if c["projection"] == abi.RD_PROJECTION_PERSPECTIVE:
            proj = {"Perspective": {"fov": c["projection_param"]}}
        else:
            proj = {"Orthographic": {"size": c["projection_param"]}}
        cam = {
            "position": _vec(c["position"]), "target": _vec(c["target"]), "up": _vec(c["up"]),
            "resolution_x": int(c["resolution_x"]), "resolution_y": int(c["resolution_y"]),
            "projection": proj, "near_clip": c["near_clip"], "far_clip": c["far_clip"],
            "view_matrix": _mat(c["view_matrix"]), "proj_matrix": _mat(c["proj_matrix"]),
            "inverse_view_matrix": _mat(c["inverse_view_matrix"]),
            "inverse_proj_matrix": _mat(c["inverse_proj_matrix"]),
        }
        w = self.world
        if w["kind"] == abi.RD_WORLD_SKY_COLOR:
            world = {"SkyColor": {"top_color": _vec(w["top_color"]),
                                  "bottom_color": _vec(w["bottom_color"])}}
        elif w["kind"] == abi.RD_WORLD_SOLID_COLOR:
            world = {"SolidColor": _vec(w["solid_color"])}
        else:
            world = "Transparent"
        objects = []
        for o in self.objects:
            key, sz = ("Sphere", "radius") if o["kind"] == abi.RD_SPHERE else ("Cube", "side_length")
            objects.append({
                "geometry": {key: {"center": _vec(o["center"]), sz: float(o["size"])}},
                "material": {
                    "albedo": _vec(o["albedo"]), "roughness": float(o["roughness"]),
                    "metallic": float(o["metallic"]),
                    "emission_color": _vec(o["emission_color"]),
                    "emission_strength": float(o["emission_strength"]),
                    "transmission": float(o["transmission"]), "ior": float(o["ior"]),
                },
            })
        return _json_dump({"camera": cam, "world": world, "objects": objects})

    def save(self, path):
        with open(path, "w") as f:
            f.write(self.to_json())

    # ------------------------------------------------------------- C ABI ---
    @property
    def width(self):
        return int(self.camera["resolution_x"])

    @property
    def height(self):
        return int(self.camera["resolution_y"])

    def as_rd_camera(self):
        """The camera as an abi.RdCamera."""
        c = self.camera
        cam = abi.RdCamera()
        cam.position[:] = c["position"]
        cam.target[:] = c["target"]
        cam.up[:] = c["up"]
        cam.resolution_x = c["resolution_x"]
        cam.resolution_y = c["resolution_y"]
        cam.projection = c["projection"]
        cam.projection_param = c["projection_param"]
        cam.near_clip = c["near_clip"]
        cam.far_clip = c["far_clip"]
        for k in ("view_matrix", "proj_matrix", "inverse_view_matrix", "inverse_proj_matrix"):
            getattr(cam, k)[:] = c[k]
        return cam

    def as_rd(self):
        """Return an abi.RdScene pointing at this scene's object array.

        The returned struct borrows `self.objects`; keep the Scene alive while
        the struct is in use."""
        s = abi.RdScene()
        s.camera = self.as_rd_camera()
        w = self.world
        s.world.kind = w["kind"]
        s.world.top_color[:] = w["top_color"]
        s.world.bottom_color[:] = w["bottom_color"]
        s.world.solid_color[:] = w["solid_color"]
        s.object_count = len(self.objects)
        s.objects = self.objects.ctypes.data_as(abi.C.POINTER(abi.RdObject))
        return s

    @classmethod
    def from_rd(cls, s):
        """Copy an abi.RdScene (e.g. one returned by rd_scene_load_rscn)."""
        cam = s.camera
        camera = dict(
            position=list(cam.position), target=list(cam.target), up=list(cam.up),
            resolution_x=cam.resolution_x, resolution_y=cam.resolution_y,
            projection=cam.projection, projection_param=cam.projection_param,
            near_clip=cam.near_clip, far_clip=cam.far_clip,
            view_matrix=list(cam.view_matrix), proj_matrix=list(cam.proj_matrix),
            inverse_view_matrix=list(cam.inverse_view_matrix),
            inverse_proj_matrix=list(cam.inverse_proj_matrix),
        )
        world = dict(kind=s.world.kind, top_color=list(s.world.top_color),
                     bottom_color=list(s.world.bottom_color),
                     solid_color=list(s.world.solid_color))
        n = s.object_count
        if n:
            buf = abi.C.string_at(s.objects, n * 64)
            objs = np.frombuffer(buf, dtype=abi.OBJECT_DTYPE).copy()
        else:
            objs = np.zeros(0, dtype=abi.OBJECT_DTYPE)
        return cls(camera, world, objs)


# ---------------------------------------------------------------- camera ---
def look_at_lh(eye, target, up):
    """cgmath 0.18 Matrix4::look_at_lh (camera.rs:211), column-major, f64."""
    eye, target, up = (np.asarray(v, dtype=np.float64) for v in (eye, target, up))
    f = eye - target
    f = f / np.linalg.norm(f)
    s = np.cross(f, up)
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    cols = [
        [s[0], u[0], -f[0], 0.0], [s[1], u[1], -f[1], 0.0], [s[2], u[2], -f[2], 0.0],
        [-eye.dot(s), -eye.dot(u), eye.dot(f), 1.0],
    ]
    return np.array(cols, dtype=np.float64).reshape(16)


def perspective(fov_deg, aspect, near, far):
    """cgmath 0.18 perspective(Deg(fov), aspect, near, far) (camera.rs:217)."""
    f = 1.0 / math.tan(math.radians(fov_deg) / 2.0)
    m = np.zeros(16)
    m[0] = f / aspect
    m[5] = f
    m[10] = (far + near) / (near - far)
    m[11] = -1.0
    m[14] = (2.0 * far * near) / (near - far)
    return m


def ortho(left, right, bottom, top, near, far):
    """cgmath 0.18 ortho (camera.rs:218-225)."""
    m = np.zeros(16)
    m[0] = 2.0 / (right - left)
    m[5] = 2.0 / (top - bottom)
    m[10] = -2.0 / (far - near)
    m[12] = -(right + left) / (right - left)
    m[13] = -(top + bottom) / (top - bottom)
    m[14] = -(far + near) / (far - near)
    m[15] = 1.0
    return m


def _inv(m):
    a = np.asarray(m, dtype=np.float64).reshape(4, 4).T  # row-major view of column-major data
    return np.linalg.inv(a).T.reshape(16)


def make_camera(position, target, up, width, height, fov_deg=None, ortho_size=None,
                near=0.01, far=1000.0):
    """Camera::new + update_matrices (camera.rs:33-60, 210-231)."""
    aspect = float(_F32(width) / _F32(height))
    view = look_at_lh(position, target, up)
    if ortho_size is None:
        proj = perspective(fov_deg, aspect, near, far)
        pk, pp = abi.RD_PROJECTION_PERSPECTIVE, float(fov_deg)
    else:
        s = ortho_size
        proj = ortho(-s * aspect, s * aspect, -s, s, near, far)
        pk, pp = abi.RD_PROJECTION_ORTHOGRAPHIC, float(ortho_size)
    f32 = lambda v: [float(_F32(x)) for x in v]  # noqa: E731
    return dict(
        position=f32(position), target=f32(target), up=f32(up),
        resolution_x=int(width), resolution_y=int(height), projection=pk,
        projection_param=float(_F32(pp)), near_clip=float(_F32(near)),
        far_clip=float(_F32(far)), view_matrix=f32(view), proj_matrix=f32(proj),
        inverse_view_matrix=f32(_inv(view)), inverse_proj_matrix=f32(_inv(proj)),
    )


def sky_world(top=(0.5, 0.7, 1.0), bottom=(1.0, 1.0, 1.0)):
    return dict(kind=abi.RD_WORLD_SKY_COLOR, top_color=[float(_F32(v)) for v in top],
                bottom_color=[float(_F32(v)) for v in bottom], solid_color=[0.0] * 3)


def solid_world(color):
    return dict(kind=abi.RD_WORLD_SOLID_COLOR, top_color=[0.0] * 3, bottom_color=[0.0] * 3,
                solid_color=[float(_F32(v)) for v in color])


def make_object(kind, center, size, albedo=(0.8, 0.8, 0.8), roughness=0.5, metallic=0.0,
                emission_color=(0.0, 0.0, 0.0), emission_strength=0.0, transmission=0.0,
                ior=1.5):
    """An Object; defaults are Material::default() (material.rs:15-27)."""
    o = np.zeros((), dtype=abi.OBJECT_DTYPE)
    o["kind"] = kind
    o["center"] = center
    o["size"] = size
    o["albedo"] = albedo
    o["roughness"] = roughness
    o["metallic"] = metallic
    o["emission_color"] = emission_color
    o["emission_strength"] = emission_strength
    o["transmission"] = transmission
    o["ior"] = ior
    return o
