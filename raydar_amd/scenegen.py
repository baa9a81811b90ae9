"""Seeded synthetic scenes for the benchmark workloads.

raydar ships no benchmark generator; these recipes produce .rscn-compatible
scenes of the shapes the performance workloads name.  Every draw comes from
SplitMix64(seed), U[a,b] = a + (b-a) * (x >> 40) * 2^-24, so a scene is a pure
function of its seed.

Material classes map onto raydar's single parametric Material
(src/scene/material.rs:4-13) as the scatter code reads it (src/renderer/cpu.rs:262-314):
  Diffuse  roughness 1, metallic 0, transmission 0, albedo U[0.1,0.9]^3
  Specular roughness 0, metallic 0, transmission 0, albedo 1
  Metal    metallic 1, roughness sqrt(fuzz) (offset scale = fuzz), albedo U[0.5,1]^3
  Glass    transmission 1, ior 1.5, roughness 0, albedo 1
Lighting is raydar's sky gradient, World::SkyColor (src/scene/world.rs:8-11),
bottom (1,1,1) to top (0.5,0.7,1.0).
"""
import math

import numpy as np

from . import abi
from .scene import Scene, make_camera, make_object, sky_world

MASK64 = (1 << 64) - 1

DIFFUSE, SPECULAR, METAL, GLASS = range(4)
RESOLUTIONS_BATCH = ((128, 128), (256, 256), (512, 512), (1024, 768), (1920, 1080))
RESOLUTIONS_SERVE = ((256, 256), (512, 512), (1024, 768), (1920, 1080))


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood, OOPSLA 2014)."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 40) * 2.0 ** -24)

    def choice(self, n):
        return int(self.uniform() * n)

    def log_uniform(self, a, b):
        return math.exp(self.uniform(math.log(a), math.log(b)))

    def normal(self):
        # Box-Muller on two draws; 1-u keeps the log argument in (0, 1]
        u1, u2 = 1.0 - self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def material(rng, cls):
    if cls == DIFFUSE:
        alb = [rng.uniform(0.1, 0.9) for _ in range(3)]
        return dict(albedo=alb, roughness=1.0, metallic=0.0, transmission=0.0, ior=1.5)
    if cls == SPECULAR:
        return dict(albedo=(1.0, 1.0, 1.0), roughness=0.0, metallic=0.0, transmission=0.0, ior=1.5)
    if cls == METAL:
        alb = [rng.uniform(0.5, 1.0) for _ in range(3)]
        fuzz = rng.uniform(0.0, 0.3)
        return dict(albedo=alb, roughness=math.sqrt(fuzz), metallic=1.0, transmission=0.0, ior=1.5)
    return dict(albedo=(1.0, 1.0, 1.0), roughness=0.0, metallic=0.0, transmission=1.0, ior=1.5)


def ground():
    return make_object(abi.RD_SPHERE, (0.0, -1000.0, 0.0), 1000.0, albedo=(0.5, 0.5, 0.5),
                       roughness=1.0, metallic=0.0, transmission=0.0)


def small_camera(width=256, height=256):
    return make_camera((0.0, 2.0, 8.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), width, height,
                       fov_deg=40.0)


def cluster_camera(width=1920, height=1080):
    return make_camera((0.0, 25.0, 75.0), (0.0, 10.0, 0.0), (0.0, 1.0, 0.0), width, height,
                       fov_deg=40.0)


def plain_small(seed=0):
    """256x256, 9 objects: ground + 8 spheres r~U[0.3,1], x,z~U[-4,4], y=r."""
    rng = SplitMix64(seed)
    objs = [ground()]
    for _ in range(8):
        r = rng.uniform(0.3, 1.0)
        x, z = rng.uniform(-4, 4), rng.uniform(-4, 4)
        m = material(rng, rng.choice(4))
        objs.append(make_object(abi.RD_SPHERE, (x, r, z), r, **m))
    return Scene(small_camera(), sky_world(), np.array(objs))


def plain_1080p(seed=1, width=1920, height=1080):
    """Ground + 64 spheres + 32 cubes, sizes U[0.2,1.2], x,z~U[-10,10], no overlaps."""
    rng = SplitMix64(seed)
    objs = [ground()]
    placed = []  # (x, y, z, bounding radius)
    for kind, count in ((abi.RD_SPHERE, 64), (abi.RD_CUBE, 32)):
        for _ in range(count):
            while True:
                size = rng.uniform(0.2, 1.2)
                x, z = rng.uniform(-10, 10), rng.uniform(-10, 10)
                if kind == abi.RD_SPHERE:
                    y, br = size, size
                else:
                    y, br = size / 2.0, size * math.sqrt(3.0) / 2.0
                if all((x - px) ** 2 + (y - py) ** 2 + (z - pz) ** 2 >= (br + pr) ** 2
                       for px, py, pz, pr in placed):
                    break
            placed.append((x, y, z, br))
            m = material(rng, rng.choice(4))
            objs.append(make_object(kind, (x, y, z), size, **m))
    return Scene(small_camera(width, height), sky_world(), np.array(objs))


def _cluster_class(rng):
    u = rng.uniform()
    return DIFFUSE if u < 0.4 else SPECULAR if u < 0.6 else METAL if u < 0.8 else GLASS


def clustered(seed=2, n_prims=4096, width=1920, height=1080):
    """n_prims (every 4th a cube), sizes log-uniform [0.05,1], 16 Gaussian
    clusters (sigma 3) in x,z in [-30,30], y in [0,20]; no ground."""
    rng = SplitMix64(seed)
    centres = [(rng.uniform(-30, 30), rng.uniform(0, 20), rng.uniform(-30, 30)) for _ in range(16)]
    lo, hi = (-30.0, 0.0, -30.0), (30.0, 20.0, 30.0)
    objs = []
    for i in range(n_prims):
        kind = abi.RD_CUBE if i % 4 == 3 else abi.RD_SPHERE
        size = rng.log_uniform(0.05, 1.0)
        c = centres[rng.choice(16)]
        p = [min(max(c[a] + 3.0 * rng.normal(), lo[a]), hi[a]) for a in range(3)]
        m = material(rng, _cluster_class(rng))
        objs.append(make_object(kind, p, size, **m))
    return Scene(cluster_camera(width, height), sky_world(), np.array(objs))


class Job:
    """One render job of a batch or of the serving stream."""

    def __init__(self, scene, spp, bounces, seed):
        self.scene, self.spp, self.bounces, self.seed = scene, spp, bounces, seed

    @property
    def paths(self):
        return self.scene.width * self.scene.height * self.spp


def batch_jobs(first_seed=1000, count=64, resolutions=RESOLUTIONS_BATCH, spps=(8, 32, 128),
               bounces=(4, 8, 12)):
    """Scenes as `clustered` with N ~ log-uniform[16,2048] and drawn
    resolution / spp / bounce limits."""
    jobs = []
    for i in range(count):
        seed = first_seed + i
        rng = SplitMix64(seed ^ 0xB47C4)
        n = int(round(rng.log_uniform(16, 2048)))
        w, h = resolutions[rng.choice(len(resolutions))]
        spp = spps[rng.choice(len(spps))]
        nb = bounces[rng.choice(len(bounces))]
        jobs.append(Job(clustered(seed, n, w, h), spp, nb, seed))
    return jobs


def serve_arrivals(rate=20.0, duration=60.0, seed=7, first_scene_seed=2000):
    """Poisson arrivals: list of (arrival time s, Job)."""
    rng = SplitMix64(seed)
    t, out, i = 0.0, [], 0
    while True:
        t += -math.log(1.0 - rng.uniform()) / rate
        if t >= duration:
            break
        s = first_scene_seed + i
        srng = SplitMix64(s ^ 0x5E4E)
        n = int(round(srng.log_uniform(16, 2048)))
        w, h = RESOLUTIONS_SERVE[srng.choice(len(RESOLUTIONS_SERVE))]
        spp = (32, 64, 128)[srng.choice(3)]
        out.append((t, Job(clustered(s, n, w, h), spp, 12, s)))
        i += 1
    return out


WORKLOADS = {
    "plain_small": lambda: Job(plain_small(0), 16, 12, 0),
    "plain_1080p": lambda: Job(plain_1080p(1), 64, 12, 1),
    "clustered_1080p": lambda: Job(clustered(2), 256, 12, 2),
}
